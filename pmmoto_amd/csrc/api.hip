// api.hip -- error state, version and small whole-array utilities.
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace pmx {

static thread_local std::string g_last_error;

void set_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

// ---- per-kernel timing --------------------------------------------------
namespace {
struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};
struct ProfState {
    bool on = false;
    std::vector<ProfRec> recs;
    std::vector<hipEvent_t> pool;
    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};
ProfState g_prof;
std::mutex g_prof_mu;
}  // namespace

void prof_mark(const char *name, hipStream_t s, bool start)
{
    if (!g_prof.on) return;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (start) {
        ProfRec r{name, g_prof.get(), nullptr};
        if (r.a) (void)hipEventRecord(r.a, s);
        g_prof.recs.push_back(r);
    } else {
        for (auto it = g_prof.recs.rbegin(); it != g_prof.recs.rend(); ++it)
            if (it->b == nullptr && it->name == name) {
                it->b = g_prof.get();
                if (it->b) (void)hipEventRecord(it->b, s);
                break;
            }
    }
}

namespace {

// Copies own voxels into the pads of one periodic axis: pad voxel at
// index i < lo takes own voxel i + n_own, pad voxel i >= hi takes i - n_own
// (update_buffer's self-neighbour copy, communication.py:101-118,278-285).
template <class T>
__global__ __launch_bounds__(256) void k_pad_axis(T *a, int64_t n0, int64_t n1, int64_t n2, int axis,
                                                  int64_t lo, int64_t hi)
{
    const int64_t n_own = hi - lo;
    const int64_t ext[3] = {n0, n1, n2};
    const int64_t npad = ext[axis] - n_own;
    // the pad slab: npad layers along `axis`, full extent on the others
    int64_t other = 1;
    for (int a = 0; a < 3; ++a)
        if (a != axis) other *= ext[a];
    const int64_t total = npad * other;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t layer = i / other, rest = i - layer * other;
        const int64_t pos = layer < lo ? layer : hi + (layer - lo);
        const int64_t srcpos = pos < lo ? pos + n_own : pos - n_own;
        int64_t c[3];
        int k = 0;
        int64_t r = rest;
        int64_t o[2];
        for (int a = 0; a < 3; ++a)
            if (a != axis) o[k++] = ext[a];
        const int64_t c2 = r % o[1];
        const int64_t c1 = r / o[1];
        k = 0;
        for (int a = 0; a < 3; ++a) c[a] = 0;
        for (int a = 0; a < 3; ++a)
            if (a != axis) c[a] = (k++ == 0) ? c1 : c2;
        c[axis] = pos;
        const int64_t dst = (c[0] * n1 + c[1]) * n2 + c[2];
        c[axis] = srcpos;
        const int64_t src = (c[0] * n1 + c[1]) * n2 + c[2];
        a[dst] = a[src];
    }
}

template <class T>
int pad_all(T *a, const int64_t *shape, const int64_t *pad, const int32_t *periodic, hipStream_t s)
{
    for (int ax = 0; ax < 3; ++ax) {
        if (!periodic[ax]) continue;
        const int64_t lo = pad[2 * ax], hi = shape[ax] - pad[2 * ax + 1];
        if (lo == 0 && hi == shape[ax]) continue;
        if (hi - lo < std::max(pad[2 * ax], pad[2 * ax + 1]))
            return fail(PMX_ERR_INVALID, "axis %d: pads wider than the own extent", ax);
        hipLaunchKernelGGL(k_pad_axis<T>, dim3(1024), dim3(256), 0, s, a, shape[0], shape[1],
                           shape[2], ax, lo, hi);
        if (int e = check_launch("periodic pad fill")) return e;
    }
    return PMX_OK;
}

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_profile_begin(void)
{
    std::lock_guard<std::mutex> lk(pmx::g_prof_mu);
    pmx::g_prof.on = true;
    return PMX_OK;
}

int pmx_profile_end(char *names, int names_len, double *ms, int64_t *calls, int max_kernels,
                    int *n_kernels)
{
    pmx::clear_error();
    std::lock_guard<std::mutex> lk(pmx::g_prof_mu);
    pmx::g_prof.on = false;
    std::map<std::string, std::pair<double, int64_t>> agg;
    std::vector<std::string> order;
    for (auto &r : pmx::g_prof.recs) {
        float t = 0.f;
        if (r.a && r.b && hipEventSynchronize(r.b) == hipSuccess &&
            hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
            if (!agg.count(r.name)) order.push_back(r.name);
            agg[r.name].first += t;
            agg[r.name].second += 1;
        }
        if (r.a) pmx::g_prof.pool.push_back(r.a);
        if (r.b) pmx::g_prof.pool.push_back(r.b);
    }
    pmx::g_prof.recs.clear();
    int k = 0, pos = 0;
    for (auto &nm : order) {
        if (k >= max_kernels) break;
        if (ms) ms[k] = agg[nm].first;
        if (calls) calls[k] = agg[nm].second;
        if (names && pos + (int)nm.size() + 1 < names_len) {
            std::memcpy(names + pos, nm.c_str(), nm.size());
            pos += (int)nm.size();
            names[pos++] = '\n';
            names[pos] = 0;
        }
        ++k;
    }
    if (n_kernels) *n_kernels = k;
    return PMX_OK;
}

int pmx_version(void) { return 100; }

const char *pmx_last_error(void) { return pmx::g_last_error.c_str(); }

int pmx_fill_periodic_pads(void *arr, int elem, const int64_t *shape, const int64_t *pad,
                           const int32_t *periodic, hipStream_t stream)
{
    pmx::clear_error();
    if (!arr || !shape || !pad || !periodic) return pmx::fail(PMX_ERR_INVALID, "NULL argument");
    for (int a = 0; a < 3; ++a)
        if (shape[a] < 1 || pad[2 * a] < 0 || pad[2 * a + 1] < 0 ||
            pad[2 * a] + pad[2 * a + 1] >= shape[a])
            return pmx::fail(PMX_ERR_INVALID, "axis %d: bad shape/pad", a);
    switch (elem) {
    case 1: return pmx::pad_all(static_cast<uint8_t *>(arr), shape, pad, periodic, stream);
    case 4: return pmx::pad_all(static_cast<uint32_t *>(arr), shape, pad, periodic, stream);
    case 8: return pmx::pad_all(static_cast<uint64_t *>(arr), shape, pad, periodic, stream);
    default: return pmx::fail(PMX_ERR_INVALID, "elem must be 1, 4 or 8");
    }
}

}  // extern "C"
