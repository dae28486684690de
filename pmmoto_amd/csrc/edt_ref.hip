// edt_ref.hip -- the reference's float32 EDT arithmetic on the GPU.
//
// The exact integer kernels (edt.hip) cover voxel-unit domains without
// periodic faces.  PMMoTo's default domain is the unit box, so most callers
// run at resolution 1/N, where the reference's float32 results depend on its
// exact sequence of roundings, and on periodic axes the reference keeps at
// most five boundary parabolas per face.  This path reproduces both, per
// line, so outputs are bit-identical at any resolution:
//
//   pass 0 (axis a0)  squared_edt_1d<uint8_t> (_distance.hpp:135-182):
//                     label-aware forward sweep d += res, corrector seeding
//                     and capping, backward fminf(d, d_next + res), square,
//                     INF -> FLT_MAX (to_finite, :22-32).  Periodic
//                     correctors: edt3d's N - last_zero / first_zero + 1
//                     (_distance.pyx:240-279), or corrected_edt's
//                     pad-skipping face distances on a padded subdomain
//                     (euclidean_distance_transform.py:304-447).
//   passes a1, a2     squared_edt_1d_parabolic (:184-272) with float
//                     numerators and the double-precision division of
//                     intersect_parabolas (:60-72), pop test s <= range
//                     (:89-112), evaluation aniso * sq(i - v) + h; periodic
//                     lines first get up to num_hull + 1 = 5 boundary
//                     parabolas from each face (return_boundary_hull,
//                     :296-356, segments as get_boundary_hull cuts them,
//                     euclidean_distance_transform.py:246-298 / :450-528),
//                     all taken from the pass input before any line changes.
//
// The float operations and their order are the reference's because
// bit-parity requires them; the GPU formulation is this file's own:
//
//   * one lane per line, grids sized to the line count;
//   * a parabolic pass streams its line through a stack in LDS (one wave per
//     workgroup, entry k of lane l at [k][l]).  A parabola of height 0 (a
//     solid voxel) is never popped: a later parabola r meets it at
//     s >= (r + p) / 2 > p >= its range.  So every member below it is final,
//     the positions before it are evaluated and written, and the stack
//     restarts at that entry.  The stack then holds only the members between
//     two solids;
//   * passes read one array and write another (the caller's output and a
//     workspace array in turn), so a line whose stack outgrows the LDS is
//     recomputed from its untouched input by a fallback kernel with stacks
//     in the workspace; the boundary hulls likewise;
//   * the last pass writes sqrtf when asked (np.sqrt of float32 rounds
//     correctly, as sqrtf does).
//
// Compiled with contraction off (the reference's x86-64 build has no FMA),
// so a*b+c rounds twice.
#include "common.hpp"

namespace pmx {
namespace {

constexpr float kFltMax = 3.40282347e+38f;
constexpr int kNumHull = 4;            // euclidean_distance_transform.py:275,284
constexpr int kHullCap = kNumHull + 1; // members returned per face
#ifndef PMX_REF_LDS_CAP
#define PMX_REF_LDS_CAP 64
#endif
// LDS stack entries per lane: 64 leave about 0.1 % of the lines of a
// porosity-0.5 sphere pack to the fallback (32: 11 %, 48: 1 %)
constexpr int kLdsCap = PMX_REF_LDS_CAP;
constexpr int kFbMaxThreads = 32768;   // persistent lanes of the fallback kernels, at most

struct RefHull {
    int vertex;
    float height;
    float range;
};

// Lines along `dim` of a box n[3]; a line is numbered by its two other
// coordinates (d1, d2 ascending, the last fastest), as the reference's
// get_dimensions orders them.  Element strides per axis for the uint8
// image, the pass input and the pass output (z stride 1 for the image).
struct LineGeom {
    int64_t n[3];
    int64_t is[3], fs[3], os[3];
    int dim, d1, d2;
    __host__ __device__ int64_t lines() const { return n[d1] * n[d2]; }
    __device__ __forceinline__ void base(int64_t line, int64_t *ib, int64_t *fb, int64_t *ob) const
    {
        const int64_t a = line / n[d2], b = line - a * n[d2];
        *ib = a * is[d1] + b * is[d2];
        *fb = a * fs[d1] + b * fs[d2];
        *ob = a * os[d1] + b * os[d2];
    }
};

// Per-axis search skips of the padded-subdomain mode (0 for edt3d).
struct Skips {
    int p_lo, p_hi;  // pad widths of the axis; padded mode when padded != 0
    int padded;
};

// first (forward) / last index of value 0 within [lo_skip, n - hi_skip), or -1
// (_voxels.hpp:60-79 through _voxels.pyx:231-283)
__device__ int nearest_zero(const uint8_t *img, int n, int64_t stride, bool forward, int lo_skip, int hi_skip)
{
    if (forward) {
        for (int i = lo_skip; i < n - hi_skip; ++i)
            if (img[(int64_t)i * stride] == 0) return i;
    } else {
        for (int i = n - hi_skip - 1; i >= lo_skip; --i)
            if (img[(int64_t)i * stride] == 0) return i;
    }
    return -1;
}

// pass 0 along g.dim from the image into out (squared, INF -> FLT_MAX)
__global__ __launch_bounds__(256) void k_ref_pass0(const uint8_t *__restrict__ img, float *__restrict__ out,
                                                   LineGeom g, float res, int periodic, Skips sk)
{
    const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= g.lines()) return;
    const int n = (int)g.n[g.dim];
    const int64_t ist = g.is[g.dim], ost = g.os[g.dim];
    int64_t ib, fb, ob;
    g.base(line, &ib, &fb, &ob);
    const uint8_t *im = img + ib;
    float *o = out + ob;
    float lo = INFINITY, hi = INFINITY;
    if (periodic && !sk.padded) {
        // edt3d: correctors of the whole line (_distance.pyx:240-279),
        // float64 arrays converted to float at the call
        const int first = nearest_zero(im, n, ist, true, 0, 0);
        const int last = nearest_zero(im, n, ist, false, 0, 0);
        lo = last != -1 ? (float)((double)n - (double)last) : INFINITY;
        hi = first != -1 ? (float)((double)first + 1.0) : INFINITY;
    } else if (periodic) {
        // corrected_edt: the lower face searches [2 p_lo, n), the upper face
        // [2 p_lo, n - 2 p_hi) (the skip variables of
        // get_nearest_boundary_distance persist from the lower face); the
        // line start takes the upper face's distance, the end the lower's
        const int f = nearest_zero(im, n, ist, true, 2 * sk.p_lo, 0);
        const int l = nearest_zero(im, n, ist, false, 2 * sk.p_lo, 2 * sk.p_hi);
        lo = l != -1 ? (float)(n - l - sk.p_hi - 1) : INFINITY;
        hi = f != -1 ? (float)(f - sk.p_lo) : INFINITY;
    }
    // forward sweep, multi-label restarts (_distance.hpp:146-166)
    uint8_t seg = im[0];
    float prev = seg == 0 ? 0.0f : lo * res;
    o[0] = prev;
    uint8_t vprev = seg;
    for (int i = 1; i < n; ++i) {
        const uint8_t v = im[(int64_t)i * ist];
        float cur;
        if (v == 0) {
            cur = 0.0f;
        } else if (v == seg) {
            cur = prev + res;
        } else {
            cur = res;
            o[(int64_t)(i - 1) * ost] = (float)(vprev != 0) * res;
            seg = v;
        }
        o[(int64_t)i * ost] = cur;
        prev = cur;
        vprev = v;
    }
    // cap, backward min, square, to_finite (:168-181)
    float next = o[(int64_t)(n - 1) * ost];
    if (next > hi * res) next = hi * res;
    {
        const float sq = next * next;
        o[(int64_t)(n - 1) * ost] = sq == INFINITY ? kFltMax : sq;
    }
    for (int i = n - 2; i >= 0; --i) {
        const float cur = fminf(o[(int64_t)i * ost], next + res);
        const float sq = cur * cur;
        o[(int64_t)i * ost] = sq == INFINITY ? kFltMax : sq;
        next = cur;
    }
}

__device__ __forceinline__ float intersect(float f_r, float r, float f_q, float q, float aniso)
{
    const float factor1 = (r - q) * aniso;
    const float factor2 = r + q;
    const float num = f_r - f_q + factor1 * factor2;
    return (float)((double)num / (2.0 * (double)factor1));
}

// A lane's stack: entry k at base[k * step] (LDS: step = lanes of the
// wave; workspace: step = lanes of the grid).
struct Stack {
    int *hv;
    float *hh, *rg;
    int64_t step;
    __device__ __forceinline__ int &V(int k) const { return hv[(int64_t)k * step]; }
    __device__ __forceinline__ float &H(int k) const { return hh[(int64_t)k * step]; }
    __device__ __forceinline__ float &R(int k) const { return rg[(int64_t)k * step]; }
};

// update_hull (_distance.hpp:89-112); false when entry k+1 would reach cap
__device__ __forceinline__ bool update_hull(const Stack &S, int &k, int i, float fv, float aniso, int cap)
{
    float s = intersect(fv, (float)i, S.H(k), (float)S.V(k), aniso);
    while (k > 0 && s <= S.R(k)) {
        --k;
        s = intersect(fv, (float)i, S.H(k), (float)S.V(k), aniso);
    }
    ++k;
    if (k + 1 >= cap) return false;
    S.V(k) = i;
    S.H(k) = fv;
    S.R(k) = s;
    S.R(k + 1) = INFINITY;
    return true;
}

// return_boundary_hull (_distance.hpp:296-356) on n values at p (stride
// st); writes <= kHullCap members to out and returns the count, or -1 when
// the envelope outgrows cap.  A one-voxel segment with no finite value reads
// past the reference's vectors there (undefined behaviour); this path then
// returns the zero-initialised slot, as the oracle does.
__device__ int boundary_hull(const float *p, int n, int64_t st, float res, int index_corrector, bool forward,
                             const Stack &S, int cap, RefHull *out)
{
    if (n == 1 && p[0] < kFltMax) {
        out[0] = {index_corrector, p[0], INFINITY};
        return 1;
    }
    const float aniso = res * res;
    S.V(0) = 0;
    S.H(0) = p[0];
    S.V(1) = 0;
    S.H(1) = 0.0f;
    S.R(0) = -INFINITY;
    S.R(1) = INFINITY;
    int k = 1;
    for (int i = 1; i < n; ++i)
        if (!update_hull(S, k, i, p[(int64_t)i * st], aniso, cap)) return -1;
    const int num_hull = min(kNumHull, k);
    int cnt = 0, kk = forward ? 0 : k;
    const int step = forward ? 1 : -1;
    while (cnt <= num_hull && kk >= 0 && kk <= k) {
        if (S.H(kk) < kFltMax) {
            out[cnt++] = {S.V(kk) + index_corrector, S.H(kk), S.R(kk) + (float)index_corrector};
            if (S.H(kk) == 0.9f) break;
        }
        kk += step;
    }
    return cnt;
}

// Both face hulls of one line (get_boundary_hull, _distance.pyx:544-618):
// side 0 forward (from the lower face to the first zero), side 1 backward.
// Returns false when a segment outgrows cap.
__device__ bool line_hulls(const uint8_t *og, const float *f, const LineGeom &g, float res, Skips sk,
                           const Stack &S, int cap, RefHull *hulls, int *counts)
{
    const int n = (int)g.n[g.dim];
    const int64_t st = g.fs[g.dim], ist = g.is[g.dim];
    for (int side = 0; side < 2; ++side) {
        const bool fw = side == 0;
        const int lsk = sk.padded && fw ? 2 * sk.p_lo : 0;
        const int usk = sk.padded && !fw ? 2 * sk.p_hi : 0;
        const int bound = nearest_zero(og, n, ist, fw, lsk, usk);
        int start, nv, ic;
        if (bound == -1) {
            start = lsk;
            nv = n - usk - lsk;
            ic = fw ? 0 : -nv;
        } else if (fw) {
            start = lsk;
            nv = bound - lsk + 1;
            ic = 0;
        } else {
            start = bound;
            nv = n - bound - usk;
            ic = -nv;
        }
        int c = 0;
        if (nv > 0) {
            c = boundary_hull(f + (int64_t)start * st, nv, st, res, ic, fw, S, cap, hulls + side * kHullCap);
            if (c < 0) return false;
        }
        counts[side] = c;
    }
    return true;
}

// hulls[line][side][kHullCap], counts[line][side]; lines whose segments
// outgrow the LDS stack go to the overflow list
__global__ __launch_bounds__(kWave) void k_ref_hulls(const uint8_t *__restrict__ og, const float *__restrict__ f,
                                                     LineGeom g, float res, Skips sk, RefHull *__restrict__ hulls,
                                                     int *__restrict__ counts, int *__restrict__ ovf,
                                                     int *__restrict__ n_ovf)
{
    __shared__ int sv[kLdsCap * kWave];
    __shared__ float sh[kLdsCap * kWave], sr[kLdsCap * kWave];
    const int64_t line = (int64_t)blockIdx.x * kWave + threadIdx.x;
    if (line >= g.lines()) return;
    Stack S{sv + threadIdx.x, sh + threadIdx.x, sr + threadIdx.x, kWave};
    int64_t ib, fb, ob;
    g.base(line, &ib, &fb, &ob);
    if (!line_hulls(og + ib, f + fb, g, res, sk, S, kLdsCap, hulls + line * 2 * kHullCap, counts + line * 2))
        ovf[atomicAdd(n_ovf, 1)] = (int)line;
}

__global__ __launch_bounds__(256) void k_ref_hulls_fallback(const uint8_t *__restrict__ og,
                                                            const float *__restrict__ f, LineGeom g, float res,
                                                            Skips sk, RefHull *__restrict__ hulls,
                                                            int *__restrict__ counts, const int *__restrict__ ovf,
                                                            const int *__restrict__ n_ovf, int *stk, int cap)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    Stack S{stk + tid, reinterpret_cast<float *>(stk) + (int64_t)cap * T + tid,
            reinterpret_cast<float *>(stk) + (int64_t)2 * cap * T + tid, T};
    const int m = *n_ovf;
    for (int64_t j = tid; j < m; j += T) {
        const int64_t line = ovf[j];
        int64_t ib, fb, ob;
        g.base(line, &ib, &fb, &ob);
        line_hulls(og + ib, f + fb, g, res, sk, S, cap, hulls + line * 2 * kHullCap, counts + line * 2);
    }
}

// squared_edt_1d_parabolic (_distance.hpp:184-272) on one line, from fin
// to fout, with the injected hulls (lower pushed in reverse, upper at
// vertex + n); flushes the final members before every zero-height
// parabola.  Returns false when the stack outgrows cap (fout is then
// partly written; the caller recomputes the line).
__device__ bool parabolic_line(const float *fin, float *fout, int n, int64_t ist, int64_t ost, float aniso,
                               const RefHull *lower, int nl, const RefHull *upper, int nu, const Stack &S, int cap,
                               bool take_sqrt)
{
    int k = 0, loop_start = 1;
    S.R(0) = -INFINITY;
    S.R(1) = INFINITY;
    if (nl > 0) {
        if (nl + 1 >= cap) return false;
        loop_start = 0;
        for (int m = nl - 1; m >= 0; --m) {
            S.V(k) = lower[m].vertex;
            S.H(k) = lower[m].height;
            S.R(k) = lower[m].range;
            S.R(k + 1) = INFINITY;
            ++k;
        }
        --k;
    } else {
        S.V(0) = 0;
        S.H(0) = fin[0];
    }
    int from = 0;  // first position not yet written
    auto flush = [&](int to) {
        int kk = 0;
        for (int i = from; i < to; ++i) {
            while (S.R(kk + 1) < (float)i) ++kk;
            const float di = (float)(i - S.V(kk));
            const float v = aniso * (di * di) + S.H(kk);
            fout[(int64_t)i * ost] = take_sqrt ? sqrtf(v) : v;
        }
        from = to;
    };
    for (int i = loop_start; i < n; ++i) {
        const float fi = fin[(int64_t)i * ist];
        if (!update_hull(S, k, i, fi, aniso, cap)) return false;
        if (fi == 0.0f && k > 0) {
            // entries 0..k are final for every position before i
            flush(i);
            S.V(0) = S.V(k);
            S.H(0) = S.H(k);
            S.R(0) = S.R(k);
            S.R(1) = INFINITY;
            k = 0;
        }
    }
    for (int m = 0; m < nu; ++m)
        if (!update_hull(S, k, upper[m].vertex + n, upper[m].height, aniso, cap)) return false;
    flush(n);
    return true;
}

__global__ __launch_bounds__(kWave) void k_ref_parabolic(const float *__restrict__ fin, float *__restrict__ fout,
                                                         LineGeom g, float res, const RefHull *__restrict__ hulls,
                                                         const int *__restrict__ counts, int take_sqrt,
                                                         int *__restrict__ ovf, int *__restrict__ n_ovf)
{
    __shared__ int sv[kLdsCap * kWave];
    __shared__ float sh[kLdsCap * kWave], sr[kLdsCap * kWave];
    const int64_t line = (int64_t)blockIdx.x * kWave + threadIdx.x;
    if (line >= g.lines()) return;
    Stack S{sv + threadIdx.x, sh + threadIdx.x, sr + threadIdx.x, kWave};
    int64_t ib, fb, ob;
    g.base(line, &ib, &fb, &ob);
    const RefHull *lower = hulls ? hulls + (line * 2 + 1) * kHullCap : nullptr;
    const RefHull *upper = hulls ? hulls + (line * 2 + 0) * kHullCap : nullptr;
    const int nl = hulls ? counts[line * 2 + 1] : 0;
    const int nu = hulls ? counts[line * 2 + 0] : 0;
    if (!parabolic_line(fin + fb, fout + ob, (int)g.n[g.dim], g.fs[g.dim], g.os[g.dim], res * res, lower, nl, upper,
                        nu, S, kLdsCap, take_sqrt != 0))
        ovf[atomicAdd(n_ovf, 1)] = (int)line;
}

__global__ __launch_bounds__(256) void k_ref_parabolic_fallback(const float *__restrict__ fin,
                                                                float *__restrict__ fout, LineGeom g, float res,
                                                                const RefHull *__restrict__ hulls,
                                                                const int *__restrict__ counts, int take_sqrt,
                                                                const int *__restrict__ ovf,
                                                                const int *__restrict__ n_ovf, int *stk, int cap)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    Stack S{stk + tid, reinterpret_cast<float *>(stk) + (int64_t)cap * T + tid,
            reinterpret_cast<float *>(stk) + (int64_t)2 * cap * T + tid, T};
    const int m = *n_ovf;
    for (int64_t j = tid; j < m; j += T) {
        const int64_t line = ovf[j];
        int64_t ib, fb, ob;
        g.base(line, &ib, &fb, &ob);
        const RefHull *lower = hulls ? hulls + (line * 2 + 1) * kHullCap : nullptr;
        const RefHull *upper = hulls ? hulls + (line * 2 + 0) * kHullCap : nullptr;
        const int nl = hulls ? counts[line * 2 + 1] : 0;
        const int nu = hulls ? counts[line * 2 + 0] : 0;
        parabolic_line(fin + fb, fout + ob, (int)g.n[g.dim], g.fs[g.dim], g.os[g.dim], res * res, lower, nl, upper,
                       nu, S, cap, take_sqrt != 0);
    }
}

__global__ void k_zero_ints(int *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

LineGeom make_line_geom(const int64_t *n, const int64_t *is, const int64_t *fs, const int64_t *os, int dim)
{
    LineGeom g;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = n[a];
        g.is[a] = is[a];
        g.fs[a] = fs[a];
        g.os[a] = os[a];
    }
    g.dim = dim;
    int o[2], c = 0;
    for (int a = 0; a < 3; ++a)
        if (a != dim) o[c++] = a;
    g.d1 = o[0];
    g.d2 = o[1];
    return g;
}

// stack capacity of the fallback lanes: the reference's vectors hold
// n + 2 * kHullCap entries plus the ranges' sentinel
int fb_cap(int64_t nmax) { return (int)(nmax + 2 * kHullCap + 2); }

int64_t max_lines(const int64_t *n)
{
    int64_t lines = 0;
    for (int a = 0; a < 3; ++a) lines = std::max(lines, n[0] * n[1] * n[2] / n[a]);
    return lines;
}

// fallback lanes: one per line up to kFbMaxThreads, whole blocks of 256
int fb_threads(const int64_t *n) { return (int)std::min<int64_t>(kFbMaxThreads, (max_lines(n) + 255) / 256 * 256); }

struct RefWs {
    float *tmp;
    int *stk;
    RefHull *hulls;
    int *counts, *ovf, *n_ovf;
};

RefWs carve_ref(Carver &c, const int64_t *n)
{
    const int64_t nmax = std::max(n[0], std::max(n[1], n[2]));
    const int64_t lines = max_lines(n);
    RefWs w;
    w.tmp = c.take<float>((size_t)(n[0] * n[1] * n[2]));
    w.stk = c.take<int>((size_t)3 * fb_cap(nmax) * fb_threads(n));
    w.hulls = c.take<RefHull>((size_t)lines * 2 * kHullCap);
    w.counts = c.take<int>((size_t)lines * 2);
    w.ovf = c.take<int>((size_t)lines);
    w.n_ovf = c.take<int>(4);
    return w;
}

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_edt_ref_workspace_size(const int64_t *n, size_t *bytes)
{
    pmx::clear_error();
    if (!n || !bytes) return pmx::fail(PMX_ERR_INVALID, "NULL argument");
    for (int a = 0; a < 3; ++a)
        if (n[a] < 1 || n[a] > 16384) return pmx::fail(PMX_ERR_INVALID, "extent n[%d] outside [1, 16384]", a);
    if (n[0] * n[1] * n[2] / std::min(n[0], std::min(n[1], n[2])) > INT32_MAX)
        return pmx::fail(PMX_ERR_INVALID, "more than 2^31 lines along one axis");
    pmx::Carver c(nullptr, SIZE_MAX);
    pmx::carve_ref(c, n);
    *bytes = c.need();
    return PMX_OK;
}

int pmx_edt_ref(const uint8_t *img, const int64_t *n, const int64_t *in_s, float *d, const int64_t *out_s,
                int first_axis, const int32_t *periodic, const int64_t *pads, const float *resolution, int take_sqrt,
                void *workspace, size_t workspace_bytes, hipStream_t s)
{
    using namespace pmx;
    clear_error();
    if (!img || !n || !in_s || !d || !out_s || !periodic || !resolution)
        return fail(PMX_ERR_INVALID, "NULL argument");
    if (first_axis < 0 || first_axis > 2) return fail(PMX_ERR_INVALID, "first_axis must be 0, 1 or 2");
    if (pads) {
        if (first_axis != 0) return fail(PMX_ERR_INVALID, "padded-subdomain mode is 3-D (first_axis 0)");
        for (int a = 0; a < 3; ++a) {
            if (pads[2 * a] < 0 || pads[2 * a + 1] < 0)
                return fail(PMX_ERR_INVALID, "negative pad on axis %d", a);
            if (periodic[a] && 2 * (pads[2 * a] + pads[2 * a + 1]) >= n[a])
                return fail(PMX_ERR_INVALID, "axis %d: pads %lld + %lld leave no own voxels to search", a,
                            (long long)pads[2 * a], (long long)pads[2 * a + 1]);
        }
    }
    size_t need = 0;
    if (int e = pmx_edt_ref_workspace_size(n, &need)) return e;
    if (!workspace || workspace_bytes < need)
        return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    Carver c(workspace, workspace_bytes);
    const RefWs w = carve_ref(c, n);
    const int cap = fb_cap(std::max(n[0], std::max(n[1], n[2])));
    const int64_t is[3] = {in_s[0], in_s[1], 1}, os[3] = {out_s[0], out_s[1], 1};
    const int64_t ts[3] = {n[1] * n[2], n[2], 1};
    auto skips = [&](int a) {
        Skips k{0, 0, 0};
        if (pads) {
            k.p_lo = (int)pads[2 * a];
            k.p_hi = (int)pads[2 * a + 1];
            k.padded = 1;
        }
        return k;
    };
    // pass order: first_axis with the 1-D sweep, then the remaining axes in
    // ascending order (edt3d: 0 then 1, 2; edt2d on a (1, X, Y) box: 2 then
    // 1, 0); pass 0 writes d, the parabolic passes go d -> tmp -> d
    {
        ProfScope ps("edt_ref_pass0", s);
        LineGeom g = make_line_geom(n, is, os, os, first_axis);
        const int64_t nb = (g.lines() + 255) / 256;
        hipLaunchKernelGGL(k_ref_pass0, dim3((unsigned)nb), dim3(256), 0, s, img, d, g, resolution[first_axis],
                           periodic[first_axis] ? 1 : 0, skips(first_axis));
        if (int e = check_launch("reference pass 0")) return e;
    }
    int step = 0;
    for (int dim = 0; dim < 3; ++dim) {
        if (dim == first_axis) continue;
        const bool last = step == 1;
        const float *fin = step == 0 ? d : w.tmp;
        float *fout = step == 0 ? w.tmp : d;
        const int64_t *fs = step == 0 ? os : ts;
        const int64_t *fo = step == 0 ? ts : os;
        LineGeom g = make_line_geom(n, is, fs, fo, dim);
        const int64_t lines = g.lines();
        const unsigned nb = (unsigned)((lines + kWave - 1) / kWave);
        const int sq = last && take_sqrt ? 1 : 0;
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, w.n_ovf, 2);
        if (periodic[dim]) {
            {
                ProfScope ps("edt_ref_hulls", s);
                hipLaunchKernelGGL(k_ref_hulls, dim3(nb), dim3(kWave), 0, s, img, fin, g, resolution[dim],
                                   skips(dim), w.hulls, w.counts, w.ovf, w.n_ovf);
            }
            if (int e = check_launch("reference hulls")) return e;
            {
                ProfScope ps("edt_ref_hulls_fallback", s);
                hipLaunchKernelGGL(k_ref_hulls_fallback, dim3(fb_threads(n) / 256), dim3(256), 0, s, img, fin, g,
                                   resolution[dim], skips(dim), w.hulls, w.counts, w.ovf, w.n_ovf, w.stk, cap);
            }
            if (int e = check_launch("reference hulls fallback")) return e;
        }
        const RefHull *hl = periodic[dim] ? w.hulls : nullptr;
        {
            ProfScope ps(dim == 2 ? "edt_ref_parabolic_rows" : "edt_ref_parabolic", s);
            hipLaunchKernelGGL(k_ref_parabolic, dim3(nb), dim3(kWave), 0, s, fin, fout, g, resolution[dim], hl,
                               w.counts, sq, w.ovf, w.n_ovf + 1);
        }
        if (int e = check_launch("reference parabolic")) return e;
        {
            ProfScope ps("edt_ref_parabolic_fallback", s);
            hipLaunchKernelGGL(k_ref_parabolic_fallback, dim3(fb_threads(n) / 256), dim3(256), 0, s, fin, fout, g,
                               resolution[dim], hl, w.counts, sq, w.ovf, w.n_ovf + 1, w.stk, cap);
        }
        if (int e = check_launch("reference parabolic fallback")) return e;
        ++step;
    }
    return PMX_OK;
}

}  // extern "C"
