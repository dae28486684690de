// batch.hip -- many unevenly sized domains per call, and the EDT tile
// autotuner.
//
// pmx_psd_sweep_batch runs the fused pore-size-distribution sweep of every
// domain of a batch on a set of caller streams ("lanes"): domains go to
// lanes by longest-processing-time order on their voxel counts, each lane
// runs its domains one after another in its own slice of the workspace,
// and nothing synchronises with the host, so a caller can capture the
// whole batch in one graph (fork the lanes from the capturing stream and
// join them back) and replay it.
//
// pmx_edt_autotune times the windowed EDT passes with 32- and 64-row tiles
// on a domain and keeps the faster choice per (shape, boundary, clamp
// class) for the rest of the process.
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "psd_internal.hpp"

namespace pmx {
namespace {

// domain -> lane assignment, longest processing time first on voxel counts
std::vector<int> assign_lanes(int n, const pmx_psd_domain *d, int n_lanes)
{
    std::vector<int> idx(n), lane(n, 0);
    for (int i = 0; i < n; ++i) idx[i] = i;
    auto vox = [&](int i) { return d[i].g.n[0] * d[i].g.n[1] * d[i].g.n[2]; };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return vox(a) > vox(b); });
    std::vector<int64_t> load(n_lanes, 0);
    for (int i : idx) {
        int best = 0;
        for (int l = 1; l < n_lanes; ++l)
            if (load[l] < load[best]) best = l;
        lane[i] = best;
        load[best] += vox(i);
    }
    return lane;
}

// per-lane workspace slices: the largest sweep workspace among the lane's domains
std::vector<size_t> lane_bytes(int n, const pmx_psd_domain *d, int n_lanes, const std::vector<int> &lane)
{
    std::vector<size_t> b(n_lanes, 0);
    for (int i = 0; i < n; ++i) b[lane[i]] = std::max(b[lane[i]], align_up(psd_workspace_bytes(&d[i].g), 256));
    return b;
}

int check_batch(int n, const pmx_psd_domain *d, int n_lanes)
{
    if (n < 1 || !d) return fail(PMX_ERR_INVALID, "empty batch");
    if (n_lanes < 1 || n_lanes > 64) return fail(PMX_ERR_INVALID, "n_lanes %d outside [1, 64]", n_lanes);
    for (int i = 0; i < n; ++i) {
        if (int e = validate_geom(&d[i].g)) return fail(e, "domain %d: %s", i, pmx_last_error());
        if (d[i].tile_rows != 0 && d[i].tile_rows != 32 && d[i].tile_rows != 64)
            return fail(PMX_ERR_INVALID, "domain %d: tile_rows %d: use 0, 32 or 64", i, d[i].tile_rows);
    }
    return PMX_OK;
}

struct TuneKey {
    int64_t n[3];
    int32_t bc[6];
    int invert, wide;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(n[0], n[1], n[2], bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], invert, wide) <
               std::tie(o.n[0], o.n[1], o.n[2], o.bc[0], o.bc[1], o.bc[2], o.bc[3], o.bc[4], o.bc[5], o.invert, o.wide);
    }
};

std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;  // read-mostly: written once per key

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_psd_batch_workspace_size(int n_domains, const pmx_psd_domain *domains, int n_lanes, size_t *bytes)
{
    using namespace pmx;
    clear_error();
    if (!bytes) return fail(PMX_ERR_INVALID, "bytes is NULL");
    if (int e = check_batch(n_domains, domains, n_lanes)) return e;
    const auto lane = assign_lanes(n_domains, domains, n_lanes);
    size_t total = 256;  // alignment slack
    for (size_t b : lane_bytes(n_domains, domains, n_lanes, lane)) total += b;
    *bytes = total;
    return PMX_OK;
}

int pmx_psd_sweep_batch(int n_domains, const pmx_psd_domain *domains, int n_levels, const int64_t *t_erode,
                        const int64_t *t_dilate, int n_lanes, const hipStream_t *lanes, void *workspace,
                        size_t workspace_bytes)
{
    using namespace pmx;
    clear_error();
    if (int e = check_batch(n_domains, domains, n_lanes)) return e;
    if (!lanes) return fail(PMX_ERR_INVALID, "lanes is NULL");
    for (int i = 0; i < n_domains; ++i) {
        const pmx_psd_domain &d = domains[i];
        if (int e = check_psd(d.img, &d.g, d.has_inlet ? d.inlet_lo : nullptr, d.has_inlet ? d.inlet_ext : nullptr,
                              n_levels, t_erode, t_dilate, d.code_map))
            return fail(e, "domain %d: %s", i, pmx_last_error());
    }
    const auto lane = assign_lanes(n_domains, domains, n_lanes);
    const auto lb = lane_bytes(n_domains, domains, n_lanes, lane);
    size_t need = 256;
    for (size_t b : lb) need += b;
    if (!workspace || workspace_bytes < need)
        return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    std::vector<char *> base(n_lanes);
    char *p = reinterpret_cast<char *>(align_up(reinterpret_cast<uintptr_t>(workspace), 256));
    for (int l = 0; l < n_lanes; ++l) {
        base[l] = p;
        p += lb[l];
    }
    for (int i = 0; i < n_domains; ++i) {
        const pmx_psd_domain &d = domains[i];
        const int l = lane[i];
        if (int e = run_psd(d.img, &d.g, d.has_inlet ? d.inlet_lo : nullptr, d.has_inlet ? d.inlet_ext : nullptr,
                            n_levels, t_erode, t_dilate, d.wall_mask, d.code_map, base[l], lb[l], lanes[l],
                            d.tile_rows))
            return fail(e, "domain %d: %s", i, pmx_last_error());
    }
    return PMX_OK;
}

int pmx_edt_autotune_workspace_size(const pmx_geom *g, size_t *bytes)
{
    using namespace pmx;
    clear_error();
    if (int e = validate_geom(g)) return e;
    if (!bytes) return fail(PMX_ERR_INVALID, "bytes is NULL");
    EdtPlan p;
    *bytes = plan_edt(g, nullptr, &p) + align_up((size_t)(g->n[0] * g->n[1] * g->n[2]), 256) + 512;
    return PMX_OK;
}

int pmx_edt_autotune(const uint8_t *img, const pmx_geom *g, int64_t cap, int invert, int32_t *tile_rows,
                     void *workspace, size_t workspace_bytes, hipStream_t s)
{
    using namespace pmx;
    clear_error();
    if (int e = validate_geom(g)) return e;
    if (!img || !tile_rows) return fail(PMX_ERR_INVALID, "NULL argument");
    EdtPlan plan;
    const size_t edt_bytes = plan_edt(g, nullptr, &plan);
    const size_t need = edt_bytes + align_up((size_t)(g->n[0] * g->n[1] * g->n[2]), 256) + 512;
    if (!workspace || workspace_bytes < need)
        return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    const int64_t ccap = std::min<int64_t>(std::max<int64_t>(cap, 0), int64_t(1) << 30);
    TuneKey key{{g->n[0], g->n[1], g->n[2]}, {g->bc[0], g->bc[1], g->bc[2], g->bc[3], g->bc[4], g->bc[5]},
                invert != 0, ccap < (int64_t(1) << 30) && ccap > 17 * 17};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            *tile_rows = it->second;
            return PMX_OK;
        }
    }
    // time each tile choice on this domain: a thresholded EDT into scratch
    uint8_t *scratch = reinterpret_cast<uint8_t *>(
        align_up(reinterpret_cast<uintptr_t>(workspace), 256) + align_up(edt_bytes, 256));
    if ((size_t)(scratch - reinterpret_cast<uint8_t *>(workspace)) + (size_t)(g->n[0] * g->n[1] * g->n[2]) >
        workspace_bytes)
        return fail(PMX_ERR_WORKSPACE, "workspace too small for the autotune scratch");
    PassXOut o{};
    o.u8 = scratch;
    o.sx = g->n[1] * g->n[2];
    o.sy = g->n[2];
    o.t = (int32_t)std::min<int64_t>(ccap - 1, INT32_MAX - 1);
    o.v_le = 1;
    pmx_geom gc = *g;
    gc.out_s[0] = o.sx;
    gc.out_s[1] = o.sy;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return fail(PMX_ERR_HIP, "event creation failed");
    const int choices[2] = {32, 64};
    float best_ms = 0.f;
    int best = 64;
    int rc = PMX_OK;
    for (int c = 0; c < 2 && rc == PMX_OK; ++c) {
        rc = run_edt(img, &gc, invert != 0, o, ccap, workspace, edt_bytes + 256, s, nullptr, choices[c]);  // warm-up
        if (rc) break;
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < 3 && rc == PMX_OK; ++r)
            rc = run_edt(img, &gc, invert != 0, o, ccap, workspace, edt_bytes + 256, s, nullptr, choices[c]);
        (void)hipEventRecord(e1, s);
        float ms = 0.f;
        if (rc == PMX_OK && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = fail(PMX_ERR_HIP, "timing failed");
        if (rc == PMX_OK && (c == 0 || ms < best_ms)) {
            best_ms = ms;
            best = choices[c];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune.emplace(key, best);
    }
    *tile_rows = best;
    return PMX_OK;
}

}  // extern "C"
