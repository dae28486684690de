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
// pmx_psd_batch_create / pmx_psd_batch_run: the fused batch.  Every pass of
// every level is one launch over all domains (each block finds its domain
// in a table of per-domain arguments), the erosion EDT is each domain's
// full exact EDT, and a replay issues a few hundred launches in place of a
// few tens of thousands.
//
// pmx_edt_autotune times the windowed EDT passes with 32- and 64-row tiles
// on a domain and keeps the faster choice per (shape, boundary, clamp
// class) for the rest of the process (and, through pmx_tune_cache_load /
// pmx_tune_cache_save, across processes).
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "edt_kernels.hpp"
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


// ---------------------------------------------------------------------------
// the fused batch
// ---------------------------------------------------------------------------

constexpr int kBatchMaxLevels = 250;
constexpr int kBatchFallbackLanes = 4096;  // per domain: deferred lines are rare

// per-domain arguments of the batched sweep setup: levels, the code map and
// the first-invasion state (one pass over the full EDT and the image)
struct PsdArgs {
    const int32_t *d2;  // the full EDT, contiguous box
    const uint8_t *img;
    int64_t in_s0, in_s1, out_s0, out_s1;
    int n1, n2;
    int64_t nvox;
    uint8_t *level, *state;
    int32_t *code_map;
    unsigned int *counts, *any;
};

// counts of every domain cleared (one block per domain)
__global__ __launch_bounds__(256) void k_psd_clear_batch(const PsdArgs *__restrict__ args, int n_levels)
{
    const PsdArgs &a = args[blockIdx.x];
    for (int k = threadIdx.x; k <= n_levels; k += blockDim.x) a.counts[k] = 0u;
}

// level[i] = first k with d2 > te[k] (n_levels if none), state = (img == 1),
// code map 0, per-level counts through an LDS histogram per block
__global__ __launch_bounds__(256) void k_psd_setup_batch(const PsdArgs *__restrict__ args,
                                                         const int64_t *__restrict__ first, int n,
                                                         const int64_t *__restrict__ te, int n_levels)
{
    __shared__ int64_t ste[kBatchMaxLevels];
    __shared__ unsigned int hist[kBatchMaxLevels];
    const int64_t gb = blockIdx.x;
    const int k = batch_slot(first, n, gb);
    const PsdArgs &a = args[k];
    for (int q = threadIdx.x; q < n_levels; q += blockDim.x) {
        ste[q] = te[q];
        hist[q] = 0;
    }
    __syncthreads();
    const int64_t nblk = first[k + 1] - first[k];
    // a box holds fewer than 2^31 voxels: 32-bit index arithmetic; contiguous boxes skip it
    const bool dense = a.in_s1 == a.n2 && a.out_s1 == a.n2 && a.in_s0 == (int64_t)a.n1 * a.n2 &&
                       a.out_s0 == (int64_t)a.n1 * a.n2;
    for (int64_t i = (gb - first[k]) * blockDim.x + threadIdx.x; i < a.nvox; i += nblk * blockDim.x) {
        const int64_t v = a.d2[i];
        int lo = 0, hi = n_levels;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (v > ste[mid])
                hi = mid;
            else
                lo = mid + 1;
        }
        a.level[i] = (uint8_t)lo;
        if (lo < n_levels) atomicAdd(hist + lo, 1u);
        int64_t ii = i, io = i;
        if (!dense) {
            const int i32 = (int)i, z = i32 % a.n2, t = i32 / a.n2, y = t % a.n1, x = t / a.n1;
            ii = x * a.in_s0 + y * a.in_s1 + z;
            io = x * a.out_s0 + y * a.out_s1 + z;
        }
        a.state[i] = a.img[ii] == 1 ? 1 : 0;
        a.code_map[io] = 0;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < n_levels; q += blockDim.x)
        if (hist[q]) atomicAdd(a.counts + q, hist[q]);
}

// any[k] = level k has voxels: without an inlet requirement they are the
// level's new centres (one block per domain)
__global__ __launch_bounds__(256) void k_psd_any_batch(const PsdArgs *__restrict__ args, int n_levels)
{
    const PsdArgs &a = args[blockIdx.x];
    for (int k = threadIdx.x; k < n_levels; k += blockDim.x) a.any[k] = a.counts[k] != 0u;
}

// one batched launch: its kernel's per-domain argument table and block prefix
struct BatchGroup {
    int n = 0;                 // domains
    int64_t total = 0;         // blocks
    const int64_t *first = nullptr;
    const void *args = nullptr;
};

}  // namespace
}  // namespace pmx

struct pmx_psd_batch {
    int n_domains = 0, n_levels = 0;
    std::vector<int64_t> t_dilate;
    const int64_t *te = nullptr;  // device
    // pass Z by lanes per row (16, 32, 64) for the full EDTs and the dilations
    pmx::BatchGroup z_full[3], z_dil[3];
    pmx::BatchGroup clear_bits;  // the long-row domains' deferred-line bitmaps
    // windowed passes by tile rows (32, 64)
    int tile[2] = {32, 64};
    pmx::BatchGroup y16[2], y32[2], x16[2], x32[2], ydil[2], xdil[2];
    pmx::BatchGroup fy16, fy32, fx16, fx32;
    pmx::BatchGroup setup;            // blocks over each domain's voxels
    const pmx::PsdArgs *psd = nullptr;  // one entry per domain (clear / any kernels)
};

namespace pmx {
namespace {

// Host side of the plan: each table is built as a vector and placed in one
// host buffer, uploaded once.
struct TableBuf {
    std::vector<char> bytes;
    template <class T> size_t put(const std::vector<T> &v)
    {
        const size_t off = align_up(bytes.size(), 256);
        bytes.resize(off + v.size() * sizeof(T));
        if (!v.empty()) memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

struct DomPlan {
    EdtPlan e;
    uint8_t *level = nullptr, *state = nullptr;
    int32_t *d2 = nullptr;
    unsigned int *counts = nullptr, *any = nullptr;
};

// the per-domain slices of the fused batch's workspace (sizes only when base
// is null), then the shared thresholds and the argument tables
size_t carve_fused(int n, const pmx_psd_domain *d, int32_t *const *edt_d2, int n_levels, size_t table_bytes,
                   void *base, std::vector<DomPlan> *plans, int64_t **te, char **tables)
{
    Carver c(base, SIZE_MAX);
    for (int i = 0; i < n; ++i) {
        const pmx_geom &g = d[i].g;
        const size_t nv = (size_t)(g.n[0] * g.n[1] * g.n[2]);
        DomPlan dp;
        const size_t eb = plan_edt(&g, nullptr, &dp.e, kBatchFallbackLanes);
        char *eslice = c.take<char>(eb);
        if (base) plan_edt(&g, eslice, &dp.e, kBatchFallbackLanes);
        dp.level = c.take<uint8_t>(nv + 16);  // slack: 8-byte loads past the end
        dp.state = c.take<uint8_t>(nv);
        dp.counts = c.take<unsigned int>(kBatchMaxLevels + 1);
        dp.any = c.take<unsigned int>(kBatchMaxLevels);
        int32_t *own = (edt_d2 && edt_d2[i]) ? nullptr : c.take<int32_t>(nv);
        dp.d2 = (edt_d2 && edt_d2[i]) ? edt_d2[i] : own;
        if (plans) plans->push_back(dp);
    }
    int64_t *t = c.take<int64_t>((size_t)std::max(n_levels, 1));
    char *tb = c.take<char>(table_bytes);
    if (te) *te = t;
    if (tables) *tables = tb;
    return c.need();
}

int lane_class(int64_t nz) { return nz <= 256 ? 0 : nz <= 512 ? 1 : 2; }  // 16, 32, 64 lanes per row

int check_fused(int n, const pmx_psd_domain *d, int n_levels, const int64_t *t_erode, const int64_t *t_dilate)
{
    if (n < 1 || !d) return fail(PMX_ERR_INVALID, "empty batch");
    if (n_levels < 1 || n_levels > kBatchMaxLevels)
        return fail(PMX_ERR_INVALID, "n_levels must be in [1, %d]", kBatchMaxLevels);
    for (int i = 0; i < n; ++i) {
        if (int e = check_psd(d[i].img, &d[i].g, nullptr, nullptr, n_levels, t_erode, t_dilate, d[i].code_map))
            return fail(e, "domain %d: %s", i, pmx_last_error());
        if (d[i].has_inlet)
            return fail(PMX_ERR_INVALID, "domain %d: the fused batch runs sweeps without an inlet requirement "
                                         "(pmx_psd_sweep_batch runs the others)", i);
        if (d[i].tile_rows != 0 && d[i].tile_rows != 32 && d[i].tile_rows != 64)
            return fail(PMX_ERR_INVALID, "domain %d: tile_rows %d: use 0, 32 or 64", i, d[i].tile_rows);
    }
    for (int k = 0; k < n_levels; ++k) {
        // no dilation may defer a line: its fallbacks are not part of the fused batch
        const int64_t cap = std::min<int64_t>(std::max<int64_t>(t_dilate[k], -1), INT32_MAX - 1) + 1;
        if (cap >= (kReach + 1) * (kReach + 1))
            return fail(PMX_ERR_INVALID, "t_dilate[%d] = %lld: the fused batch takes thresholds below %d",
                        k, (long long)t_dilate[k], (kReach + 1) * (kReach + 1) - 1);
    }
    return PMX_OK;
}

template <class T>
void launch_group(const BatchGroup &g, hipStream_t s, T launch)
{
    if (g.n > 0 && g.total > 0) launch(g);
}

bool aligned16(const void *p, int64_t s0, int64_t s1)
{
    return (uintptr_t)p % 16 == 0 && s0 % 16 == 0 && s1 % 16 == 0;
}

// Builds the fused plan for the workspace at base (null: sizes only): the
// argument tables go to tb (device addresses inside the table region of the
// workspace), the launch groups to b, the workspace size to *bytes.
void build_fused(int n, const pmx_psd_domain *d, int32_t *const *edt_d2, int n_levels, void *base, size_t *bytes,
                 TableBuf *tb, pmx_psd_batch *b)
{
    std::vector<DomPlan> plans;
    int64_t *te = nullptr;
    char *tables = nullptr;
    carve_fused(n, d, edt_d2, n_levels, 0, base, &plans, &te, &tables);
    b->te = te;
    auto dev = [&](size_t off) { return tables ? (const void *)(tables + off) : nullptr; };
    // a group from its per-domain arguments and block counts
    auto group = [&](auto &args, const std::vector<int64_t> &blocks, BatchGroup &g) {
        if (args.empty()) return;
        std::vector<int64_t> first(blocks.size() + 1, 0);
        for (size_t i = 0; i < blocks.size(); ++i) first[i + 1] = first[i] + blocks[i];
        g.n = (int)args.size();
        g.total = first.back();
        g.args = dev(tb->put(args));
        g.first = static_cast<const int64_t *>(dev(tb->put(first)));
    };
    // pass Z by lanes per row
    for (int c = 0; c < 3; ++c) {
        const int L = 16 << c;
        std::vector<PassZArgs> full, dil;
        std::vector<ClearArgs> clr;
        std::vector<int64_t> blocks, cblocks;
        for (int i = 0; i < n; ++i) {
            const pmx_geom &g = d[i].g;
            if (lane_class(g.n[2]) != c) continue;
            const DomPlan &dp = plans[i];
            const int64_t rows = g.n[0] * g.n[1];
            PassZArgs a{};
            a.img = d[i].img;
            a.in_sx = g.in_s[0];
            a.in_sy = g.in_s[1];
            a.ny = (int)g.n[1];
            a.nz = (int)g.n[2];
            a.rows = rows;
            a.bc_lo = g.bc[4];
            a.bc_hi = g.bc[5];
            a.vec_in = aligned16(a.img, a.in_sx, a.in_sy) ? 1 : 0;
            a.pitch = dp.e.pitch;
            a.out = dp.e.dz;
            a.counters = dp.e.counters;
            a.bits = dp.e.bits_y;
            a.n_bits = (int64_t)dp.e.n_bits;
            full.push_back(a);
            // the dilations: distance to the level's voxels of the contiguous level map; WALL
            // faces see nothing beyond them (run_psd's dilation geometry)
            a.img = dp.level;
            a.in_sx = g.n[1] * g.n[2];
            a.in_sy = g.n[2];
            a.bc_lo = g.bc[4] == PMX_BC_WALL ? PMX_BC_END : g.bc[4];
            a.bc_hi = g.bc[5] == PMX_BC_WALL ? PMX_BC_END : g.bc[5];
            a.vec_in = aligned16(a.img, a.in_sx, a.in_sy) ? 1 : 0;
            a.proceed = dp.any;
            dil.push_back(a);
            blocks.push_back(c == 2 ? (rows + 3) / 4 : (rows + 4 * (kWave / L) - 1) / (4 * (kWave / L)));
            if (c == 2) {
                clr.push_back(ClearArgs{dp.e.bits_y, (int64_t)dp.e.n_bits});
                cblocks.push_back(std::max<int64_t>(1, ((int64_t)dp.e.n_bits + 255) / 256));
            }
        }
        group(full, blocks, b->z_full[c]);
        group(dil, blocks, b->z_dil[c]);
        if (c == 2) group(clr, cblocks, b->clear_bits);
    }
    // windowed passes by tile rows
    for (int ti = 0; ti < 2; ++ti) {
        const int T = b->tile[ti];
        std::vector<WinArgs> y16, y32, x16, x32, ydil, xdil;
        std::vector<int64_t> by, bx;
        for (int i = 0; i < n; ++i) {
            const int tr = d[i].tile_rows == 0 ? 64 : d[i].tile_rows;
            if (tr != T) continue;
            const pmx_geom &g = d[i].g;
            const DomPlan &dp = plans[i];
            const int nx = (int)g.n[0], ny = (int)g.n[1], nz = (int)g.n[2];
            const int64_t pitch = dp.e.pitch;
            auto bce = [&](int f) { return g.bc[f] == PMX_BC_WALL ? PMX_BC_END : g.bc[f]; };
            WinGeom gy{ny, pitch, ny * pitch, nx, nz, g.bc[2], g.bc[3], kBig, pitch, pitch, nx * pitch};
            WinGeom gx{nx, pitch, nx * pitch, ny, nz, g.bc[0], g.bc[1], kBig, pitch, 0, 0};
            WinGeom gyd{ny, pitch, ny * pitch, nx, nz, bce(2), bce(3), kBig, pitch, pitch, nx * pitch};
            WinGeom gxd{nx, pitch, nx * pitch, ny, nz, bce(0), bce(1), kBig, pitch, 0, 0};
            by.push_back(window_blocks(gy, T, gy.tpb));
            bx.push_back(window_blocks(gx, T, gx.tpb));
            PassXOut full{};
            full.d2 = dp.d2;
            full.sx = g.n[1] * g.n[2];
            full.sy = g.n[2];
            PassXOut psd{};
            psd.sx = g.out_s[0];
            psd.sy = g.out_s[1];
            psd.wall_mask = d[i].wall_mask;
            psd.psd_state = dp.state;
            psd.ssx = g.n[1] * g.n[2];
            psd.ssy = g.n[2];
            psd.psd_map = d[i].code_map;
            const unsigned int *over = dp.e.counters + 2, *none = dp.e.counters + 3;
            y16.push_back(WinArgs{dp.e.dz, dp.e.b, PassXOut{}, gy, nx, ny, dp.e.bits_y, dp.e.counters + 0, dp.e.ovf_list,
                                  nullptr});
            y32.push_back(WinArgs{dp.e.dz, dp.e.b, PassXOut{}, gy, nx, ny, dp.e.bits_y, dp.e.counters + 0, dp.e.ovf_list,
                                  over});
            x16.push_back(WinArgs{dp.e.b, nullptr, full, gx, nx, ny, dp.e.bits_x, dp.e.counters + 1, dp.e.ovf_list,
                                  none});
            x32.push_back(WinArgs{dp.e.b, nullptr, full, gx, nx, ny, dp.e.bits_x, dp.e.counters + 1, dp.e.ovf_list,
                                  over});
            ydil.push_back(WinArgs{dp.e.dz, dp.e.b, PassXOut{}, gyd, nx, ny, dp.e.bits_y, dp.e.counters + 0,
                                   dp.e.ovf_list, dp.any});
            xdil.push_back(WinArgs{dp.e.b, nullptr, psd, gxd, nx, ny, dp.e.bits_x, dp.e.counters + 1, dp.e.ovf_list,
                                   dp.any});
        }
        group(y16, by, b->y16[ti]);
        group(y32, by, b->y32[ti]);
        group(x16, bx, b->x16[ti]);
        group(x32, bx, b->x32[ti]);
        group(ydil, by, b->ydil[ti]);
        group(xdil, bx, b->xdil[ti]);
    }
    // the full EDTs' fallbacks, and the sweep setup
    {
        std::vector<FallbackArgs> fy16, fy32, fx16, fx32;
        std::vector<PsdArgs> psd;
        std::vector<int64_t> fb, vb;
        for (int i = 0; i < n; ++i) {
            const pmx_geom &g = d[i].g;
            const DomPlan &dp = plans[i];
            const int nx = (int)g.n[0], ny = (int)g.n[1], nz = (int)g.n[2];
            const int64_t pitch = dp.e.pitch;
            PassXOut full{};
            full.d2 = dp.d2;
            full.sx = g.n[1] * g.n[2];
            full.sy = g.n[2];
            const unsigned int *over = dp.e.counters + 2, *none = dp.e.counters + 3;
            FallbackArgs y{dp.e.dz, dp.e.b, PassXOut{}, nx, ny, nz, pitch, nx * pitch, pitch, g.bc[2], g.bc[3],
                           dp.e.counters + 0, dp.e.ovf_list, dp.e.stk, dp.e.cap, nullptr};
            fy16.push_back(y);
            y.proceed = over;
            fy32.push_back(y);
            FallbackArgs x{dp.e.b, nullptr, full, nx, ny, nz, pitch, nx * pitch, pitch, g.bc[0], g.bc[1],
                           dp.e.counters + 1, dp.e.ovf_list, dp.e.stk, dp.e.cap, none};
            fx16.push_back(x);
            x.proceed = over;
            fx32.push_back(x);
            fb.push_back(dp.e.fallback_lanes / 256);
            const int64_t nv = g.n[0] * g.n[1] * g.n[2];
            psd.push_back(PsdArgs{dp.d2, d[i].img, g.in_s[0], g.in_s[1], g.out_s[0], g.out_s[1], (int)g.n[1],
                                  (int)g.n[2], nv, dp.level, dp.state, d[i].code_map, dp.counts, dp.any});
            vb.push_back(std::max<int64_t>(1, (nv + 256 * 16 - 1) / (256 * 16)));
        }
        group(fy16, fb, b->fy16);
        group(fy32, fb, b->fy32);
        group(fx16, fb, b->fx16);
        group(fx32, fb, b->fx32);
        group(psd, vb, b->setup);
        b->psd = static_cast<const PsdArgs *>(b->setup.args);
    }
    *bytes = carve_fused(n, d, edt_d2, n_levels, tb->bytes.size(), nullptr, nullptr, nullptr, nullptr);
}

int run_fused(const pmx_psd_batch *b, hipStream_t s)
{
    auto grid = [](int64_t blocks) { return dim3((unsigned)blocks); };
    // the full exact EDT of every domain (16 bits first, int32 passes where a value saturated)
    {
        ProfScope ps("batch_edt", s);
        launch_group(b->clear_bits, s, [&](const BatchGroup &g) {
            hipLaunchKernelGGL(k_clear_words_batch, grid(g.total), dim3(256), 0, s,
                               static_cast<const ClearArgs *>(g.args), g.first, g.n);
        });
#define PMX_ZB(INV, MT, L, G, MATCH, GATE) PMX_ZBD(INV, MT, L, uint16_t, G, MATCH, GATE)
#define PMX_ZBD(INV, MT, L, D, G, MATCH, GATE)                                                                \
    launch_group(G, s, [&](const BatchGroup &g) {                                                             \
        hipLaunchKernelGGL((k_pass_z_batch<INV, MT, L, D>), grid(g.total), dim3(256), 0, s,                   \
                           static_cast<const PassZArgs *>(g.args), g.first, g.n, (uint32_t)(MATCH), GATE);   \
    })
        PMX_ZB(false, false, 16, b->z_full[0], 0, -1);
        PMX_ZB(false, false, 32, b->z_full[1], 0, -1);
        PMX_ZB(false, false, 64, b->z_full[2], 0, -1);
#define PMX_WB(T, W, E, O, D2, G, GATE, CAP, THR, CODE)                                                       \
    launch_group(G, s, [&](const BatchGroup &g) {                                                             \
        hipLaunchKernelGGL((k_window_batch<T, W, E, O, D2>), grid(g.total), dim3(kWinThreads), 0, s,          \
                           static_cast<const WinArgs *>(g.args), g.first, g.n, g.total, GATE, CAP, THR, CODE); \
    })
// pass X of a dilation (input B: Sq16, or Sq8 under a cap below 256): the
// first-invasion update is its only output
#define PMX_WBP(T, W, B, G, GATE, CAP, THR, CODE)                                                             \
    launch_group(G, s, [&](const BatchGroup &g) {                                                             \
        hipLaunchKernelGGL((k_window_batch<T, W, B, int32_t, false, true>), grid(g.total), dim3(kWinThreads),   \
                           0, s, static_cast<const WinArgs *>(g.args), g.first, g.n, g.total, GATE, CAP, THR, CODE); \
    })
        PMX_WB(32, 16, uint16_t, Sq16, false, b->y16[0], -1, -1, 0, -1);
        PMX_WB(64, 16, uint16_t, Sq16, false, b->y16[1], -1, -1, 0, -1);
        launch_group(b->fy16, s, [&](const BatchGroup &g) {
            hipLaunchKernelGGL(k_fallback_y_batch<Sq16>, grid(g.total), dim3(256), 0, s,
                               static_cast<const FallbackArgs *>(g.args), g.first, g.n, kBig);
        });
        PMX_WB(32, 16, uint16_t, int32_t, false, b->y32[0], -1, -1, 0, -1);
        PMX_WB(64, 16, uint16_t, int32_t, false, b->y32[1], -1, -1, 0, -1);
        launch_group(b->fy32, s, [&](const BatchGroup &g) {
            hipLaunchKernelGGL(k_fallback_y_batch<int32_t>, grid(g.total), dim3(256), 0, s,
                               static_cast<const FallbackArgs *>(g.args), g.first, g.n, kBig);
        });
        PMX_WB(32, 16, Sq16, int32_t, true, b->x16[0], -1, -1, 0, -1);
        PMX_WB(64, 16, Sq16, int32_t, true, b->x16[1], -1, -1, 0, -1);
        launch_group(b->fx16, s, [&](const BatchGroup &g) {
            hipLaunchKernelGGL(k_fallback_x_batch<Sq16>, grid(g.total), dim3(256), 0, s,
                               static_cast<const FallbackArgs *>(g.args), g.first, g.n, kBig);
        });
        PMX_WB(32, 16, int32_t, int32_t, true, b->x32[0], -1, -1, 0, -1);
        PMX_WB(64, 16, int32_t, int32_t, true, b->x32[1], -1, -1, 0, -1);
        launch_group(b->fx32, s, [&](const BatchGroup &g) {
            hipLaunchKernelGGL(k_fallback_x_batch<int32_t>, grid(g.total), dim3(256), 0, s,
                               static_cast<const FallbackArgs *>(g.args), g.first, g.n, kBig);
        });
    }
    if (int e = check_launch("batched EDT")) return e;
    {
        ProfScope ps("batch_levels", s);
        hipLaunchKernelGGL(k_psd_clear_batch, dim3((unsigned)b->n_domains), dim3(256), 0, s, b->psd, b->n_levels);
        hipLaunchKernelGGL(k_psd_setup_batch, grid(b->setup.total), dim3(256), 0, s, b->psd, b->setup.first,
                           b->setup.n, b->te, b->n_levels);
        hipLaunchKernelGGL(k_psd_any_batch, dim3((unsigned)b->n_domains), dim3(256), 0, s, b->psd, b->n_levels);
    }
    if (int e = check_launch("batched levels")) return e;
    ProfScope ps("batch_dilate", s);
    for (int lv = 0; lv < b->n_levels; ++lv) {
        // the level's centres (level map == lv), dilated to d2 <= t, first-invasion update
        const int32_t thr = (int32_t)std::min<int64_t>(std::max<int64_t>(b->t_dilate[lv], -1), INT32_MAX - 1);
        const int cap = thr + 1;
        const bool wide = cap > 17 * 17;
        if (cap <= 255) {
            // 8-bit distances from pass Z
            PMX_ZBD(true, true, 16, uint8_t, b->z_dil[0], lv, lv);
            PMX_ZBD(true, true, 32, uint8_t, b->z_dil[1], lv, lv);
            PMX_ZBD(true, true, 64, uint8_t, b->z_dil[2], lv, lv);
        } else {
            PMX_ZB(true, true, 16, b->z_dil[0], lv, lv);
            PMX_ZB(true, true, 32, b->z_dil[1], lv, lv);
            PMX_ZB(true, true, 64, b->z_dil[2], lv, lv);
        }
        if (wide) {
            PMX_WB(32, 32, uint16_t, Sq16, false, b->ydil[0], lv, cap, 0, -1);
            PMX_WB(64, 32, uint16_t, Sq16, false, b->ydil[1], lv, cap, 0, -1);
            PMX_WBP(32, 32, Sq16, b->xdil[0], lv, cap, thr, lv + 1);
            PMX_WBP(64, 32, Sq16, b->xdil[1], lv, cap, thr, lv + 1);
        } else if (cap <= 255) {
            // 8-bit squared distances between passes Y and X
            PMX_WB(32, 16, Dz8, Sq8, false, b->ydil[0], lv, cap, 0, -1);
            PMX_WB(64, 16, Dz8, Sq8, false, b->ydil[1], lv, cap, 0, -1);
            PMX_WBP(32, 16, Sq8, b->xdil[0], lv, cap, thr, lv + 1);
            PMX_WBP(64, 16, Sq8, b->xdil[1], lv, cap, thr, lv + 1);
        } else {
            PMX_WB(32, 16, uint16_t, Sq16, false, b->ydil[0], lv, cap, 0, -1);
            PMX_WB(64, 16, uint16_t, Sq16, false, b->ydil[1], lv, cap, 0, -1);
            PMX_WBP(32, 16, Sq16, b->xdil[0], lv, cap, thr, lv + 1);
            PMX_WBP(64, 16, Sq16, b->xdil[1], lv, cap, thr, lv + 1);
        }
        if (int e = check_launch("batched dilation")) return e;
    }
#undef PMX_ZB
#undef PMX_ZBD
#undef PMX_WB
#undef PMX_WBP
    return PMX_OK;
}

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


int pmx_psd_batch_plan_size(int n_domains, const pmx_psd_domain *domains, int32_t *const *edt_d2, int n_levels,
                            const int64_t *t_erode, const int64_t *t_dilate, size_t *bytes)
{
    using namespace pmx;
    clear_error();
    if (!bytes) return fail(PMX_ERR_INVALID, "bytes is NULL");
    if (int e = check_fused(n_domains, domains, n_levels, t_erode, t_dilate)) return e;
    TableBuf tb;
    pmx_psd_batch b;
    build_fused(n_domains, domains, edt_d2, n_levels, nullptr, bytes, &tb, &b);
    return PMX_OK;
}

int pmx_psd_batch_create(int n_domains, const pmx_psd_domain *domains, int32_t *const *edt_d2, int n_levels,
                         const int64_t *t_erode, const int64_t *t_dilate, void *workspace, size_t workspace_bytes,
                         hipStream_t s, pmx_psd_batch **out)
{
    using namespace pmx;
    clear_error();
    if (!out) return fail(PMX_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (int e = check_fused(n_domains, domains, n_levels, t_erode, t_dilate)) return e;
    size_t need = 0;
    {
        TableBuf tb;
        pmx_psd_batch probe;
        build_fused(n_domains, domains, edt_d2, n_levels, nullptr, &need, &tb, &probe);
    }
    if (!workspace || workspace_bytes < need)
        return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    auto *b = new pmx_psd_batch;
    b->n_domains = n_domains;
    b->n_levels = n_levels;
    b->t_dilate.assign(t_dilate, t_dilate + n_levels);
    TableBuf tb;
    size_t got = 0;
    build_fused(n_domains, domains, edt_d2, n_levels, workspace, &got, &tb, b);
    // the tables and thresholds, once (synchronous: the plan is made before any capture)
    char *tables = nullptr;
    int64_t *te = nullptr;
    carve_fused(n_domains, domains, edt_d2, n_levels, tb.bytes.size(), workspace, nullptr, &te, &tables);
    if (hipMemcpyAsync(tables, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(te, t_erode, sizeof(int64_t) * (size_t)n_levels, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        delete b;
        return fail(PMX_ERR_HIP, "uploading the batch tables failed");
    }
    *out = b;
    return PMX_OK;
}

int pmx_psd_batch_run(const pmx_psd_batch *b, hipStream_t s)
{
    using namespace pmx;
    clear_error();
    if (!b) return fail(PMX_ERR_INVALID, "batch is NULL");
    return run_fused(b, s);
}

void pmx_psd_batch_destroy(pmx_psd_batch *b) { delete b; }


int pmx_tune_cache_save(const char *path)
{
    using namespace pmx;
    clear_error();
    if (!path) return fail(PMX_ERR_INVALID, "path is NULL");
    std::map<TuneKey, int> copy;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        copy = g_tune;
    }
    FILE *f = fopen(path, "w");
    if (!f) return fail(PMX_ERR_INVALID, "cannot write %s", path);
    fprintf(f, "# pmx tile choices: n0 n1 n2 bc0..bc5 invert wide tile_rows\n");
    for (const auto &kv : copy) {
        const TuneKey &k = kv.first;
        fprintf(f, "%lld %lld %lld %d %d %d %d %d %d %d %d %d\n", (long long)k.n[0], (long long)k.n[1],
                (long long)k.n[2], k.bc[0], k.bc[1], k.bc[2], k.bc[3], k.bc[4], k.bc[5], k.invert, k.wide, kv.second);
    }
    const bool ok = fclose(f) == 0;
    return ok ? PMX_OK : fail(PMX_ERR_INVALID, "writing %s failed", path);
}

int pmx_tune_cache_load(const char *path)
{
    using namespace pmx;
    clear_error();
    if (!path) return fail(PMX_ERR_INVALID, "path is NULL");
    FILE *f = fopen(path, "r");
    if (!f) return PMX_OK;  // nothing saved yet
    char line[512];
    int lineno = 0, rc = PMX_OK;
    std::vector<std::pair<TuneKey, int>> got;
    while (fgets(line, sizeof line, f)) {
        ++lineno;
        if (line[0] == '#' || line[0] == '\n') continue;
        long long n0, n1, n2;
        int bc[6], inv, wide, tr;
        if (sscanf(line, "%lld %lld %lld %d %d %d %d %d %d %d %d %d", &n0, &n1, &n2, &bc[0], &bc[1], &bc[2], &bc[3],
                   &bc[4], &bc[5], &inv, &wide, &tr) != 12 ||
            (tr != 32 && tr != 64)) {
            rc = fail(PMX_ERR_INVALID, "%s:%d: not a tile choice", path, lineno);
            break;
        }
        TuneKey k{{n0, n1, n2}, {bc[0], bc[1], bc[2], bc[3], bc[4], bc[5]}, inv != 0, wide != 0};
        got.emplace_back(k, tr);
    }
    fclose(f);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &kv : got) g_tune[kv.first] = kv.second;
    return PMX_OK;
}

}  // extern "C"
