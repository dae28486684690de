// edt.hip -- exact separable squared EDT and EDT-threshold morphology.
//
// Replaces the reference's three per-line passes (_distance.hpp:135-272,
// driven by _distance.pyx:86-127,308-353 from
// euclidean_distance_transform.py:56-101,208-303) with three kernels:
//
//   pass Z  one wavefront per z-row: 1-D distance to the nearest solid
//           from per-lane 16-voxel solid masks and wave scans; stored as
//           uint16 (distances are at most 16384).
//   pass Y  windowed minimum of parabolas along y: a workgroup stages a
//           tile of rows x 64 z-columns (plus halo rows) in LDS and scans
//           each voxel outward until k^2 >= its best value, so the work per
//           voxel is about its distance to the nearest solid.  Output
//           stored [y][x][z].
//   pass X  the same along x over the pass-Y array, writing the requested
//           outputs (int32 d2, float32 / float64 sqrt, or a u8 threshold).
//
// A voxel not settled within the tile continues its scan in global memory
// up to 128 rows; one still open after that defers its whole line to the
// exact lower envelope of parabolas (Felzenszwalb, int64 intersection
// tests), run by a persistent fallback kernel with stacks in the workspace
// (capacity 3n+4, enough for a periodic line with no solid at all, which
// needs the neighbouring periods in full).  The envelope is evaluated and
// its stack reset at every solid: a zero-height parabola dominates
// everything behind it.
//
// The result is the exact squared EDT.  At unit resolution the reference's
// float32 arithmetic yields the same integers whenever a solid is
// reachable and every axis is at most 1024 voxels (tests/test_oracle_edt.py
// checks the reference against exact integers).
//
// Boundaries are in-kernel: END = nothing beyond the face, WALL = a solid
// layer just beyond it (the reference's zero pad, subdomain.py:273-287),
// PERIODIC = wrap (the reference's correctors and boundary hulls,
// euclidean_distance_transform.py:395-528).
#include "edt_kernels.hpp"
namespace pmx {
namespace {

constexpr int kDefaultTileRows = 64;  // rows per tile of the windowed passes (32 is the other choice)
constexpr size_t kFallbackStackBytes = size_t(256) << 20;

// launches one windowed pass with tile rows T and halo W
template <int T, int W, class E, class O>
void launch_window(const E *in, O *out, const PassXOut &o, const WinGeom &wg0, int nx, int ny,
                   uint32_t *bits, uint32_t *count, uint32_t *list, const unsigned int *proceed, hipStream_t s)
{
    WinGeom wg = wg0;
    wg.tpb = tiles_per_block(wg, T);
    const int64_t blocks = window_blocks(wg, T, wg.tpb);
    constexpr bool kX = !std::is_same<E, uint16_t>::value && !std::is_same<E, Dz8>::value;  // pass X
    const bool d2only = kX && o.d2 && !o.f32sq && !o.f32 && !o.f64 && !o.u8 && !o.psd_map;
    const bool psd_only = kX && o.psd_map && !o.d2 && !o.f32sq && !o.f32 && !o.f64 && !o.u8;
    if (psd_only)
        hipLaunchKernelGGL((k_window<T, W, E, O, false, false, true>), dim3((unsigned)blocks), dim3(kWinThreads), 0, s,
                           in, out, o, wg, nx, ny, bits, count, list, proceed);
    else if (d2only)
        hipLaunchKernelGGL((k_window<T, W, E, O, true>), dim3((unsigned)blocks), dim3(kWinThreads), 0, s, in, out, o,
                           wg, nx, ny, bits, count, list, proceed);
    else
        hipLaunchKernelGGL((k_window<T, W, E, O, false>), dim3((unsigned)blocks), dim3(kWinThreads), 0, s, in, out,
                           o, wg, nx, ny, bits, count, list, proceed);
}

template <class E, class O>
void dispatch_window(int tile_rows, bool wide, const E *in, O *out, const PassXOut &o, const WinGeom &wg,
                     int nx, int ny, uint32_t *bits, uint32_t *count, uint32_t *list, const unsigned int *proceed,
                     hipStream_t s)
{
    if (tile_rows == 32) {
        if (wide)
            launch_window<32, 32>(in, out, o, wg, nx, ny, bits, count, list, proceed, s);
        else
            launch_window<32, 16>(in, out, o, wg, nx, ny, bits, count, list, proceed, s);
    } else {
        if (wide)
            launch_window<64, 32>(in, out, o, wg, nx, ny, bits, count, list, proceed, s);
        else
            launch_window<64, 16>(in, out, o, wg, nx, ny, bits, count, list, proceed, s);
    }
}

}  // namespace

size_t plan_edt(const pmx_geom *g, void *ws, EdtPlan *p, int max_lanes)
{
    Carver c(ws, SIZE_MAX);
    // working rows padded to a multiple of 8 voxels: 16-byte aligned rows
    // for both the uint16 pass-Z array and the int32 pass-Y array
    p->pitch = (g->n[2] + 7) / 8 * 8;
    const size_t vox = (size_t)(g->n[0] * g->n[1] * p->pitch);
    p->dz = c.take<uint16_t>(vox);
    p->b = c.take<int32_t>(vox);
    p->counters = c.take<uint32_t>(4);
    const size_t ly = (size_t)(g->n[0] * g->n[2]), lx = (size_t)(g->n[1] * g->n[2]);
    p->n_bits = (ly + 31) / 32 + (lx + 31) / 32;
    p->bits_y = c.take<uint32_t>(p->n_bits);
    p->bits_x = p->bits_y + (ly + 31) / 32;
    p->ovf_list = c.take<uint32_t>(std::max(ly, lx));
    p->cap = (int)(3 * std::max(g->n[0], g->n[1]) + 4);
    // persistent fallback lanes: one per deferred line up to 32768, with
    // the stacks bounded to kFallbackStackBytes
    const size_t per_lane = (size_t)(2 * p->cap + kChunk) * sizeof(int);
    size_t lanes = std::min<size_t>((size_t)std::max(max_lanes, 256) / 256 * 256, (std::max(ly, lx) + 255) / 256 * 256);
    lanes = std::min(lanes, std::max<size_t>(256, kFallbackStackBytes / per_lane / 256 * 256));
    p->fallback_lanes = (int)lanes;
    p->stk = c.take<int>((size_t)(2 * p->cap + kChunk) * lanes);
    return c.need();
}

namespace {

// The launches of passes Y and X and their fallbacks, with the intermediate
// array `b` of type B (stored [y][x][z] with the row pitch of pass Z's
// array: pass X reads rows pitch * sizeof(B) bytes apart), each gated by
// `gate` (device memory; nothing runs while it reads 0).  name: profiling
// scope.
struct YX {
    const EdtPlan &p;
    const pmx_geom *g;
    const PassXOut &o;
    int ccap, tr;
    bool wide;
    hipStream_t s;
    int nx() const { return (int)g->n[0]; }
    int ny() const { return (int)g->n[1]; }
    int nz() const { return (int)g->n[2]; }
    int64_t bsx() const { return p.pitch; }
    int64_t bsy() const { return (int64_t)nx() * p.pitch; }

    // pass Z's array: uint16 distances, or uint8 (Dz8) under a cap below 256
    template <class B, class Z = uint16_t> int pass_y(B *b, const unsigned int *gate, const char *name) const
    {
        ProfScope ps(name, s);
        const WinGeom wg{ny(), p.pitch, ny() * p.pitch, nx(), nz(), g->bc[2], g->bc[3], ccap, p.pitch, bsx(), bsy()};
        dispatch_window(tr, wide, reinterpret_cast<const Z *>(p.dz), b, o, wg, nx(), ny(), p.bits_y, p.counters + 0,
                        p.ovf_list, gate, s);
        return check_launch("pass y");
    }
    template <class B> int fallback_y(B *b, const unsigned int *gate, const char *name) const
    {
        ProfScope ps(name, s);
        hipLaunchKernelGGL(k_fallback_y<B>, dim3(p.fallback_lanes / 256), dim3(256), 0, s, p.dz, b, nx(), ny(), nz(),
                           p.pitch, bsx(), bsy(), g->bc[2], g->bc[3], ccap, p.counters + 0, p.ovf_list, p.stk, p.cap,
                           gate);
        return check_launch("pass y fallback");
    }
    template <class B> int pass_x(const B *b, const unsigned int *gate, const char *name) const
    {
        ProfScope ps(name, s);
        const WinGeom wg{nx(), bsx(), bsy(), ny(), nz(), g->bc[0], g->bc[1], ccap, p.pitch, 0, 0};
        dispatch_window(tr, wide, b, (int32_t *)nullptr, o, wg, nx(), ny(), p.bits_x, p.counters + 1, p.ovf_list, gate,
                        s);
        return check_launch("pass x");
    }
    template <class B> int fallback_x(const B *b, const unsigned int *gate, const char *name) const
    {
        ProfScope ps(name, s);
        hipLaunchKernelGGL(k_fallback_x<B>, dim3(p.fallback_lanes / 256), dim3(256), 0, s, b, o, nx(), ny(), nz(),
                           bsx(), bsy(), g->bc[0], g->bc[1], ccap, p.counters + 1, p.ovf_list, p.stk, p.cap, gate);
        return check_launch("pass x fallback");
    }
};

// Every value is <= the cap, so a scan reaching row offset kReach + 1 has
// k^2 > cap >= best: under a cap below (kReach + 1)^2 no line is ever
// deferred (the bitmaps start each EDT clear) and the fallbacks, which would
// find empty lists, are not launched.
inline bool may_defer(int ccap) { return ccap >= (kReach + 1) * (kReach + 1); }

template <class B, class Z = uint16_t>
int run_yx(const YX &yx, B *b, const unsigned int *proceed)
{
    const bool fb = may_defer(yx.ccap);
    if (int e = yx.pass_y<B, Z>(b, proceed, "edt_pass_y")) return e;
    if (fb)
        if (int e = yx.fallback_y(b, proceed, "edt_fallback_y")) return e;
    if (int e = yx.pass_x((const B *)b, proceed, "edt_pass_x")) return e;
    if (fb)
        if (int e = yx.fallback_x((const B *)b, proceed, "edt_fallback_x")) return e;
    return PMX_OK;
}

// A cap above 65535 with no caller gate (the plain EDT).  Clamping commutes
// with the passes, so a pass Y that saturates at 65535 gives, through pass
// X, min(d2, 65535) exactly: every output below 65535 is the true d2.  Pass
// Y writes 16 bits and raises a flag when it saturates anything (squared
// distances above 65535, i.e. distances above 255 voxels: rare in a porous
// medium).  Then pass Y runs again to int32 and pass X reads that; without
// the flag pass X reads the 16-bit array.  The choice is made on the device
// (the gates), so the host never waits.  The int32 array aliases the 16-bit
// one, which is dead once the flag is raised.
int run_yx_narrow_first(const YX &yx)
{
    const unsigned int *over = yx.p.counters + 2, *none = yx.p.counters + 3;
    Sq16 *b16 = reinterpret_cast<Sq16 *>(yx.p.b);
    if (int e = yx.pass_y(b16, nullptr, "edt_pass_y")) return e;
    if (int e = yx.fallback_y(b16, nullptr, "edt_fallback_y")) return e;
    if (int e = yx.pass_y(yx.p.b, over, "edt_pass_y32")) return e;
    if (int e = yx.fallback_y(yx.p.b, over, "edt_fallback_y32")) return e;
    if (int e = yx.pass_x((const Sq16 *)b16, none, "edt_pass_x")) return e;
    if (int e = yx.fallback_x((const Sq16 *)b16, none, "edt_fallback_x")) return e;
    if (int e = yx.pass_x((const int32_t *)yx.p.b, over, "edt_pass_x32")) return e;
    return yx.fallback_x((const int32_t *)yx.p.b, over, "edt_fallback_x32");
}

}  // namespace

int run_edt(const uint8_t *img, const pmx_geom *g, bool invert, const PassXOut &o, int64_t cap,
            void *ws, size_t ws_bytes, hipStream_t s, const unsigned int *proceed, int tile_rows, int match,
            int max_lanes)
{
    if (tile_rows != 0 && tile_rows != 32 && tile_rows != 64)
        return fail(PMX_ERR_INVALID, "tile_rows %d: use 0 (default), 32 or 64", tile_rows);
    const int tr = tile_rows == 0 ? kDefaultTileRows : tile_rows;
    if (int e = validate_geom(g)) return e;
    if (!img) return fail(PMX_ERR_INVALID, "img is NULL");
    EdtPlan p;
    const size_t need = plan_edt(g, nullptr, &p, max_lanes);
    if (!ws || ws_bytes < need)
        return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", ws_bytes, need);
    plan_edt(g, ws, &p, max_lanes);
    const int ccap = (int)std::min<int64_t>(std::max<int64_t>(cap, 0), kBig);
    const bool dz8 = ccap <= 255;  // 8-bit distances between passes Z and Y, and Y and X
    const int nx = (int)g->n[0], ny = (int)g->n[1], nz = (int)g->n[2];
    const int64_t rows = (int64_t)nx * ny;
    const int64_t pitch = p.pitch;

    // short rows: several rows per wave (that kernel also clears the deferred-line bitmaps)
    const int lpr = nz <= 256 ? 16 : nz <= 512 ? 32 : 64;
    if (lpr == 64)
        hipLaunchKernelGGL(k_clear_words, dim3((unsigned)std::min<int64_t>((p.n_bits + 255) / 256 + 1, 4096)),
                           dim3(256), 0, s, p.bits_y, (int64_t)p.n_bits);
    {
    ProfScope ps("edt_pass_z", s);
    const dim3 bz(256);
    const dim3 gz((unsigned)((rows + 3) / 4));
    const int vec_in = ((uintptr_t)img % 16 == 0 && g->in_s[0] % 16 == 0 && g->in_s[1] % 16 == 0) ? 1 : 0;
    const dim3 gs((unsigned)((rows + 4 * (kWave / lpr) - 1) / (4 * (kWave / lpr))));
    const uint32_t mv = match < 0 ? 0u : (uint32_t)match;
    const bool mt = match >= 0;
#define PMX_PASS_Z_SHORT(INV, MT, L, D)                                                                         \
    hipLaunchKernelGGL((k_pass_z_short<INV, MT, L, D>), gs, bz, 0, s, img, g->in_s[0], g->in_s[1], ny, nz,     \
                       rows, g->bc[4], g->bc[5], vec_in, mv, pitch, reinterpret_cast<D *>(p.dz), p.counters,   \
                       proceed, p.bits_y, (int64_t)p.n_bits)
#define PMX_PASS_Z_LONG(INV, MT, D)                                                                             \
    hipLaunchKernelGGL((k_pass_z<INV, MT, D>), gz, bz, 0, s, img, g->in_s[0], g->in_s[1], ny, nz, rows,        \
                       g->bc[4], g->bc[5], vec_in, mv, pitch, reinterpret_cast<D *>(p.dz), p.counters, proceed)
#define PMX_PASS_Z_ALL(D)                                                                                       \
    if (lpr == 16) {                                                                                            \
        if (mt) PMX_PASS_Z_SHORT(true, true, 16, D);                                                            \
        else if (invert) PMX_PASS_Z_SHORT(true, false, 16, D);                                                  \
        else PMX_PASS_Z_SHORT(false, false, 16, D);                                                             \
    } else if (lpr == 32) {                                                                                     \
        if (mt) PMX_PASS_Z_SHORT(true, true, 32, D);                                                            \
        else if (invert) PMX_PASS_Z_SHORT(true, false, 32, D);                                                  \
        else PMX_PASS_Z_SHORT(false, false, 32, D);                                                             \
    } else {                                                                                                    \
        if (mt) PMX_PASS_Z_LONG(true, true, D);                                                                 \
        else if (invert) PMX_PASS_Z_LONG(true, false, D);                                                       \
        else PMX_PASS_Z_LONG(false, false, D);                                                                  \
    }
    if (dz8) {
        PMX_PASS_Z_ALL(uint8_t)
    } else {
        PMX_PASS_Z_ALL(uint16_t)
    }
#undef PMX_PASS_Z_ALL
#undef PMX_PASS_Z_SHORT
#undef PMX_PASS_Z_LONG
    }
    if (int e = check_launch("pass z")) return e;

    // halo: 16 rows settle nearly every voxel of a porous medium; a clamped
    // EDT whose threshold reaches past 17^2 scans up to 32 rows in LDS
    const bool wide = ccap < kBig && ccap > 17 * 17;
    const YX yx{p, g, o, ccap, tr, wide, s};
    // a cap that fits in 8 (16) bits: pass Y hands pass X 8-bit (16-bit) squared
    // distances; under 8 bits pass Z hands pass Y 8-bit distances too
    if (dz8) return run_yx<Sq8, Dz8>(yx, reinterpret_cast<Sq8 *>(p.b), proceed);
    if (ccap <= 65535) return run_yx(yx, reinterpret_cast<Sq16 *>(p.b), proceed);
    // otherwise 16 bits first, int32 only where a value saturated (decided on the device)
    if (!proceed) return run_yx_narrow_first(yx);
    return run_yx(yx, p.b, proceed);
}




int validate_geom(const pmx_geom *g)
{
    if (!g) return fail(PMX_ERR_INVALID, "geometry is NULL");
    for (int a = 0; a < 3; ++a)
        if (g->n[a] < 1 || g->n[a] > 16384)
            return fail(PMX_ERR_INVALID, "extent n[%d]=%lld outside [1, 16384]", a, (long long)g->n[a]);
    for (int f = 0; f < 6; ++f)
        if (g->bc[f] < PMX_BC_END || g->bc[f] > PMX_BC_PERIODIC)
            return fail(PMX_ERR_INVALID, "bc[%d]=%d is not a boundary type", f, g->bc[f]);
    for (int a = 0; a < 3; ++a)
        if ((g->bc[2 * a] == PMX_BC_PERIODIC) != (g->bc[2 * a + 1] == PMX_BC_PERIODIC))
            return fail(PMX_ERR_INVALID, "axis %d: periodic must be set on both faces", a);
    if (g->in_s[1] < g->n[2] || g->in_s[0] < g->in_s[1] * g->n[1])
        return fail(PMX_ERR_INVALID, "input strides do not describe a C-order view");
    if (g->out_s[1] < g->n[2] || g->out_s[0] < g->out_s[1] * g->n[1])
        return fail(PMX_ERR_INVALID, "output strides do not describe a C-order view");
    if (g->n[0] * g->n[2] > UINT32_MAX || g->n[1] * g->n[2] > UINT32_MAX)
        return fail(PMX_ERR_INVALID, "too many lines");
    return PMX_OK;
}

}  // namespace pmx

extern "C" {

int pmx_edt_workspace_size(const pmx_geom *g, size_t *bytes)
{
    pmx::clear_error();
    if (int e = pmx::validate_geom(g)) return e;
    if (!bytes) return pmx::fail(PMX_ERR_INVALID, "bytes is NULL");
    pmx::EdtPlan p;
    *bytes = pmx::plan_edt(g, nullptr, &p);
    return PMX_OK;
}

int pmx_edt(const uint8_t *img, const pmx_geom *g, int32_t *d2, float *f32sq, float *f32,
            double *f64, void *workspace, size_t workspace_bytes, hipStream_t stream)
{
    pmx::clear_error();
    if (!d2 && !f32sq && !f32 && !f64) return pmx::fail(PMX_ERR_INVALID, "no output requested");
    pmx::PassXOut o{};
    o.d2 = d2;
    o.f32sq = f32sq;
    o.f32 = f32;
    o.f64 = f64;
    o.sx = g ? g->out_s[0] : 0;
    o.sy = g ? g->out_s[1] : 0;
    return pmx::run_edt(img, g, false, o, INT64_MAX, workspace, workspace_bytes, stream);
}

int pmx_morph(const uint8_t *img, const pmx_geom *g, int invert, int64_t t, uint8_t v_le,
              uint8_t v_gt, uint32_t wall_mask, uint8_t *out, void *workspace,
              size_t workspace_bytes, hipStream_t stream)
{
    pmx::clear_error();
    if (!out) return pmx::fail(PMX_ERR_INVALID, "out is NULL");
    if (wall_mask > 63u) return pmx::fail(PMX_ERR_INVALID, "wall_mask has bits above 5");
    const int32_t tt = t < 0 ? -1 : (t > (int64_t)INT32_MAX - 1 ? INT32_MAX - 1 : (int32_t)t);
    pmx::PassXOut o{};
    o.u8 = out;
    o.sx = g ? g->out_s[0] : 0;
    o.sy = g ? g->out_s[1] : 0;
    o.t = tt;
    o.v_le = v_le;
    o.v_gt = v_gt;
    o.wall_mask = wall_mask;
    // only "d2 <= t" is asked: clamp at t + 1
    return pmx::run_edt(img, g, invert != 0, o, (int64_t)tt + 1, workspace, workspace_bytes, stream);
}

}  // extern "C"
