// ccl.hip -- 26-connected multi-label connected components.
//
// Replaces cc3d.connected_components(img, return_N=True) used at
// src/pmmoto/filters/connected_components.py:52-54 and, for periodic
// faces, the boundary-label matching of connect_subdomain_boundaries
// (:65-101, core/voxels.py:167-397): the wrap is part of the neighbourhood.
//
//   local     per 8 x 8 x 64 tile, in LDS, by 2 x 2 x 2 voxel blocks: the
//             voxels of one value in a block are linked to the block's
//             first such voxel, and these representatives unite across
//             the faces, edges and corners of neighbouring blocks;
//             parent[i] = global index of the tile-local root (background
//             marked); the tile-local roots enter a root bitmap (one bit per
//             voxel)
//   border    the backward block pairs that leave a tile or wrap on
//             periodic axes, unioned globally; lock-free union by atomicCAS
//             hooking the larger root under the smaller, so every
//             component's root is its smallest C-order index
//   settle    the tile-local roots that were hooked leave the bitmap and
//             point straight at their root; roots counted per 4096-voxel
//             rank tile
//   rank      tile counts scanned; from the bitmap, each root's rank is
//             written over its parent entry as -(rank+1)
//   label     label[i] = rank of root(i) + 1 (a walk up the parents):
//             components numbered 1..N in C-order of their first voxel (the
//             canonical numbering)
#include "common.hpp"

namespace pmx {
namespace {

constexpr int kBg = INT32_MIN;
constexpr int kTile = 4096;  // voxels per rank tile: 16 chunks of 256, 64 root words

struct Box {
    int n0, n1, n2;
    int64_t is0, is1;  // input strides
    int per0, per1, per2;
};

__device__ __forceinline__ int ld(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ __forceinline__ int find_root(int *parent, int x)
{
    int p = ld(parent + x);
    while (p != x) {
        const int gp = ld(parent + p);
        if (gp != p) parent[x] = gp;  // path halving; stays within the tree
        x = p;
        p = gp;
    }
    return x;
}

__device__ __forceinline__ void unite(int *parent, int a, int b)
{
    while (true) {
        a = find_root(parent, a);
        b = find_root(parent, b);
        if (a == b) return;
        const int hi = max(a, b), lo = min(a, b);
        const int old = atomicCAS(parent + hi, hi, lo);
        if (old == hi) return;
        a = old;
        b = lo;
    }
}

// Tiled merging.  A block owns a tile of kTX x kTY x kTZ voxels (z
// fastest).  Phase 1 (k_local) unions, in LDS, every voxel with its
// backward 26-neighbours inside the tile (no wrap) and writes each voxel's
// global parent as the global index of its tile-local root: local C-order
// equals global C-order within a tile, so every parent index is <= its
// voxel's, as the global union-find requires.  Phase 2 (k_border) unions,
// globally, only the neighbour pairs phase 1 did not see: those that leave
// the tile or wrap.  Together they cover every backward pair once.
constexpr int kTX = 8, kTY = 8, kTZ = 64, kTVox = kTX * kTY * kTZ;  // 16 x 16 x 64 and 8 x 16 x 64 measured slower
constexpr int kBorderHash = 1024;  // entries of a border block's set of joined parent pairs

struct TileGeom {
    int tx, ty, tz;     // tiles per axis
};

__device__ __forceinline__ void tile_of(const Box &b, const TileGeom &tg, int64_t t, int &x0, int &y0, int &z0,
                                        int &ex, int &ey, int &ez)
{
    const int tzi = (int)(t % tg.tz);
    const int64_t r = t / tg.tz;
    const int tyi = (int)(r % tg.ty), txi = (int)(r / tg.ty);
    x0 = txi * kTX;
    y0 = tyi * kTY;
    z0 = tzi * kTZ;
    ex = min(kTX, b.n0 - x0);
    ey = min(kTY, b.n1 - y0);
    ez = min(kTZ, b.n2 - z0);
}

// backward neighbour k of 13: dx=-1 (9), dx=0,dy=-1 (3), dx=0,dy=0,dz=-1 (1)
__device__ __forceinline__ void back_nb(int k, int &dx, int &dy, int &dz)
{
    if (k < 9) {
        dx = -1;
        dy = k / 3 - 1;
        dz = k % 3 - 1;
    } else if (k < 12) {
        dx = 0;
        dy = -1;
        dz = k - 10;
    } else {
        dx = 0;
        dy = 0;
        dz = -1;
    }
}

// path halving: only non-roots are rewritten (to an ancestor), and a CAS
// only succeeds on a root, so the two never conflict
__device__ __forceinline__ int lfind(int *lp, int x)
{
    int p = lp[x];
    while (p != x) {
        const int gp = lp[p];
        if (gp != p) lp[x] = gp;
        x = p;
        p = gp;
    }
    return x;
}

__device__ __forceinline__ void lunite(int *lp, int a, int b)
{
    while (true) {
        a = lfind(lp, a);
        b = lfind(lp, b);
        if (a == b) return;
        const int hi = max(a, b), lo = min(a, b);
        const int old = atomicCAS(lp + hi, hi, lo);
        if (old == hi) return;
        a = old;
        b = lo;
    }
}

// 2 x 2 x 2 voxel blocks.  The voxels of a block are pairwise 26-adjacent,
// so its voxels of one value belong to one component: each is linked to the
// block's first voxel of that value (in C-order, so the smallest index: every
// component's smallest voxel stays its root), and the union-find only joins
// these representatives across the faces, edges and corners of neighbouring
// blocks.  Tile origins are multiples of 2 on every axis, so blocks never
// straddle a tile; a block at the high end of an odd extent is partial (its
// missing voxels read 0).  Bit k = 4 dx + 2 dy + dz of a block mask is its
// voxel (dx, dy, dz); byte k of a block word is that voxel's value.
constexpr int kBX = kTX / 2, kBY = kTY / 2, kBZ = kTZ / 2, kTBlk = kBX * kBY * kBZ;

// the block's voxels whose coordinate on `axis` is c
__device__ __forceinline__ uint32_t layer_bits(int axis, int c)
{
    constexpr uint32_t kLayer[3][2] = {{0x0fu, 0xf0u}, {0x33u, 0xccu}, {0x55u, 0xaau}};
    return kLayer[axis][c];
}

// The voxels of block A and of its neighbour B = A + d (d in {-1, 0, 1}^3)
// that face each other: per axis, d = -1 takes A's layer 0 and B's layer 1,
// d = +1 the reverse, d = 0 both layers.  Every voxel of A's set is
// 26-adjacent to every voxel of B's set.
__device__ __forceinline__ void facing(int dx, int dy, int dz, uint32_t &fa, uint32_t &fb)
{
    fa = fb = 0xffu;
    const int d[3] = {dx, dy, dz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0) continue;
        fa &= layer_bits(a, d[a] < 0 ? 0 : 1);
        fb &= layer_bits(a, d[a] < 0 ? 1 : 0);
    }
}

// bit k set where byte k of w equals v (v < 256; exact, no carry crosses a byte)
__device__ __forceinline__ uint32_t eq_mask8(uint64_t w, uint32_t v)
{
    constexpr uint64_t k7f = 0x7f7f7f7f7f7f7f7full;
    const uint64_t x = w ^ (0x0101010101010101ull * v);
    const uint64_t t = ~(((x & k7f) + k7f) | x | k7f);  // 0x80 in the bytes of x that are 0
    return (uint32_t)(((t >> 7) * 0x0102040810204080ull) >> 56);
}

__device__ __forceinline__ uint32_t byte_of(uint64_t w, int k) { return (uint32_t)(w >> (8 * k)) & 0xffu; }

// tile-local index of voxel k of the block whose first voxel is l0
__device__ __forceinline__ int blk_voxel(int l0, int k)
{
    return l0 + ((k >> 2) & 1) * (kTY * kTZ) + ((k >> 1) & 1) * kTZ + (k & 1);
}

__global__ __launch_bounds__(256) void k_local(const uint8_t *__restrict__ img, Box b, TileGeom tg,
                                               int vec16, int *__restrict__ parent, uint64_t *__restrict__ root_bits)
{
    static_assert(kTVox == 256 * 16, "one 16-byte quarter row per thread");
    static_assert(kTBlk == 2 * 256, "two blocks per thread");
    __shared__ int lp[kTVox];
    __shared__ __align__(16) uint8_t lv[kTVox];
    __shared__ uint64_t bv[kTBlk];  // each block's eight values
    int x0, y0, z0, ex, ey, ez;
    tile_of(b, tg, blockIdx.x, x0, y0, z0, ex, ey, ez);
    {
        // thread t stages bytes 16 (t % 4) .. +15 of tile row t / 4 (64 rows of 64 bytes)
        const int r = threadIdx.x >> 2, c0 = (threadIdx.x & 3) * 16;
        const int lx = r / kTY, ly = r % kTY;
        const bool row_in = lx < ex && ly < ey;
        const uint8_t *src = img + (int64_t)(x0 + lx) * b.is0 + (int64_t)(y0 + ly) * b.is1 + z0 + c0;
        uint4 q = make_uint4(0u, 0u, 0u, 0u);
        if (row_in && vec16 && c0 + 16 <= ez) {
            q = *reinterpret_cast<const uint4 *>(src);  // rows 16-byte aligned (checked on the host)
        } else if (row_in) {
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            for (int k = 0; k < 16; ++k)
                if (c0 + k < ez) w[k >> 2] |= (uint32_t)src[k] << (8 * (k & 3));
            q = make_uint4(w[0], w[1], w[2], w[3]);
        }
        *reinterpret_cast<uint4 *>(lv + r * kTZ + c0) = q;
    }
    __syncthreads();
    // blocks: values, and every voxel linked to its block's first voxel of the same value
    const int ebx = (ex + 1) >> 1, eby = (ey + 1) >> 1, ebz = (ez + 1) >> 1;  // blocks of the tile in the box
    uint64_t mv[2];
    int ml0[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int bi = threadIdx.x + 256 * s;
        const int bx = bi / (kBY * kBZ), by = (bi / kBZ) % kBY, bz = bi % kBZ;
        const int l0 = ((2 * bx) * kTY + 2 * by) * kTZ + 2 * bz;
        const uint16_t *h = reinterpret_cast<const uint16_t *>(lv);
        const uint64_t w = (uint64_t)h[l0 >> 1] | ((uint64_t)h[(l0 + kTZ) >> 1] << 16) |
                           ((uint64_t)h[(l0 + kTY * kTZ) >> 1] << 32) |
                           ((uint64_t)h[(l0 + kTY * kTZ + kTZ) >> 1] << 48);
        bv[bi] = w;
        mv[s] = w;
        ml0[s] = l0;
        const uint32_t fg = ~eq_mask8(w, 0u) & 0xffu;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int l = blk_voxel(l0, k);
            int r = l;
            if ((fg >> k) & 1u) r = blk_voxel(l0, __builtin_ctz(eq_mask8(w, byte_of(w, k))));
            lp[l] = r;
        }
    }
    __syncthreads();
    // each block joins its 13 backward neighbour blocks inside the tile: one
    // union per value the facing voxels share (others cross the tile: phase 2)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int bi = threadIdx.x + 256 * s;
        const int bx = bi / (kBY * kBZ), by = (bi / kBZ) % kBY, bz = bi % kBZ;
        const uint64_t w = mv[s];
        const uint32_t fg = ~eq_mask8(w, 0u) & 0xffu;
        if (!fg || bx >= ebx || by >= eby || bz >= ebz) continue;
        const int l0 = ml0[s];
        const uint32_t v0 = byte_of(w, __builtin_ctz(fg));
        const bool uni = eq_mask8(w, v0) == fg;  // one value (every binary image's blocks)
#pragma unroll
        for (int j = 0; j < 13; ++j) {
            int dx, dy, dz;
            back_nb(j, dx, dy, dz);
            const int nbx = bx + dx, nby = by + dy, nbz = bz + dz;
            if (nbx < 0 || nby < 0 || nby >= eby || nbz < 0 || nbz >= ebz) continue;
            uint32_t fa, fb;
            facing(dx, dy, dz, fa, fb);
            const uint32_t sa = fg & fa;
            if (!sa) continue;
            const int bj = (nbx * kBY + nby) * kBZ + nbz;
            const uint64_t wb = bv[bj];
            const int lb0 = ((2 * nbx) * kTY + 2 * nby) * kTZ + 2 * nbz;
            if (uni) {
                const uint32_t mb = eq_mask8(wb, v0);
                if (mb & fb) lunite(lp, blk_voxel(l0, __builtin_ctz(fg)), blk_voxel(lb0, __builtin_ctz(mb)));
                continue;
            }
            uint32_t todo = sa;
            while (todo) {
                const int k = __builtin_ctz(todo);
                const uint32_t v = byte_of(w, k);
                const uint32_t ma = eq_mask8(w, v);
                todo &= ~ma;  // one union per shared value
                const uint32_t mb = eq_mask8(wb, v);
                if (mb & fb) lunite(lp, blk_voxel(l0, __builtin_ctz(ma)), blk_voxel(lb0, __builtin_ctz(mb)));
            }
        }
    }
    __syncthreads();
    // each block's voxels get the global index of their group's tile-local
    // root (one root lookup per value); a group whose representative is that
    // root enters the root bitmap
    auto tile_to_global = [&](int r) {
        const int rx = r / (kTY * kTZ), ry = (r / kTZ) % kTY, rz = r % kTZ;
        return (int)(((int64_t)(x0 + rx) * b.n1 + (y0 + ry)) * b.n2 + z0 + rz);
    };
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int bi = threadIdx.x + 256 * s;
        const int bx = bi / (kBY * kBZ), by = (bi / kBZ) % kBY, bz = bi % kBZ;
        if (2 * bx >= ex || 2 * by >= ey || 2 * bz >= ez) continue;
        const uint64_t w = mv[s];
        const int l0 = ml0[s];
        uint32_t todo = ~eq_mask8(w, 0u) & 0xffu;
        int out[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) out[k] = kBg;
        while (todo) {
            const uint32_t ma = eq_mask8(w, byte_of(w, __builtin_ctz(todo)));
            todo &= ~ma;
            const int rep = blk_voxel(l0, __builtin_ctz(ma));
            const int r = lfind(lp, rep);
            const int gr = tile_to_global(r);
            if (r == rep) atomicOr(reinterpret_cast<unsigned long long *>(root_bits + (gr >> 6)), 1ull << (gr & 63));
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if ((ma >> k) & 1u) out[k] = gr;
        }
        // rows (dx, dy) of the block, two voxels each
#pragma unroll
        for (int rw = 0; rw < 4; ++rw) {
            const int lx = 2 * bx + (rw >> 1), ly = 2 * by + (rw & 1);
            if (lx >= ex || ly >= ey) continue;
            const int64_t i = ((int64_t)(x0 + lx) * b.n1 + (y0 + ly)) * b.n2 + z0 + 2 * bz;
            parent[i] = out[2 * rw];
            if (2 * bz + 1 < ez) parent[i + 1] = out[2 * rw + 1];
        }
    }
}

// A pair of parents to join (pj < 0: none), at most once per block: a pj
// this lane has just joined adds nothing, and the rest go through the
// block's set of joined parent pairs (a pair already there was joined by the
// lane that inserted it; a full probe falls back to joining).
__device__ __forceinline__ void join_pair(int *parent, unsigned long long *seen, int pi, int pj, int &last)
{
    constexpr int kProbe = 8;
    constexpr unsigned long long kEmpty = ~0ull;
    if (pj < 0 || pj == pi || pj == last) return;
    last = pj;
    const unsigned long long key = ((unsigned long long)(uint32_t)pi << 32) | (uint32_t)pj;
    const uint32_t h = ((uint32_t)pi * 0x9e3779b1u) ^ ((uint32_t)pj * 0x85ebca6bu);
    for (int q = 0; q < kProbe; ++q) {
        const unsigned long long old = atomicCAS(&seen[(h + q) & (kBorderHash - 1)], kEmpty, key);
        if (old == kEmpty) break;  // inserted: this lane joins the pair
        if (old == key) return;    // inserted before: already joined
    }
    unite(parent, pi, pj);
}

// join_pair called by every lane of the wave at once: a lane whose pair
// equals the previous lane's adds nothing (that lane, or one before it,
// joins it), which keeps lanes with one pair off the same hash slot.
__device__ __forceinline__ void join_pair_wave(int *parent, unsigned long long *seen, int pi, int pj, int &last,
                                               int lane)
{
    const int ppi = __shfl_up(pi, 1), ppj = __shfl_up(pj, 1);
    if (lane > 0 && ppj >= 0 && ppi == pi && ppj == pj) return;
    join_pair(parent, seen, pi, pj, last);
}

// coordinate c (possibly outside [0, n)) on an axis of n voxels: wrapped on a
// periodic axis, else -1 outside the box
__device__ __forceinline__ int wrap_coord(int c, int n, int per)
{
    if (c >= 0 && c < n) return c;
    if (!per) return -1;
    c %= n;
    return c < 0 ? c + n : c;
}

// Phase 2: the backward pairs k_local did not see -- those whose neighbour
// leaves the tile or wraps -- by blocks, like k_local.  The tile and a
// one-voxel shell around it (tile coordinates -1 .. kT on each axis; wrapped
// on periodic axes, 0 outside the box) are staged in LDS.  For a block A
// and a backward direction d, the neighbour voxels V are, per axis, A's
// range where d = 0 and the layer just below (d = -1) or above (d = +1) A's
// range; every voxel of V is adjacent to every voxel of A's facing set, and
// V lies in one 2 x 2 x 2 block of one tile (its same-valued voxels share a
// tile-local root).  Where V leaves the tile, each value A's facing set
// shares with V joins A's parent with the parent of V's voxel of that value.
// Parent reads go through L2: a stale parent is still an ancestor, and a
// union acting on a stale root fails its CAS and retries.
// The shell box in LDS: rows (x, y) for tile coordinates -1 .. kT, each row
// kHP bytes with z = -1 .. kTZ at byte z + 16 (z = 0 .. kTZ-1 16-byte aligned).
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHP = kTZ + 32;

__device__ __forceinline__ int hoff(int x, int y, int z) { return ((x + 1) * kHY + (y + 1)) * kHP + z + 16; }

struct BorderCtx {
    const uint8_t *hv;               // the shell box
    int *parent;
    unsigned long long *seen;        // the block's joined parent pairs
    int e[3], o[3], n[3], per[3];    // tile extents in the box, tile origin, box extents, periodic axes
    int lane;
};

// a tile block A: its range per axis (lo .. hi: one voxel on the high side
// of an odd extent), values, voxel mask, first value and its voxels, and the
// parent of that value's voxels
struct BorderBlock {
    int lo[3], hi[3];
    uint64_t w;
    uint32_t fg, v0, m0;
    int pi0;
};

// global index of the voxel at tile coordinates (x, y, z), inside the box or
// wrapped into it
__device__ __forceinline__ int gindex(const BorderCtx &c, int x, int y, int z)
{
    const int gx = wrap_coord(c.o[0] + x, c.n[0], c.per[0]), gy = wrap_coord(c.o[1] + y, c.n[1], c.per[1]),
              gz = wrap_coord(c.o[2] + z, c.n[2], c.per[2]);
    return (gx * c.n[1] + gy) * c.n[2] + gz;
}

// a voxel of value v in V = vlo .. vhi (the axes with d = 0 span A's range,
// the others one layer): its global index, or -1
template <int DX, int DY, int DZ>
__device__ __forceinline__ int find_in(const BorderCtx &c, const int (&vlo)[3], const int (&vhi)[3], uint32_t v)
{
    // (shell coordinates start at -1, so a flag, not a coordinate, says "found")
    bool found = false;
    int fx = 0, fy = 0, fz = 0;
#pragma unroll
    for (int qx = (DX == 0 ? 1 : 0); qx >= 0; --qx)
#pragma unroll
        for (int qy = (DY == 0 ? 1 : 0); qy >= 0; --qy)
#pragma unroll
            for (int qz = (DZ == 0 ? 1 : 0); qz >= 0; --qz) {
                const int x = vlo[0] + qx, y = vlo[1] + qy, z = vlo[2] + qz;
                if (x <= vhi[0] && y <= vhi[1] && z <= vhi[2] && c.hv[hoff(x, y, z)] == v) {
                    found = true;
                    fx = x;
                    fy = y;
                    fz = z;
                }
            }
    return found ? gindex(c, fx, fy, fz) : -1;
}

// Backward direction (DX, DY, DZ) of block A: V and A's facing voxels as
// above; when V leaves the tile, one union per value they share.  Every
// lane of the wave calls this; a wave none of whose lanes has such a V
// returns at once.
template <int DX, int DY, int DZ>
__device__ __forceinline__ void border_dir(const BorderCtx &c, const BorderBlock &a, int &last)
{
    constexpr int d[3] = {DX, DY, DZ};
    int vlo[3], vhi[3];
    bool out = false;
    uint32_t fa = 0xffu;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (d[q] == 0) {
            vlo[q] = a.lo[q];
            vhi[q] = a.hi[q];
        } else {
            vlo[q] = vhi[q] = d[q] < 0 ? a.lo[q] - 1 : a.hi[q] + 1;
            out |= d[q] < 0 ? vlo[q] < 0 : vlo[q] >= c.e[q];
            fa &= layer_bits(q, d[q] < 0 ? 0 : a.hi[q] - a.lo[q]);
        }
    }
    // pairs inside the tile were joined by k_local
    const bool first = out && (a.m0 & fa) != 0;
    const uint32_t rest = out ? a.fg & fa & ~a.m0 : 0u;  // the other values of blocks holding several
    if (__builtin_amdgcn_ballot_w64(first || rest != 0) == 0) return;
    int pj = -1;
    if (first) {
        const int j = find_in<DX, DY, DZ>(c, vlo, vhi, a.v0);
        if (j >= 0) pj = c.parent[j];
    }
    join_pair_wave(c.parent, c.seen, a.pi0, pj, last, c.lane);
    uint32_t todo = rest;
    while (todo) {
        const int k = __builtin_ctz(todo);
        const uint32_t v = byte_of(a.w, k);
        const uint32_t ma = eq_mask8(a.w, v);
        todo &= ~ma;  // one union per shared value
        const int j = find_in<DX, DY, DZ>(c, vlo, vhi, v);
        if (j < 0) continue;
        const int ka = __builtin_ctz(ma);
        const int pa = c.parent[gindex(c, a.lo[0] + ((ka >> 2) & 1), a.lo[1] + ((ka >> 1) & 1), a.lo[2] + (ka & 1))];
        join_pair(c.parent, c.seen, pa, c.parent[j], last);
    }
}

__global__ __launch_bounds__(256) void k_border(const uint8_t *__restrict__ img, Box b, TileGeom tg, int vec16,
                                                int *__restrict__ parent)
{
    constexpr unsigned long long kEmpty = ~0ull;
    __shared__ unsigned long long seen[kBorderHash];
    __shared__ __align__(16) uint8_t hv[kHX * kHY * kHP];
    for (int h = threadIdx.x; h < kBorderHash; h += blockDim.x) seen[h] = kEmpty;
    int x0, y0, z0, ex, ey, ez;
    tile_of(b, tg, blockIdx.x, x0, y0, z0, ex, ey, ez);
    {
        // four threads per row: 16 interior bytes each (one aligned load on a full z tile with
        // 16-byte aligned rows), the first and the last also the byte just outside the tile
        const int q = threadIdx.x & 3;
        const int zlo = wrap_coord(z0 - 1, b.n2, b.per2), zhi = wrap_coord(z0 + kTZ, b.n2, b.per2);
        for (int r = threadIdx.x >> 2; r < kHX * kHY; r += 64) {
            const int hx = r / kHY, hy = r % kHY;
            const int gx = wrap_coord(x0 + hx - 1, b.n0, b.per0), gy = wrap_coord(y0 + hy - 1, b.n1, b.per1);
            uint8_t *dst = hv + r * kHP + 16;
            if (gx < 0 || gy < 0) {
                *reinterpret_cast<uint4 *>(dst + 16 * q) = make_uint4(0u, 0u, 0u, 0u);
                if (q == 0) dst[-1] = 0;
                if (q == 3) dst[kTZ] = 0;
                continue;
            }
            const uint8_t *src = img + (int64_t)gx * b.is0 + (int64_t)gy * b.is1;
            if (vec16 && ez == kTZ) {
                *reinterpret_cast<uint4 *>(dst + 16 * q) = *reinterpret_cast<const uint4 *>(src + z0 + 16 * q);
            } else {
                // a partial z tile: byte by byte, the positions past its end wrapped like the shell
                for (int k = 16 * q; k < 16 * q + 16; ++k) {
                    const int gz = wrap_coord(z0 + k, b.n2, b.per2);
                    dst[k] = gz >= 0 ? src[gz] : (uint8_t)0;
                }
            }
            if (q == 0) dst[-1] = zlo >= 0 ? src[zlo] : (uint8_t)0;
            if (q == 3) dst[kTZ] = zhi >= 0 ? src[zhi] : (uint8_t)0;
        }
    }
    __syncthreads();
    const BorderCtx c{hv, parent, seen, {ex, ey, ez}, {x0, y0, z0}, {b.n0, b.n1, b.n2}, {b.per0, b.per1, b.per2},
                      (int)(threadIdx.x & 63)};
    // every lane runs every step below (join_pair_wave exchanges pairs between lanes)
#pragma unroll 1
    for (int s = 0; s < 2; ++s) {
        const int bi = threadIdx.x + 256 * s;
        BorderBlock a;
        a.lo[0] = 2 * (bi / (kBY * kBZ));
        a.lo[1] = 2 * ((bi / kBZ) % kBY);
        a.lo[2] = 2 * (bi % kBZ);
#pragma unroll
        for (int q = 0; q < 3; ++q) a.hi[q] = min(a.lo[q] + 1, c.e[q] - 1);
        const bool in_box = a.lo[0] < ex && a.lo[1] < ey && a.lo[2] < ez;
        // only blocks with a backward neighbour layer outside the tile
        const bool border =
            in_box && (a.lo[0] == 0 || a.lo[1] == 0 || a.hi[1] + 1 >= ey || a.lo[2] == 0 || a.hi[2] + 1 >= ez);
        a.w = 0;
        if (border) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int cx = a.lo[0] + ((k >> 2) & 1), cy = a.lo[1] + ((k >> 1) & 1), cz = a.lo[2] + (k & 1);
                if (cx <= a.hi[0] && cy <= a.hi[1] && cz <= a.hi[2]) a.w |= (uint64_t)hv[hoff(cx, cy, cz)] << (8 * k);
            }
        }
        a.fg = ~eq_mask8(a.w, 0u) & 0xffu;  // 0 unless a border block
        // A's first value and its parent: the only value of every binary image's blocks
        const int k0 = a.fg ? __builtin_ctz(a.fg) : 0;
        a.v0 = byte_of(a.w, k0);
        a.m0 = eq_mask8(a.w, a.v0) & a.fg;
        a.pi0 = -1;
        if (a.fg)
            a.pi0 = parent[gindex(c, a.lo[0] + ((k0 >> 2) & 1), a.lo[1] + ((k0 >> 1) & 1), a.lo[2] + (k0 & 1))];
        int last = -1;
        border_dir<-1, -1, -1>(c, a, last);
        border_dir<-1, -1, 0>(c, a, last);
        border_dir<-1, -1, 1>(c, a, last);
        border_dir<-1, 0, -1>(c, a, last);
        border_dir<-1, 0, 0>(c, a, last);
        border_dir<-1, 0, 1>(c, a, last);
        border_dir<-1, 1, -1>(c, a, last);
        border_dir<-1, 1, 0>(c, a, last);
        border_dir<-1, 1, 1>(c, a, last);
        border_dir<0, -1, -1>(c, a, last);
        border_dir<0, -1, 0>(c, a, last);
        border_dir<0, -1, 1>(c, a, last);
        border_dir<0, 0, -1>(c, a, last);
    }
}

// the root bitmap starts empty (k_local enters its tile-local roots)
__global__ __launch_bounds__(256) void k_clear_roots(uint64_t *__restrict__ bits, int64_t n_words)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = tid; i < n_words; i += stride) bits[i] = 0ull;
}

// One wave per rank tile of kTile voxels (one bitmap word per lane).  The
// bitmap holds the tile-local roots; those k_border hooked under another
// root leave it, and their entry is pointed straight at their root (a short
// read-only walk: no union runs any more, and an entry only ever changes to
// an ancestor), so every voxel is at most two entries from its root.  The
// tile's count of remaining roots goes to tile_counts.
__global__ __launch_bounds__(256) void k_settle_roots(int *__restrict__ parent, int64_t ntiles,
                                                      uint64_t *__restrict__ root_bits, int *__restrict__ tile_counts)
{
    static_assert(kTile == 64 * 64, "one root word per lane");
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;
    const int64_t wi = t * 64 + lane;
    uint64_t m = root_bits[wi], keep = m;
    while (m) {
        const int bit = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const int i = (int)(wi * 64 + bit);
        int r = parent[i];
        if (r == i) continue;
        for (int q = parent[r]; q != r; q = parent[r]) r = q;
        parent[i] = r;
        keep &= ~(1ull << bit);
    }
    if (keep != root_bits[wi]) root_bits[wi] = keep;
    int c = __popcll(keep);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) tile_counts[t] = c;
}

// single block: exclusive scan of tile counts in place, total to *count.
// Rounds of 8192 counts: a coalesced load into LDS, thread t scans its 8
// consecutive counts, a wave-shuffle scan plus 16 wave totals give the
// thread offsets, and a coalesced store writes the round back.
__global__ __launch_bounds__(1024) void k_scan_tiles(int *__restrict__ tile_counts, int64_t ntiles,
                                                     int64_t *__restrict__ count)
{
    constexpr int kPer = 8, kRound = 1024 * kPer;
    __shared__ int buf[kRound];
    __shared__ int64_t wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int64_t carry = 0;
    for (int64_t r0 = 0; r0 < ntiles; r0 += kRound) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int64_t i = r0 + k * 1024 + threadIdx.x;
            buf[k * 1024 + threadIdx.x] = i < ntiles ? tile_counts[i] : 0;
        }
        __syncthreads();
        int v[kPer];
        int s = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            v[k] = buf[kPer * threadIdx.x + k];
            s += v[k];
        }
        int64_t inc = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t u = __shfl_up(inc, o);
            if (lane >= o) inc += u;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        int64_t woff = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            if (k < w) woff += wsum[k];
            tot += wsum[k];
        }
        int64_t run = carry + woff + inc - s;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            buf[kPer * threadIdx.x + k] = (int)run;
            run += v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int64_t i = r0 + k * 1024 + threadIdx.x;
            if (i < ntiles) tile_counts[i] = buf[k * 1024 + threadIdx.x];
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = carry;
}

// One wave per rank tile (its kTile / 64 root words, one per lane): each
// root's entry becomes -(rank + 1), rank = its position among all roots in
// C-order.
__global__ __launch_bounds__(256) void k_rank_roots(int *__restrict__ parent, int64_t ntiles,
                                                    const uint64_t *__restrict__ root_bits,
                                                    const int *__restrict__ tile_offsets)
{
    static_assert(kTile == 64 * 64, "one root word per lane");
    const int lane = threadIdx.x & 63;
    const int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= ntiles) return;
    uint64_t m = root_bits[t * 64 + lane];
    const int c = __popcll(m);
    int inc = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    int r = tile_offsets[t] + inc - c;
    const int64_t i0 = t * kTile + 64 * lane;
    while (m) {
        const int b = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        parent[i0 + b] = -(++r);
    }
}

// label of a voxel whose parent entry is p: roots hold -(rank + 1) after
// k_rank_roots, other entries point at an ancestor (a smaller index), so the
// walk ends at the voxel's root
__device__ __forceinline__ int64_t label_of(const int *parent, int p)
{
    if (p == kBg) return 0;
    while (p >= 0) p = parent[p];
    return -(int64_t)p;
}

template <class T>
__global__ __launch_bounds__(256) void k_labels(const int *__restrict__ parent, Box b, int64_t os0,
                                                int64_t os1, T *__restrict__ out)
{
    const int64_t n = (int64_t)b.n0 * b.n1 * b.n2;
    const bool dense = os1 == b.n2 && os0 == (int64_t)b.n1 * b.n2;  // C-contiguous labels: no index split
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, stride = (int64_t)gridDim.x * blockDim.x;
    if constexpr (std::is_same<T, int32_t>::value) {
        if (dense && (uintptr_t)out % 16 == 0) {
            // four voxels per lane: one 16-byte parent load and one 16-byte label store
            // (the parent array is a 256-byte aligned workspace slice)
            const int4 *p4 = reinterpret_cast<const int4 *>(parent);
            int4 *o4 = reinterpret_cast<int4 *>(out);
            for (int64_t q = tid; q < n / 4; q += stride) {
                const int4 p = p4[q];
                o4[q] = make_int4((int)label_of(parent, p.x), (int)label_of(parent, p.y), (int)label_of(parent, p.z),
                                  (int)label_of(parent, p.w));
            }
            for (int64_t i = n / 4 * 4 + tid; i < n; i += stride) out[i] = (T)label_of(parent, parent[i]);
            return;
        }
    }
    for (int64_t i = tid; i < n; i += stride) {
        const T lab = (T)label_of(parent, parent[i]);
        if (dense) {
            out[i] = lab;
        } else {
            const int64_t z = i % b.n2, t = i / b.n2, y = t % b.n1, x = t / b.n1;
            out[x * os0 + y * os1 + z] = lab;
        }
    }
}

// flags[label] = 1 for every label inside the sub-box [lo, hi).  Labels
// outside [0, n_flags) cannot come from pmx_ccl26 with a matching count: they
// are not written (nothing lands past the flag array) but counted in *bad,
// which the caller turns into an error.
template <class T>
__global__ __launch_bounds__(256) void k_mark(const T *__restrict__ labels, int64_t s0, int64_t s1,
                                              int64_t lo0, int64_t lo1, int64_t lo2, int64_t e0,
                                              int64_t e1, int64_t e2, uint8_t *__restrict__ flags,
                                              uint64_t n_flags, unsigned int *__restrict__ bad)
{
    const int64_t n = e0 * e1 * e2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = i % e2, t = i / e2, y = t % e1, x = t / e1;
        const uint64_t l = (uint64_t)labels[(lo0 + x) * s0 + (lo1 + y) * s1 + lo2 + z];
        if (l < n_flags) {
            if (l) flags[l] = 1;
        } else if (bad) {
            atomicAdd(bad, 1u);
        }
    }
}

// out = keep(label) ? img : 0 over the whole array; out-of-range labels are
// dropped and counted in *bad, as in k_mark
template <class T>
__global__ __launch_bounds__(256) void k_keep(const uint8_t *__restrict__ img, const T *__restrict__ labels,
                                              int64_t n, const uint8_t *__restrict__ fa,
                                              const uint8_t *__restrict__ fb, uint64_t n_flags, int mode,
                                              int phase, uint8_t *__restrict__ out, unsigned int *__restrict__ bad)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t l = (uint64_t)labels[i];
        const uint8_t v = img[i];
        bool keep = false;
        if (l && l < n_flags) {
            const bool a = fa[l] != 0, bb = fb ? fb[l] != 0 : false;
            keep = mode == 0 ? a : (mode == 1 ? (a && bb) : (!a && !bb));
            if (phase >= 0 && v != phase) keep = false;
        } else if (l && bad) {
            atomicAdd(bad, 1u);
        }
        out[i] = keep ? v : 0;
    }
}

struct CclPlan {
    int *parent = nullptr, *tiles = nullptr;
    uint64_t *root_bits = nullptr;  // one bit per voxel, kTile / 64 words per rank tile
    int64_t ntiles = 0;
};

size_t plan_ccl(const pmx_geom *g, void *ws, CclPlan *p)
{
    Carver c(ws, SIZE_MAX);
    const int64_t n = g->n[0] * g->n[1] * g->n[2];
    p->parent = c.take<int>((size_t)n);
    p->ntiles = (n + kTile - 1) / kTile;
    p->tiles = c.take<int>((size_t)p->ntiles);
    p->root_bits = c.take<uint64_t>((size_t)p->ntiles * (kTile / 64));
    return c.need();
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65536); }

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_ccl_workspace_size(const pmx_geom *g, size_t *bytes)
{
    pmx::clear_error();
    if (int e = pmx::validate_geom(g)) return e;
    if (!bytes) return pmx::fail(PMX_ERR_INVALID, "bytes is NULL");
    pmx::CclPlan p;
    *bytes = pmx::plan_ccl(g, nullptr, &p);
    return PMX_OK;
}

int pmx_ccl26(const uint8_t *img, const pmx_geom *g, int32_t *labels32, uint64_t *labels64,
              int64_t *count, void *workspace, size_t workspace_bytes, hipStream_t s)
{
    using namespace pmx;
    clear_error();
    if (int e = validate_geom(g)) return e;
    if (!img || !count || (!labels32 == !labels64))
        return fail(PMX_ERR_INVALID, "need img, count and exactly one of labels32/labels64");
    const int64_t n = g->n[0] * g->n[1] * g->n[2];
    if (n >= INT32_MAX) return fail(PMX_ERR_INVALID, "box has %lld voxels (limit 2^31-1)", (long long)n);
    for (int a = 0; a < 3; ++a)
        if (g->bc[2 * a] == PMX_BC_WALL || g->bc[2 * a + 1] == PMX_BC_WALL)
            return fail(PMX_ERR_INVALID, "labelling treats faces as END or PERIODIC only");
    CclPlan p;
    const size_t need = plan_ccl(g, nullptr, &p);
    if (!workspace || workspace_bytes < need)
        return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    plan_ccl(g, workspace, &p);
    Box b{(int)g->n[0], (int)g->n[1], (int)g->n[2], g->in_s[0], g->in_s[1],
          g->bc[0] == PMX_BC_PERIODIC, g->bc[2] == PMX_BC_PERIODIC, g->bc[4] == PMX_BC_PERIODIC};
    const unsigned gr = grid_for(n);
    const TileGeom tg{(b.n0 + kTX - 1) / kTX, (b.n1 + kTY - 1) / kTY, (b.n2 + kTZ - 1) / kTZ};
    const int64_t ntile = (int64_t)tg.tx * tg.ty * tg.tz;
    if (ntile > UINT32_MAX / 2) return fail(PMX_ERR_INVALID, "too many tiles");
    const int vec16 = ((uintptr_t)img % 16 == 0 && g->in_s[0] % 16 == 0 && g->in_s[1] % 16 == 0) ? 1 : 0;
    {
        ProfScope ps("ccl_init", s);
        hipLaunchKernelGGL(k_clear_roots, dim3((unsigned)std::min<int64_t>((p.ntiles * (kTile / 64) + 255) / 256, 4096)),
                           dim3(256), 0, s, p.root_bits, p.ntiles * (kTile / 64));
        hipLaunchKernelGGL(k_local, dim3((unsigned)ntile), dim3(256), 0, s, img, b, tg, vec16, p.parent, p.root_bits);
    }
    if (int e = check_launch("ccl local")) return e;
    {
        ProfScope ps("ccl_merge", s);
        hipLaunchKernelGGL(k_border, dim3((unsigned)ntile), dim3(256), 0, s, img, b, tg, vec16, p.parent);
    }
    if (int e = check_launch("ccl merge")) return e;
    {
        ProfScope ps("ccl_settle", s);
        hipLaunchKernelGGL(k_settle_roots, dim3((unsigned)((p.ntiles + 3) / 4)), dim3(256), 0, s, p.parent, p.ntiles,
                           p.root_bits, p.tiles);
    }
    if (int e = check_launch("ccl settle")) return e;
    {
        ProfScope ps("ccl_rank", s);
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(1024), 0, s, p.tiles, p.ntiles, count);
        hipLaunchKernelGGL(k_rank_roots, dim3((unsigned)((p.ntiles + 3) / 4)), dim3(256), 0, s, p.parent, p.ntiles,
                           p.root_bits, p.tiles);
    }
    if (int e = check_launch("ccl rank")) return e;
    {
        ProfScope ps("ccl_labels", s);
        if (labels32)
            hipLaunchKernelGGL(k_labels<int32_t>, dim3(gr), dim3(256), 0, s, p.parent, b, g->out_s[0],
                               g->out_s[1], labels32);
        else
            hipLaunchKernelGGL(k_labels<uint64_t>, dim3(gr), dim3(256), 0, s, p.parent, b, g->out_s[0],
                               g->out_s[1], labels64);
    }
    return check_launch("ccl labels");
}

int pmx_mark_labels(const void *labels, int label_bytes, const int64_t *strides, const int64_t *lo,
                    const int64_t *extent, uint8_t *flags, uint64_t n_flags, unsigned int *bad, hipStream_t s)
{
    using namespace pmx;
    clear_error();
    if (!labels || !strides || !lo || !extent || !flags) return fail(PMX_ERR_INVALID, "NULL argument");
    const int64_t n = extent[0] * extent[1] * extent[2];
    if (n <= 0) return PMX_OK;
    if (label_bytes == 4)
        hipLaunchKernelGGL(k_mark<int32_t>, dim3(grid_for(n)), dim3(256), 0, s,
                           static_cast<const int32_t *>(labels), strides[0], strides[1], lo[0], lo[1],
                           lo[2], extent[0], extent[1], extent[2], flags, n_flags, bad);
    else if (label_bytes == 8)
        hipLaunchKernelGGL(k_mark<uint64_t>, dim3(grid_for(n)), dim3(256), 0, s,
                           static_cast<const uint64_t *>(labels), strides[0], strides[1], lo[0], lo[1],
                           lo[2], extent[0], extent[1], extent[2], flags, n_flags, bad);
    else
        return fail(PMX_ERR_INVALID, "label_bytes must be 4 or 8");
    return check_launch("mark labels");
}

int pmx_keep_labels(const uint8_t *img, const void *labels, int label_bytes, int64_t n,
                    const uint8_t *flags_a, const uint8_t *flags_b, uint64_t n_flags, int mode, int phase,
                    uint8_t *out, unsigned int *bad, hipStream_t s)
{
    using namespace pmx;
    clear_error();
    if (!img || !labels || !flags_a || !out || n < 0) return fail(PMX_ERR_INVALID, "NULL argument");
    if (mode < 0 || mode > 2 || (mode > 0 && !flags_b)) return fail(PMX_ERR_INVALID, "bad mode");
    if (phase > 255) return fail(PMX_ERR_INVALID, "phase must be < 256 (or < 0 for any)");
    if (n == 0) return PMX_OK;
    if (label_bytes == 4)
        hipLaunchKernelGGL(k_keep<int32_t>, dim3(grid_for(n)), dim3(256), 0, s, img,
                           static_cast<const int32_t *>(labels), n, flags_a, flags_b, n_flags, mode, phase, out, bad);
    else if (label_bytes == 8)
        hipLaunchKernelGGL(k_keep<uint64_t>, dim3(grid_for(n)), dim3(256), 0, s, img,
                           static_cast<const uint64_t *>(labels), n, flags_a, flags_b, n_flags, mode, phase, out, bad);
    else
        return fail(PMX_ERR_INVALID, "label_bytes must be 4 or 8");
    return check_launch("keep labels");
}

}  // extern "C"
