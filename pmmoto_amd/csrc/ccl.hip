// ccl.hip -- 26-connected multi-label connected components.
//
// Replaces cc3d.connected_components(img, return_N=True) used at
// src/pmmoto/filters/connected_components.py:52-54 and, for periodic
// faces, the boundary-label matching of connect_subdomain_boundaries
// (:65-101, core/voxels.py:167-397): the wrap is part of the neighbourhood.
//
//   local     per 8 x 8 x 64 tile, in LDS: each voxel unions with its 13
//             backward 26-neighbours of the same value inside the tile;
//             parent[i] = global index of the tile-local root (background
//             marked)
//   border    the backward pairs that leave a tile or wrap on periodic
//             axes, unioned globally; lock-free union by atomicCAS hooking
//             the larger root under the smaller, so every component's root
//             is its smallest C-order index
//   compress  parent[i] = root(i); per 4096-voxel rank tile, its roots as a
//             bitmap and their count
//   rank      tile counts scanned; from the bitmap, each root's rank is
//             written over its parent entry as -(rank+1)
//   label     label[i] = rank of root(i) + 1: components numbered 1..N in
//             C-order of their first voxel (the canonical numbering)
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

__global__ __launch_bounds__(256) void k_local(const uint8_t *__restrict__ img, Box b, TileGeom tg,
                                               int vec16, int *__restrict__ parent)
{
    static_assert(kTVox == 256 * 16, "one 16-byte quarter row per thread");
    __shared__ int lp[kTVox];
    __shared__ __align__(16) uint8_t lv[kTVox];
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
    for (int l = threadIdx.x; l < kTVox; l += blockDim.x) lp[l] = l;
    __syncthreads();
    for (int l = threadIdx.x; l < kTVox; l += blockDim.x) {
        const uint8_t v = lv[l];
        if (!v) continue;
        const int lx = l / (kTY * kTZ), ly = (l / kTZ) % kTY, lz = l % kTZ;
        int last = -1;  // neighbours pointing at the node already joined add nothing
#pragma unroll
        for (int k = 0; k < 13; ++k) {
            int dx, dy, dz;
            back_nb(k, dx, dy, dz);
            const int nx = lx + dx, ny = ly + dy, nz = lz + dz;
            if (nx < 0 || ny < 0 || ny >= ey || nz < 0 || nz >= ez) continue;  // leaves the tile: phase 2
            const int m = (nx * kTY + ny) * kTZ + nz;
            if (lv[m] != v) continue;
            const int pm = lp[m];
            if (pm == last) continue;
            lunite(lp, l, m);
            last = pm;
        }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < kTVox; l += blockDim.x) {
        const int lx = l / (kTY * kTZ), ly = (l / kTZ) % kTY, lz = l % kTZ;
        if (lx >= ex || ly >= ey || lz >= ez) continue;
        const int64_t i = ((int64_t)(x0 + lx) * b.n1 + (y0 + ly)) * b.n2 + z0 + lz;
        if (!lv[l]) {
            parent[i] = kBg;
            continue;
        }
        const int r = lfind(lp, l);
        const int rx = r / (kTY * kTZ), ry = (r / kTZ) % kTY, rz = r % kTZ;
        parent[i] = (int)(((int64_t)(x0 + rx) * b.n1 + (y0 + ry)) * b.n2 + z0 + rz);
    }
}

// coordinate c + d on an axis of n voxels: wrapped on a periodic axis,
// else -1 outside the box
__device__ __forceinline__ int step(int c, int d, int n, int per)
{
    int r = c + d;
    if (r < 0 || r >= n) {
        if (!per) return -1;
        r = r < 0 ? r + n : r - n;
    }
    return r;
}

// One candidate pair per lane (pj < 0: none).  A lane whose (parent of i,
// parent of j) pair equals the previous lane's adds nothing (that lane, or
// one before it, joins them), nor does one whose pj it has just joined;
// the rest go through the block's set of joined parent pairs: a pair
// already there was joined by the lane that inserted it, and a full probe
// falls back to joining.  Every lane of the wave calls this (shuffles).
__device__ __forceinline__ void join_pair(int *parent, unsigned long long *seen, int pi, int pj, int &last, int lane)
{
    constexpr int kProbe = 8;
    constexpr unsigned long long kEmpty = ~0ull;
    const int ppi = __shfl_up(pi, 1), ppj = __shfl_up(pj, 1);
    const bool dup = lane > 0 && ppj >= 0 && ppi == pi && ppj == pj;
    if (pj < 0 || dup || pj == last) return;
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

// Phase 2: the backward pairs k_local did not see, in two disjoint groups.
//  (a) Face rows: a z-row of the tile with a backward row R (offsets
//      (-1,-1), (-1,0), (-1,+1), (0,-1) in x, y) outside the tile's x/y
//      footprint takes every dz of that R (wrapped on periodic axes).  A
//      wave holds the row, lane = z: R is loaded once, with one halo voxel
//      at each end, and dz = -1 / +1 come from lane shuffles.
//  (b) Z faces: for backward rows inside the footprint (and the voxel's own
//      row for dz = -1), the pairs whose dz leaves the tile's z range: the
//      lz = 0 voxels with dz = -1 and the lz = ez-1 voxels with dz = +1, one
//      lane per row.
// Parent reads go through L2: a stale parent is still an ancestor, and a
// union acting on a stale root fails its CAS and retries.
__global__ __launch_bounds__(256) void k_border(const uint8_t *__restrict__ img, Box b, TileGeom tg,
                                                int *__restrict__ parent)
{
    static_assert(kTX * kTY == 64 && kTZ == 64, "a z-row per wave, a z face per wave");
    constexpr unsigned long long kEmpty = ~0ull;
    __shared__ unsigned long long seen[kBorderHash];
    for (int h = threadIdx.x; h < kBorderHash; h += blockDim.x) seen[h] = kEmpty;
    __syncthreads();
    int x0, y0, z0, ex, ey, ez;
    tile_of(b, tg, blockIdx.x, x0, y0, z0, ex, ey, ez);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

    // (a) face rows
    for (int row = wv; row < ex * ey; row += 4) {
        const int lx = row / ey, ly = row - lx * ey;
        if (lx != 0 && ly != 0 && ly != ey - 1) continue;  // every backward row inside the footprint
        const int x = x0 + lx, y = y0 + ly, z = z0 + lane;
        const bool in_z = lane < ez;
        const uint8_t v = in_z ? img[(int64_t)x * b.is0 + (int64_t)y * b.is1 + z] : (uint8_t)0;
        const int pi = v ? parent[((int64_t)x * b.n1 + y) * b.n2 + z] : -1;
        const int zlo = lane == 0 ? step(z0, -1, b.n2, b.per2) : -1;
        const int zhi = lane == ez - 1 ? step(z0 + ez - 1, 1, b.n2, b.per2) : -1;
        int last = -1;
        for (int r = 0; r < 4; ++r) {
            const int dx = r < 3 ? -1 : 0, dy = r < 3 ? r - 1 : -1;
            if (lx + dx >= 0 && ly + dy >= 0 && ly + dy < ey) continue;  // k_local or the z faces
            const int nx = step(x, dx, b.n0, b.per0), ny = step(y, dy, b.n1, b.per1);
            if (nx < 0 || ny < 0) continue;
            const uint8_t *irow = img + (int64_t)nx * b.is0 + (int64_t)ny * b.is1;
            const int *prow = parent + ((int64_t)nx * b.n1 + ny) * b.n2;
            const uint8_t cv = in_z ? irow[z] : (uint8_t)0;
            const int cp = cv ? prow[z] : -1;
            int lov = __shfl_up((int)cv, 1), lop = __shfl_up(cp, 1);
            int hiv = __shfl_down((int)cv, 1), hip = __shfl_down(cp, 1);
            if (lane == 0) {
                lov = zlo >= 0 ? irow[zlo] : 0;
                lop = lov ? prow[zlo] : -1;
            }
            if (lane == ez - 1) {
                hiv = zhi >= 0 ? irow[zhi] : 0;
                hip = hiv ? prow[zhi] : -1;
            }
            join_pair(parent, seen, pi, v && lov == v ? lop : -2, last, lane);
            join_pair(parent, seen, pi, v && cv == v ? cp : -2, last, lane);
            join_pair(parent, seen, pi, v && hiv == v ? hip : -2, last, lane);
        }
    }

    // (b) z faces: wave 2 the lz = 0 face (dz = -1), wave 3 the lz = ez-1 face (dz = +1)
    if (wv >= 2) {
        const bool lo = wv == 2;
        const int lx = lane / kTY, ly = lane % kTY;
        const int x = x0 + lx, y = y0 + ly, z = z0 + (lo ? 0 : ez - 1);
        const int nz = step(z, lo ? -1 : 1, b.n2, b.per2);  // the same for every lane
        if (nz < 0) return;
        const uint8_t v = lx < ex && ly < ey ? img[(int64_t)x * b.is0 + (int64_t)y * b.is1 + z] : (uint8_t)0;
        const int pi = v ? parent[((int64_t)x * b.n1 + y) * b.n2 + z] : -1;
        int last = -1;
        for (int r = 0; r < 5; ++r) {
            const int dx = r < 3 ? -1 : 0, dy = r < 3 ? r - 1 : (r == 3 ? -1 : 0);
            // rows outside the footprint are the face rows' (a); the own row only for dz = -1
            const bool inside = lx + dx >= 0 && ly + dy >= 0 && ly + dy < ey && (r < 4 || lo);
            const int64_t m = (int64_t)(x + dx) * b.is0 + (int64_t)(y + dy) * b.is1 + nz;
            const uint8_t nv = v && inside ? img[m] : (uint8_t)0;
            const int pj = v && nv == v ? parent[((int64_t)(x + dx) * b.n1 + (y + dy)) * b.n2 + nz] : -2;
            join_pair(parent, seen, pi, pj, last, lane);
        }
    }
}

// The root walks here only read.  Compression must not halve other voxels'
// paths: a halving write computed from a stale read could overwrite a
// pointer another lane has just set to the root with a non-root ancestor,
// and the label gather relies on every parent being a root.  No union runs
// any more, so plain (L2) reads are enough: an entry only ever changes to
// its root.
//
// One block per rank tile of kTile voxels, in chunks of 256 consecutive
// voxels: parent[i] = root(i) (each lane writes only its own entry, so roots
// and pointers already set to a root stay as written), the tile's roots as
// one bit per voxel (a ballot per 64 voxels), and its root count.
__global__ __launch_bounds__(256) void k_compress(int *__restrict__ parent, int64_t n,
                                                  uint64_t *__restrict__ root_bits, int *__restrict__ tile_counts)
{
    __shared__ int sh[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kTile;
    int c = 0;
    // four chunks at a time: their four root walks advance in lockstep, so a
    // lane keeps four independent parent loads in flight
    constexpr int kWays = 4;
    for (int k = 0; k < kTile / 256; k += kWays) {
        int64_t i[kWays];
        int p[kWays], r[kWays];
        bool root[kWays];
#pragma unroll
        for (int s = 0; s < kWays; ++s) {
            i[s] = base + 256 * (k + s) + threadIdx.x;
            p[s] = i[s] < n ? parent[i[s]] : kBg;
            root[s] = p[s] == (int)i[s];
            r[s] = p[s];
        }
        // r[s] walks from p[s] to its root (the entry of a root is itself); background and
        // roots need no walk
        bool open[kWays];
#pragma unroll
        for (int s = 0; s < kWays; ++s) open[s] = p[s] != kBg && !root[s];
        while (open[0] || open[1] || open[2] || open[3]) {
            int q[kWays];
#pragma unroll
            for (int s = 0; s < kWays; ++s) q[s] = open[s] ? parent[r[s]] : r[s];
#pragma unroll
            for (int s = 0; s < kWays; ++s) {
                if (q[s] == r[s]) open[s] = false;
                r[s] = q[s];
            }
        }
#pragma unroll
        for (int s = 0; s < kWays; ++s) {
            if (p[s] != kBg && !root[s] && r[s] != p[s]) parent[i[s]] = r[s];
            const uint64_t m = __ballot(root[s]);
            if (lane == 0) root_bits[(base + 256 * (k + s)) / 64 + wv] = m;
            c += root[s];
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) sh[wv] = c;
    __syncthreads();
    if (threadIdx.x == 0) tile_counts[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
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

__device__ __forceinline__ int64_t label_of(const int *parent, int p)
{
    return p == kBg ? 0 : (p < 0 ? -(int64_t)p : -(int64_t)parent[p]);
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
    {
        ProfScope ps("ccl_init", s);
        const int vec16 = ((uintptr_t)img % 16 == 0 && g->in_s[0] % 16 == 0 && g->in_s[1] % 16 == 0) ? 1 : 0;
        hipLaunchKernelGGL(k_local, dim3((unsigned)ntile), dim3(256), 0, s, img, b, tg, vec16, p.parent);
    }
    if (int e = check_launch("ccl local")) return e;
    {
        ProfScope ps("ccl_merge", s);
        hipLaunchKernelGGL(k_border, dim3((unsigned)ntile), dim3(256), 0, s, img, b, tg, p.parent);
    }
    if (int e = check_launch("ccl merge")) return e;
    {
        ProfScope ps("ccl_compress", s);
        hipLaunchKernelGGL(k_compress, dim3((unsigned)p.ntiles), dim3(256), 0, s, p.parent, n, p.root_bits, p.tiles);
    }
    if (int e = check_launch("ccl compress")) return e;
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
