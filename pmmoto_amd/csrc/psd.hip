// psd.hip -- the whole pore-size-distribution sweep in one call.
//
// pore_size_distribution(radii, inlet=True, mode="distance")
// (src/pmmoto/filters/porosimetry.py:163-227) runs, per radius r (in
// descending order), porosimetry: erosion (EDT of the image <= r -> 0),
// the inlet-connected part of the eroded set (cc3d + inlet labels), its
// dilation by r, and the first-invasion update of the radius map.  Here:
//
//  * one exact EDT of the image; the erosion at radius i is d2 > Te[i];
//  * voxels are bucketed by the first level at which they survive the
//    erosion (Te decreases with i, so the eroded sets grow), and a single
//    union-find over the sweep absorbs each level's new voxels: unions only
//    ever merge, so the labelling of level i is incremental, and inlet
//    connectivity is monotone (a root flagged once stays flagged).  Its
//    nodes are 2 x 2 x 2 voxel blocks: the voxels of a block are pairwise
//    26-adjacent, so a block's active voxels are always in one component;
//  * per level: inlet-slab voxels flag their roots, blocks whose root is
//    flagged join the inlet-connected set (their active voxels, and later
//    ones as they activate, take the level in in_c), and one exact EDT of
//    the level's joiners (distance to them) feeds the first-invasion update
//    directly from its last pass.
#include "psd_internal.hpp"

namespace pmx {
namespace {

// in_c of a voxel not (yet) in the inlet-connected set (levels are < 250):
constexpr uint32_t kInactive = 0xff;  // pore voxel not yet joined
constexpr uint32_t kNever = 0xfe;     // in no eroded set of the sweep (solid, or too thin)
constexpr int kNoLevel = 256;          // above every level: no neighbour there
constexpr uint8_t kNoJoin = 0xff;      // bjoin of a block not yet in the inlet-connected set

__device__ __forceinline__ int ld(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Parent reads at device scope (past the per-XCD L2) by default.  The
// merge of one level reads through L2 (kCached): within one kernel a stale
// parent is still an ancestor (pointers only move to smaller indices, so
// every walk ends), and a union acting on a stale root fails its CAS, which
// is atomic at device scope, and retries from the value it returns.
template <bool kCached = false>
__device__ __forceinline__ int ld_parent(const int *p)
{
    if constexpr (kCached) return *p;
    return ld(p);
}

template <bool kCached = false>
__device__ __forceinline__ int find_root(int *parent, int x)
{
    int p = ld_parent<kCached>(parent + x);
    while (p != x) {
        const int gp = ld_parent<kCached>(parent + p);
        if (gp != p) parent[x] = gp;
        x = p;
        p = gp;
    }
    return x;
}

// joins the sets of a and b; returns the root of the joined set as it was
// at the join (a later join may hook it under a smaller root, so it is an
// element of the set, a short walk from the root)
template <bool kCached = false>
__device__ __forceinline__ int unite(int *parent, int a, int b)
{
    while (true) {
        a = find_root<kCached>(parent, a);
        b = find_root<kCached>(parent, b);
        if (a == b) return a;
        const int hi = max(a, b), lo = min(a, b);
        const int old = atomicCAS(parent + hi, hi, lo);
        if (old == hi) return lo;
        a = old;
        b = lo;
    }
}

struct Dims {
    int n0, n1, n2;
    int per0, per1, per2;
    int nb0, nb1, nb2;  // 2 x 2 x 2 blocks per axis (the last one partial on an odd extent)
};

__device__ __forceinline__ int block_of(const Dims &d, int x, int y, int z)
{
    return ((x >> 1) * d.nb1 + (y >> 1)) * d.nb2 + (z >> 1);
}

constexpr int kMaxLevels = 250;
constexpr int kBucketItems = 16;  // voxels per thread between bucket reservations

// level[v] = first i with d2 > te[i] (the erosion at radius i keeps v),
// or n_levels if none (te does not increase, so a binary search finds
// it); counts[level] += per-block totals from an LDS histogram.
__global__ __launch_bounds__(256) void k_levels(const int32_t *__restrict__ d2, int64_t n, const int64_t *__restrict__ te,
                                                int n_levels, uint8_t *__restrict__ level,
                                                unsigned int *__restrict__ counts, uint8_t *__restrict__ in_c)
{
    __shared__ int64_t ste[kMaxLevels];
    __shared__ unsigned int hist[kMaxLevels];
    for (int k = threadIdx.x; k < n_levels; k += blockDim.x) {
        ste[k] = te[k];
        hist[k] = 0;
    }
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = d2[i];
        int lo = 0, hi = n_levels;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (v > ste[mid])
                hi = mid;
            else
                lo = mid + 1;
        }
        level[i] = (uint8_t)lo;
        in_c[i] = lo < n_levels ? kInactive : kNever;
        if (lo < n_levels) atomicAdd(hist + lo, 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n_levels; k += blockDim.x)
        if (hist[k]) atomicAdd(counts + k, hist[k]);
}

// exclusive scan of counts (n_levels + 1 entries) into offsets
__global__ void k_level_offsets(const unsigned int *counts, int n_levels, unsigned int *offsets, unsigned int *cursor)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned int run = 0;
        for (int k = 0; k <= n_levels; ++k) {
            offsets[k] = run;
            cursor[k] = run;
            run += (k < n_levels) ? counts[k] : 0;
        }
    }
}

// order[] = voxel indices grouped by level.  A block ranks a chunk of
// 256 * kBucketItems voxels per level in LDS, then reserves each level's
// range with one global atomic.
__global__ __launch_bounds__(256) void k_bucket(const uint8_t *__restrict__ level, int64_t n, int n_levels,
                                                unsigned int *__restrict__ cursor, int *__restrict__ order)
{
    __shared__ unsigned int cnt[kMaxLevels], base[kMaxLevels];
    const int64_t chunk = (int64_t)blockDim.x * kBucketItems;
    for (int64_t c0 = (int64_t)blockIdx.x * chunk; c0 < n; c0 += (int64_t)gridDim.x * chunk) {
        for (int k = threadIdx.x; k < n_levels; k += blockDim.x) cnt[k] = 0;
        __syncthreads();
        int lv[kBucketItems];
        unsigned int rk[kBucketItems];
#pragma unroll
        for (int j = 0; j < kBucketItems; ++j) {
            const int64_t i = c0 + (int64_t)j * blockDim.x + threadIdx.x;
            lv[j] = i < n ? level[i] : n_levels;
            rk[j] = lv[j] < n_levels ? atomicAdd(cnt + lv[j], 1u) : 0u;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < n_levels; k += blockDim.x)
            if (cnt[k]) base[k] = atomicAdd(cursor + k, cnt[k]);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kBucketItems; ++j)
            if (lv[j] < n_levels) order[base[lv[j]] + rk[j]] = (int)(c0 + (int64_t)j * blockDim.x + threadIdx.x);
        __syncthreads();
    }
}

// Per 2 x 2 x 2 block: blevel = the smallest level of its voxels (the level
// at which the block becomes a node of the union-find; n_levels: never),
// bjoin = kNoJoin, bflag = 0.  The parent array needs no initial value: a
// block's entry is read only once blevel says it is active, after
// k_activate.
__global__ __launch_bounds__(256) void k_block_init(const uint8_t *__restrict__ level, Dims d, int n_levels,
                                                    uint8_t *__restrict__ blevel, uint8_t *__restrict__ bjoin,
                                                    uint8_t *__restrict__ bflag)
{
    const int64_t nb = (int64_t)d.nb0 * d.nb1 * d.nb2;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x) {
        const int bz = (int)(b % d.nb2);
        const int64_t t = b / d.nb2;
        const int by = (int)(t % d.nb1), bx = (int)(t / d.nb1);
        int lo = n_levels;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                const int x = 2 * bx + i, y = 2 * by + j, z = 2 * bz;
                if (x >= d.n0 || y >= d.n1) continue;
                const int64_t v = ((int64_t)x * d.n1 + y) * d.n2 + z;
                lo = min(lo, (int)level[v]);
                if (z + 1 < d.n2) lo = min(lo, (int)level[v + 1]);
            }
        blevel[b] = (uint8_t)lo;
        bjoin[b] = kNoJoin;
        bflag[b] = 0;
    }
}

// Level lv's voxels are order[offsets[lv] .. offsets[lv+1]): the range is
// read on the device, so the sweep issues a fixed sequence of launches
// with no host synchronisation (it can be captured in a graph).  A block
// whose first voxels arrive at this level becomes a node (every such voxel
// writes the same entry).
__global__ __launch_bounds__(256) void k_activate(const int *__restrict__ order, const unsigned int *__restrict__ offsets,
                                                  int lv, Dims d, const uint8_t *__restrict__ blevel,
                                                  int *__restrict__ bparent)
{
    const unsigned int begin = offsets[lv], end = offsets[lv + 1];
    for (int64_t k = (int64_t)begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < end;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int v = order[k];
        const int z = v % d.n2, t = v / d.n2, y = t % d.n1, x = t / d.n1;
        const int b = block_of(d, x, y, z);
        if (blevel[b] == lv) bparent[b] = b;
    }
}

// each new voxel joins its block with the blocks of its active 26-neighbours
// (wrapping on periodic axes).  Activity comes from the level map (active
// at level lv <=> level <= lv), a read-only byte array that stays in L2.  A
// pair of new voxels is joined once, by the one with the larger index.
// Each union starts from the root the previous one returned (an element of
// the block's set), not from the block, and from the neighbour block's
// parent; a neighbour whose block is the voxel's own, or whose parent is
// that root or the parent just joined, is already in the set.
__global__ __launch_bounds__(256) void k_merge_new(const int *__restrict__ order, const unsigned int *__restrict__ offsets,
                                                   int lv, Dims d, const uint8_t *__restrict__ level,
                                                   int *__restrict__ bparent)
{
    const unsigned int begin = offsets[lv], end = offsets[lv + 1];
    for (int64_t k = (int64_t)begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < end;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int v = order[k];
        const int z = v % d.n2, t = v / d.n2, y = t % d.n1, x = t / d.n1;
        const int bv = block_of(d, x, y, z);
        int rv = bv, last = -1;
        for (int dx = -1; dx <= 1; ++dx) {
            int nx = x + dx;
            if (nx < 0 || nx >= d.n0) {
                if (!d.per0) continue;
                nx = nx < 0 ? nx + d.n0 : nx - d.n0;
            }
            for (int dy = -1; dy <= 1; ++dy) {
                int ny = y + dy;
                if (ny < 0 || ny >= d.n1) {
                    if (!d.per1) continue;
                    ny = ny < 0 ? ny + d.n1 : ny - d.n1;
                }
                const int row = (nx * d.n1 + ny) * d.n2;
                // the levels at z-1, z, z+1: away from the z faces one aligned 8-byte load
                // holds all three (two when they straddle an 8-byte boundary; the level map is
                // a 256-byte aligned workspace slice with slack after it)
                int lz[3] = {kNoLevel, kNoLevel, kNoLevel}, nzs[3] = {-1, -1, -1};
                if (z >= 1 && z + 1 < d.n2) {
                    const int a = row + z - 1, off = a & 7;
                    const uint64_t *q = reinterpret_cast<const uint64_t *>(level + (a - off));
                    uint64_t w = q[0] >> (8 * off);
                    if (off > 5) w |= q[1] << (64 - 8 * off);
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) {
                        lz[kk] = (int)((w >> (8 * kk)) & 0xffu);
                        nzs[kk] = z + kk - 1;
                    }
                } else {
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) {
                        int nz = z + kk - 1;
                        if (nz < 0 || nz >= d.n2) {
                            if (!d.per2) continue;
                            nz = nz < 0 ? nz + d.n2 : nz - d.n2;
                        }
                        nzs[kk] = nz;
                        lz[kk] = level[row + nz];
                    }
                }
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    if (!dx && !dy && kk == 1) continue;
                    const int nz = nzs[kk], lu = lz[kk];
                    if (nz < 0 || lu > lv || (lu == lv && row + nz > v)) continue;
                    const int bu = block_of(d, nx, ny, nz);
                    if (bu == bv) continue;
                    const int pu = bparent[bu];
                    if (pu == rv || pu == last) continue;
                    last = pu;
                    rv = unite<true>(bparent, rv, pu);
                }
            }
        }
    }
}

// active voxels of the inlet slab flag their blocks' current roots.  Here
// and in k_join_blocks no union runs, so the forest only changes by path
// halving (a pointer moves to an ancestor) and its reads go through L2.
__global__ __launch_bounds__(256) void k_flag_inlet(Dims d, int64_t lo0, int64_t lo1, int64_t lo2, int64_t e0,
                                                    int64_t e1, int64_t e2, const uint8_t *__restrict__ level, int lv,
                                                    int *__restrict__ bparent, uint8_t *__restrict__ bflag)
{
    const int64_t n = e0 * e1 * e2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = i % e2, t = i / e2, y = t % e1, x = t / e1;
        const int gx = (int)(lo0 + x), gy = (int)(lo1 + y), gz = (int)(lo2 + z);
        if (level[((int64_t)gx * d.n1 + gy) * d.n2 + gz] <= lv) {
            const int r = find_root<true>(bparent, block_of(d, gx, gy, gz));
            if (!bflag[r]) bflag[r] = 1;  // read first: a large component's blocks share one root
        }
    }
}

// 1 in each byte of w equal to the byte in pat, else 0 (exact: no carry
// crosses a byte, (b & 0x7f) + 0x7f <= 0xfe)
__device__ __forceinline__ uint32_t eq_bytes(uint32_t w, uint32_t pat)
{
    const uint32_t x = w ^ pat;
    return ~(((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x | 0x7f7f7f7fu) >> 7;
}

// 1 in each byte of w that is <= the byte in pat (unsigned).  Per byte, with
// x = pat and y = w: where the top bits differ, x >= y iff x's top bit is
// set; where they agree, iff the low seven bits compare so, which is the top
// bit of (x | 0x80) - (y & 0x7f) (at least 1 per byte: no borrow crosses)
__device__ __forceinline__ uint32_t le_bytes(uint32_t w, uint32_t pat)
{
    constexpr uint32_t kH = 0x80808080u;
    const uint32_t low = (pat | kH) - (w & ~kH);
    return (((pat & ~w) | (~(pat ^ w) & low)) & kH) >> 7;
}

// The inlet-connected set at level lv.  A block joins when its root is
// flagged (bjoin = lv): its active voxels (level <= lv) take in_c = lv.  A
// member stays a member, so only active blocks not yet joined look up their
// root; 16 blocks per lane step (bjoin and blevel are 256-byte aligned
// workspace slices, levels < 128), and the candidates of a group walk to
// their roots in lockstep, each entry then set to its root (no union runs
// here, so the root is a valid ancestor).  any: set when the level has a
// new member (otherwise its dilation is skipped on the GPU).
__global__ __launch_bounds__(256) void k_join_blocks(Dims d, int lv, int *__restrict__ bparent,
                                                     const uint8_t *__restrict__ bflag,
                                                     const uint8_t *__restrict__ blevel, uint8_t *__restrict__ bjoin,
                                                     const uint8_t *__restrict__ level, uint8_t *__restrict__ in_c,
                                                     unsigned int *__restrict__ any)
{
    __shared__ int blk_seen;
    if (threadIdx.x == 0) blk_seen = 0;
    __syncthreads();
    bool seen = false;
    const int64_t nb = (int64_t)d.nb0 * d.nb1 * d.nb2;
    const int64_t ng = (nb + 15) / 16, stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t none4 = kNoJoin * 0x01010101u, lv4 = (uint32_t)lv * 0x01010101u;
    // writes the level into in_c for block b's voxels of level <= lv
    auto join_voxels = [&](int64_t b) {
        const int bz = (int)(b % d.nb2);
        const int64_t t = b / d.nb2;
        const int by = (int)(t % d.nb1), bx = (int)(t / d.nb1);
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                const int x = 2 * bx + i, y = 2 * by + j, z = 2 * bz;
                if (x >= d.n0 || y >= d.n1) continue;
                const int64_t v = ((int64_t)x * d.n1 + y) * d.n2 + z;
                if (level[v] <= lv) in_c[v] = (uint8_t)lv;
                if (z + 1 < d.n2 && level[v + 1] <= lv) in_c[v + 1] = (uint8_t)lv;
            }
    };
    for (int64_t q = tid; q < ng; q += stride) {
        uint32_t cand = 0;
        if (16 * q + 16 <= nb) {
            const uint4 jw = reinterpret_cast<const uint4 *>(bjoin)[q];
            const uint4 lw = reinterpret_cast<const uint4 *>(blevel)[q];
            const uint32_t jj[4] = {jw.x, jw.y, jw.z, jw.w}, ll[4] = {lw.x, lw.y, lw.z, lw.w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const uint32_t m = eq_bytes(jj[h], none4) & le_bytes(ll[h], lv4);
                // one bit per byte: bits 0, 8, 16, 24 -> 4 consecutive bits
                cand |= ((m | (m >> 7) | (m >> 14) | (m >> 21)) & 0xfu) << (4 * h);
            }
        } else {
            for (int e = 0; 16 * q + e < nb; ++e)
                if (bjoin[16 * q + e] == kNoJoin && blevel[16 * q + e] <= lv) cand |= 1u << e;
        }
        if (!cand) continue;
        int r[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) r[e] = (int)(16 * q + e);
        uint32_t walking = cand;
        while (walking) {
            int nx[16];
#pragma unroll
            for (int e = 0; e < 16; ++e) nx[e] = (walking >> e) & 1u ? bparent[r[e]] : r[e];
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                if (nx[e] == r[e]) walking &= ~(1u << e);
                r[e] = nx[e];
            }
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            if (!((cand >> e) & 1u)) continue;
            const int64_t b = 16 * q + e;
            if (bparent[b] != r[e]) bparent[b] = r[e];
            if (bflag[r[e]]) {
                bjoin[b] = (uint8_t)lv;
                join_voxels(b);
                seen = true;
            }
        }
    }
    if (__any(seen) && (threadIdx.x & 63) == 0) blk_seen = 1;
    __syncthreads();
    // one lane per block, and only while the flag still reads 0: the shared word is
    // read at device scope (past the per-XCD L2), so few reads and fewer atomics
    if (threadIdx.x == 0 && blk_seen && ld(reinterpret_cast<const int *>(any)) == 0) atomicOr(any, 1u);
}

// the level's new voxels in blocks that joined at an earlier level join now
__global__ __launch_bounds__(256) void k_join_new(const int *__restrict__ order, const unsigned int *__restrict__ offsets,
                                                  int lv, Dims d, const uint8_t *__restrict__ bjoin,
                                                  uint8_t *__restrict__ in_c, unsigned int *__restrict__ any)
{
    const unsigned int begin = offsets[lv], end = offsets[lv + 1];
    bool seen = false;
    for (int64_t k = (int64_t)begin + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < end;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int v = order[k];
        const int z = v % d.n2, t = v / d.n2, y = t % d.n1, x = t / d.n1;
        const uint32_t j = bjoin[block_of(d, x, y, z)];
        if (j < (uint32_t)lv) {
            in_c[v] = (uint8_t)lv;
            seen = true;
        }
    }
    if (__any(seen) && (threadIdx.x & 63) == 0 && ld(reinterpret_cast<const int *>(any)) == 0) atomicOr(any, 1u);
}

// the erosion thresholds travel as a kernel argument (a graph-capturable
// upload: no host-to-device copy from pageable memory)
struct Thresholds {
    int64_t t[kMaxLevels];
};

__global__ void k_store_thresholds(Thresholds th, int n_levels, int64_t *te)
{
    for (int i = threadIdx.x; i < n_levels; i += blockDim.x) te[i] = th.t[i];
}

// Without an inlet requirement the eroded sets themselves are the centres:
// level lv's new centres are the voxels whose erosion level is lv.
// any[k] = 1 where level k has voxels (its new centres without an inlet requirement)
__global__ void k_any_from_counts(const unsigned int *__restrict__ counts, int n_levels, unsigned int *__restrict__ any)
{
    for (int k = threadIdx.x; k < n_levels; k += blockDim.x) any[k] = counts[k] != 0u;
}

__global__ void k_clear_u32(unsigned int *a, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = 0;
}

// code map = 0 over the box; state (contiguous box) = 1 on pore voxels
// (img == 1), the ones the first-invasion update may still assign
__global__ __launch_bounds__(256) void k_zero_map(int32_t *__restrict__ map, const uint8_t *__restrict__ img,
                                                  uint8_t *__restrict__ state, Dims d, int64_t s0, int64_t s1)
{
    // a row (x, y) per block step: one division per row, none per voxel
    const int64_t rows = (int64_t)d.n0 * d.n1;
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const int64_t x = r / d.n1, y = r - x * d.n1;
        const int64_t jr = x * s0 + y * s1, ir = r * d.n2;
        for (int z = threadIdx.x; z < d.n2; z += blockDim.x) {
            map[jr + z] = 0;
            state[ir + z] = img[jr + z] == 1 ? 1 : 0;
        }
    }
}

struct PsdPlan {
    int32_t *d2 = nullptr;
    uint8_t *level = nullptr, *in_c = nullptr, *state = nullptr;
    uint8_t *blevel = nullptr, *bjoin = nullptr, *bflag = nullptr;  // per 2 x 2 x 2 block
    int *bparent = nullptr, *order = nullptr;
    unsigned int *counts = nullptr, *offsets = nullptr, *cursor = nullptr, *any = nullptr;
    int64_t *te = nullptr;
    char *edt_ws = nullptr;
    size_t edt_bytes = 0;
};

size_t plan_psd(const pmx_geom *g, void *ws, PsdPlan *p)
{
    Carver c(ws, SIZE_MAX);
    const size_t n = (size_t)(g->n[0] * g->n[1] * g->n[2]);
    p->d2 = c.take<int32_t>(n);
    const size_t nb = (size_t)((g->n[0] + 1) / 2) * ((g->n[1] + 1) / 2) * ((g->n[2] + 1) / 2);
    p->level = c.take<uint8_t>(n);
    p->in_c = c.take<uint8_t>(n);
    p->state = c.take<uint8_t>(n);
    p->blevel = c.take<uint8_t>(nb);
    p->bjoin = c.take<uint8_t>(nb);
    p->bflag = c.take<uint8_t>(nb);
    p->bparent = c.take<int>(nb);
    p->order = c.take<int>(n);
    p->counts = c.take<unsigned int>(kMaxLevels + 1);
    p->offsets = c.take<unsigned int>(kMaxLevels + 1);
    p->cursor = c.take<unsigned int>(kMaxLevels + 1);
    p->any = c.take<unsigned int>(kMaxLevels);
    p->te = c.take<int64_t>(kMaxLevels);
    EdtPlan e;
    p->edt_bytes = plan_edt(g, nullptr, &e);
    p->edt_ws = c.take<char>(p->edt_bytes);
    return c.need();
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65536); }

}  // namespace

size_t psd_workspace_bytes(const pmx_geom *g)
{
    PsdPlan p;
    return plan_psd(g, nullptr, &p);
}

// argument checks shared by the single and batched sweeps
int check_psd(const uint8_t *img, const pmx_geom *g, const int64_t *inlet_lo, const int64_t *inlet_ext, int n_levels,
              const int64_t *t_erode, const int64_t *t_dilate, const int32_t *code_map)
{
    if (int e = validate_geom(g)) return e;
    if (!img || !t_erode || !t_dilate || !code_map) return fail(PMX_ERR_INVALID, "NULL argument");
    if (n_levels < 1 || n_levels > kMaxLevels) return fail(PMX_ERR_INVALID, "n_levels must be in [1, %d]", kMaxLevels);
    const int64_t n = g->n[0] * g->n[1] * g->n[2];
    if (n >= INT32_MAX) return fail(PMX_ERR_INVALID, "box has %lld voxels (limit 2^31-1)", (long long)n);
    // the first-invasion update reads the image at the code map's index
    if (g->out_s[0] != g->in_s[0] || g->out_s[1] != g->in_s[1])
        return fail(PMX_ERR_INVALID, "the sweep needs the same strides for the image and the code map");
    for (int k = 1; k < n_levels; ++k)
        if (t_erode[k] > t_erode[k - 1])
            return fail(PMX_ERR_INVALID, "erosion thresholds must not increase (radii in descending order)");
    const bool has_inlet = inlet_lo && inlet_ext;
    if (has_inlet)
        for (int a = 0; a < 3; ++a)
            if (inlet_lo[a] < 0 || inlet_ext[a] < 0 || inlet_lo[a] + inlet_ext[a] > g->n[a])
                return fail(PMX_ERR_INVALID, "inlet slab outside the box on axis %d", a);
    return PMX_OK;
}

int run_psd(const uint8_t *img, const pmx_geom *g, const int64_t *inlet_lo, const int64_t *inlet_ext, int n_levels,
            const int64_t *t_erode, const int64_t *t_dilate, uint32_t wall_mask, int32_t *code_map, void *workspace,
            size_t workspace_bytes, hipStream_t s, int tile_rows)
{
    const int64_t n = g->n[0] * g->n[1] * g->n[2];
    const bool has_inlet = inlet_lo && inlet_ext;
    PsdPlan p;
    const size_t need = plan_psd(g, nullptr, &p);
    if (!workspace || workspace_bytes < need)
        return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    plan_psd(g, workspace, &p);
    Dims d{(int)g->n[0], (int)g->n[1], (int)g->n[2], g->bc[0] == PMX_BC_PERIODIC, g->bc[2] == PMX_BC_PERIODIC,
           g->bc[4] == PMX_BC_PERIODIC, (int)(g->n[0] + 1) / 2, (int)(g->n[1] + 1) / 2, (int)(g->n[2] + 1) / 2};
    const int64_t nblk = (int64_t)d.nb0 * d.nb1 * d.nb2;
    const unsigned gr = grid_for(n);
    // the erosion EDT reads the image with its strides and writes the
    // internal contiguous d2; the dilations read the internal contiguous mask
    pmx_geom gc = *g;
    gc.out_s[0] = g->n[1] * g->n[2];
    gc.out_s[1] = g->n[2];
    // A WALL face's layer is solid: a feature of the erosion EDT (distance
    // to solid) but not of the dilations (distance to the centres), which
    // see nothing beyond the face (morphological_operators.py:103-172 pads
    // with solid and dilates from the pore-phase centres only).
    pmx_geom gd = gc;
    gd.in_s[0] = gc.out_s[0];
    gd.in_s[1] = gc.out_s[1];
    for (int f = 0; f < 6; ++f)
        if (gd.bc[f] == PMX_BC_WALL) gd.bc[f] = PMX_BC_END;

    // 1. one exact EDT of the image; levels; buckets
    {
        ProfScope ps("psd_edt", s);
        PassXOut o{};
        o.d2 = p.d2;
        o.sx = gc.out_s[0];
        o.sy = gc.out_s[1];
        // only comparisons with the erosion thresholds follow: clamp above the largest
        if (int e = run_edt(img, &gc, false, o, t_erode[0] < INT64_MAX ? t_erode[0] + 1 : INT64_MAX, p.edt_ws, p.edt_bytes, s,
                            nullptr, tile_rows)) return e;
    }
    {
        Thresholds th{};
        for (int k = 0; k < n_levels; ++k) th.t[k] = t_erode[k];
        hipLaunchKernelGGL(k_store_thresholds, dim3(1), dim3(256), 0, s, th, n_levels, p.te);
    }
    {
        ProfScope ps("psd_levels", s);
        hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(256), 0, s, p.counts, kMaxLevels + 1);
        hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(256), 0, s, p.any, kMaxLevels);
        // few blocks: each adds its per-level totals with one global atomic per level
        hipLaunchKernelGGL(k_levels, dim3(std::min<unsigned>(gr, 2048u)), dim3(256), 0, s, p.d2, n, p.te, n_levels,
                           p.level, p.counts, p.in_c);
        if (!has_inlet) hipLaunchKernelGGL(k_any_from_counts, dim3(1), dim3(256), 0, s, p.counts, n_levels, p.any);
        if (has_inlet) {  // the union-find absorbs each level's voxels from these buckets
            hipLaunchKernelGGL(k_level_offsets, dim3(1), dim3(64), 0, s, p.counts, n_levels, p.offsets, p.cursor);
            hipLaunchKernelGGL(k_bucket, dim3(grid_for((n + kBucketItems - 1) / kBucketItems)), dim3(256), 0, s,
                               p.level, n, n_levels, p.cursor, p.order);
            hipLaunchKernelGGL(k_block_init, dim3(grid_for(nblk)), dim3(256), 0, s, p.level, d, n_levels, p.blevel,
                               p.bjoin, p.bflag);
        }
        const unsigned grow = (unsigned)std::min<int64_t>((int64_t)d.n0 * d.n1, 65536);
        hipLaunchKernelGGL(k_zero_map, dim3(grow), dim3(256), 0, s, code_map, img, p.state, d, g->out_s[0],
                           g->out_s[1]);
    }
    if (int e = check_launch("psd setup")) return e;

    const unsigned gu = std::min<unsigned>(gr, 2048u);  // grid-stride over each level's bucket
    for (int lv = 0; lv < n_levels; ++lv) {
        // without an inlet the level's new centres are the voxels of level lv:
        // the dilation matches them in the level map
        if (has_inlet) {
            {
                ProfScope ps("psd_union", s);
                hipLaunchKernelGGL(k_activate, dim3(gu), dim3(256), 0, s, p.order, p.offsets, lv, d, p.blevel,
                                   p.bparent);
                hipLaunchKernelGGL(k_merge_new, dim3(gu), dim3(256), 0, s, p.order, p.offsets, lv, d, p.level,
                                   p.bparent);
            }
            ProfScope ps("psd_mask", s);
            const int64_t ni = inlet_ext[0] * inlet_ext[1] * inlet_ext[2];
            hipLaunchKernelGGL(k_flag_inlet, dim3(grid_for(ni)), dim3(256), 0, s, d, inlet_lo[0], inlet_lo[1],
                               inlet_lo[2], inlet_ext[0], inlet_ext[1], inlet_ext[2], p.level, lv, p.bparent, p.bflag);
            hipLaunchKernelGGL(k_join_blocks, dim3(grid_for(nblk / 16 + 1)), dim3(256), 0, s, d, lv, p.bparent,
                               p.bflag, p.blevel, p.bjoin, p.level, p.in_c, p.any + lv);
            hipLaunchKernelGGL(k_join_new, dim3(gu), dim3(256), 0, s, p.order, p.offsets, lv, d, p.bjoin, p.in_c,
                               p.any + lv);
        }
        if (int rc = check_launch("psd level")) return rc;
        // dilation of the level's new centres with the first-invasion update
        PassXOut o{};
        o.sx = g->out_s[0];
        o.sy = g->out_s[1];
        o.t = (int32_t)std::min<int64_t>(std::max<int64_t>(t_dilate[lv], -1), INT32_MAX - 1);
        o.wall_mask = wall_mask;
        o.psd_state = p.state;
        o.ssx = g->n[1] * g->n[2];
        o.ssy = g->n[2];
        o.psd_map = code_map;
        o.psd_code = lv + 1;
        ProfScope ps("psd_dilate", s);
        // skipped on the GPU when the level has no new centre
        if (int rc = run_edt(has_inlet ? p.in_c : p.level, &gd, true, o, (int64_t)o.t + 1, p.edt_ws, p.edt_bytes, s,
                             p.any + lv, tile_rows, lv)) return rc;
    }
    return PMX_OK;
}

}  // namespace pmx

extern "C" {

int pmx_psd_workspace_size(const pmx_geom *g, size_t *bytes)
{
    pmx::clear_error();
    if (int e = pmx::validate_geom(g)) return e;
    if (!bytes) return pmx::fail(PMX_ERR_INVALID, "bytes is NULL");
    pmx::PsdPlan p;
    *bytes = pmx::plan_psd(g, nullptr, &p);
    return PMX_OK;
}

int pmx_psd_sweep(const uint8_t *img, const pmx_geom *g, const int64_t *inlet_lo, const int64_t *inlet_ext,
                  int n_levels, const int64_t *t_erode, const int64_t *t_dilate, uint32_t wall_mask,
                  int32_t *code_map, void *workspace, size_t workspace_bytes, hipStream_t s)
{
    using namespace pmx;
    clear_error();
    if (int e = check_psd(img, g, inlet_lo, inlet_ext, n_levels, t_erode, t_dilate, code_map)) return e;
    return run_psd(img, g, inlet_lo, inlet_ext, n_levels, t_erode, t_dilate, wall_mask, code_map, workspace,
                   workspace_bytes, s, 0);
}

}  // extern "C"
