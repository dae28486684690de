// edt_kernels.hpp -- the device code of the exact EDT (pass Z, the
// windowed passes Y and X, the envelope fallbacks), shared by the
// single-domain entry points (edt.hip) and the batched sweep (batch.hip).
// Internal linkage: each translation unit that includes it instantiates its
// own kernels.
#pragma once

#include "edt_internal.hpp"

namespace pmx {
namespace {

constexpr int kInf = PMX_D2_INF;
constexpr int kNone = 1 << 29;  // "no solid on this side" position marker

// ---------------------------------------------------------------------------
// pass Z
// ---------------------------------------------------------------------------

// Wave scans on the data-parallel-primitive paths: Hillis-Steele steps
// within each row of 16 lanes (row_shr / row_shl), then the row totals
// (three readlanes) folded into the later (or earlier) rows.
__device__ __forceinline__ int wave_incl_max(int v, int lane)
{
    v = max(v, __builtin_amdgcn_update_dpp(-kNone, v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(-kNone, v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(-kNone, v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(-kNone, v, 0x118, 0xf, 0xf, false));  // row_shr:8
    const int r0 = __builtin_amdgcn_readlane(v, 15), r1 = __builtin_amdgcn_readlane(v, 31),
              r2 = __builtin_amdgcn_readlane(v, 47);
    const int row = lane >> 4;
    const int s01 = max(r0, r1), s012 = max(s01, r2);
    const int add = row == 0 ? -kNone : row == 1 ? r0 : row == 2 ? s01 : s012;
    return max(v, add);
}

__device__ __forceinline__ int wave_incl_min_rev(int v, int lane)
{
    v = min(v, __builtin_amdgcn_update_dpp(kNone, v, 0x101, 0xf, 0xf, false));  // row_shl:1
    v = min(v, __builtin_amdgcn_update_dpp(kNone, v, 0x102, 0xf, 0xf, false));  // row_shl:2
    v = min(v, __builtin_amdgcn_update_dpp(kNone, v, 0x104, 0xf, 0xf, false));  // row_shl:4
    v = min(v, __builtin_amdgcn_update_dpp(kNone, v, 0x108, 0xf, 0xf, false));  // row_shl:8
    const int r1 = __builtin_amdgcn_readlane(v, 16), r2 = __builtin_amdgcn_readlane(v, 32),
              r3 = __builtin_amdgcn_readlane(v, 48);
    const int row = lane >> 4;
    const int s23 = min(r2, r3), s123 = min(r1, s23);
    const int add = row == 0 ? s123 : row == 1 ? s23 : row == 2 ? r3 : kNone;
    return min(v, add);
}

// One wavefront per row of up to 16 * 1024 voxels.  Lane l owns voxels
// [s*1024 + 16l, s*1024 + 16l + 16) of super-chunk s: one 16-byte load and
// four 16-byte stores per lane per super-chunk when the row is aligned.
// A 16-bit solid mask per lane and super-chunk gives the in-block nearest
// solids; wave prefix-max / suffix-min scans over the lanes' last / first
// solid give the carries from other blocks.
constexpr int kSup = 16;

// The voxels the distance is measured to: zeros (the solid of a pore
// image), nonzeros (kInvert: the centres of a dilation), or, with kMatch,
// the bytes equal to `match` (a dilation from one level of a level map).
template <bool kInvert, bool kMatch>
__device__ __forceinline__ bool is_site(uint32_t v, uint32_t match)
{
    if constexpr (kMatch) return v == match;
    return kInvert ? (v != 0) : (v == 0);
}

template <bool kInvert, bool kMatch>
__device__ __forceinline__ uint32_t bytes_mask(const uint32_t w[4], uint32_t match)
{
    uint32_t m = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t v = (w[k] >> (8 * b)) & 0xffu;
            if (is_site<kInvert, kMatch>(v, match)) m |= 1u << (4 * k + b);
        }
    return m;
}

// Solid mask of row bytes e0 .. e0+15 (bits past nz stay 0).  Aligned rows
// take one 16-byte load; rows at any other byte offset read the five
// aligned dwords around the chunk and shift them into place, as long as
// those dwords lie inside the row; the ends of such rows go byte by byte.
template <bool kInvert, bool kMatch>
__device__ __forceinline__ uint32_t solid_mask16(const uint8_t *src, int e0, int nz, bool vec, uint32_t match)
{
    if (vec && e0 + 16 <= nz) {
        const uint4 q = *reinterpret_cast<const uint4 *>(src + e0);
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
        return bytes_mask<kInvert, kMatch>(w, match);
    }
    const int sh = (int)((uintptr_t)(src + e0) & 3u);
    if (!vec && e0 - sh >= 0 && e0 - sh + 20 <= nz) {
        const uint32_t *a = reinterpret_cast<const uint32_t *>(src + e0 - sh);
        const uint32_t d0 = a[0], d1 = a[1], d2 = a[2], d3 = a[3], d4 = a[4];
        const uint32_t w[4] = {__builtin_amdgcn_alignbyte(d1, d0, sh), __builtin_amdgcn_alignbyte(d2, d1, sh),
                               __builtin_amdgcn_alignbyte(d3, d2, sh), __builtin_amdgcn_alignbyte(d4, d3, sh)};
        return bytes_mask<kInvert, kMatch>(w, match);
    }
    uint32_t m = 0;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        const int e = e0 + b;
        if (e < nz) {
            const uint8_t v = src[e];
            if (is_site<kInvert, kMatch>(v, match)) m |= 1u << b;
        }
    }
    return m;
}

constexpr uint16_t kNoSolid = 0xffff;  // pass-Z output where no solid is reachable

// Pass Z: 1-D distance along z (not squared; kNoSolid where none), as
// uint16 -- exact, since a distance is at most 16384.
//
// One wavefront per row; lane l owns bytes 16l .. 16l+15 of each 1024-byte
// super-chunk: one 16-byte load (its solid mask is kept in LDS for sweep
// 2) and two 16-byte stores per lane when the row is aligned.  A 16-bit solid mask per lane gives the in-lane nearest
// solids; wave prefix-max / suffix-min scans over the lanes' last / first
// solid give the carries from other lanes, and sweep 1's per-super-chunk
// summaries the carries from other super-chunks.
// the deferred-line bitmaps of passes Y and X start each EDT empty; the
// short-row pass Z clears them on the way (small domains, where a launch
// is a large share of the EDT); long rows keep a separate clear (folded
// into the long-row kernel it cost more than the launch at 1024^3)
__device__ __forceinline__ void clear_bits(uint32_t *bits, int64_t n, int64_t blk, int64_t nblk)
{
    const int64_t base = blk * blockDim.x;
    if (base >= n) return;  // block-uniform: most blocks of a large grid skip at once
    for (int64_t i = base + threadIdx.x; i < n; i += nblk * blockDim.x) bits[i] = 0u;
}

// A lane's 16 distances (row positions e0 .. e0 + 15) into the row at dst:
// uint16 (kNoSolid where no solid is reachable), or uint8 clamped at 255 for a
// clamped EDT under a cap below 256, which reads every distance above 15 as
// beyond the cap.  uint16 rows are 16-byte aligned, uint8 rows 8-byte.
template <class D>
__device__ __forceinline__ void store_dist16(D *__restrict__ dst, int e0, int nz, const uint32_t (&v)[16])
{
    if constexpr (std::is_same<D, uint16_t>::value) {
        uint32_t dv[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) dv[b] = v[2 * b] | (v[2 * b + 1] << 16);
        if (e0 + 16 <= nz) {
            uint4 *q = reinterpret_cast<uint4 *>(dst + e0);
            q[0] = make_uint4(dv[0], dv[1], dv[2], dv[3]);
            q[1] = make_uint4(dv[4], dv[5], dv[6], dv[7]);
        } else {
#pragma unroll
            for (int b = 0; b < 16; ++b)
                if (e0 + b < nz) dst[e0 + b] = (uint16_t)v[b];
        }
    } else {
        uint32_t dv[4];
#pragma unroll
        for (int b = 0; b < 4; ++b)
            dv[b] = min(v[4 * b], 255u) | (min(v[4 * b + 1], 255u) << 8) | (min(v[4 * b + 2], 255u) << 16) |
                    (min(v[4 * b + 3], 255u) << 24);
        if (e0 + 16 <= nz) {
            uint2 *q = reinterpret_cast<uint2 *>(dst + e0);
            q[0] = make_uint2(dv[0], dv[1]);
            q[1] = make_uint2(dv[2], dv[3]);
        } else {
#pragma unroll
            for (int b = 0; b < 16; ++b)
                if (e0 + b < nz) dst[e0 + b] = (uint8_t)min(v[b], 255u);
        }
    }
}

template <bool kInvert, bool kMatch, class D = uint16_t>
__device__ __forceinline__ void pass_z_body(int64_t blk, const uint8_t *__restrict__ img, int64_t in_sx,
                                            int64_t in_sy, int ny, int nz, int64_t rows, int bc_lo, int bc_hi,
                                            int vec_in, uint32_t match, int64_t pitch, D *__restrict__ out,
                                            uint32_t *__restrict__ counters)
{
    __shared__ int sc[4][2][kSup];           // per wave: each super-chunk's last / first solid
    __shared__ uint32_t smask[4][kSup][kWave];  // per wave: each lane's masks (the row is read once)
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    const int64_t row = blk * (blockDim.x / kWave) + w;
    // counters: deferred-line counts of passes Y and X, then the 16-bit flags of pass Y
    // ([2] = 1: a value above 65535 was saturated; [3] = 1: none was)
    if (blk == 0 && threadIdx.x < 4) counters[threadIdx.x] = threadIdx.x == 3 ? 1u : 0u;
    if (row >= rows) return;
    const int64_t x = row / ny, y = row - x * ny;
    const uint8_t *src = img + x * in_sx + y * in_sy;
    D *dst = out + row * pitch;
    const int nsup = (nz + 1023) / 1024;
    const bool vin = vec_in != 0;

    // sweep 1: each super-chunk's solid masks, last and first solid (the
    // totals only matter across super-chunks or periodic faces)
    const bool totals = nsup > 1 || bc_lo == PMX_BC_PERIODIC;
    for (int s = 0; s < nsup; ++s) {
        const int e0 = s * 1024 + lane * 16;
        const uint32_t m = solid_mask16<kInvert, kMatch>(src, e0, nz, vin, match);
        smask[w][s][lane] = m;
        int l = -kNone, f = kNone;
        if (totals) {
            l = __builtin_amdgcn_readlane(wave_incl_max(m ? e0 + 31 - __clz(m) : -kNone, lane), 63);
            f = __builtin_amdgcn_readlane(wave_incl_min_rev(m ? e0 + __ffs(m) - 1 : kNone, lane), 0);
        }
        if (lane == 0) {
            sc[w][0][s] = l;
            sc[w][1][s] = f;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    int row_last = -kNone, row_first = kNone;
    for (int s = 0; s < nsup; ++s) {
        row_last = max(row_last, sc[w][0][s]);
        row_first = min(row_first, sc[w][1][s]);
    }
    int prev0, next0;  // the solids beyond the row's ends
    if (bc_lo == PMX_BC_PERIODIC) {
        prev0 = row_last > -kNone ? row_last - nz : -kNone;
        next0 = row_first < kNone ? row_first + nz : kNone;
    } else {
        prev0 = bc_lo == PMX_BC_WALL ? -1 : -kNone;
        next0 = bc_hi == PMX_BC_WALL ? nz : kNone;
    }
    // sweep 2: distances, super-chunk by super-chunk
    int pcar = prev0;
    for (int s = 0; s < nsup; ++s) {
        int ncar = next0;
        for (int t = nsup - 1; t > s; --t) ncar = min(ncar, sc[w][1][t]);
        const int e0 = s * 1024 + lane * 16;
        const uint32_t m = smask[w][s][lane];
        const int l = m ? e0 + 31 - __clz(m) : -kNone;
        const int f = m ? e0 + __ffs(m) - 1 : kNone;
        const int inc_l = wave_incl_max(l, lane);
        int pc = __builtin_amdgcn_update_dpp(-kNone, inc_l, 0x138, 0xf, 0xf, false);  // wave_shr:1
        pc = max(pc, pcar);
        const int inc_f = wave_incl_min_rev(f, lane);
        int nc = __builtin_amdgcn_update_dpp(kNone, inc_f, 0x130, 0xf, 0xf, false);  // wave_shl:1
        nc = min(nc, ncar);
        // running distances, no bit searches: bm[b] = -1 where element b is solid; the
        // distance to the previous solid restarts at 0 there and otherwise grows by one
        // (forward from element -1, pc < e0), likewise to the next solid (backward from
        // element 16, nc >= e0 + 16 except past a partial row's end, where nothing is
        // stored); min(.., kNoSolid) turns the "no solid on either side" sentinels
        // (about kNone) into kNoSolid, real distances being below 16 * 1024
        int bm[16], dl[16];
#pragma unroll
        for (int b = 0; b < 16; ++b) bm[b] = (int)(m << (31 - b)) >> 31;
        int cur = e0 - 1 - pc;
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            cur = ~bm[b] & (cur + 1);
            dl[b] = cur;
        }
        uint32_t dv[16];
        int curr = nc - (e0 + 16);
#pragma unroll
        for (int b = 15; b >= 0; --b) {
            curr = ~bm[b] & (curr + 1);
            dv[b] = (uint32_t)min(min(dl[b], curr), (int)kNoSolid);
        }
        store_dist16(dst, e0, nz, dv);
        pcar = max(pcar, sc[w][0][s]);
    }
}

template <bool kInvert, bool kMatch, class D = uint16_t>
__global__ __launch_bounds__(256) void k_pass_z(const uint8_t *__restrict__ img, int64_t in_sx,
                                                int64_t in_sy, int ny, int nz, int64_t rows,
                                                int bc_lo, int bc_hi, int vec_in, uint32_t match, int64_t pitch,
                                                D *__restrict__ out,
                                                uint32_t *__restrict__ counters,
                                                const unsigned int *__restrict__ proceed)
{
    if (proceed && *proceed == 0u) return;
    pass_z_body<kInvert, kMatch, D>(blockIdx.x, img, in_sx, in_sy, ny, nz, rows, bc_lo, bc_hi, vec_in, match, pitch,
                                    out, counters);
}

// Pass Z for short rows (nz <= 16 L): L lanes per row, 64 / L rows per
// wave, the scans segmented at the 16-lane DPP rows (L = 16: no cross-row
// step; L = 32: rows 1 and 3 take rows 0 and 2).  Same arithmetic as
// k_pass_z with one super-chunk.
template <int L>
__device__ __forceinline__ int seg_incl_max(int v, int lane)
{
    v = max(v, __builtin_amdgcn_update_dpp(-kNone, v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = max(v, __builtin_amdgcn_update_dpp(-kNone, v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = max(v, __builtin_amdgcn_update_dpp(-kNone, v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = max(v, __builtin_amdgcn_update_dpp(-kNone, v, 0x118, 0xf, 0xf, false));  // row_shr:8
    if constexpr (L == 32) {
        const int r0 = __builtin_amdgcn_readlane(v, 15), r2 = __builtin_amdgcn_readlane(v, 47);
        const int row = lane >> 4;
        v = max(v, row == 1 ? r0 : row == 3 ? r2 : -kNone);
    }
    return v;
}

template <int L>
__device__ __forceinline__ int seg_incl_min_rev(int v, int lane)
{
    v = min(v, __builtin_amdgcn_update_dpp(kNone, v, 0x101, 0xf, 0xf, false));  // row_shl:1
    v = min(v, __builtin_amdgcn_update_dpp(kNone, v, 0x102, 0xf, 0xf, false));  // row_shl:2
    v = min(v, __builtin_amdgcn_update_dpp(kNone, v, 0x104, 0xf, 0xf, false));  // row_shl:4
    v = min(v, __builtin_amdgcn_update_dpp(kNone, v, 0x108, 0xf, 0xf, false));  // row_shl:8
    if constexpr (L == 32) {
        const int r1 = __builtin_amdgcn_readlane(v, 16), r3 = __builtin_amdgcn_readlane(v, 48);
        const int row = lane >> 4;
        v = min(v, row == 0 ? r1 : row == 2 ? r3 : kNone);
    }
    return v;
}

// the value of lane seg * L + pos, for every lane of segment seg
template <int L>
__device__ __forceinline__ int seg_lane(int v, int seg, int pos)
{
    if constexpr (L == 16) {
        const int a = __builtin_amdgcn_readlane(v, pos), b = __builtin_amdgcn_readlane(v, 16 + pos),
                  c = __builtin_amdgcn_readlane(v, 32 + pos), d = __builtin_amdgcn_readlane(v, 48 + pos);
        return seg == 0 ? a : seg == 1 ? b : seg == 2 ? c : d;
    } else {
        const int a = __builtin_amdgcn_readlane(v, pos), b = __builtin_amdgcn_readlane(v, 32 + pos);
        return seg == 0 ? a : b;
    }
}

template <bool kInvert, bool kMatch, int L, class D = uint16_t>
__device__ __forceinline__ void pass_z_short_body(int64_t blk, int64_t nblk, const uint8_t *__restrict__ img,
                                                  int64_t in_sx, int64_t in_sy, int ny, int nz, int64_t rows,
                                                  int bc_lo, int bc_hi, int vec_in, uint32_t match, int64_t pitch,
                                                  D *__restrict__ out, uint32_t *__restrict__ counters,
                                                  uint32_t *__restrict__ bits, int64_t n_bits)
{
    clear_bits(bits, n_bits, blk, nblk);
    constexpr int R = kWave / L;  // rows per wave
    const int lane = threadIdx.x & (kWave - 1);
    const int seg = lane / L, sl = lane % L;
    const int64_t row0 = (blk * (blockDim.x / kWave) + threadIdx.x / kWave) * R;
    // counters: deferred-line counts of passes Y and X, then the 16-bit flags of pass Y
    // ([2] = 1: a value above 65535 was saturated; [3] = 1: none was)
    if (blk == 0 && threadIdx.x < 4) counters[threadIdx.x] = threadIdx.x == 3 ? 1u : 0u;
    if (row0 >= rows) return;  // wave-uniform: the scans below need every lane of the wave
    const int64_t row = row0 + seg;
    const bool valid = row < rows;
    const int64_t rr = valid ? row : row0;
    const int64_t x = rr / ny, y = rr - x * ny;
    const uint8_t *src = img + x * in_sx + y * in_sy;
    const int e0 = sl * 16;
    const uint32_t m = valid && e0 < nz ? solid_mask16<kInvert, kMatch>(src, e0, nz, vec_in != 0, match) : 0u;
    const int l = m ? e0 + 31 - __clz(m) : -kNone;
    const int f = m ? e0 + __ffs(m) - 1 : kNone;
    const int inc_l = seg_incl_max<L>(l, lane);
    const int inc_f = seg_incl_min_rev<L>(f, lane);
    int prev0, next0;  // the solids beyond the row's ends
    if (bc_lo == PMX_BC_PERIODIC) {
        const int row_last = seg_lane<L>(inc_l, seg, L - 1), row_first = seg_lane<L>(inc_f, seg, 0);
        prev0 = row_last > -kNone ? row_last - nz : -kNone;
        next0 = row_first < kNone ? row_first + nz : kNone;
    } else {
        prev0 = bc_lo == PMX_BC_WALL ? -1 : -kNone;
        next0 = bc_hi == PMX_BC_WALL ? nz : kNone;
    }
    int pc = __builtin_amdgcn_update_dpp(-kNone, inc_l, 0x138, 0xf, 0xf, false);  // wave_shr:1
    if (sl == 0) pc = -kNone;
    pc = max(pc, prev0);
    int nc = __builtin_amdgcn_update_dpp(kNone, inc_f, 0x130, 0xf, 0xf, false);  // wave_shl:1
    if (sl == L - 1) nc = kNone;
    nc = min(nc, next0);
    // running distances as in k_pass_z
    int bm[16], dl[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) bm[b] = (int)(m << (31 - b)) >> 31;
    int cur = e0 - 1 - pc;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
        cur = ~bm[b] & (cur + 1);
        dl[b] = cur;
    }
    uint32_t dv[16];
    int curr = nc - (e0 + 16);
#pragma unroll
    for (int b = 15; b >= 0; --b) {
        curr = ~bm[b] & (curr + 1);
        dv[b] = (uint32_t)min(min(dl[b], curr), (int)kNoSolid);
    }
    if (!valid || e0 >= nz) return;
    store_dist16(out + row * pitch, e0, nz, dv);
}

template <bool kInvert, bool kMatch, int L, class D = uint16_t>
__global__ __launch_bounds__(256) void k_pass_z_short(const uint8_t *__restrict__ img, int64_t in_sx,
                                                      int64_t in_sy, int ny, int nz, int64_t rows, int bc_lo,
                                                      int bc_hi, int vec_in, uint32_t match, int64_t pitch,
                                                      D *__restrict__ out, uint32_t *__restrict__ counters,
                                                      const unsigned int *__restrict__ proceed,
                                                      uint32_t *__restrict__ bits, int64_t n_bits)
{
    if (proceed && *proceed == 0u) return;
    pass_z_short_body<kInvert, kMatch, L, D>(blockIdx.x, gridDim.x, img, in_sx, in_sy, ny, nz, rows, bc_lo, bc_hi,
                                             vec_in, match, pitch, out, counters, bits, n_bits);
}

// ---------------------------------------------------------------------------
// passes Y and X: lower envelope of parabolas along strided lines
// ---------------------------------------------------------------------------

struct Stack {
    int *v, *h;
    int64_t step;
    int cap;
    int *buf;  // kChunk staging slots of this lane (stride `step`)
    __device__ __forceinline__ int &V(int k) const { return v[k * step]; }
    __device__ __forceinline__ int &H(int k) const { return h[k * step]; }
};

// Streams positions p0..p1 (unrolled coordinates; the value at p is
// load(p mod n)) through body(p, f) in chunks of kChunk: the next chunk's
// loads are issued into registers before the current chunk, parked in the
// lane's LDS slots `buf` (stride `bstep`), is processed, so every lane
// keeps kChunk loads in flight.  Stops early if body returns false.
constexpr int kChunk = 16;

template <class Load, class Body>
__device__ __forceinline__ bool stream_line(int p0, int p1, int n, Load load, Body body, int *buf,
                                            int64_t bstep)
{
    auto wrap = [n](int p) { return p < 0 ? p + n : (p >= n ? p - n : p); };
    int nxt[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) nxt[j] = (p0 + j <= p1) ? load(wrap(p0 + j)) : 0;
    for (int base = p0; base <= p1; base += kChunk) {
#pragma unroll
        for (int j = 0; j < kChunk; ++j) buf[j * bstep] = nxt[j];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            const int q = base + kChunk + j;
            nxt[j] = (q <= p1) ? load(wrap(q)) : 0;
        }
        const int m = min(kChunk, p1 - base + 1);
        for (int j = 0; j < m; ++j)
            if (!body(base + j, buf[j * bstep])) return false;
    }
    return true;
}

// Envelope of one line.  `load(i)` returns the input at position i in
// [0, n), `emit(i, d2)` receives the output at position i.  Returns false
// if the stack would overflow (the caller then defers the line).
template <class Load, class Emit>
__device__ bool envelope_line(int n, int bc_lo, int bc_hi, const Stack &st, bool allow_triple,
                              Load load, Emit emit)
{
    int top = -1;
    // push (r, hr), popping members whose range became empty; false on overflow
    auto push = [&](int r, int hr) -> bool {
        while (top >= 1) {
            const int t = st.V(top), ht = st.H(top);
            const int u = st.V(top - 1), hu = st.H(top - 1);
            const int64_t num_new = (int64_t)hr - ht + (int64_t)r * r - (int64_t)t * t;
            const int64_t num_top = (int64_t)ht - hu + (int64_t)t * t - (int64_t)u * u;
            if (num_new * (int64_t)(t - u) <= num_top * (int64_t)(r - t))
                --top;
            else
                break;
        }
        if (top + 1 >= st.cap) return false;
        ++top;
        st.V(top) = r;
        st.H(top) = hr;
        return true;
    };
    // evaluate positions lo..hi (unrolled coordinates) from the stack
    auto flush = [&](int lo, int hi) {
        if (top < 0) {
            for (int x = lo; x <= hi; ++x) emit(x % n, kInf);
            return;
        }
        int j = 0;
        int vj = st.V(0), hj = st.H(0);
        for (int x = lo; x <= hi; ++x) {
            int dv = x - vj;
            int best = dv * dv + hj;
            while (j < top) {
                const int vn = st.V(j + 1), hn = st.H(j + 1);
                const int dn = x - vn;
                const int cand = dn * dn + hn;
                if (cand > best) break;
                ++j;
                vj = vn;
                hj = hn;
                best = cand;
            }
            emit(x >= n ? x - n : x, best);
        }
    };
    int done = -1;
    // a parabola per finite value; at a solid evaluate everything behind it
    // and restart the stack from that solid
    auto body = [&](int p, int f) -> bool {
        if (f == kInf) return true;
        if (!push(p, f)) return false;
        if (f == 0) {
            flush(done + 1, p);
            done = p;
            top = 0;
            st.V(0) = p;
            st.H(0) = 0;
        }
        return true;
    };

    if (bc_lo != PMX_BC_PERIODIC) {
        if (bc_lo == PMX_BC_WALL) push(-1, 0);
        if (!stream_line(0, n - 1, n, load, body, st.buf, st.step)) return false;
        if (bc_hi == PMX_BC_WALL && !push(n, 0)) return false;
        if (done < n - 1) flush(done + 1, n - 1);
        return true;
    }

    // periodic: rotate the line to start at its first solid b and run one
    // full period to the same solid at b + n
    int b = -1;
    stream_line(0, n - 1, n, load, [&](int p, int f) -> bool {
        if (f == 0) {
            b = p;
            return false;
        }
        return true;
    }, st.buf, st.step);
    if (b >= 0) {
        top = 0;
        st.V(0) = b;
        st.H(0) = 0;
        done = b - 1;
        return stream_line(b + 1, b + n, n, load, body, st.buf, st.step);
    }
    if (!allow_triple) return false;
    // no solid on the line: parabolas of the previous, own and next period
    if (!stream_line(-n, 2 * n - 1, n, load, [&](int p, int f) -> bool {
            return f == kInf ? true : push(p, f);
        }, st.buf, st.step))
        return false;
    flush(0, n - 1);
    return true;
}

// The first-invasion update alone (the porosimetry dilation's pass X: no
// other output): a covered pore voxel not yet invaded (state 1) takes the
// level's code (porosimetry.py:220-225).
__device__ __forceinline__ void psd_update(const PassXOut &o, int64_t x, int64_t y, int64_t z, int v)
{
    if (v > o.t) return;
    const int64_t si = x * o.ssx + y * o.ssy + z;
    if (o.psd_state[si] == 1) {
        o.psd_state[si] = 2;
        o.psd_map[x * o.sx + y * o.sy + z] = o.psd_code;
    }
}

__device__ __forceinline__ void emit_x(const PassXOut &o, int64_t x, int64_t y, int64_t z,
                                       int nx, int ny, int nz, int v)
{
    const int64_t idx = x * o.sx + y * o.sy + z;
    if (o.d2) o.d2[idx] = v;
    if (o.f32sq) o.f32sq[idx] = v == kInf ? 3.40282347e+38f : (float)v;
    if (o.f32) o.f32[idx] = v == kInf ? sqrtf(3.40282347e+38f) : sqrtf((float)v);
    if (o.f64) o.f64[idx] = v == kInf ? sqrt((double)3.40282347e+38f) : sqrt((double)v);
    if (o.psd_map && v <= o.t) {
        // first invasion (porosimetry.py:220-225): covered pore voxels not yet assigned; one
        // state byte says both (1: pore, not invaded), so a covered voxel costs one byte read
        const int64_t si = x * o.ssx + y * o.ssy + z;
        if (o.psd_state[si] == 1) {
            o.psd_state[si] = 2;
            o.psd_map[idx] = o.psd_code;
        }
    }
    if (o.u8) {
        uint8_t r = v <= o.t ? o.v_le : o.v_gt;
        const uint32_t w = o.wall_mask;
        if (w) {
            if (((w & 1u) && x == 0) || ((w & 2u) && x == nx - 1) || ((w & 4u) && y == 0) ||
                ((w & 8u) && y == ny - 1) || ((w & 16u) && z == 0) || ((w & 32u) && z == nz - 1))
                r = 0;
        }
        o.u8[idx] = r;
    }
}


// ---------------------------------------------------------------------------
// passes Y and X: windowed minimum over an LDS tile
// ---------------------------------------------------------------------------
//
// d(y) = min_k f(y + k) + k^2.  Scanning k = 1, 2, ... outward from y can
// stop as soon as k^2 >= best: no farther candidate can win.  A solid
// (f = 0) stops the scan at its distance, so the work per voxel is about
// its distance to the nearest solid (a few voxels in a porous medium).
// A workgroup owns a strip of 64 z-columns of one line bundle and walks
// up to kTilesPerBlock row tiles of T rows (fewer on small boxes, so that a
// launch still has enough blocks to fill the chip), each staged in LDS with W halo rows
// on each side; one lane per z-column.  The next tile's loads are issued
// before the current tile is scanned.  Voxels not settled within W
// continue the scan in global memory up to kReach rows; a voxel still
// open after that flags its line for the envelope fallback.
//
// Values are clamped at `cap` on load.  Clamping commutes with the
// passes: min(d, cap) of the whole EDT equals the EDT of the clamped
// input, so a caller that only compares against a threshold t passes
// cap = t + 1 and no scan ever runs past sqrt(t + 1).

constexpr int kBig = 1 << 30;   // "no solid" inside the windowed passes
constexpr int kTileZ = 64;
constexpr int kWinThreads = 256;
constexpr int kReach = 128;     // farthest row offset scanned before a line is deferred
constexpr int kRowWall = -1, kRowNone = -2;
constexpr int kTilesPerBlock = 8;
constexpr int64_t kFillBlocks = 1024;  // fewer tiles per block below this many blocks per launch

// row r of a line of n rows under the boundary rule: an in-range row,
// kRowWall (the solid layer just beyond a WALL face) or kRowNone
__device__ __forceinline__ int map_row(int r, int n, int bc_lo, int bc_hi)
{
    if (r >= 0 && r < n) return r;
    if (bc_lo == PMX_BC_PERIODIC) {
        r %= n;
        return r < 0 ? r + n : r;
    }
    if ((r == -1 && bc_lo == PMX_BC_WALL) || (r == n && bc_hi == PMX_BC_WALL)) return kRowWall;
    return kRowNone;
}


struct WinGeom {
    int n;          // rows per line (the pass axis)
    int64_t rs;     // input element stride between rows
    int64_t os;     // input element stride between outer indices (x in pass Y, y in pass X)
    int nouter, nz;
    int bc_lo, bc_hi;
    int cap;        // values are clamped to [0, cap], cap <= kBig
    int64_t pitch;  // input row length in memory: a multiple of 8 >= nz, 16-byte aligned rows
    int64_t out_os, out_rs;  // output strides of the outer index and of a row (pass Y)
    int tpb = kTilesPerBlock;  // row tiles per block (a power of two <= kTilesPerBlock)
    // truncated periodic wrap (the windowed kernels' kHull form): per line,
    // the wrapped rows that may serve as candidates (see hull_row_kept); a
    // wrapped row r < 0 reads row r + wrap_lo, a row r >= n row r - wrap_hi
    // (both n for a whole periodic line; less on a padded subdomain, whose
    // period skips the pads)
    const uint32_t *hmask = nullptr;
    int wrap_lo = 0, wrap_hi = 0;
};

// Truncated wrap of a periodic line: a row r < 0 or r >= n is a candidate
// only if hmask marks it.  Words 0..3 of a line's eight mark the rows -1 - b
// (bit b), words 4..7 the rows n + b, b < kHullMaskRows; rows farther out are
// never candidates (a scan stops at kReach rows, so none is read).
constexpr int kHullMaskRows = 128;
constexpr int kHullMaskWords = 8;

// map_row, or (kHull) the row a wrapped row reads under the truncated wrap
template <bool kHull>
__device__ __forceinline__ int map_row_g(int r, const WinGeom &g)
{
    if (kHull && (r < 0 || r >= g.n)) {
        const int m = r < 0 ? r + g.wrap_lo : r - g.wrap_hi;
        return m >= 0 && m < g.n ? m : kRowNone;
    }
    return map_row(r, g.n, g.bc_lo, g.bc_hi);
}
__device__ __forceinline__ bool hull_row_kept(const uint32_t *__restrict__ m, int64_t line, int r, int n)
{
    const int b = r < 0 ? -1 - r : r - n;
    if (b >= kHullMaskRows) return false;
    const uint32_t w = m[line * kHullMaskWords + (r < 0 ? 0 : 4) + (b >> 5)];
    return (w >> (b & 31)) & 1u;
}

// blocks of a windowed pass with T-row tiles and tpb tiles per block
__host__ __device__ __forceinline__ int64_t window_blocks(const WinGeom &g, int T, int tpb)
{
    const int nzt = (g.nz + kTileZ - 1) / kTileZ, nrt = (g.n + T - 1) / T;
    return (int64_t)g.nouter * nzt * ((nrt + tpb - 1) / tpb);
}

// tiles per block: halved from kTilesPerBlock while the launch has fewer
// than kFillBlocks blocks (small boxes; each halving re-reads 2W halo rows
// per extra block, which matters less there than idle compute units)
inline int tiles_per_block(const WinGeom &g, int T)
{
    int tpb = kTilesPerBlock;
    while (tpb > 1 && window_blocks(g, T, tpb) < kFillBlocks) tpb /= 2;
    return tpb;
}

// A squared distance already clamped to a cap <= 65535, stored in 16 bits:
// the pass-Y output of a clamped EDT (morphology, porosimetry), which
// halves the traffic between passes Y and X.
struct Sq16 {
    uint16_t v;
};

// The same clamped to a cap <= 255, in 8 bits: the pass-Y output of a
// dilation of radius below 16 (most levels of a pore-size sweep), which
// halves that traffic again.
struct Sq8 {
    uint8_t v;
};
// Pass Z's distance in 8 bits (clamped at 255) for a clamped EDT under a cap
// below 256: any distance above 15 squares past the cap.
struct Dz8 {
    uint8_t v;
};
template <class T> struct IsSq : std::false_type {};
template <> struct IsSq<Sq16> : std::true_type {};
template <> struct IsSq<Sq8> : std::true_type {};

// two 16-bit values in one register (v_pk_* integer operations)
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// Input element -> clamped squared distance: pass Y reads pass Z's uint16
// distances (kNoSolid: none), pass X the squared distances of pass Y (int32,
// or Sq16 when the cap fits in 16 bits).
__device__ __forceinline__ int decode(uint16_t d, int cap)
{
    // kNoSolid^2 = 0xfffe0001 fits in 32 bits and exceeds every cap (<= kBig), so one
    // unsigned minimum also maps "no solid" to cap
    return (int)min((uint32_t)d * (uint32_t)d, (uint32_t)cap);
}
__device__ __forceinline__ int decode(int32_t v, int cap) { return min(v, cap); }
__device__ __forceinline__ int decode(Sq16 v, int cap) { return min((int)v.v, cap); }
__device__ __forceinline__ int decode(Sq8 v, int cap) { return min((int)v.v, cap); }
__device__ __forceinline__ int decode(Dz8 d, int cap) { return min((int)d.v * (int)d.v, cap); }

// input at row r of the line starting at `line` (number line_id), under the
// boundary rule; kHull: wrapped rows only where the line's mask keeps them
template <bool kHull, class E>
__device__ __forceinline__ int fetch(const E *line, const WinGeom &g, int r, int64_t line_id)
{
    if (kHull && (r < 0 || r >= g.n) && !hull_row_kept(g.hmask, line_id, r, g.n)) return g.cap;
    const int m = map_row_g<kHull>(r, g);
    if (m >= 0) return decode(line[(int64_t)m * g.rs], g.cap);
    return m == kRowWall ? 0 : g.cap;
}

// Staging of tile rows: thread (q, row0) = (t % 16, t / 16) owns columns
// 4q .. 4q+3 of rows row0 + 16 i, loaded as one 8-byte (uint16 input) or
// 16-byte (int32 input) vector.  Loads are issued (addresses clamped into
// the line) well before their commit.
template <class E> struct Quad;
template <> struct Quad<uint16_t> { using type = uint2; };
template <> struct Quad<int32_t> { using type = int4; };
template <> struct Quad<Sq16> { using type = uint2; };
template <> struct Quad<Sq8> { using type = uint32_t; };
template <> struct Quad<Dz8> { using type = uint32_t; };

template <int N, class E, bool kHull = false>
__device__ __forceinline__ void issue_rows(const E *__restrict__ base, const WinGeom &g, int first, int row0,
                                           int c4, int (&mrow)[N], typename Quad<E>::type (&raw)[N])
{
    using V = typename Quad<E>::type;
#pragma unroll
    for (int i = 0; i < N; ++i) mrow[i] = map_row_g<kHull>(first + row0 + 16 * i, g);
#pragma unroll
    for (int i = 0; i < N; ++i) raw[i] = *reinterpret_cast<const V *>(base + (int64_t)max(mrow[i], 0) * g.rs + c4);
}

// four loaded elements -> four clamped squared distances
template <class E> __device__ __forceinline__ int4 unpack(typename Quad<E>::type r, int cap);
template <> __device__ __forceinline__ int4 unpack<uint16_t>(uint2 r, int cap)
{
    return make_int4(decode((uint16_t)(r.x & 0xffffu), cap), decode((uint16_t)(r.x >> 16), cap),
                     decode((uint16_t)(r.y & 0xffffu), cap), decode((uint16_t)(r.y >> 16), cap));
}
template <> __device__ __forceinline__ int4 unpack<Sq16>(uint2 r, int cap)
{
    return make_int4(min((int)(r.x & 0xffffu), cap), min((int)(r.x >> 16), cap), min((int)(r.y & 0xffffu), cap),
                     min((int)(r.y >> 16), cap));
}
template <> __device__ __forceinline__ int4 unpack<Dz8>(uint32_t r, int cap)
{
    return make_int4(decode(Dz8{(uint8_t)r}, cap), decode(Dz8{(uint8_t)(r >> 8)}, cap),
                     decode(Dz8{(uint8_t)(r >> 16)}, cap), decode(Dz8{(uint8_t)(r >> 24)}, cap));
}
template <> __device__ __forceinline__ int4 unpack<Sq8>(uint32_t r, int cap)
{
    return make_int4(min((int)(r & 0xffu), cap), min((int)((r >> 8) & 0xffu), cap), min((int)((r >> 16) & 0xffu), cap),
                     min((int)(r >> 24), cap));
}
template <> __device__ __forceinline__ int4 unpack<int32_t>(int4 r, int cap)
{
    return make_int4(min(r.x, cap), min(r.y, cap), min(r.z, cap), min(r.w, cap));
}

// The LDS tile holds int32 values, or uint16 ones (half the LDS, so more
// blocks per CU) when every value the pass needs is at most 65535: a pass X
// over 16-bit input, and a pass Y with 16-bit output.  A uint16 tile
// saturates at 65535.  For pass Y that leaves its 16-bit output as it was:
// every candidate s(f) + k^2 lies between min(f + k^2, 65535) and f + k^2, so
// the windowed minimum equals the true one below 65535 and is >= 65535
// (saturated and flagged) otherwise.
template <class L> struct TileIO;
template <> struct TileIO<int> {
    static __device__ __forceinline__ void put4(int *p, int4 v) { *reinterpret_cast<int4 *>(p) = v; }
    static __device__ __forceinline__ int4 get4(const int *p) { return *reinterpret_cast<const int4 *>(p); }
};
template <> struct TileIO<uint16_t> {
    static __device__ __forceinline__ void put4(uint16_t *p, int4 v)
    {
        const uint32_t a = min((uint32_t)v.x, 65535u) | (min((uint32_t)v.y, 65535u) << 16);
        const uint32_t b = min((uint32_t)v.z, 65535u) | (min((uint32_t)v.w, 65535u) << 16);
        *reinterpret_cast<uint2 *>(p) = make_uint2(a, b);
    }
    static __device__ __forceinline__ int4 get4(const uint16_t *p)
    {
        const uint2 r = *reinterpret_cast<const uint2 *>(p);
        return make_int4((int)(r.x & 0xffffu), (int)(r.x >> 16), (int)(r.y & 0xffffu), (int)(r.y >> 16));
    }
};

// clamp (columns past nz hold 0, settled at once; rows beyond the line the
// boundary value) and store at tile rows dst + row0 + 16 i; acc keeps the
// per-column minimum.  first: the line row of the first staged row (as
// issue_rows takes it); line0: the line number of this thread's first
// column (kHull: wrapped rows go through the line masks).
template <bool kHull, class E, int N, class V, class L>
__device__ __forceinline__ void commit_rows(L *tile, const WinGeom &g, int dst, int row0, int q, int zl,
                                            const int (&mrow)[N], const V (&raw)[N], int4 &acc, bool inside,
                                            int first, int64_t line0)
{
    // inside (wave-uniform): every staged row lies in the line; with a full quad of columns
    // the values go to the tile as loaded
    if (inside && zl >= 3) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const int4 v = unpack<E>(raw[i], g.cap);
            acc.x = min(acc.x, v.x);
            acc.y = min(acc.y, v.y);
            acc.z = min(acc.z, v.z);
            acc.w = min(acc.w, v.w);
            TileIO<L>::put4(tile + (dst + row0 + 16 * i) * kTileZ + 4 * q, v);
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool in_line = mrow[i] >= 0;
        const int c = mrow[i] == kRowWall ? 0 : g.cap;
        int4 u = unpack<E>(raw[i], g.cap);
        if constexpr (kHull) {
            const int r = first + row0 + 16 * i;
            if (r < 0 || r >= g.n) {
                if (zl >= 0 && !hull_row_kept(g.hmask, line0 + 0, r, g.n)) u.x = g.cap;
                if (zl >= 1 && !hull_row_kept(g.hmask, line0 + 1, r, g.n)) u.y = g.cap;
                if (zl >= 2 && !hull_row_kept(g.hmask, line0 + 2, r, g.n)) u.z = g.cap;
                if (zl >= 3 && !hull_row_kept(g.hmask, line0 + 3, r, g.n)) u.w = g.cap;
            }
        }
        int4 v;
        v.x = in_line ? (zl >= 0 ? u.x : 0) : c;
        v.y = in_line ? (zl >= 1 ? u.y : 0) : c;
        v.z = in_line ? (zl >= 2 ? u.z : 0) : c;
        v.w = in_line ? (zl >= 3 ? u.w : 0) : c;
        acc.x = min(acc.x, v.x);
        acc.y = min(acc.y, v.y);
        acc.z = min(acc.z, v.z);
        acc.w = min(acc.w, v.w);
        TileIO<L>::put4(tile + (dst + row0 + 16 * i) * kTileZ + 4 * q, v);
    }
}

// E = uint16_t or Dz8: pass Y (input: pass Z's distances, output: the array `out`
// of O = int32_t, or Sq16 under a cap <= 65535); E = int32_t or Sq16: pass X
// (input: that array, output: PassXOut).
//
// The tile is a sliding window: after tile i is scanned its last 2W rows
// move to the top and only the T new rows are loaded (issued while tile i
// is scanned), so each row is loaded once per block.  D2ONLY (pass X):
// the int32 squared distance is the only output, written without the
// per-row tests of the other outputs.
template <int T, int W, class E, class O, bool D2ONLY, bool kHull = false, bool kPsd = false>
__device__ __forceinline__ void window_body(int64_t t, const E *__restrict__ in, O *__restrict__ out,
                                            const PassXOut &o, const WinGeom &g, int nx, int ny,
                                            uint32_t *__restrict__ line_bits, uint32_t *__restrict__ ovf_count,
                                            uint32_t *__restrict__ ovf_list)
{
    constexpr bool kY = std::is_same<E, uint16_t>::value || std::is_same<E, Dz8>::value;
    using V = typename Quad<E>::type;
    static_assert(W % 8 == 0 && T % 16 == 0, "tile layout");
    constexpr int kRows = T + 2 * W;
    constexpr int kWaves = kWinThreads / kWave;
    static_assert(kWaves == 4 && kWinThreads == 256, "four waves, 16 threads per row of four columns");
    constexpr int kHalo = 2 * W / 16;  // rows of four columns per thread in the halo move
    constexpr bool kNarrow = kY ? IsSq<O>::value : IsSq<E>::value;
    using L = typename std::conditional<kNarrow, uint16_t, int>::type;
    __shared__ __align__(16) L tile[kRows * kTileZ];
    __shared__ __align__(16) int part[kWaves][kTileZ];

    const int nzt = (g.nz + kTileZ - 1) / kTileZ, nrt = (g.n + T - 1) / T;
    const int nchunk = (nrt + g.tpb - 1) / g.tpb;
    const int chunk = (int)(t % nchunk);
    const int64_t t2 = t / nchunk;
    const int zt = (int)(t2 % nzt);
    const int outer = (int)(t2 / nzt);
    const int rt_end = min(nrt, (chunk + 1) * g.tpb);
    const int z0 = zt * kTileZ;

    // the wave index is uniform: keep it (and every row index derived from
    // it) in scalar registers, so per-row address arithmetic runs on the
    // scalar unit
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
    const int z = z0 + lane;
    const L *col = tile + lane;
    const int64_t line = (int64_t)outer * g.nz + z;
    const int q = threadIdx.x % (kTileZ / 4), row0 = threadIdx.x / (kTileZ / 4);
    const int zl = g.nz - 1 - z0 - 4 * q;  // valid columns of this thread's quad: 0 .. zl
    // rows are padded to a multiple of 8 columns: a quad load at a clamped
    // column stays inside the row (columns >= nz are dropped)
    const int c4 = min(4 * q, (int)(g.pitch - 4 - z0));
    const E *base = in + (int64_t)outer * g.os + z0;
    // pass X with the int32 output only: two neighbouring columns go out as one 8-byte store
    // when every pair starts 8-byte aligned (even strides, aligned array)
    const bool d2_pairs = D2ONLY && o.d2 && ((o.sx | o.sy) & 1) == 0 && ((uintptr_t)o.d2 & 7) == 0;

    int4 acc = make_int4(kBig, kBig, kBig, kBig);
    int r0 = chunk * g.tpb * T;
    int xm[T / 16];
    V xr[T / 16];
    {
        // the first tile: all kRows rows
        int fm[kRows / 16];
        V fr[kRows / 16];
        issue_rows<kRows / 16, E, kHull>(base, g, r0 - W, row0, c4, fm, fr);
        commit_rows<kHull, E, kRows / 16>(tile, g, 0, row0, q, zl, fm, fr, acc, r0 - W >= 0 && r0 - W + kRows <= g.n,
                                          r0 - W, (int64_t)outer * g.nz + z0 + 4 * q);
    }
    for (int rt = chunk * g.tpb; rt < rt_end; ++rt, r0 += T) {
#pragma unroll
        for (int s = 16; s < kWave; s <<= 1) {
            acc.x = min(acc.x, __shfl_xor(acc.x, s, kWave));
            acc.y = min(acc.y, __shfl_xor(acc.y, s, kWave));
            acc.z = min(acc.z, __shfl_xor(acc.z, s, kWave));
            acc.w = min(acc.w, __shfl_xor(acc.w, s, kWave));
        }
        if (lane < 16) *reinterpret_cast<int4 *>(&part[wv][4 * lane]) = acc;
        __syncthreads();
        // every tile row of a column is >= its minimum cm: a scan can stop once cm + k^2 >= best
        int cm = kBig, cm_lo = kBig, cm_hi = kBig;
        if constexpr (kNarrow) {
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                cm_lo = min(cm_lo, part[w][2 * (lane & 31)]);
                cm_hi = min(cm_hi, part[w][2 * (lane & 31) + 1]);
            }
        } else {
            cm = part[0][lane];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) cm = min(cm, part[w][lane]);
        }
        const bool more = rt + 1 < rt_end;
        if (more) issue_rows<T / 16, E, kHull>(base, g, r0 + T + W, row0, c4, xm, xr);

        const int jend = min(T, g.n - r0);  // wave-uniform: the row loop needs no exec mask
        // the rest of a row once its tile scan is done (wave-uniform calls: it holds a ballot)
        auto finish_row = [&](int j, int best) {
            const int y = r0 + j;
            // the tile is done; rows beyond it are at least W + 1 away
            int k = W + 1;
            bool ok = k * k >= best;
            // rare: one ballot keeps the common case free of exec masks
            if (__builtin_amdgcn_ballot_w64(!ok && z < g.nz) != 0 && !ok && z < g.nz) {
                // beyond the tile: scan the line in global memory, unless the
                // line is already deferred
                const uint32_t bit = 1u << (line & 31);
                if (!(line_bits[line >> 5] & bit)) {
                    const E *ln = in + (int64_t)outer * g.os + z;
                    // kHull: candidates lie in [-n, 2n), no shorter way round
                    const int kmax = kHull ? 2 * g.n : g.bc_lo == PMX_BC_PERIODIC ? g.n / 2 : max(y + 1, g.n - y);
                    for (; k <= kReach && k <= kmax; ++k) {
                        if (k * k >= best) break;
                        best = min(best, min(fetch<kHull>(ln, g, y - k, line), fetch<kHull>(ln, g, y + k, line)) + k * k);
                    }
                    ok = k * k >= best || k > kmax;
                }
                if (!ok && !(atomicOr(line_bits + (line >> 5), bit) & bit))
                    ovf_list[atomicAdd(ovf_count, 1u)] = (uint32_t)line;
            }
            if (!ok || z >= g.nz) return;
            if constexpr (kY) {
                const int32_t v = best >= kBig ? kInf : best;
                if constexpr (std::is_same<O, Sq16>::value) {
                    // 16-bit output saturates; a saturated value sends the EDT to its int32 pass
                    // Y and pass X (the flags after the two deferred-line counts)
                    if (v >= 65535) {  // >=: with a uint16 tile a larger true value reads 65535
                        ovf_count[2] = 1u;
                        ovf_count[3] = 0u;
                    }
                    out[(int64_t)outer * g.out_os + (int64_t)y * g.out_rs + z] = Sq16{(uint16_t)min(v, 65535)};
                } else if constexpr (std::is_same<O, Sq8>::value) {
                    out[(int64_t)outer * g.out_os + (int64_t)y * g.out_rs + z] = Sq8{(uint8_t)min(v, 255)};
                } else {
                    out[(int64_t)outer * g.out_os + (int64_t)y * g.out_rs + z] = v;
                }
            } else if constexpr (D2ONLY) {
                o.d2[(int64_t)y * o.sx + (int64_t)outer * o.sy + z] = best >= kBig ? kInf : best;
            } else if constexpr (kPsd) {
                psd_update(o, y, outer, z, best);
            } else {
                emit_x(o, y, outer, z, nx, ny, g.nz, best >= kBig ? kInf : best);
            }
        };
        // The rest of a lane's two columns once their tile scan is done (wave-uniform calls;
        // y: the row of this lane's half).  The common case -- every column settled within
        // the tile -- costs one ballot; the global-memory scan and the deferral are the rare
        // path, column by column.
        auto finish_pair = [&](int y, int zp, bool act, u16x2 best) {
            int bv0 = best.x, bv1 = best.y;
            bool ok0 = true, ok1 = true;
            constexpr int kTileReach = (W + 1) * (W + 1);
            if (__builtin_amdgcn_ballot_w64(act && max(bv0, bv1) > kTileReach) != 0) {
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    int &bv = c ? bv1 : bv0;
                    bool &ok = c ? ok1 : ok0;
                    const int zc = zp + c;
                    if (!act || zc >= g.nz || bv <= kTileReach) continue;
                    const int64_t ln_id = (int64_t)outer * g.nz + zc;
                    const uint32_t bit = 1u << (ln_id & 31);
                    int k = W + 1;
                    ok = false;
                    if (!(line_bits[ln_id >> 5] & bit)) {
                        const E *ln = in + (int64_t)outer * g.os + zc;
                        const int kmax = kHull ? 2 * g.n : g.bc_lo == PMX_BC_PERIODIC ? g.n / 2 : max(y + 1, g.n - y);
                        for (; k <= kReach && k <= kmax; ++k) {
                            if (k * k >= bv) break;
                            bv = min(bv, min(fetch<kHull>(ln, g, y - k, ln_id), fetch<kHull>(ln, g, y + k, ln_id)) + k * k);
                        }
                        ok = k * k >= bv || k > kmax;
                    }
                    if (!ok && !(atomicOr(line_bits + (ln_id >> 5), bit) & bit))
                        ovf_list[atomicAdd(ovf_count, 1u)] = (uint32_t)ln_id;
                }
            }
            if constexpr (kY && std::is_same<O, Sq8>::value) {
                // 8-bit output of both columns (values are at most the cap, below 256)
                if (act && zp < g.pitch)
                    *reinterpret_cast<uint16_t *>(out + (int64_t)outer * g.out_os + (int64_t)y * g.out_rs + zp) =
                        (uint16_t)(min(bv0, 255) | (min(bv1, 255) << 8));
            } else if constexpr (kY) {
                // 16-bit output of both columns in one store (a deferred column's value is
                // rewritten by the fallback; columns past nz land in the row padding); a
                // settled value of 65535 sends the EDT to its int32 passes
                const bool sat = act && ((ok0 && bv0 >= 65535 && zp < g.nz) || (ok1 && bv1 >= 65535 && zp + 1 < g.nz));
                if (__builtin_amdgcn_ballot_w64(sat) != 0 && lane == 0) {
                    ovf_count[2] = 1u;
                    ovf_count[3] = 0u;
                }
                if (act && zp < g.pitch)
                    *reinterpret_cast<uint32_t *>(out + (int64_t)outer * g.out_os + (int64_t)y * g.out_rs + zp) =
                        (uint32_t)min(bv0, 65535) | ((uint32_t)min(bv1, 65535) << 16);
            } else if constexpr (D2ONLY) {
                // every value of this path is below 65535: no "none" marker can occur
                int32_t *d = o.d2 + (int64_t)y * o.sx + (int64_t)outer * o.sy + zp;
                if (act && ok0 && ok1 && zp + 1 < g.nz) {
                    if (d2_pairs)
                        *reinterpret_cast<int2 *>(d) = make_int2(bv0, bv1);
                    else {
                        d[0] = bv0;
                        d[1] = bv1;
                    }
                } else {
                    if (act && ok0 && zp < g.nz) d[0] = bv0;
                    if (act && ok1 && zp + 1 < g.nz) d[1] = bv1;
                }
            } else if constexpr (kPsd) {
                if (act && ok0 && zp < g.nz) psd_update(o, y, outer, zp, bv0);
                if (act && ok1 && zp + 1 < g.nz) psd_update(o, y, outer, zp + 1, bv1);
            } else {
                if (act && ok0 && zp < g.nz) emit_x(o, y, outer, zp, nx, ny, g.nz, bv0 >= kBig ? kInf : bv0);
                if (act && ok1 && zp + 1 < g.nz) emit_x(o, y, outer, zp + 1, nx, ny, g.nz, bv1 >= kBig ? kInf : bv1);
            }
        };
        if constexpr (kNarrow) {
            // Packed 16-bit scan: every tile value is at most 65535, so a lane takes two
            // neighbouring columns (2p, 2p + 1) in one 32-bit register: one LDS read and one
            // v_pk_min_u16 / v_pk_add_u16 (saturating) per offset serve both.  Lanes 0..31 take
            // row j, lanes 32..63 row j + kWaves (the rows the scalar scan pairs).  A candidate
            // min(f, 65535) + k^2 saturating at 65535 leaves the minimum unchanged below 65535
            // and at 65535 (flagged) otherwise, as the uint16 tile does.
            const int p = lane & 31, h = lane >> 5;
            const int zp = z0 + 2 * p;  // first column of this lane's pair
            const uint32_t *pcol = reinterpret_cast<const uint32_t *>(tile) + p;
            const u16x2 cmv = {(unsigned short)min(cm_lo, 65535), (unsigned short)min(cm_hi, 65535)};
            for (int j = wv; j < jend; j += 2 * kWaves) {
                const bool act = h == 0 || j + kWaves < jend;
                const int jr = act ? j + kWaves * h : j;
                const uint32_t *c = pcol + (W + jr) * (kTileZ / 2);
                // an idle half (the last row of an odd count) scans nothing: best 0
                u16x2 best = __builtin_bit_cast(u16x2, act ? c[0] : 0u);
#pragma unroll
                for (int k = 1; k <= W; k += 4) {
                    // a column can still improve while best - cm > k^2
                    const u16x2 gap = __builtin_elementwise_sub_sat(best, cmv);
                    if (__builtin_amdgcn_ballot_w64(max((int)gap.x, (int)gap.y) > k * k) == 0) break;
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int kk = k + u;
                        const u16x2 a = __builtin_bit_cast(u16x2, c[-kk * (kTileZ / 2)]);
                        const u16x2 b = __builtin_bit_cast(u16x2, c[kk * (kTileZ / 2)]);
                        const u16x2 k2 = {(unsigned short)(kk * kk), (unsigned short)(kk * kk)};
                        best = __builtin_elementwise_min(best, __builtin_elementwise_add_sat(__builtin_elementwise_min(a, b), k2));
                    }
                }
                finish_pair(r0 + jr, zp, act, best);
            }
        } else {
            // two rows per step (j and j + kWaves): their scans run in one loop, so a wave keeps
            // the LDS reads of both in flight; the loop ends when neither can improve
            for (int j = wv; j < jend; j += 2 * kWaves) {
                const int j2 = j + kWaves;
                const bool two = j2 < jend;  // wave-uniform
                const L *c = col + (W + j) * kTileZ;
                const L *c2 = col + (W + (two ? j2 : j)) * kTileZ;
                int best = c[0], best2 = c2[0];
                // blocks of four offsets while any lane of the wave can still improve; a lane that
                // is done only folds in candidates >= k^2 >= best, which leave it unchanged, so the
                // lanes need no exec mask and the loop exit is one wave-wide ballot
    #pragma unroll
                for (int k = 1; k <= W; k += 4) {
                    if (__builtin_amdgcn_ballot_w64(cm + k * k < max(best, best2)) == 0) break;
    #pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int kk = k + u;
                        best = min(best, min(c[-kk * kTileZ], c[kk * kTileZ]) + kk * kk);
                        best2 = min(best2, min(c2[-kk * kTileZ], c2[kk * kTileZ]) + kk * kk);
                    }
                }
                finish_row(j, best);
                if (two) finish_row(j2, best2);
            }
        }
        if (!more) break;
        // slide: rows T .. T+2W-1 become rows 0 .. 2W-1, the new rows follow
        int4 h[kHalo];
#pragma unroll
        for (int i = 0; i < kHalo; ++i) h[i] = TileIO<L>::get4(tile + (T + row0 + 16 * i) * kTileZ + 4 * q);
        __syncthreads();
        acc = make_int4(kBig, kBig, kBig, kBig);
#pragma unroll
        for (int i = 0; i < kHalo; ++i) {
            TileIO<L>::put4(tile + (row0 + 16 * i) * kTileZ + 4 * q, h[i]);
            acc.x = min(acc.x, h[i].x);
            acc.y = min(acc.y, h[i].y);
            acc.z = min(acc.z, h[i].z);
            acc.w = min(acc.w, h[i].w);
        }
        commit_rows<kHull, E, T / 16>(tile, g, 2 * W, row0, q, zl, xm, xr, acc,
                                      r0 + T + W >= 0 && r0 + 2 * T + W <= g.n, r0 + T + W,
                                      (int64_t)outer * g.nz + z0 + 4 * q);
    }
}

template <int T, int W, class E, class O, bool D2ONLY = false, bool kHull = false, bool kPsd = false>
__global__ __launch_bounds__(kWinThreads) void k_window(const E *__restrict__ in, O *__restrict__ out,
                                                        PassXOut o, WinGeom g, int nx, int ny,
                                                        uint32_t *__restrict__ line_bits,
                                                        uint32_t *__restrict__ ovf_count,
                                                        uint32_t *__restrict__ ovf_list,
                                                        const unsigned int *__restrict__ proceed)
{
    if (proceed && *proceed == 0u) return;
    window_body<T, W, E, O, D2ONLY, kHull, kPsd>(xcd_tile(blockIdx.x, window_blocks(g, T, g.tpb)), in, out, o, g, nx, ny,
                                    line_bits, ovf_count, ovf_list);
}

// Fallback for deferred lines: the exact envelope, persistent threads,
// stacks in the workspace; results clamped like the windowed passes.
__device__ __forceinline__ int finish(int v, int cap)
{
    v = min(v, cap);
    return v >= kBig ? kInf : v;
}

template <class O>
__device__ __forceinline__ void fallback_y_body(int64_t blk, int64_t nblk, const uint16_t *__restrict__ in,
                                                O *__restrict__ out, int nx, int ny, int nz, int64_t pitch,
                                                int64_t out_sx, int64_t out_sy, int bc_lo, int bc_hi, int cap,
                                                uint32_t *__restrict__ ovf_count,
                                                const uint32_t *__restrict__ ovf_list, int *__restrict__ stk,
                                                int stk_cap)
{
    const uint32_t count = *ovf_count;
    const int64_t tid = blk * blockDim.x + threadIdx.x;
    const int64_t nthreads = nblk * blockDim.x;
    Stack st{stk + tid, stk + (int64_t)stk_cap * nthreads + tid, nthreads, stk_cap,
             stk + (int64_t)2 * stk_cap * nthreads + tid};
    for (int64_t i = tid; i < (int64_t)count; i += nthreads) {
        const int64_t line = ovf_list[i];
        const int64_t x = line / nz, z = line - x * nz;
        const uint16_t *src = in + x * ny * pitch + z;
        O *dst = out + x * out_sx + z;
        envelope_line(
            ny, bc_lo, bc_hi, st, true,
            [&](int j) {
                const uint16_t d = src[(int64_t)j * pitch];
                return d == kNoSolid ? kInf : (int)d * (int)d;
            },
            [&](int j, int v) {
                const int f = finish(v, cap);
                if constexpr (std::is_same<O, Sq16>::value) {
                    if (f >= 65535) {  // saturated (or exactly 65535): the EDT takes its int32 passes
                        ovf_count[2] = 1u;
                        ovf_count[3] = 0u;
                    }
                    dst[(int64_t)j * out_sy] = Sq16{(uint16_t)min(f, 65535)};
                } else if constexpr (std::is_same<O, Sq8>::value) {
                    dst[(int64_t)j * out_sy] = Sq8{(uint8_t)min(f, 255)};
                } else {
                    dst[(int64_t)j * out_sy] = f;
                }
            });
    }
}

template <class O>
__global__ __launch_bounds__(256) void k_fallback_y(const uint16_t *__restrict__ in, O *__restrict__ out,
                                                    int nx, int ny, int nz, int64_t pitch, int64_t out_sx,
                                                    int64_t out_sy, int bc_lo, int bc_hi, int cap,
                                                    uint32_t *__restrict__ ovf_count,
                                                    const uint32_t *__restrict__ ovf_list,
                                                    int *__restrict__ stk, int stk_cap,
                                                    const unsigned int *__restrict__ proceed)
{
    if (proceed && *proceed == 0u) return;
    fallback_y_body<O>(blockIdx.x, gridDim.x, in, out, nx, ny, nz, pitch, out_sx, out_sy, bc_lo, bc_hi, cap,
                       ovf_count, ovf_list, stk, stk_cap);
}

template <class E>
__device__ __forceinline__ void fallback_x_body(int64_t blk, int64_t nblk, const E *__restrict__ in,
                                                const PassXOut &o, int nx, int ny, int nz, int64_t in_sx,
                                                int64_t in_sy, int bc_lo, int bc_hi, int cap,
                                                const uint32_t *__restrict__ ovf_count,
                                                const uint32_t *__restrict__ ovf_list, int *__restrict__ stk,
                                                int stk_cap)
{
    const uint32_t count = *ovf_count;
    const int64_t tid = blk * blockDim.x + threadIdx.x;
    const int64_t nthreads = nblk * blockDim.x;
    Stack st{stk + tid, stk + (int64_t)stk_cap * nthreads + tid, nthreads, stk_cap,
             stk + (int64_t)2 * stk_cap * nthreads + tid};
    for (int64_t i = tid; i < (int64_t)count; i += nthreads) {
        const int64_t line = ovf_list[i];
        const int64_t y = line / nz, z = line - y * nz;
        const E *src = in + y * in_sy + z;
        envelope_line(
            nx, bc_lo, bc_hi, st, true, [&](int j) { return decode(src[(int64_t)j * in_sx], kInf); },
            [&](int j, int v) { emit_x(o, j, y, z, nx, ny, nz, finish(v, cap)); });
    }
}

template <class E>
__global__ __launch_bounds__(256) void k_fallback_x(const E *__restrict__ in, PassXOut o,
                                                    int nx, int ny, int nz, int64_t in_sx, int64_t in_sy,
                                                    int bc_lo, int bc_hi, int cap,
                                                    const uint32_t *__restrict__ ovf_count,
                                                    const uint32_t *__restrict__ ovf_list,
                                                    int *__restrict__ stk, int stk_cap,
                                                    const unsigned int *__restrict__ proceed)
{
    if (proceed && *proceed == 0u) return;
    fallback_x_body<E>(blockIdx.x, gridDim.x, in, o, nx, ny, nz, in_sx, in_sy, bc_lo, bc_hi, cap, ovf_count,
                       ovf_list, stk, stk_cap);
}

__global__ __launch_bounds__(256) void k_clear_words(uint32_t *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}


// ---------------------------------------------------------------------------
// batched launches: one launch runs a kernel over many domains
// ---------------------------------------------------------------------------
//
// Domain k of a batched launch owns the launch's blocks first[k] ..
// first[k+1]-1 (for the windowed passes after the XCD-aware reordering, so
// each XCD takes a contiguous run of domains) and reads its arguments from
// args[k].  gate >= 0 offsets each domain's proceed pointer (per-level
// flags: a domain whose flag reads 0 does nothing); the other overrides
// (>= 0) replace the per-level parts of the arguments, which every domain of
// the batch shares.

__device__ __forceinline__ int batch_slot(const int64_t *__restrict__ first, int n, int64_t gb)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= gb)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ bool batch_gate(const unsigned int *p, int gate)
{
    if (p && gate >= 0) p += gate;
    return !p || *p != 0u;
}

struct PassZArgs {
    const uint8_t *img;
    int64_t in_sx, in_sy;
    int ny, nz;
    int64_t rows;
    int bc_lo, bc_hi, vec_in;
    int64_t pitch;
    uint16_t *out;
    uint32_t *counters;
    const unsigned int *proceed;
    uint32_t *bits;  // cleared by the short-row kernel
    int64_t n_bits;
};

// L = 64: one row per wave (rows longer than 512 voxels), else L lanes per row
template <bool kInvert, bool kMatch, int L, class D = uint16_t>
__global__ __launch_bounds__(256) void k_pass_z_batch(const PassZArgs *__restrict__ args,
                                                      const int64_t *__restrict__ first, int n, uint32_t match,
                                                      int gate)
{
    const int64_t gb = blockIdx.x;
    const int k = batch_slot(first, n, gb);
    const PassZArgs a = args[k];
    if (!batch_gate(a.proceed, gate)) return;
    const int64_t blk = gb - first[k], nblk = first[k + 1] - first[k];
    D *out = reinterpret_cast<D *>(a.out);  // uint8 distances share the uint16 array's memory
    if constexpr (L == 64)
        pass_z_body<kInvert, kMatch, D>(blk, a.img, a.in_sx, a.in_sy, a.ny, a.nz, a.rows, a.bc_lo, a.bc_hi,
                                        a.vec_in, match, a.pitch, out, a.counters);
    else
        pass_z_short_body<kInvert, kMatch, L, D>(blk, nblk, a.img, a.in_sx, a.in_sy, a.ny, a.nz, a.rows, a.bc_lo,
                                                 a.bc_hi, a.vec_in, match, a.pitch, out, a.counters, a.bits,
                                                 a.n_bits);
}

struct WinArgs {
    const void *in;
    void *out;
    PassXOut o;
    WinGeom g;
    int nx, ny;
    uint32_t *bits, *count, *list;
    const unsigned int *proceed;
};

template <int T, int W, class E, class O, bool D2ONLY, bool kPsd = false>
__global__ __launch_bounds__(kWinThreads) void k_window_batch(const WinArgs *__restrict__ args,
                                                              const int64_t *__restrict__ first, int n,
                                                              int64_t total, int gate, int cap, int32_t thr,
                                                              int32_t code)
{
    const int64_t gb = xcd_tile(blockIdx.x, total);
    const int k = batch_slot(first, n, gb);
    const WinArgs &a = args[k];
    if (!batch_gate(a.proceed, gate)) return;
    WinGeom g = a.g;
    PassXOut o = a.o;
    if (cap >= 0) g.cap = cap;
    if (code >= 0) {
        o.t = thr;
        o.psd_code = code;
    }
    window_body<T, W, E, O, D2ONLY, false, kPsd>(gb - first[k], static_cast<const E *>(a.in), static_cast<O *>(a.out), o, g,
                                    a.nx, a.ny, a.bits, a.count, a.list);
}

struct FallbackArgs {
    const void *in;
    void *out;        // pass Y: its output array
    PassXOut o;       // pass X: its outputs
    int nx, ny, nz;
    int64_t s0, s1;   // pass Y: output strides of x and y; pass X: input strides of x and y
    int64_t pitch;    // pass Y: row pitch of its input
    int bc_lo, bc_hi;
    uint32_t *count;
    const uint32_t *list;
    int *stk;
    int stk_cap;
    const unsigned int *proceed;
};

template <class O>
__global__ __launch_bounds__(256) void k_fallback_y_batch(const FallbackArgs *__restrict__ args,
                                                          const int64_t *__restrict__ first, int n, int cap)
{
    const int64_t gb = blockIdx.x;
    const int k = batch_slot(first, n, gb);
    const FallbackArgs &a = args[k];
    if (!batch_gate(a.proceed, -1)) return;
    fallback_y_body<O>(gb - first[k], first[k + 1] - first[k], static_cast<const uint16_t *>(a.in),
                       static_cast<O *>(a.out), a.nx, a.ny, a.nz, a.pitch, a.s0, a.s1, a.bc_lo, a.bc_hi, cap,
                       a.count, a.list, a.stk, a.stk_cap);
}

template <class E>
__global__ __launch_bounds__(256) void k_fallback_x_batch(const FallbackArgs *__restrict__ args,
                                                          const int64_t *__restrict__ first, int n, int cap)
{
    const int64_t gb = blockIdx.x;
    const int k = batch_slot(first, n, gb);
    const FallbackArgs &a = args[k];
    if (!batch_gate(a.proceed, -1)) return;
    fallback_x_body<E>(gb - first[k], first[k + 1] - first[k], static_cast<const E *>(a.in), a.o, a.nx, a.ny,
                       a.nz, a.s0, a.s1, a.bc_lo, a.bc_hi, cap, a.count, a.list, a.stk, a.stk_cap);
}

// the deferred-line bitmaps of the long-row domains (the short-row pass Z
// clears its own): domain k's words are cleared by its blocks
struct ClearArgs {
    uint32_t *p;
    int64_t n;
};

__global__ __launch_bounds__(256) void k_clear_words_batch(const ClearArgs *__restrict__ args,
                                                           const int64_t *__restrict__ first, int n)
{
    const int64_t gb = blockIdx.x;
    const int k = batch_slot(first, n, gb);
    const ClearArgs a = args[k];
    const int64_t nblk = first[k + 1] - first[k];
    for (int64_t i = (gb - first[k]) * blockDim.x + threadIdx.x; i < a.n; i += nblk * blockDim.x) a.p[i] = 0u;
}

}  // namespace
}  // namespace pmx
