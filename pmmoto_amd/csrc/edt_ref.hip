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
#include "edt_kernels.hpp"

namespace pmx {
namespace {

constexpr float kFltMax = 3.40282347e+38f;
constexpr int kNumHull = 4;            // euclidean_distance_transform.py:275,284
constexpr int kHullCap = kNumHull + 1; // members returned per face
#ifndef PMX_REF_LDS_CAP
#define PMX_REF_LDS_CAP 64
#endif
// LDS stack entries per lane of the boundary-hull kernel
constexpr int kLdsCap = PMX_REF_LDS_CAP;
#ifndef PMX_REF_RING
#define PMX_REF_RING 32
#endif
// LDS ring entries per lane of the streaming parabolic pass
constexpr int kRingSlots = PMX_REF_RING;
constexpr int kStage = 8;        // input positions staged per lane at a time
constexpr int kEvalPerStep = 2;  // positions evaluated per input position
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

// Calls f(i, p[i * st]) for i = i0, i0 + step, ... while i is in [lo, hi)
// (step +1 from i0 = lo, or -1 from i0 = hi - 1), stopping when f returns
// false.  One lane walks its line in order, so the loads run kScan values
// ahead of the calls: a line pays one memory latency per kScan values
// instead of one per value.
constexpr int kScan = 16;
template <int kStep, class T, class F>
__host__ __device__ __forceinline__ void stream_values(const T *p, int64_t st, int lo, int hi, F &&f)
{
    static_assert(kStep == 1 || kStep == -1, "unit steps");
    const int count = hi - lo;
    if (count <= 0) return;
    const int i0 = kStep > 0 ? lo : hi - 1;
    auto load = [&](T (&buf)[kScan], int j0) {
#pragma unroll
        for (int c = 0; c < kScan; ++c)
            buf[c] = j0 + c < count ? p[(int64_t)(i0 + kStep * (j0 + c)) * st] : T();
    };
    T cur[kScan], nxt[kScan];
    load(cur, 0);
    for (int j0 = 0; j0 < count; j0 += kScan) {
        const bool more = j0 + kScan < count;
        if (more) load(nxt, j0 + kScan);
#pragma unroll
        for (int c = 0; c < kScan; ++c)
            if (j0 + c < count && !f(i0 + kStep * (j0 + c), cur[c])) return;
        if (more) {
#pragma unroll
            for (int c = 0; c < kScan; ++c) cur[c] = nxt[c];
        }
    }
}

// first (forward) / last index of value 0 within [lo_skip, n - hi_skip), or -1
// (_voxels.hpp:60-79 through _voxels.pyx:231-283)
__device__ int nearest_zero(const uint8_t *img, int n, int64_t stride, bool forward, int lo_skip, int hi_skip)
{
    int found = -1;
    auto hit = [&](int i, uint8_t v) {
        if (v != 0) return true;
        found = i;
        return false;
    };
    if (forward)
        stream_values<1>(img, stride, lo_skip, n - hi_skip, hit);
    else
        stream_values<-1>(img, stride, lo_skip, n - hi_skip, hit);
    return found;
}

// pass 0 along g.dim from the image into out (squared, INF -> FLT_MAX)
__global__ __launch_bounds__(256) void k_ref_pass0(const uint8_t *__restrict__ img, float *__restrict__ out,
                                                   LineGeom g, float res, int periodic, Skips sk,
                                                   const unsigned *__restrict__ gate)
{
    if (gate && *gate == 0u) return;
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
    stream_values<1>(im, ist, 1, n, [&](int i, uint8_t v) {
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
        return true;
    });
    // cap, backward min, square, to_finite (:168-181)
    float next = o[(int64_t)(n - 1) * ost];
    if (next > hi * res) next = hi * res;
    {
        const float sq = next * next;
        o[(int64_t)(n - 1) * ost] = sq == INFINITY ? kFltMax : sq;
    }
    stream_values<-1>(o, ost, 0, n - 1, [&](int i, float v) {
        const float cur = fminf(v, next + res);
        const float sq = cur * cur;
        o[(int64_t)i * ost] = sq == INFINITY ? kFltMax : sq;
        next = cur;
        return true;
    });
}

// intersect_parabolas divides the float numerator by 2.0 * factor1 in
// double and rounds to float.  A quotient of two floats rounded to double
// (53 bits) and then to float (24 bits) equals the correctly rounded float
// quotient, since 53 >= 2 * 24 + 2 (double rounding is innocuous for
// division at that margin), and 2 * factor1 is exact in float below 2^127; so
// the float division (correctly rounded in HIP) gives the same bits, with the
// double route kept for a factor1 near the top of the float range.
// kFloatDiv: the caller knows |factor1| < 2^125 (vertices less than 3n apart,
// 3n * aniso below it), so the check is left out.
template <bool kFloatDiv = false>
__host__ __device__ __forceinline__ float intersect(float f_r, float r, float f_q, float q, float aniso)
{
    const float factor1 = (r - q) * aniso;
    const float factor2 = r + q;
    const float num = f_r - f_q + factor1 * factor2;
    if (kFloatDiv || fabsf(factor1) < 0x1p125f) return num / (2.0f * factor1);
    return (float)((double)num / (2.0 * (double)factor1));
}

// A lane's stack: entry k at base[k * step] (LDS: step = lanes of the
// wave; workspace: step = lanes of the grid).
struct LineStack {
    int *hv;
    float *hh, *rg;
    int64_t step;
    __host__ __device__ __forceinline__ int &V(int k) const { return hv[(int64_t)k * step]; }
    __host__ __device__ __forceinline__ float &H(int k) const { return hh[(int64_t)k * step]; }
    __host__ __device__ __forceinline__ float &R(int k) const { return rg[(int64_t)k * step]; }
};

// update_hull (_distance.hpp:89-112); false when entry k+1 would reach cap
__host__ __device__ __forceinline__ bool update_hull(const LineStack &S, int &k, int i, float fv, float aniso, int cap)
{
    float s = intersect<>(fv, (float)i, S.H(k), (float)S.V(k), aniso);
    while (k > 0 && s <= S.R(k)) {
        --k;
        s = intersect<>(fv, (float)i, S.H(k), (float)S.V(k), aniso);
    }
    ++k;
    if (k + 1 >= cap) return false;
    S.V(k) = i;
    S.H(k) = fv;
    S.R(k) = s;
    S.R(k + 1) = INFINITY;
    return true;
}

// Pass inputs as the reference's float32 heights: the float array itself,
// or (the integer form of the reference arithmetic at a power-of-two
// resolution) a 1-D distance c or a squared distance m in voxel units, whose
// float heights (c * res)^2 = aniso * c^2 and aniso * m are exact.  A line
// with no zero along the first axis holds kNoSolid, the reference's
// infinity, which get_initial_envelope turns into FLT_MAX (_distance.pyx:351).
struct AsFloat {
    __host__ __device__ __forceinline__ float operator()(float v) const { return v; }
};
struct FromDist {
    float aniso;
    __device__ __forceinline__ float operator()(uint16_t c) const
    {
        return c == kNoSolid ? kFltMax : aniso * (float)((uint32_t)c * c);
    }
};
struct FromSq {
    float aniso;
    __device__ __forceinline__ float operator()(uint16_t m) const { return aniso * (float)m; }
};

// return_boundary_hull (_distance.hpp:296-356) on n values at p (stride
// st); writes <= kHullCap members to out and returns the count, or -1 when
// the envelope outgrows cap.  A one-voxel segment with no finite value reads
// past the reference's vectors there (undefined behaviour); this path then
// returns the zero-initialised slot, as the oracle does.
template <class T, class Cv>
__host__ __device__ int boundary_hull(const T *p, Cv cv, int n, int64_t st, float res, int index_corrector,
                                      bool forward, const LineStack &S, int cap, RefHull *out)
{
    const float p0 = cv(p[0]);
    if (n == 1 && p0 < kFltMax) {
        out[0] = {index_corrector, p0, INFINITY};
        return 1;
    }
    const float aniso = res * res;
    S.V(0) = 0;
    S.H(0) = p0;
    S.V(1) = 0;
    S.H(1) = 0.0f;
    S.R(0) = -INFINITY;
    S.R(1) = INFINITY;
    int k = 1;
    bool fits = true;
    stream_values<1>(p, st, 1, n, [&](int i, T v) { return fits = update_hull(S, k, i, cv(v), aniso, cap); });
    if (!fits) return -1;
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
template <class T, class Cv>
__device__ bool line_hulls(const uint8_t *og, const T *f, Cv cv, const LineGeom &g, float res, Skips sk,
                           const LineStack &S, int cap, RefHull *hulls, int *counts)
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
            c = boundary_hull(f + (int64_t)start * st, cv, nv, st, res, ic, fw, S, cap, hulls + side * kHullCap);
            if (c < 0) return false;
        }
        counts[side] = c;
    }
    return true;
}

// The wrapped rows a line's hulls keep, as the windowed kernels read them
// (hull_row_kept): lower-hull vertices v < 0 (side 1) mark row v, upper-hull
// vertices v >= 0 (side 0, pushed at v + n) row n + v.
__device__ __forceinline__ void write_hull_mask(const RefHull *h, const int *counts, uint32_t *m)
{
    uint32_t w[kHullMaskWords] = {};
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int c = counts[side];
#pragma unroll
        for (int j = 0; j < kHullCap; ++j) {
            if (j >= c) break;
            const int v = h[side * kHullCap + j].vertex;
            const int b = side ? -1 - v : v;
            if (b < 0 || b >= kHullMaskRows) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if ((b >> 5) == q) w[(side ? 0 : 4) + q] |= 1u << (b & 31);
        }
    }
#pragma unroll
    for (int q = 0; q < kHullMaskWords; ++q) m[q] = w[q];
}

// hulls[line][side][kHullCap], counts[line][side]; lines whose segments
// outgrow the LDS stack go to the overflow list.  masks: also the wrapped
// rows each line keeps (write_hull_mask).  Nothing runs while *gate == 0.
template <class T, class Cv>
__global__ __launch_bounds__(kWave) void k_ref_hulls(const uint8_t *__restrict__ og, const T *__restrict__ f, Cv cv,
                                                     LineGeom g, float res, Skips sk, RefHull *__restrict__ hulls,
                                                     int *__restrict__ counts, int *__restrict__ ovf,
                                                     int *__restrict__ n_ovf, uint32_t *__restrict__ masks,
                                                     const unsigned *__restrict__ gate)
{
    if (gate && *gate == 0u) return;
    __shared__ int sv[kLdsCap * kWave];
    __shared__ float sh[kLdsCap * kWave], sr[kLdsCap * kWave];
    const int64_t line = (int64_t)blockIdx.x * kWave + threadIdx.x;
    if (line >= g.lines()) return;
    LineStack S{sv + threadIdx.x, sh + threadIdx.x, sr + threadIdx.x, kWave};
    int64_t ib, fb, ob;
    g.base(line, &ib, &fb, &ob);
    RefHull *h = hulls + line * 2 * kHullCap;
    if (!line_hulls(og + ib, f + fb, cv, g, res, sk, S, kLdsCap, h, counts + line * 2))
        ovf[atomicAdd(n_ovf, 1)] = (int)line;
    else if (masks)
        write_hull_mask(h, counts + line * 2, masks + line * kHullMaskWords);
}

// A lane's stack of cap entries for the run and hull fallbacks: dynamic LDS
// of [cap][lanes of the workgroup] ints, floats, floats (kLds), else the
// workspace `gstk` as [cap][lanes of the grid].
template <bool kLds>
__device__ __forceinline__ LineStack lane_stack(int *dyn, int *gstk, int cap)
{
    if (kLds) {
        const int lanes = blockDim.x;
        float *f = reinterpret_cast<float *>(dyn);
        return LineStack{dyn + threadIdx.x, f + (size_t)cap * lanes + threadIdx.x,
                     f + (size_t)2 * cap * lanes + threadIdx.x, lanes};
    }
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, T = (int64_t)gridDim.x * blockDim.x;
    float *f = reinterpret_cast<float *>(gstk);
    return LineStack{gstk + tid, f + (int64_t)cap * T + tid, f + (int64_t)2 * cap * T + tid, T};
}

// the overflow list of k_ref_hulls, stacks of cap entries (a whole line)
template <bool kLds, class T, class Cv>
__global__ __launch_bounds__(kWave) void k_ref_hulls_fallback(const uint8_t *__restrict__ og, const T *__restrict__ f,
                                                              Cv cv, LineGeom g, float res, Skips sk,
                                                              RefHull *__restrict__ hulls, int *__restrict__ counts,
                                                              const int *__restrict__ ovf,
                                                              const int *__restrict__ n_ovf, int *gstk, int cap,
                                                              uint32_t *__restrict__ masks,
                                                              const unsigned *__restrict__ gate)
{
    if (gate && *gate == 0u) return;
    extern __shared__ int dyn[];
    const LineStack S = lane_stack<kLds>(dyn, gstk, cap);
    const int64_t nt = (int64_t)gridDim.x * blockDim.x;
    const int m = *n_ovf;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += nt) {
        const int64_t line = ovf[j];
        int64_t ib, fb, ob;
        g.base(line, &ib, &fb, &ob);
        RefHull *h = hulls + line * 2 * kHullCap;
        line_hulls(og + ib, f + fb, cv, g, res, sk, S, cap, h, counts + line * 2);
        if (masks) write_hull_mask(h, counts + line * 2, masks + line * kHullMaskWords);
    }
}

// A lane's stack for the streaming pass: entry k lives at slot (k & mask),
// slot s at base[s * lanes]; with a power-of-two ring (mask = slots - 1)
// entries below the evaluation cursor are overwritten as the line advances.
// Vertices are held as v + voff in 16 bits (voff = n: lower-hull vertices
// lie in [-n, 0), upper-hull ones in [n, 2n), so v + n < 3n < 2^16).
struct RingStack {
    unsigned short *v;
    float *h, *r;
    int lanes;
    unsigned mask;
    int voff;
    __host__ __device__ __forceinline__ int at(int k) const { return (int)((unsigned)k & mask) * lanes; }
    __host__ __device__ __forceinline__ int V(int o) const { return (int)v[o] - voff; }
    __host__ __device__ __forceinline__ void set(int o, int vert, float hv, float rv) const
    {
        v[o] = (unsigned short)(vert + voff);
        h[o] = hv;
        r[o] = rv;
    }
};

// Positions [p1, p2) of a line, between zero-height parabolas at p1 and p2;
// p1 = -1 starts at the line start (lower hull or position 0), p2 = n runs
// to the line end (upper hull).
//
// Such a run is a problem of its own.  The zero at p1 is never popped (a
// later parabola r meets it at s >= (r + p1) / 2 > p1 >= its range), every
// entry below it has a range that ends before p1, and update_hull never
// reaches below it; so the envelope over [p1, p2] and the values there are
// those of a stack that starts with the zero at p1 alone.
struct Chord {
    int line, p1, p2;
};

// squared_edt_1d_parabolic (_distance.hpp:184-272) on one line, streaming:
//
//   * the envelope's top entry stays in registers; an entry is written to
//     the ring when pushed and read back only when the entries above it are
//     popped;
//   * a position is evaluated once a zero-height parabola lies after it
//     (its covering entry and that entry's range are then final), at most
//     kEvalPerStep positions per input position, so the lanes of a wave do
//     the same work per step; the rest is evaluated after the upper hull;
//   * the input arrives kStage positions at a time through `stage` (this
//     lane's column of an LDS tile, stride `lanes`), the next chunk's loads
//     in flight while the current one is processed;
//   * a run whose live entries outgrow the ring (after everything final is
//     evaluated) is handed on whole, emit(p1, p2), and the lane resumes with
//     a fresh stack at the zero that ends it.  emit returns false when it
//     took the whole line instead; the lane then stops.
//
// The covering entry of position i is the first kk with R(kk + 1) >= i, and
// its value aniso * (i - v)^2 + h, as in the reference, so each value is
// bit-identical.  All lanes of a wave run the call together (the chunk loop
// is uniform); a lane with alive = false follows along without effect.
template <bool kSqrt, bool kFloatDiv, class Emit>
__host__ __device__ void parabolic_stream(const float *fin, float *fout, int n, int64_t ist, int64_t ost,
                                          float aniso, const RefHull *lower, int nl, const RefHull *upper, int nu,
                                          const RingStack &S, int cap, bool alive, float *stage, int lanes,
                                          Emit &&emit)
{
    int k = 0, loop_start = 1;
    int tv = 0, bv = 0;  // top entry k and the entry k - 1 below it
    float th = 0.0f, tr = -INFINITY, bh = 0.0f, br = -INFINITY;
    if (alive) {
        if (nl > 0) {
            loop_start = 0;
            for (int m = nl - 1, j = 0; m >= 0; --m, ++j) S.set(S.at(j), lower[m].vertex, lower[m].height, lower[m].range);
            k = nl - 1;
            tv = lower[0].vertex;
            th = lower[0].height;
            tr = lower[0].range;
            if (nl > 1) {
                bv = lower[1].vertex;
                bh = lower[1].height;
                br = lower[1].range;
            }
        } else {
            th = fin[0];
            S.set(S.at(0), 0, th, -INFINITY);
        }
    }
    // evaluation cursor: next position pe (output op), its covering
    // candidate kk (ev, eh) and R(kk + 1) in er; positions below zl (the
    // last zero) are final
    int kk = 0, pe = 0, zl = -1;
    int ev = nl > 0 ? lower[nl - 1].vertex : 0;
    float eh = nl > 0 ? lower[nl - 1].height : th, er = INFINITY;
    float *op = fout;
    auto eval = [&]() {
        while (er < (float)pe) {
            ++kk;
            const int o = S.at(kk);
            ev = S.V(o);
            eh = S.h[o];
            er = kk < k ? S.r[S.at(kk + 1)] : INFINITY;
        }
        const float di = (float)(pe - ev);
        const float val = aniso * (di * di) + eh;
        *op = kSqrt ? sqrtf(val) : val;
        op += ost;
        ++pe;
    };
    // update_hull (_distance.hpp:89-112).  A pop takes the entry below from
    // registers and refills that copy from the ring, so the next pop test
    // does not wait for the ring.  (The copy can hold an overwritten slot
    // only when the top is the last zero's entry, which is never popped.)
    auto pop = [&]() {
        --k;
        tv = bv;
        th = bh;
        tr = br;
        if (k > 0) {
            const int o = S.at(k - 1);
            bv = S.V(o);
            bh = S.h[o];
            br = S.r[o];
        }
    };
    // room for one more entry, after evaluating what is final if needed
    auto room = [&]() {
        if (k + 1 - kk >= cap)
            while (pe < zl && k + 1 - kk >= cap) eval();
        return k + 1 - kk < cap;
    };
    auto push = [&](int i, float fv, float s) {
        ++k;
        bv = tv;
        bh = th;
        br = tr;
        tv = i;
        th = fv;
        tr = s;
        S.set(S.at(k), i, fv, s);
    };
    // false when the ring is full
    auto insert = [&](int i, float fv) -> bool {
        float s = intersect<kFloatDiv>(fv, (float)i, th, (float)tv, aniso);
        while (k > 0 && s <= tr) {
            pop();
            s = intersect<kFloatDiv>(fv, (float)i, th, (float)tv, aniso);
        }
        if (!room()) return false;
        push(i, fv, s);
        return true;
    };
    // a fresh stack holding the zero at i alone
    auto restart = [&](int i) {
        k = kk = 0;
        S.set(S.at(0), i, 0.0f, -INFINITY);
        tv = ev = i;
        th = eh = 0.0f;
        tr = -INFINITY;
        er = INFINITY;
        pe = zl = i;
        op = fout + (int64_t)i * ost;
    };
    bool skip = false;  // inside a run handed on, which starts at p1
    int p1 = -1;
    float pre[kStage];
#pragma unroll
    for (int c = 0; c < kStage; ++c) pre[c] = c < n ? fin[(int64_t)c * ist] : 0.0f;
    for (int i0 = 0; i0 < n; i0 += kStage) {
#pragma unroll
        for (int c = 0; c < kStage; ++c) stage[c * lanes] = pre[c];
        if (i0 + kStage < n) {
#pragma unroll
            for (int c = 0; c < kStage; ++c) {
                const int i = i0 + kStage + c;
                pre[c] = i < n ? fin[(int64_t)i * ist] : 0.0f;
            }
        }
        const int m = n - i0 < kStage ? n - i0 : kStage;
        float fnext = stage[0];
        for (int c = 0; c < m; ++c) {
            const int i = i0 + c;
            const float fv = fnext;
            if (c + 1 < m) fnext = stage[(c + 1) * lanes];
            if (alive && i >= loop_start) {
                if (skip) {
                    if (fv == 0.0f) {
                        skip = false;
                        alive = emit(p1, i);
                        restart(i);
                    }
                } else if (!insert(i, fv)) {
                    if (fv == 0.0f) {
                        alive = emit(zl, i);
                        restart(i);
                    } else {
                        skip = true;
                        p1 = zl;
                    }
                } else if (fv == 0.0f) {
                    // entries 0..k are final; R(kk + 1) may have changed
                    // since it was read (kk + 1 > the previous zero's entry)
                    zl = i;
                    er = kk < k ? S.r[S.at(kk + 1)] : INFINITY;
                }
            }
            if (alive && !skip) {
#pragma unroll
                for (int e = 0; e < kEvalPerStep; ++e)
                    if (pe < zl) eval();
            }
        }
    }
    if (alive && !skip) {
        for (int m = 0; m < nu; ++m) {
            if (!insert(upper[m].vertex + n, upper[m].height)) {
                skip = true;
                p1 = zl;
                break;
            }
        }
    }
    if (alive && skip) {
        emit(p1, n);
    } else if (alive) {
        zl = n;
        er = kk < k ? S.r[S.at(kk + 1)] : INFINITY;
        while (pe < n) eval();
    }
}

// Stack entries a run can need: its positions, the hull entries at the line
// ends, and update_hull's range sentinel.
__host__ __device__ __forceinline__ int chord_entries(const Chord &c, int n)
{
    const int first = c.p1 < 0 ? 0 : c.p1, last = c.p2 < n ? c.p2 : n - 1;
    return last - first + 1 + 2 * kHullCap + 2;
}

// One run on its own (see Chord): the envelope over its positions (and the
// hulls at the line ends), then its positions evaluated against it.  Returns
// false when the stack outgrows cap.
template <bool kSqrt>
__host__ __device__ bool chord_line(const float *fin, float *fout, int n, int64_t ist, int64_t ost, float aniso,
                                    const RefHull *lower, int nl, const RefHull *upper, int nu, Chord c,
                                    const LineStack &S, int cap)
{
    int k = 0, i0;
    S.R(0) = -INFINITY;
    S.R(1) = INFINITY;
    if (c.p1 >= 0) {
        S.V(0) = c.p1;
        S.H(0) = 0.0f;
        i0 = c.p1 + 1;
    } else if (nl > 0) {
        if (nl + 1 >= cap) return false;
        for (int m = nl - 1; m >= 0; --m) {
            S.V(k) = lower[m].vertex;
            S.H(k) = lower[m].height;
            S.R(k) = lower[m].range;
            S.R(k + 1) = INFINITY;
            ++k;
        }
        --k;
        i0 = 0;
    } else {
        S.V(0) = 0;
        S.H(0) = fin[0];
        i0 = 1;
    }
    const int last = c.p2 < n ? c.p2 : n - 1;
    bool fits = true;
    stream_values<1>(fin, ist, i0, last + 1, [&](int i, float v) { return fits = update_hull(S, k, i, v, aniso, cap); });
    if (!fits) return false;
    if (c.p2 >= n)
        for (int m = 0; m < nu; ++m)
            if (!update_hull(S, k, upper[m].vertex + n, upper[m].height, aniso, cap)) return false;
    const int e0 = c.p1 < 0 ? 0 : c.p1, e1 = c.p2 < n ? c.p2 : n;
    float *op = fout + (int64_t)e0 * ost;
    int kk = 0;
    for (int i = e0; i < e1; ++i, op += ost) {
        while (S.R(kk + 1) < (float)i) ++kk;
        const float di = (float)(i - S.V(kk));
        const float val = aniso * (di * di) + S.H(kk);
        *op = kSqrt ? sqrtf(val) : val;
    }
    return true;
}

struct LineHulls {
    const RefHull *lower, *upper;
    int nl, nu;
};

__device__ __forceinline__ LineHulls hulls_of(const RefHull *hulls, const int *counts, int64_t line)
{
    LineHulls h{nullptr, nullptr, 0, 0};
    if (hulls) {
        h.lower = hulls + (line * 2 + 1) * kHullCap;
        h.upper = hulls + (line * 2 + 0) * kHullCap;
        h.nl = counts[line * 2 + 1];
        h.nu = counts[line * 2 + 0];
    }
    return h;
}

// Run lists between the parabolic kernels.  Each workgroup of the ring pass
// collects its runs in its own segment of kSegChords slots (counted in LDS,
// so the lanes do not contend for one counter), then reserves a range of the
// dense list `a` with one atomic and copies them there.  `c` takes, with one
// counter, the runs too long for the second kernel and the whole lines of a
// workgroup whose segment is full; c_cap = lines + every slot of `a`, so `c`
// cannot fill up.
constexpr int kSegChords = kWave * 4;
struct ChordLists {
    Chord *seg, *a, *c;
    int *n_a, *n_c;
    int a_cap, c_cap;
};

// one wave per workgroup, one line per lane, stacks in an LDS ring of
// kRingSlots entries per lane
template <int kRingSlots, bool kSqrt, bool kFloatDiv>
__global__ __launch_bounds__(kWave) void k_ref_parabolic(const float *__restrict__ fin, float *__restrict__ fout,
                                                         LineGeom g, float res, const RefHull *__restrict__ hulls,
                                                         const int *__restrict__ counts, ChordLists L,
                                                         const unsigned *__restrict__ gate)
{
    if (gate && *gate == 0u) return;
    static_assert((kRingSlots & (kRingSlots - 1)) == 0, "ring slots must be a power of two");
    __shared__ unsigned short sv[kRingSlots * kWave];
    __shared__ float sh[kRingSlots * kWave], sr[kRingSlots * kWave], st[kStage * kWave];
    const int lane = threadIdx.x;
    const int64_t lines = g.lines();
    int64_t line = (int64_t)blockIdx.x * kWave + lane;
    const bool real = line < lines;
    if (!real) line = lines - 1;  // keeps the tail lanes' loads in bounds
    int64_t ib, fb, ob;
    g.base(line, &ib, &fb, &ob);
    const LineHulls h = hulls_of(hulls, counts, line);
    const int n = (int)g.n[g.dim];
    const RingStack S{sv + lane, sh + lane, sr + lane, kWave, (unsigned)(kRingSlots - 1), n};
    __shared__ int seg_n;
    if (lane == 0) seg_n = 0;
    __syncthreads();
    Chord *seg = L.seg + (size_t)blockIdx.x * kSegChords;
    auto emit = [&](int p1, int p2) -> bool {
        const int j = atomicAdd(&seg_n, 1);
        if (j < kSegChords) {
            seg[j] = Chord{(int)line, p1, p2};
            return true;
        }
        L.c[atomicAdd(L.n_c, 1)] = Chord{(int)line, -1, n};
        return false;
    };
    parabolic_stream<kSqrt, kFloatDiv>(fin + fb, fout + ob, n, g.fs[g.dim], g.os[g.dim], res * res, h.lower, h.nl, h.upper,
                            h.nu, S, kRingSlots, real, st + lane, kWave, emit);
    __syncthreads();  // the segment's writes are visible to the workgroup
    __shared__ int base;
    const int cnt = min(seg_n, kSegChords);
    if (lane == 0) base = cnt ? atomicAdd(L.n_a, cnt) : 0;
    __syncthreads();
    for (int t = lane; t < cnt; t += kWave) L.a[base + t] = seg[t];
}

// The runs of a list (n_in[0] of them, at most in_cap), one per lane, each
// with a stack of `cap` entries (see lane_stack).  Runs that need more than
// cap entries go to `out` (non-null whenever cap is below a whole line's).
template <bool kSqrt, bool kLds>
__global__ __launch_bounds__(kWave) void k_ref_chords(const float *__restrict__ fin, float *__restrict__ fout,
                                                      LineGeom g, float res, const RefHull *__restrict__ hulls,
                                                      const int *__restrict__ counts, const Chord *__restrict__ in,
                                                      const int *__restrict__ n_in, int in_cap, Chord *out,
                                                      int *n_out, int cap, int *gstk,
                                                      const unsigned *__restrict__ gate)
{
    if (gate && *gate == 0u) return;
    extern __shared__ int dyn[];
    const LineStack S = lane_stack<kLds>(dyn, gstk, cap);
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    const int n = (int)g.n[g.dim];
    const int m = min(*n_in, in_cap);
    const float aniso = res * res;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < m; j += T) {
        const Chord c = in[j];
        int64_t ib, fb, ob;
        g.base(c.line, &ib, &fb, &ob);
        const LineHulls h = hulls_of(hulls, counts, c.line);
        if (chord_entries(c, n) > cap ||
            !chord_line<kSqrt>(fin + fb, fout + ob, n, g.fs[g.dim], g.os[g.dim], aniso, h.lower, h.nl, h.upper,
                               h.nu, c, S, cap))
            if (out) out[atomicAdd(n_out, 1)] = c;  // cap below a whole line only with an out list
    }
}

__global__ void k_zero_ints(int *p, int n, const unsigned *__restrict__ gate)
{
    if (gate && *gate == 0u) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// The integer form of the reference arithmetic (edt3d, or corrected_edt on a
// padded subdomain, at a power-of-two resolution)
// ---------------------------------------------------------------------------
//
// Let res = 2^e on every axis, aniso = res^2, and every extent be at most
// kIntMaxN.  Then every float of edt3d is aniso times an integer below 2^24,
// computed exactly: pass 0's d += res and d * d (distances c * res, c <= n),
// intersect_parabolas' numerator aniso * (m_r - m_q + (r - q)(r + q)) with
// vertices in [-n, 2n), and the evaluation aniso * (j^2 + m).  Only the
// intersections s = (m_r - m_q + r^2 - q^2) / (2 (r - q)) round.  Such an s,
// when not an integer, lies at least 1 / (6n) from every integer, farther
// than its rounding error (at most 2^-23 n for |s| <= 2n, twice that for a
// hull range shifted by its index corrector, whose vertices lie less than n
// apart: 1 / (2n) there); n <= 1024 keeps both below the gap.  So no range
// rounds across an integer position, and a float pop decision that differs
// from the exact one only moves envelope breakpoints inside one float's
// rounding interval, which holds no position: at every position the
// reference selects a parabola of least value, and its output is aniso times
// the exact minimum over the parabolas it is given.
//
// Which parabolas those are does depend on the rounding: on a periodic axis
// the line gets the first / last kHullCap members of the float envelope of
// its head / tail segment (return_boundary_hull).  Those are computed with
// the reference's float operations (k_ref_hulls over the integer pass input)
// and kept as masks of the wrapped rows they occupy.  The passes are then
// integer minima: pass 0 the 1-D distance, pass 1 the windowed kernel over
// y with the masks (the exact EDT's pass Y), pass 2 a scan along z with the
// masks, which writes aniso * d2 (or its sqrtf).
//
// A line along x with no zero holds the reference's infinity (FLT_MAX after
// get_initial_envelope), which every finite parabola of a later line beats
// in the float envelope too (its intersections come out as +-huge or
// infinite, so it is popped or never selected).  A case outside the form (a
// squared distance above 65534 after pass 1, which includes
// a line along y with no finite height, a scan past kIntWrap rows) raises
// ctl[kCtlBail] or the windowed pass's flags; the float kernels then run in
// full, gated on ctl[kCtlGate].
constexpr int kIntMaxN = 1024;
constexpr int kIntWrap = kHullMaskRows;  // rows past a line end a scan may read
// control words: the bail flag, the windowed pass's counters (deferred
// lines [0], saturation [2]), the gate of the float kernels (0: the integer
// result stands, 1: it is recomputed, 2: the integer form did not apply),
// then the windowed pass's deferred-line bitmap
constexpr int kCtlBail = 0, kCtlLabels = 1, kCtlWin = 4, kCtlGate = 8, kCtlBits = 16;
bool g_int_form = true;

// pass 0 (axis 0) in integers: the 1-D distance c to the nearest zero along
// x at dz[x][y][z] (row pitch `pitch`), one lane per (y, z) with its
// column's zeros as a bit mask in LDS.  Past the line's last zero the next
// one is vnext, before its first the previous one vprev: edt3d's wrap
// (last - n, first + n; _distance.pyx:240-279) with wrap == 1, corrected_edt's
// pad-skipping face distances with wrap == 2 (first zero f at or after
// 2 p_lo and last zero l before n - 2 p_hi, l >= 2 p_lo: l + p_hi + 1 - n and
// n - 1 + f - p_lo, the virtual zeros of its correctors; :304-447), none
// (END) with wrap == 0.  A byte above 1 (a multi-label image) raises
// ctl[kCtlLabels], and the kLabels instance, which returns at once unless
// that flag is set, writes the
// label-aware pass: squared_edt_1d (_distance.hpp:136-182) resets to one
// resolution step on both voxels of a change between two different non-zero
// bytes, so each such edge voxel is a source of value 1 and a voxel's value
// is min(distance to the nearest zero or virtual zero, 1 + distance to the
// nearest edge voxel of its line); the correctors' virtual zeros come from
// the zeros alone (get_nearest_boundary_index with label 0), and
// corrected_edt runs the same first pass over the whole padded line.
template <bool kLabels>
__global__ __launch_bounds__(256) void k_int_pass0(const uint8_t *__restrict__ img, int64_t is0, int64_t is1, int nx,
                                                   int ny, int nz, int wrap, int plo, int phi,
                                                   uint16_t *__restrict__ dz, int64_t pitch, uint32_t *__restrict__ ctl)
{
    __shared__ uint32_t bits[(kIntMaxN / 32) * 256];
    __shared__ uint32_t ebits[kLabels ? (kIntMaxN / 32) * 256 : 1];
    if constexpr (kLabels) {
        if (ctl[kCtlLabels] == 0u) return;  // the image holds no byte above 1
    }
    const int t = threadIdx.x;
    const int64_t col = (int64_t)blockIdx.x * 256 + t;
    if (col >= (int64_t)ny * nz) return;  // no barriers below: each lane owns its LDS column
    const int y = (int)(col / nz), z = (int)(col - (int64_t)y * nz);
    const uint8_t *p = img + (int64_t)y * is1 + z;
    const int nw = (nx + 31) >> 5;
    uint32_t bad = 0;
    uint8_t last = 0;  // the byte before the chunk (none before x = 0)
    for (int x0 = 0; x0 < nx; x0 += 32) {
        uint8_t v[32];
#pragma unroll
        for (int j = 0; j < 32; ++j) v[j] = x0 + j < nx ? p[(int64_t)(x0 + j) * is0] : (uint8_t)1;
        uint32_t word = 0;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            word |= (uint32_t)(v[j] == 0) << j;
            bad |= v[j] > 1;
        }
        bits[(x0 >> 5) * 256 + t] = word;
        if constexpr (kLabels) {
            // bit x: a label change between x - 1 and x (two different non-zero bytes)
            uint32_t chg = 0;
#pragma unroll
            for (int j = 0; j < 32; ++j) {
                const uint8_t pv = j ? v[j - 1] : last;
                chg |= (uint32_t)(pv != 0 && v[j] != 0 && pv != v[j]) << j;
            }
            last = v[31];
            const int left = nx - x0;
            ebits[(x0 >> 5) * 256 + t] = left >= 32 ? chg : chg & ((1u << left) - 1u);
        }
    }
    if constexpr (!kLabels) {
        if (bad) ctl[kCtlLabels] = 1u;
    } else {
        // the edge voxels: a change on either side (x, or x + 1)
        for (int q = 0; q < nw; ++q) {
            const uint32_t c = ebits[q * 256 + t], cn = q + 1 < nw ? ebits[(q + 1) * 256 + t] : 0u;
            ebits[q * 256 + t] = c | (c >> 1) | (cn << 31);
        }
    }
    // first set bit at or after a / last set bit before b of a column mask, or -1
    auto first_from = [&](const uint32_t *m, int a) -> int {
        if (a >= nx) return -1;
        int q = a >> 5;
        uint32_t wq = m[q * 256 + t] & (~0u << (a & 31));
        while (true) {
            if (wq) return q * 32 + __builtin_ctz(wq);
            if (++q >= nw) return -1;
            wq = m[q * 256 + t];
        }
    };
    auto last_before = [&](const uint32_t *m, int b) -> int {
        if (b <= 0) return -1;
        int q = (b - 1) >> 5;
        uint32_t wq = m[q * 256 + t] & (0xffffffffu >> (31 - ((b - 1) & 31)));
        while (true) {
            if (wq) return q * 32 + 31 - __builtin_clz(wq);
            if (--q < 0) return -1;
            wq = m[q * 256 + t];
        }
    };
    uint16_t *o = dz + (int64_t)y * pitch + z;
    const int64_t xs = (int64_t)ny * pitch;
    const int first = first_from(bits, 0);
    const int efirst = kLabels ? first_from(ebits, 0) : -1;
    if (first < 0 && efirst < 0) {
        // no zero (and no label edge): the reference's infinity (FLT_MAX in its
        // later passes), which loses to every finite parabola of a line
        for (int x = 0; x < nx; ++x) o[(int64_t)x * xs] = kNoSolid;
        return;
    }
    constexpr int kFar = 1 << 20;
    int vprev = -kFar, vnext = kFar;
    if (first >= 0) {
        if (wrap == 1) {
            vprev = last_before(bits, nx) - nx;
            vnext = first + nx;
        } else if (wrap == 2) {
            const int f = first_from(bits, 2 * plo), l = last_before(bits, nx - 2 * phi);
            if (l >= 2 * plo) vprev = l + phi + 1 - nx;
            if (f >= 0) vnext = nx - 1 + f - plo;
        }
    }
    int prev = vprev, next = first >= 0 ? first : vnext;
    int eprev = -kFar, enext = efirst >= 0 ? efirst : kFar;
    for (int x = 0; x < nx; ++x) {
        int d;
        if (x == next) {
            d = 0;
            prev = x;
            const int nn = first_from(bits, x + 1);
            next = nn >= 0 ? nn : vnext;
        } else {
            d = min(x - prev, next - x);
        }
        if constexpr (kLabels) {
            // an edge voxel holds 1 (squared_edt_1d's reset at a label change), and the
            // distance grows by one per voxel from it
            if (x == enext) {
                eprev = x;
                const int nn = first_from(ebits, x + 1);
                enext = nn >= 0 ? nn : kFar;
            }
            d = min(d, min(x - eprev, enext - x) + 1);
        }
        o[(int64_t)x * xs] = (uint16_t)min(d, 0xfffe);
    }
}

// pass 2 (axis 2, contiguous) in integers: per voxel the minimum over rows q
// of (z - q)^2 + m(q), scanning outward until k^2 reaches the best value, on
// lines staged in LDS with kIntWrap rows past each end (the wrapped rows the
// line's hulls keep on a periodic axis, none otherwise: row r < 0 reads
// r + wrap_lo, row r >= nz reads r - wrap_hi); written as the float
// aniso * d2, or its sqrtf.  One wave per two lines of b[y][x][z] (line
// numbers 2p, 2p + 1 of x * ny + y), held as pairs of uint16 in one word, so
// one LDS read and packed minimum / saturating add serve both (values are at
// most 65534, a missing row reads 65535, and a sum saturating at 65535 never
// undercuts a best value).
template <bool kSqrt>
__global__ __launch_bounds__(256) void k_int_pass2(const uint16_t *__restrict__ b, int nx, int ny, int nz,
                                                   int64_t pitch, int periodic, int wrap_lo, int wrap_hi,
                                                   const uint32_t *__restrict__ masks,
                                                   float *__restrict__ out, int64_t os0, int64_t os1, float aniso,
                                                   uint32_t *__restrict__ ctl)
{
    constexpr int kExt = kIntMaxN + 2 * kIntWrap;
    __shared__ __align__(16) uint32_t ext[4][kExt];
    __shared__ uint32_t msk[4][2 * kHullMaskWords];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t lines = (int64_t)nx * ny;
    const int64_t l0 = ((int64_t)blockIdx.x * 4 + w) * 2, l1 = l0 + 1;
    const bool live0 = l0 < lines, live1 = l1 < lines;
    const int x0 = live0 ? (int)(l0 / ny) : 0, y0 = live0 ? (int)(l0 - (int64_t)x0 * ny) : 0;
    const int x1 = live1 ? (int)(l1 / ny) : x0, y1 = live1 ? (int)(l1 - (int64_t)x1 * ny) : y0;
    const uint16_t *s0 = b + (int64_t)y0 * nx * pitch + (int64_t)x0 * pitch;
    const uint16_t *s1 = b + (int64_t)y1 * nx * pitch + (int64_t)x1 * pitch;
    uint32_t *e = ext[w];
    if (live0) {
        if (periodic && lane < 2 * kHullMaskWords) {
            const int64_t ln = lane < kHullMaskWords ? l0 : (live1 ? l1 : l0);
            msk[w][lane] = masks[ln * kHullMaskWords + (lane & (kHullMaskWords - 1))];
        }
        // 8 values of each line per lane (rows are 16-byte aligned and padded
        // to the pitch; values past nz are overwritten by the wrap rows below)
        for (int j = lane * 8; j < nz; j += kWave * 8) {
            const uint4 a = *reinterpret_cast<const uint4 *>(s0 + j);
            const uint4 c = live1 ? *reinterpret_cast<const uint4 *>(s1 + j) : make_uint4(0, 0, 0, 0);
            const uint32_t av[4] = {a.x, a.y, a.z, a.w}, cv[4] = {c.x, c.y, c.z, c.w};
            uint32_t pr[8];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                pr[2 * t] = (av[t] & 0xffffu) | (cv[t] << 16);
                pr[2 * t + 1] = (av[t] >> 16) | (cv[t] & 0xffff0000u);
            }
            *reinterpret_cast<uint4 *>(e + kIntWrap + j) = make_uint4(pr[0], pr[1], pr[2], pr[3]);
            *reinterpret_cast<uint4 *>(e + kIntWrap + j + 4) = make_uint4(pr[4], pr[5], pr[6], pr[7]);
        }
    }
    __syncthreads();
    if (live0) {
        // rows -1 - j and nz + j: the wrapped copies the masks keep, else 65535
        for (int j = lane; j < kIntWrap; j += kWave) {
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                uint32_t v = 0xffffffffu;
                const int r = side ? nz + j : -1 - j;
                const int m = side ? r - wrap_hi : r + wrap_lo;  // the row a wrapped row reads
                if (periodic && m >= 0 && m < nz) {
                    const bool inb = j < kHullMaskRows;
                    const int wi = (side ? 4 : 0) + (j >> 5);
                    const bool k0 = inb && ((msk[w][wi] >> (j & 31)) & 1u);
                    const bool k1 = inb && ((msk[w][kHullMaskWords + wi] >> (j & 31)) & 1u);
                    const uint32_t own = e[kIntWrap + m];
                    v = (k0 ? (own & 0xffffu) : 0xffffu) | (k1 ? (own & 0xffff0000u) : 0xffff0000u);
                }
                if (side) e[kIntWrap + nz + j] = v;
                else e[kIntWrap - 1 - j] = v;
            }
        }
    }
    __syncthreads();
    if (!live0) return;
    bool over = false;
    float *o0 = out + (int64_t)x0 * os0 + (int64_t)y0 * os1;
    float *o1 = out + (int64_t)x1 * os0 + (int64_t)y1 * os1;
    for (int z = lane; z < nz; z += kWave) {
        const uint32_t *c = e + kIntWrap + z;
        u16x2 best = __builtin_bit_cast(u16x2, c[0]);
        if (!live1) best.y = 0;
        int k = 1;
        for (; k <= kIntWrap; k += 4) {
            if (__builtin_amdgcn_ballot_w64(max((int)best.x, (int)best.y) > k * k) == 0) break;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int kk = k + u;
                const u16x2 lo = __builtin_bit_cast(u16x2, c[-kk]), hi = __builtin_bit_cast(u16x2, c[kk]);
                const u16x2 k2 = {(unsigned short)(kk * kk), (unsigned short)(kk * kk)};
                best = __builtin_elementwise_min(best, __builtin_elementwise_add_sat(__builtin_elementwise_min(lo, hi), k2));
            }
        }
        // offsets up to kIntWrap are in; a value above (kIntWrap + 1)^2 may need more
        over |= max((int)best.x, (int)best.y) > (kIntWrap + 1) * (kIntWrap + 1);
        const float v0 = aniso * (float)best.x, v1 = aniso * (float)best.y;
        o0[z] = kSqrt ? sqrtf(v0) : v0;
        if (live1) o1[z] = kSqrt ? sqrtf(v1) : v1;
    }
    if (over) ctl[kCtlBail] = 1u;
}

// Stack entries per lane: a first kernel with a small stack keeps many
// waves in flight; the lines it cannot hold go to a second one with a large
// stack (few lines, so few waves), and the rare lines beyond that to the
// whole-line float fallback.
#ifndef PMX_INT_HULL_CAP
#define PMX_INT_HULL_CAP 16
#endif
constexpr int kIntHullCap = PMX_INT_HULL_CAP;
constexpr int kIntHullCapBig = 128;
constexpr int kIntHullListWaves = 1024;  // persistent waves of the second kernel
constexpr uint32_t kIntInfM = (1u << 22) - 1;

// one hull of one line into its half of the line's mask; false when the
// envelope outgrows kCap entries (then nothing is written)
template <bool kDist, int kCap>
__device__ __forceinline__ bool int_hull_side(const uint16_t *__restrict__ f, const LineGeom &g, float aniso,
                                              int hs, int te, int64_t line, int side, uint32_t *S,
                                              uint32_t *__restrict__ masks, uint32_t *__restrict__ ctl)
{
    int64_t ib, fb, ob;
    g.base(line, &ib, &fb, &ob);
    const int n = (int)g.n[g.dim];
    const int64_t st = g.fs[g.dim];
    const uint16_t *p = f + fb;
    auto mval = [](uint16_t v) -> uint32_t {
        return kDist ? (v == kNoSolid ? kIntInfM : (uint32_t)v * v) : (uint32_t)v;
    };
    auto hf = [aniso](uint32_t m) { return m == kIntInfM ? kFltMax : aniso * (float)m; };
    auto ent = [](int v, uint32_t m) { return ((uint32_t)v << 22) | m; };
    // the envelope of the segment's values, calls of add(i, m) in order
    // (i from 0); top entry and the one below it kept in registers
    int k = 0, tv = 0, bv = 0;
    uint32_t tm = 0, bm = 0;
    float th = 0.0f, tr = -INFINITY, bh = 0.0f;
    bool fits = true;
    auto add = [&](int i, uint32_t m) {
        const float h = hf(m);
        if (i == 0) {
            k = 0;
            tv = 0;
            tm = m;
            th = h;
            tr = -INFINITY;
            S[0] = ent(0, m);
            return;
        }
        float sv = intersect<true>(h, (float)i, th, (float)tv, aniso);
        while (k > 0 && sv <= tr) {
            // pop: the entry below becomes the top, its range recomputed from
            // the entry below it
            --k;
            tv = bv;
            tm = bm;
            th = bh;
            if (k > 0) {
                const uint32_t e = S[(k - 1) * kWave];
                bv = (int)(e >> 22);
                bm = e & kIntInfM;
                bh = hf(bm);
                tr = intersect<true>(th, (float)tv, bh, (float)bv, aniso);
            } else {
                tr = -INFINITY;
            }
            sv = intersect<true>(h, (float)i, th, (float)tv, aniso);
        }
        ++k;
        if (k + 1 >= kCap) {
            fits = false;
            return;
        }
        S[k * kWave] = ent(i, m);
        bv = tv;
        bm = tm;
        bh = th;
        tv = i;
        tm = m;
        th = h;
        tr = sv;
    };
    uint32_t w[4] = {};
    // the reference's selection: from the start (forward) or the top, the
    // finite entries, at most min(kNumHull, k) + 1 of them; mark row
    // bit(local vertex)
    auto pick = [&](bool forward, int nv) {
        const int num = min(kNumHull, k);
        int cnt = 0;
        for (int j = 0; j <= k && cnt <= num; ++j) {
            const uint32_t e = S[(forward ? j : k - j) * kWave];
            if ((e & kIntInfM) == kIntInfM) continue;
            ++cnt;
            const int v = (int)(e >> 22);
            const int b = forward ? v : nv - 1 - v;  // rows n + v, or -1 - (v - nv)
            if (b >= kHullMaskRows) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if ((b >> 5) == q) w[q] |= 1u << (b & 31);
        }
    };
    // side 0, the head: from hs through the first zero (to the end if none);
    // side 1, the tail: from the last zero before te (from 0 if none) to te.
    // edt3d: hs = 0, te = n; corrected_edt skips twice the pads: hs = 2 p_lo,
    // te = n - 2 p_hi (get_boundary_hull's lower_skip / upper_skip,
    // _distance.pyx:587-616)
    int nv;
    if (side == 0) {
        int first = -1;
        stream_values<1>(p + (int64_t)hs * st, st, 0, n - hs, [&](int i, uint16_t v) {
            add(i, mval(v));
            if (v == 0) {
                first = i;
                return false;
            }
            return fits;
        });
        nv = first < 0 ? n - hs : first + 1;
    } else {
        int last = -1;
        stream_values<-1>(p, st, 0, te, [&](int i, uint16_t v) {
            if (v != 0) return true;
            last = i;
            return false;
        });
        const int t0 = last < 0 ? 0 : last;
        nv = te - t0;
        stream_values<1>(p + (int64_t)t0 * st, st, 0, nv, [&](int i, uint16_t v) {
            add(i, mval(v));
            return fits;
        });
    }
    if (!fits) return false;
    if (nv == 1 && tm == kIntInfM) {
        // a one-voxel segment with no finite height: the reference's
        // selection reads its placeholder entry there; the float kernels
        // reproduce that
        ctl[kCtlBail] = 1u;
        return true;
    }
    pick(side == 0, nv);
    // side 0 (upper hull, rows n + b) -> words 4..7, side 1 -> words 0..3
    uint32_t *m = masks + line * kHullMaskWords + (side == 0 ? 4 : 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) m[q] = w[q];
    return true;
}

// Both hulls of every line, or the (line, side) items of `list` (2 line +
// side), one per lane; an item whose envelope outgrows kCap entries goes to
// `out`: as an item (out_items) or as a line for the whole-line fallback.
template <bool kDist, int kCap>
__global__ __launch_bounds__(kWave) void k_int_hulls(const uint16_t *__restrict__ f, LineGeom g, float aniso, int hs,
                                                     int te, uint32_t *__restrict__ masks, const int *__restrict__ list,
                                                     const int *__restrict__ list_n, int *__restrict__ out,
                                                     int *__restrict__ out_n, int out_items,
                                                     uint32_t *__restrict__ ctl)
{
    __shared__ uint32_t stk[kCap * kWave];
    uint32_t *S = stk + threadIdx.x;
    auto item = [&](int64_t line, int side) {
        if (!int_hull_side<kDist, kCap>(f, g, aniso, hs, te, line, side, S, masks, ctl))
            out[atomicAdd(out_n, 1)] = out_items ? (int)(2 * line + side) : (int)line;
    };
    // no barriers below: each lane owns its LDS column
    if (!list) {
        const int64_t line = (int64_t)blockIdx.x * kWave + threadIdx.x;
        if (line < g.lines()) {
            item(line, 0);
            item(line, 1);
        }
        return;
    }
    const int m = *list_n;
    for (int j = blockIdx.x * kWave + threadIdx.x; j < m; j += gridDim.x * kWave) item(list[j] >> 1, list[j] & 1);
}

// the float kernels' gate from the integer passes' flags
__global__ void k_int_decide(uint32_t *ctl)
{
    if (threadIdx.x == 0)
        ctl[kCtlGate] = (ctl[kCtlBail] | ctl[kCtlWin + 0] | ctl[kCtlWin + 2]) != 0u ? 1u : 0u;
}

__global__ void k_int_mark(uint32_t *ctl, uint32_t v)
{
    if (threadIdx.x == 0) ctl[kCtlGate] = v;
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

// Runs per lane: stacks of cap entries (12 bytes each) in kChordLdsBytes
// of LDS per workgroup, as many lanes (up to a wave) as fit, or 0
constexpr size_t kChordLdsBytes = 48 * 1024;
size_t chord_lds_bytes(int cap, int lanes) { return (size_t)lanes * (size_t)cap * 12; }
int chord_lds_lanes(int cap) { return (int)std::min<size_t>(kWave, kChordLdsBytes / chord_lds_bytes(cap, 1)); }
constexpr int kChordCap = 128;       // stack entries per lane of the second parabolic kernel
constexpr int kChordMaxBlocks = 4096;
constexpr int kIntFbBlocks = 256;   // persistent blocks of the integer form's hull fallback

struct RefWs {
    float *tmp;
    int *stk;
    RefHull *hulls;
    int *counts, *ovf, *n_ovf, *ovf2;
    Chord *cseg, *ca, *cc;
    int a_cap, c_cap;
    // the integer form: its two uint16 arrays share tmp (the float kernels
    // run only after the integer result is set aside), masks, control words
    uint16_t *dz, *b;
    int64_t pitch;
    uint32_t *masks, *ctl, *wlist;
    int64_t ctl_words;
};

RefWs carve_ref(Carver &c, const int64_t *n)
{
    const int64_t nmax = std::max(n[0], std::max(n[1], n[2]));
    const int64_t lines = max_lines(n);
    RefWs w;
    w.pitch = (n[2] + 7) / 8 * 8;
    const size_t int_words = (size_t)(n[0] * n[1] * w.pitch);  // two uint16 arrays = int_words floats
    w.tmp = c.take<float>(std::max((size_t)(n[0] * n[1] * n[2]), int_words));
    w.dz = reinterpret_cast<uint16_t *>(w.tmp);
    w.b = w.dz + int_words;
    w.stk = c.take<int>((size_t)3 * fb_cap(nmax) * fb_threads(n));
    w.hulls = c.take<RefHull>((size_t)lines * 2 * kHullCap);
    w.counts = c.take<int>((size_t)lines * 2);
    w.ovf = c.take<int>((size_t)lines * 2);  // a line may be listed once per side
    w.n_ovf = c.take<int>(4);
    w.a_cap = (int)((lines + kWave - 1) / kWave * kSegChords);
    w.c_cap = (int)(lines + w.a_cap);
    w.cseg = c.take<Chord>((size_t)w.a_cap);
    w.ca = c.take<Chord>((size_t)w.a_cap);
    w.cc = c.take<Chord>((size_t)w.c_cap);
    w.masks = c.take<uint32_t>((size_t)lines * kHullMaskWords);
    w.ovf2 = c.take<int>((size_t)lines * 2);
    w.ctl_words = kCtlBits + (lines + 31) / 32;
    w.ctl = c.take<uint32_t>((size_t)w.ctl_words);
    w.wlist = c.take<uint32_t>((size_t)lines);
    return w;
}

// whether the integer form covers these arguments (the data-dependent
// conditions are checked on the device)
bool int_form_applies(const int64_t *n, int first_axis, const int64_t *pads, const float *res)
{
    if (!g_int_form || first_axis != 0) return false;
    if (!(res[0] == res[1] && res[1] == res[2])) return false;
    int e = 0;
    if (std::frexp(res[0], &e) != 0.5f || e < -40 || e > 40) return false;
    for (int a = 0; a < 3; ++a)
        if (n[a] > kIntMaxN) return false;
    return true;
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
    // boundary hulls of the lines along g.dim (float path and integer form):
    // the LDS kernel, then the lines whose segments outgrew it
    auto hulls = [&](const auto *f, auto cv, const LineGeom &g, float r, Skips sk, uint32_t *masks,
                     const unsigned *gate) -> int {
        using T = std::remove_cv_t<std::remove_pointer_t<decltype(f)>>;
        using Cv = decltype(cv);
        const int64_t lines = g.lines();
        {
            ProfScope ps("edt_ref_hulls", s);
            hipLaunchKernelGGL((k_ref_hulls<T, Cv>), dim3((unsigned)((lines + kWave - 1) / kWave)), dim3(kWave), 0, s,
                               img, f, cv, g, r, sk, w.hulls, w.counts, w.ovf, w.n_ovf, masks, gate);
        }
        if (int e = check_launch("reference hulls")) return e;
        {
            ProfScope ps("edt_ref_hulls_fallback", s);
            const int lanes = chord_lds_lanes(cap);
            if (lanes > 0) {
                const int64_t blocks = std::min<int64_t>((lines + lanes - 1) / lanes, kChordMaxBlocks);
                hipLaunchKernelGGL((k_ref_hulls_fallback<true, T, Cv>), dim3((unsigned)blocks), dim3(lanes),
                                   chord_lds_bytes(cap, lanes), s, img, f, cv, g, r, sk, w.hulls, w.counts, w.ovf,
                                   w.n_ovf, nullptr, cap, masks, gate);
            } else {
                hipLaunchKernelGGL((k_ref_hulls_fallback<false, T, Cv>), dim3(fb_threads(n) / kWave), dim3(kWave), 0,
                                   s, img, f, cv, g, r, sk, w.hulls, w.counts, w.ovf, w.n_ovf, w.stk, cap, masks,
                                   gate);
            }
        }
        return check_launch("reference hulls");
    };

    // the integer form where it applies; the float kernels below then run
    // only if it set its result aside (ctl[kCtlGate])
    const unsigned *gate = nullptr;
    if (int_form_applies(n, first_axis, pads, resolution)) {
        const int nx = (int)n[0], ny = (int)n[1], nz = (int)n[2];
        const float aniso = resolution[0] * resolution[0];
        const int64_t pitch = w.pitch;
        // pad widths of the periodic axes (corrected_edt), 0 for edt3d
        auto pad = [&](int a, int hi) { return pads && periodic[a] ? (int)pads[2 * a + hi] : 0; };
        hipLaunchKernelGGL(k_clear_words, dim3((unsigned)std::min<int64_t>((w.ctl_words + 255) / 256, 1024)),
                           dim3(256), 0, s, w.ctl, w.ctl_words);
        {
            ProfScope ps("edt_int_pass0", s);
            const dim3 g0((unsigned)(((int64_t)ny * nz + 255) / 256));
            const int wrap0 = periodic[0] ? (pads ? 2 : 1) : 0;
            hipLaunchKernelGGL(k_int_pass0<false>, g0, dim3(256), 0, s, img, is[0], is[1], nx, ny, nz, wrap0,
                               pad(0, 0), pad(0, 1), w.dz, pitch, w.ctl);
            // a multi-label image: the label-aware pass over again (returns at once otherwise)
            hipLaunchKernelGGL(k_int_pass0<true>, g0, dim3(256), 0, s, img, is[0], is[1], nx, ny, nz, wrap0,
                               pad(0, 0), pad(0, 1), w.dz, pitch, w.ctl);
        }
        if (int e = check_launch("integer pass 0")) return e;
        // the hull masks of the lines along dim, then the lines whose
        // envelope outgrew the packed stack (few: whole-line stacks)
        auto int_hulls = [&](const uint16_t *f, auto cv, int dim, const int64_t *fs) -> int {
            using Cv = decltype(cv);
            constexpr bool kDist = std::is_same<Cv, FromDist>::value;
            const LineGeom g = make_line_geom(n, is, fs, fs, dim);
            const int64_t lines = g.lines();
            const int hs = 2 * pad(dim, 0), te = (int)n[dim] - 2 * pad(dim, 1);
            hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, w.n_ovf, 4, nullptr);
            {
                // small stacks for every line, large ones for the lines that
                // outgrew them (listed in w.ovf2), the rest into w.ovf
                ProfScope ps("edt_int_hulls", s);
                hipLaunchKernelGGL((k_int_hulls<kDist, kIntHullCap>), dim3((unsigned)((lines + kWave - 1) / kWave)),
                                   dim3(kWave), 0, s, f, g, aniso, hs, te, w.masks, nullptr, nullptr, w.ovf2,
                                   w.n_ovf + 1, 1, w.ctl);
                hipLaunchKernelGGL((k_int_hulls<kDist, kIntHullCapBig>),
                                   dim3((unsigned)std::min<int64_t>(kIntHullListWaves, (lines + kWave - 1) / kWave)),
                                   dim3(kWave), 0, s, f, g, aniso, hs, te, w.masks, w.ovf2, w.n_ovf + 1, w.ovf,
                                   w.n_ovf, 0, w.ctl);
            }
            if (int e = check_launch("integer hulls")) return e;
            ProfScope ps("edt_int_hulls_fallback", s);
            const int lanes = chord_lds_lanes(cap);
            if (lanes > 0) {
                const int64_t blocks = std::min<int64_t>((lines + lanes - 1) / lanes, kIntFbBlocks);
                hipLaunchKernelGGL((k_ref_hulls_fallback<true, uint16_t, Cv>), dim3((unsigned)blocks), dim3(lanes),
                                   chord_lds_bytes(cap, lanes), s, img, f, cv, g, resolution[dim], skips(dim), w.hulls,
                                   w.counts, w.ovf, w.n_ovf, nullptr, cap, w.masks, nullptr);
            } else {
                hipLaunchKernelGGL((k_ref_hulls_fallback<false, uint16_t, Cv>), dim3(fb_threads(n) / kWave),
                                   dim3(kWave), 0, s, img, f, cv, g, resolution[dim], skips(dim), w.hulls,
                                   w.counts, w.ovf, w.n_ovf, w.stk, cap, w.masks, nullptr);
            }
            return check_launch("integer hulls");
        };
        // pass 1 over y: dz [x][y][z] -> b [y][x][z] (uint16 squared distances)
        const int64_t fs_dz[3] = {ny * pitch, pitch, 1}, fs_b[3] = {pitch, nx * pitch, 1};
        if (periodic[1])
            if (int e = int_hulls(w.dz, FromDist{aniso}, 1, fs_dz)) return e;
        {
            ProfScope ps("edt_int_pass1", s);
            const int bc = periodic[1] ? PMX_BC_PERIODIC : PMX_BC_END;
            WinGeom wg{ny, pitch, ny * pitch, nx, nz, bc, bc, kBig, pitch, pitch, nx * pitch};
            wg.hmask = w.masks;
            wg.wrap_lo = ny - 2 * pad(1, 1);
            wg.wrap_hi = ny - 2 * pad(1, 0);
            wg.tpb = tiles_per_block(wg, 64);
            const unsigned blocks = (unsigned)window_blocks(wg, 64, wg.tpb);
            Sq16 *b16 = reinterpret_cast<Sq16 *>(w.b);
            if (periodic[1])
                hipLaunchKernelGGL((k_window<64, 16, uint16_t, Sq16, false, true>), dim3(blocks), dim3(kWinThreads),
                                   0, s, w.dz, b16, PassXOut{}, wg, nx, ny, w.ctl + kCtlBits, w.ctl + kCtlWin,
                                   w.wlist, nullptr);
            else
                hipLaunchKernelGGL((k_window<64, 16, uint16_t, Sq16, false, false>), dim3(blocks), dim3(kWinThreads),
                                   0, s, w.dz, b16, PassXOut{}, wg, nx, ny, w.ctl + kCtlBits, w.ctl + kCtlWin,
                                   w.wlist, nullptr);
        }
        if (int e = check_launch("integer pass 1")) return e;
        if (periodic[2])
            if (int e = int_hulls(w.b, FromSq{aniso}, 2, fs_b)) return e;
        {
            ProfScope ps("edt_int_pass2", s);
            const unsigned blocks = (unsigned)(((int64_t)nx * ny + 7) / 8);
            if (take_sqrt)
                hipLaunchKernelGGL(k_int_pass2<true>, dim3(blocks), dim3(256), 0, s, w.b, nx, ny, nz, pitch,
                                   periodic[2] ? 1 : 0, nz - 2 * pad(2, 1), nz - 2 * pad(2, 0), w.masks, d, os[0],
                                   os[1], aniso, w.ctl);
            else
                hipLaunchKernelGGL(k_int_pass2<false>, dim3(blocks), dim3(256), 0, s, w.b, nx, ny, nz, pitch,
                                   periodic[2] ? 1 : 0, nz - 2 * pad(2, 1), nz - 2 * pad(2, 0), w.masks, d, os[0],
                                   os[1], aniso, w.ctl);
        }
        if (int e = check_launch("integer pass 2")) return e;
        hipLaunchKernelGGL(k_int_decide, dim3(1), dim3(64), 0, s, w.ctl);
        gate = w.ctl + kCtlGate;
    } else {
        hipLaunchKernelGGL(k_int_mark, dim3(1), dim3(64), 0, s, w.ctl, 2u);
    }
    if (int e = check_launch("integer form")) return e;

    // pass order: first_axis with the 1-D sweep, then the remaining axes in
    // ascending order (edt3d: 0 then 1, 2; edt2d on a (1, X, Y) box: 2 then
    // 1, 0); pass 0 writes d, the parabolic passes go d -> tmp -> d
    {
        ProfScope ps("edt_ref_pass0", s);
        LineGeom g = make_line_geom(n, is, os, os, first_axis);
        const int64_t nb = (g.lines() + 255) / 256;
        hipLaunchKernelGGL(k_ref_pass0, dim3((unsigned)nb), dim3(256), 0, s, img, d, g, resolution[first_axis],
                           periodic[first_axis] ? 1 : 0, skips(first_axis), gate);
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
        hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, w.n_ovf, 4, gate);
        if (periodic[dim])
            if (int e = hulls(fin, AsFloat{}, g, resolution[dim], skips(dim), nullptr, gate)) return e;
        const RefHull *hl = periodic[dim] ? w.hulls : nullptr;
        const float r = resolution[dim];
        const ChordLists L{w.cseg, w.ca, w.cc, w.n_ovf + 1, w.n_ovf + 2, w.a_cap, w.c_cap};
        {
            ProfScope ps(dim == 2 ? "edt_ref_parabolic_rows" : "edt_ref_parabolic", s);
            // vertices lie in [-n, 2n): float division needs 3n * r^2 < 2^125
            const bool fdiv = 3.0 * (double)g.n[dim] * (double)(r * r) < 0x1p125;
            auto launch = [&](auto kernel) {
                hipLaunchKernelGGL(kernel, dim3(nb), dim3(kWave), 0, s, fin, fout, g, r, hl, w.counts, L, gate);
            };
            if (sq)
                fdiv ? launch(k_ref_parabolic<kRingSlots, true, true>) : launch(k_ref_parabolic<kRingSlots, true, false>);
            else
                fdiv ? launch(k_ref_parabolic<kRingSlots, false, true>)
                     : launch(k_ref_parabolic<kRingSlots, false, false>);
        }
        if (int e = check_launch("reference parabolic")) return e;
        // runs that outgrew the ring: stacks of kChordCap entries, longer
        // runs (and whole lines from a full list) with stacks of a line
        auto chords = [&](const Chord *in, const int *n_in, int in_cap, Chord *out, int *n_out, int ccap) {
            const int lanes = chord_lds_lanes(ccap);
            const int64_t slots = in_cap;
            if (lanes > 0) {
                const int64_t blocks = std::min<int64_t>((slots + lanes - 1) / lanes, kChordMaxBlocks);
                const size_t lds = chord_lds_bytes(ccap, lanes);
                if (sq)
                    hipLaunchKernelGGL((k_ref_chords<true, true>), dim3((unsigned)blocks), dim3(lanes), lds, s, fin,
                                       fout, g, r, hl, w.counts, in, n_in, in_cap, out, n_out, ccap,
                                       nullptr, gate);
                else
                    hipLaunchKernelGGL((k_ref_chords<false, true>), dim3((unsigned)blocks), dim3(lanes), lds, s, fin,
                                       fout, g, r, hl, w.counts, in, n_in, in_cap, out, n_out, ccap,
                                       nullptr, gate);
            } else {
                // fb_threads(n) lanes of the grid share w.stk (stacks of a line)
                const int64_t blocks = fb_threads(n) / kWave;
                if (sq)
                    hipLaunchKernelGGL((k_ref_chords<true, false>), dim3((unsigned)blocks), dim3(kWave), 0, s, fin,
                                       fout, g, r, hl, w.counts, in, n_in, in_cap, out, n_out, ccap,
                                       w.stk, gate);
                else
                    hipLaunchKernelGGL((k_ref_chords<false, false>), dim3((unsigned)blocks), dim3(kWave), 0, s, fin,
                                       fout, g, r, hl, w.counts, in, n_in, in_cap, out, n_out, ccap,
                                       w.stk, gate);
            }
        };
        {
            ProfScope ps("edt_ref_chords", s);
            chords(L.a, L.n_a, L.a_cap, L.c, L.n_c, std::min(kChordCap, cap));
        }
        if (int e = check_launch("reference parabolic runs")) return e;
        {
            ProfScope ps("edt_ref_chords_long", s);
            chords(L.c, L.n_c, L.c_cap, nullptr, nullptr, cap);
        }
        if (int e = check_launch("reference parabolic long runs")) return e;
        ++step;
    }
    return PMX_OK;
}

int pmx_edt_ref_int_form(int enable)
{
    const int prev = pmx::g_int_form ? 1 : 0;
    pmx::g_int_form = enable != 0;
    return prev;
}

int pmx_edt_ref_last_form(const int64_t *n, const void *workspace, size_t workspace_bytes, hipStream_t s, int *form)
{
    using namespace pmx;
    clear_error();
    if (!n || !workspace || !form) return fail(PMX_ERR_INVALID, "NULL argument");
    size_t need = 0;
    if (int e = pmx_edt_ref_workspace_size(n, &need)) return e;
    if (workspace_bytes < need) return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    Carver c(const_cast<void *>(workspace), workspace_bytes);
    const RefWs w = carve_ref(c, n);
    uint32_t v = 0;
    if (hipStreamSynchronize(s) != hipSuccess || hipMemcpy(&v, w.ctl + kCtlGate, 4, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(PMX_ERR_HIP, "reading the form flag failed");
    *form = (int)v;
    return PMX_OK;
}

}  // extern "C"
