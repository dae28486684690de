// edt_ref.hip -- the reference's float32 EDT arithmetic on the GPU.
//
// The exact integer kernels (edt.hip) cover voxel-unit domains.  PMMoTo's
// default domain is the unit box, so most callers run at resolution 1/N,
// where the reference's float32 results depend on its exact sequence of
// roundings.  This path reproduces that sequence per line, one lane per
// line, so outputs are bit-identical at any resolution:
//
//   pass 0 (axis a0)  squared_edt_1d<uint8_t> (_distance.hpp:135-182):
//                     label-aware forward sweep d += res, corrector seeding
//                     and capping, backward fminf(d, d_next + res), square,
//                     INF -> FLT_MAX (to_finite, :22-32); periodic
//                     correctors N - last_zero / first_zero + 1
//                     (_distance.pyx:240-279).
//   passes a1, a2     squared_edt_1d_parabolic (:184-272) with float
//                     numerators and the double-precision division of
//                     intersect_parabolas (:60-72), pop test s <= range
//                     (:89-112), evaluation aniso * sq(i - v) + h; periodic
//                     lines first get up to num_hull + 1 = 5 boundary
//                     parabolas from each face (return_boundary_hull,
//                     :296-356, via euclidean_distance_transform.py:246-298),
//                     all taken from the pass input before any line changes.
//
// The steps are the reference's because bit-parity requires the same float
// operations in the same order; the code is a per-line GPU formulation
// with stacks in the caller's workspace.  Compiled with contraction off
// (the reference's x86-64 build has no FMA), so a*b+c rounds twice.
#include "common.hpp"

namespace pmx {
namespace {

constexpr float kFltMax = 3.40282347e+38f;
constexpr int kNumHull = 4;            // euclidean_distance_transform.py:275,284
constexpr int kHullCap = kNumHull + 1; // members returned per face
constexpr int kRefThreads = 8192;      // persistent lanes of the line kernels

struct RefHull {
    int vertex;
    float height;
    float range;
};

struct LineGeom {
    int64_t n[3];
    int64_t s[3];  // element strides of the float array (z stride 1)
    int64_t is[3]; // element strides of the uint8 image
    int dim, d1, d2;
    __host__ __device__ int64_t lines() const { return n[d1] * n[d2]; }
};

__device__ __forceinline__ void line_base(const LineGeom &g, int64_t line, int64_t *fb, int64_t *ib)
{
    const int64_t a = line / g.n[g.d2], b = line - a * g.n[g.d2];
    *fb = a * g.s[g.d1] + b * g.s[g.d2];
    *ib = a * g.is[g.d1] + b * g.is[g.d2];
}

// first (forward) / last index of value 0 on the line, or -1
__device__ int nearest_zero(const uint8_t *img, int n, int64_t stride, bool forward)
{
    if (forward) {
        for (int i = 0; i < n; ++i)
            if (img[(int64_t)i * stride] == 0) return i;
    } else {
        for (int i = n - 1; i >= 0; --i)
            if (img[(int64_t)i * stride] == 0) return i;
    }
    return -1;
}

// pass 0 along g.dim, in place on `d` (one lane per line)
__global__ __launch_bounds__(256) void k_ref_pass0(const uint8_t *__restrict__ img, float *__restrict__ d,
                                                   LineGeom g, float res, int periodic)
{
    const int n = (int)g.n[g.dim];
    const int64_t st = g.s[g.dim], ist = g.is[g.dim];
    for (int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; line < g.lines();
         line += (int64_t)gridDim.x * blockDim.x) {
        int64_t fb, ib;
        line_base(g, line, &fb, &ib);
        const uint8_t *im = img + ib;
        float *o = d + fb;
        float lo = INFINITY, hi = INFINITY;
        if (periodic) {
            const int first = nearest_zero(im, n, ist, true);
            const int last = nearest_zero(im, n, ist, false);
            lo = last != -1 ? (float)((double)n - (double)last) : INFINITY;
            hi = first != -1 ? (float)((double)first + 1.0) : INFINITY;
        }
        uint8_t seg = im[0];
        o[0] = seg == 0 ? 0.0f : lo * res;
        for (int i = 1; i < n; ++i) {
            const uint8_t v = im[(int64_t)i * ist];
            if (v == 0) {
                o[(int64_t)i * st] = 0.0f;
            } else if (v == seg) {
                o[(int64_t)i * st] = o[(int64_t)(i - 1) * st] + res;
            } else {
                o[(int64_t)i * st] = res;
                o[(int64_t)(i - 1) * st] = (float)(im[(int64_t)(i - 1) * ist] != 0) * res;
                seg = v;
            }
        }
        if (o[(int64_t)(n - 1) * st] > hi * res) o[(int64_t)(n - 1) * st] = hi * res;
        for (int i = n - 2; i >= 0; --i)
            o[(int64_t)i * st] = fminf(o[(int64_t)i * st], o[(int64_t)(i + 1) * st] + res);
        for (int i = 0; i < n; ++i) {
            float v = o[(int64_t)i * st];
            v = v * v;
            o[(int64_t)i * st] = v == INFINITY ? kFltMax : v;
        }
    }
}

__device__ __forceinline__ float intersect(float f_r, float r, float f_q, float q, float aniso)
{
    const float factor1 = (r - q) * aniso;
    const float factor2 = r + q;
    const float num = f_r - f_q + factor1 * factor2;
    return (float)((double)num / (2.0 * (double)factor1));
}

// A lane's stack in the workspace: entry k at base[k * step].
struct RefStack {
    int *hv;
    float *hh, *rg, *ff;
    int64_t step;
    __device__ __forceinline__ int &V(int k) const { return hv[(int64_t)k * step]; }
    __device__ __forceinline__ float &H(int k) const { return hh[(int64_t)k * step]; }
    __device__ __forceinline__ float &R(int k) const { return rg[(int64_t)k * step]; }
    __device__ __forceinline__ float &F(int k) const { return ff[(int64_t)k * step]; }
};

__device__ __forceinline__ void update_hull(const RefStack &S, int &k, int i, float fv, float aniso)
{
    float s = intersect(fv, (float)i, S.H(k), (float)S.V(k), aniso);
    while (k > 0 && s <= S.R(k)) {
        --k;
        s = intersect(fv, (float)i, S.H(k), (float)S.V(k), aniso);
    }
    ++k;
    S.V(k) = i;
    S.H(k) = fv;
    S.R(k) = s;
    S.R(k + 1) = INFINITY;
}

// return_boundary_hull on n values at p (stride st); writes <= kHullCap
// members to out, returns the count.  A one-voxel segment with no finite
// value reads past the reference's vectors there (undefined behaviour);
// this path returns the zero-initialised slot, as the oracle does.
__device__ int boundary_hull(const float *p, int n, int64_t st, float res, int index_corrector,
                             bool forward, const RefStack &S, RefHull *out)
{
    if (n == 1 && p[0] < kFltMax) {
        out[0] = {index_corrector, p[0], INFINITY};
        return 1;
    }
    const float aniso = res * res;
    for (int i = 0; i < n; ++i) S.F(i) = p[(int64_t)i * st];
    S.V(0) = 0;
    S.H(0) = S.F(0);
    S.V(1) = 0;
    S.H(1) = 0.0f;
    S.R(0) = -INFINITY;
    S.R(1) = INFINITY;
    int k = 1;
    for (int i = 1; i < n; ++i) update_hull(S, k, i, S.F(i), aniso);
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

// hulls[line][side][kHullCap], counts[line][side]; side 0 = forward
// (segment from the lower face to the first zero), 1 = backward
__global__ __launch_bounds__(256) void k_ref_hulls(const uint8_t *__restrict__ og, const float *__restrict__ d,
                                                   LineGeom g, float res, int *__restrict__ stk, int cap,
                                                   RefHull *__restrict__ hulls, int *__restrict__ counts)
{
    const int n = (int)g.n[g.dim];
    const int64_t st = g.s[g.dim], ist = g.is[g.dim];
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    RefStack S{stk + tid, reinterpret_cast<float *>(stk) + (int64_t)cap * T + tid,
               reinterpret_cast<float *>(stk) + (int64_t)2 * cap * T + tid,
               reinterpret_cast<float *>(stk) + (int64_t)3 * cap * T + tid, T};
    for (int64_t line = tid; line < g.lines(); line += T) {
        int64_t fb, ib;
        line_base(g, line, &fb, &ib);
        for (int side = 0; side < 2; ++side) {
            const bool fw = side == 0;
            const int bound = nearest_zero(og + ib, n, ist, fw);
            int start, nv, ic;
            if (bound == -1) {
                start = 0;
                nv = n;
                ic = fw ? 0 : -nv;
            } else if (fw) {
                start = 0;
                nv = bound + 1;
                ic = 0;
            } else {
                start = bound;
                nv = n - bound;
                ic = -nv;
            }
            counts[line * 2 + side] = boundary_hull(d + fb + (int64_t)start * st, nv, st, res, ic, fw, S,
                                                    hulls + (line * 2 + side) * kHullCap);
        }
    }
}

// squared_edt_1d_parabolic on each line, with the swapped face hulls when
// `hulls` is given (lower <- backward hull, upper <- forward hull)
__global__ __launch_bounds__(256) void k_ref_parabolic(float *__restrict__ d, LineGeom g, float res,
                                                       int *__restrict__ stk, int cap,
                                                       const RefHull *__restrict__ hulls,
                                                       const int *__restrict__ counts)
{
    const int n = (int)g.n[g.dim];
    const int64_t st = g.s[g.dim];
    const float aniso = res * res;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = (int64_t)gridDim.x * blockDim.x;
    RefStack S{stk + tid, reinterpret_cast<float *>(stk) + (int64_t)cap * T + tid,
               reinterpret_cast<float *>(stk) + (int64_t)2 * cap * T + tid,
               reinterpret_cast<float *>(stk) + (int64_t)3 * cap * T + tid, T};
    for (int64_t line = tid; line < g.lines(); line += T) {
        int64_t fb, ib;
        line_base(g, line, &fb, &ib);
        float *o = d + fb;
        for (int i = 0; i < n; ++i) S.F(i) = o[(int64_t)i * st];
        const RefHull *lower = hulls ? hulls + (line * 2 + 1) * kHullCap : nullptr;
        const RefHull *upper = hulls ? hulls + (line * 2 + 0) * kHullCap : nullptr;
        const int nl = hulls ? counts[line * 2 + 1] : 0;
        const int nu = hulls ? counts[line * 2 + 0] : 0;
        int k = 0, loop_start = 1;
        S.R(0) = -INFINITY;
        S.R(1) = INFINITY;
        if (nl > 0) {
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
            S.H(0) = S.F(0);
        }
        for (int i = loop_start; i < n; ++i) update_hull(S, k, i, S.F(i), aniso);
        for (int m = 0; m < nu; ++m) update_hull(S, k, upper[m].vertex + n, upper[m].height, aniso);
        k = 0;
        for (int i = 0; i < n; ++i) {
            while (S.R(k + 1) < (float)i) ++k;
            const float di = (float)(i - S.V(k));
            o[(int64_t)i * st] = aniso * (di * di) + S.H(k);
        }
    }
}

__global__ __launch_bounds__(256) void k_sqrt_inplace(float *d, int64_t n0, int64_t n1, int64_t n2, int64_t s0,
                                                      int64_t s1)
{
    const int64_t total = n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = i % n2, t = i / n2, y = t % n1, x = t / n1;
        float *p = d + x * s0 + y * s1 + z;
        *p = sqrtf(*p);
    }
}

LineGeom make_line_geom(const int64_t *n, const int64_t *fs, const int64_t *is, int dim)
{
    LineGeom g;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = n[a];
        g.s[a] = fs[a];
        g.is[a] = is[a];
    }
    g.dim = dim;
    int o[2], c = 0;
    for (int a = 0; a < 3; ++a)
        if (a != dim) o[c++] = a;
    g.d1 = o[0];
    g.d2 = o[1];
    return g;
}

size_t ref_stack_ints(int64_t nmax) { return (size_t)4 * (size_t)(nmax + 2 * kHullCap + 2) * kRefThreads; }

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_edt_ref_workspace_size(const int64_t *n, size_t *bytes)
{
    pmx::clear_error();
    if (!n || !bytes) return pmx::fail(PMX_ERR_INVALID, "NULL argument");
    int64_t nmax = 0, lines = 0;
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1 || n[a] > 16384) return pmx::fail(PMX_ERR_INVALID, "extent n[%d] outside [1, 16384]", a);
        nmax = std::max(nmax, n[a]);
        lines = std::max(lines, n[0] * n[1] * n[2] / n[a]);
    }
    pmx::Carver c(nullptr, SIZE_MAX);
    c.take<int>(pmx::ref_stack_ints(nmax));
    c.take<pmx::RefHull>((size_t)lines * 2 * pmx::kHullCap);
    c.take<int>((size_t)lines * 2);
    *bytes = c.need();
    return PMX_OK;
}

int pmx_edt_ref(const uint8_t *img, const int64_t *n, const int64_t *in_s, float *d, const int64_t *out_s,
                int first_axis, const int32_t *periodic, const float *resolution, int take_sqrt,
                void *workspace, size_t workspace_bytes, hipStream_t s)
{
    using namespace pmx;
    clear_error();
    if (!img || !n || !in_s || !d || !out_s || !periodic || !resolution)
        return fail(PMX_ERR_INVALID, "NULL argument");
    if (first_axis < 0 || first_axis > 2) return fail(PMX_ERR_INVALID, "first_axis must be 0, 1 or 2");
    size_t need = 0;
    if (int e = pmx_edt_ref_workspace_size(n, &need)) return e;
    if (!workspace || workspace_bytes < need)
        return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    int64_t nmax = std::max(n[0], std::max(n[1], n[2]));
    int64_t lines = 0;
    for (int a = 0; a < 3; ++a) lines = std::max(lines, n[0] * n[1] * n[2] / n[a]);
    Carver c(workspace, workspace_bytes);
    int *stk = c.take<int>(ref_stack_ints(nmax));
    RefHull *hulls = c.take<RefHull>((size_t)lines * 2 * kHullCap);
    int *counts = c.take<int>((size_t)lines * 2);
    const int cap = (int)(nmax + 2 * kHullCap + 2);
    const int64_t fs[3] = {out_s[0], out_s[1], 1}, is[3] = {in_s[0], in_s[1], 1};
    const dim3 blk(256), grd(kRefThreads / 256);
    // pass order: first_axis with the 1-D sweep, then the remaining axes in
    // ascending order (edt3d: 0 then 1, 2; edt2d on a (1, X, Y) box: 2 then 1, 0)
    {
        ProfScope ps("edt_ref_pass0", s);
        LineGeom g = make_line_geom(n, fs, is, first_axis);
        hipLaunchKernelGGL(k_ref_pass0, grd, blk, 0, s, img, d, g, resolution[first_axis],
                           periodic[first_axis] ? 1 : 0);
    }
    if (int e = check_launch("reference pass 0")) return e;
    for (int dim = 0; dim < 3; ++dim) {
        if (dim == first_axis) continue;
        LineGeom g = make_line_geom(n, fs, is, dim);
        ProfScope ps("edt_ref_parabolic", s);
        if (periodic[dim]) {
            hipLaunchKernelGGL(k_ref_hulls, grd, blk, 0, s, img, d, g, resolution[dim], stk, cap, hulls, counts);
            if (int e = check_launch("reference hulls")) return e;
        }
        hipLaunchKernelGGL(k_ref_parabolic, grd, blk, 0, s, d, g, resolution[dim], stk, cap,
                           periodic[dim] ? hulls : nullptr, counts);
        if (int e = check_launch("reference parabolic")) return e;
    }
    if (take_sqrt) {
        hipLaunchKernelGGL(k_sqrt_inplace, dim3(4096), blk, 0, s, d, n[0], n[1], n[2], out_s[0], out_s[1]);
        return check_launch("reference sqrt");
    }
    return PMX_OK;
}

}  // extern "C"
