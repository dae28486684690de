// spheres.hip -- seeded overlapping-sphere packs rasterised on the GPU.
//
// The reference rasterises spheres by testing voxel centroids against
// each sphere (gen_pm_spheres_domain, domain_generation.py:125-155, with
// Sphere::contains, particles/spheres.hpp:42-48: dx^2+dy^2+dz^2 <= r^2).
// Packs are grown by adding spheres in order until the porosity falls to
// a target, which is sequential as written; here it is made parallel and
// exact by recording, for every voxel, the index of the FIRST sphere that
// covers it.  The porosity after k spheres is then the fraction of voxels
// whose first cover is >= k, so the stopping index follows from a
// histogram of first covers, and the image is (first >= K).
#include "common.hpp"

namespace pmx {
namespace {

struct SphereBox {
    int64_t lo[3], hi[3];
    bool full[3];
};

// Bounding box of the voxels whose centroids can lie inside the sphere,
// clamped on walls, whole period on a periodic axis it would overlap.
__device__ __forceinline__ SphereBox sphere_box(const double *sp, const int64_t *n,
                                                const int32_t *periodic)
{
    SphereBox b;
    const double r = sp[3];
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = (int64_t)floor(sp[a] - r - 0.5) - 1;
        b.hi[a] = (int64_t)ceil(sp[a] + r - 0.5) + 1;
        b.full[a] = false;
        if (!periodic[a]) {
            if (b.lo[a] < 0) b.lo[a] = 0;
            if (b.hi[a] > n[a] - 1) b.hi[a] = n[a] - 1;
        } else if (b.hi[a] - b.lo[a] + 1 >= n[a]) {
            b.lo[a] = 0;
            b.hi[a] = n[a] - 1;
            b.full[a] = true;
        }
    }
    return b;
}

// Offset of voxel coordinate c (unwrapped) from the sphere centre, and
// its wrapped index; a whole-period axis uses the minimum image.
__device__ __forceinline__ double axis_offset(int64_t c, double centre, int64_t n, bool full,
                                              int64_t *wrapped)
{
    const int64_t w = ((c % n) + n) % n;
    *wrapped = w;
    if (full) {
        double d = (double)w + 0.5 - centre;
        return d - (double)n * rint(d / (double)n);
    }
    return (double)c + 0.5 - centre;
}

struct Dims {
    int64_t n[3];
    int32_t periodic[3];
};

template <bool kCount>
__global__ __launch_bounds__(256) void k_spheres(const double *__restrict__ spheres, Dims d,
                                                 int32_t *__restrict__ first,
                                                 int64_t *__restrict__ hist)
{
    const int64_t s = blockIdx.x;
    const double *sp = spheres + 4 * s;
    const SphereBox b = sphere_box(sp, d.n, d.periodic);
    const int64_t ex = b.hi[0] - b.lo[0] + 1, ey = b.hi[1] - b.lo[1] + 1, ez = b.hi[2] - b.lo[2] + 1;
    const double r2 = sp[3] * sp[3];
    int64_t count = 0;
    if (ex > 0 && ey > 0 && ez > 0) {
        const int64_t total = ex * ey * ez;
        for (int64_t i = threadIdx.x; i < total; i += blockDim.x) {
            const int64_t cz = i % ez, t = i / ez, cy = t % ey, cx = t / ey;
            int64_t wx, wy, wz;
            const double dx = axis_offset(b.lo[0] + cx, sp[0], d.n[0], b.full[0], &wx);
            const double dy = axis_offset(b.lo[1] + cy, sp[1], d.n[1], b.full[1], &wy);
            const double dz = axis_offset(b.lo[2] + cz, sp[2], d.n[2], b.full[2], &wz);
            const double dxy = dx * dx + dy * dy;
            if (dxy + dz * dz <= r2) {
                const int64_t v = (wx * d.n[1] + wy) * d.n[2] + wz;
                if (kCount)
                    count += first[v] == (int32_t)s;
                else
                    atomicMin(first + v, (int32_t)s);
            }
        }
    }
    if (kCount) {
        __shared__ int64_t red[256];
        red[threadIdx.x] = count;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) hist[s] = red[0];
    }
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t *a, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

__global__ __launch_bounds__(256) void k_sphere_image(const int32_t *__restrict__ first, int64_t n,
                                                      int32_t k, uint8_t *__restrict__ img)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        img[i] = first[i] >= k ? 1 : 0;
}

// Spheres in the domain's own coordinates on a grid of voxel centroids
// (gen_pm_shape, _domain_generation.pyx:30-70, with Sphere::contains,
// particles/spheres.hpp:42-48): img starts at `outside` and every voxel
// whose centroid (x[i], y[j], z[k]) satisfies (cx-x)^2 + (cy-y)^2 + (cz-z)^2
// <= r*r in float64, summed in that order, becomes 1 - outside.  One
// workgroup per sphere over the candidate index box, from the nominal
// uniform grid with a one-voxel margin; the test uses the coordinate arrays.
struct Grid {
    const double *c[3];
    int64_t n[3];
    double lo[3], h[3];  // centroid i sits near lo + (i + 0.5) h
};

__device__ __forceinline__ void index_range(double cc, double r, int a, const Grid &g, int64_t *lo, int64_t *hi)
{
    // clamped to [-2, n + 2] before the integer conversion (far spheres stay in range)
    const double top = (double)g.n[a] + 2.0;
    const double f0 = fmin(fmax((cc - r - g.lo[a]) / g.h[a] - 0.5, -2.0), top);
    const double f1 = fmin(fmax((cc + r - g.lo[a]) / g.h[a] - 0.5, -2.0), top);
    *lo = max((int64_t)floor(f0) - 1, (int64_t)0);
    *hi = min((int64_t)ceil(f1) + 1, g.n[a] - 1);
}

__global__ __launch_bounds__(256) void k_spheres_grid(const double *__restrict__ spheres, int64_t n_spheres, Grid g,
                                                      uint8_t inside, uint8_t *__restrict__ img)
{
    for (int64_t s = blockIdx.x; s < n_spheres; s += gridDim.x) {
        const double cx = spheres[4 * s], cy = spheres[4 * s + 1], cz = spheres[4 * s + 2], r = spheres[4 * s + 3];
        const double r2 = r * r;
        int64_t lo[3], hi[3];
        index_range(cx, r, 0, g, &lo[0], &hi[0]);
        index_range(cy, r, 1, g, &lo[1], &hi[1]);
        index_range(cz, r, 2, g, &lo[2], &hi[2]);
        if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) continue;
        const int64_t ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
        const int64_t total = ex * ey * ez;
        for (int64_t t = threadIdx.x; t < total; t += blockDim.x) {
            const int64_t k = lo[2] + t % ez, j = lo[1] + (t / ez) % ey, i = lo[0] + t / (ez * ey);
            const double dx = cx - g.c[0][i], dy = cy - g.c[1][j], dz = cz - g.c[2][k];
            if (dx * dx + dy * dy + dz * dz <= r2) img[(i * g.n[1] + j) * g.n[2] + k] = inside;
        }
    }
}

__global__ __launch_bounds__(256) void k_fill_u8(uint8_t *a, int64_t n, uint8_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

int check_dims(const int64_t *shape, const int32_t *periodic, Dims *d)
{
    if (!shape || !periodic) return fail(PMX_ERR_INVALID, "NULL shape/periodic");
    for (int a = 0; a < 3; ++a) {
        if (shape[a] < 1 || shape[a] > (1 << 20)) return fail(PMX_ERR_INVALID, "bad shape[%d]", a);
        d->n[a] = shape[a];
        d->periodic[a] = periodic[a] ? 1 : 0;
    }
    return PMX_OK;
}

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_sphere_first_cover(const double *spheres, int64_t n_spheres, const int64_t *shape,
                           const int32_t *periodic, int32_t *first, int64_t *counts,
                           hipStream_t stream)
{
    pmx::clear_error();
    pmx::Dims d;
    if (int e = pmx::check_dims(shape, periodic, &d)) return e;
    if (!spheres || !first || !counts) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    if (n_spheres < 1 || n_spheres > (1 << 24))
        return pmx::fail(PMX_ERR_INVALID, "n_spheres=%lld out of range", (long long)n_spheres);
    const int64_t nv = d.n[0] * d.n[1] * d.n[2];
    hipLaunchKernelGGL(pmx::k_fill_i32, dim3(4096), dim3(256), 0, stream, first, nv, INT32_MAX);
    if (int e = pmx::check_launch("first-cover fill")) return e;
    hipLaunchKernelGGL(pmx::k_spheres<false>, dim3((unsigned)n_spheres), dim3(256), 0, stream,
                       spheres, d, first, counts);
    if (int e = pmx::check_launch("sphere first cover")) return e;
    hipLaunchKernelGGL(pmx::k_spheres<true>, dim3((unsigned)n_spheres), dim3(256), 0, stream,
                       spheres, d, first, counts);
    return pmx::check_launch("sphere cover counts");
}

int pmx_sphere_image(const int32_t *first, const int64_t *shape, int32_t k, uint8_t *img,
                     hipStream_t stream)
{
    pmx::clear_error();
    if (!first || !img || !shape) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    const int64_t nv = shape[0] * shape[1] * shape[2];
    if (nv < 1) return pmx::fail(PMX_ERR_INVALID, "empty shape");
    hipLaunchKernelGGL(pmx::k_sphere_image, dim3(4096), dim3(256), 0, stream, first, nv, k, img);
    return pmx::check_launch("sphere image");
}

int pmx_spheres_on_grid(const double *spheres, int64_t n_spheres, const double *x, const double *y,
                        const double *z, const int64_t *shape, const double *grid_lo, const double *spacing,
                        int invert, uint8_t *img, hipStream_t stream)
{
    pmx::clear_error();
    if (!x || !y || !z || !shape || !grid_lo || !spacing || !img) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    if (n_spheres < 0 || (n_spheres > 0 && !spheres)) return pmx::fail(PMX_ERR_INVALID, "bad sphere list");
    pmx::Grid g;
    const double *c[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        if (shape[a] < 1 || shape[a] > (1 << 20)) return pmx::fail(PMX_ERR_INVALID, "bad shape[%d]", a);
        if (!(spacing[a] > 0.0)) return pmx::fail(PMX_ERR_INVALID, "spacing must be positive");
        g.c[a] = c[a];
        g.n[a] = shape[a];
        g.lo[a] = grid_lo[a];
        g.h[a] = spacing[a];
    }
    const int64_t nv = g.n[0] * g.n[1] * g.n[2];
    const uint8_t outside = invert ? 0 : 1;
    hipLaunchKernelGGL(pmx::k_fill_u8, dim3(4096), dim3(256), 0, stream, img, nv, outside);
    if (int e = pmx::check_launch("sphere grid fill")) return e;
    if (n_spheres == 0) return PMX_OK;
    hipLaunchKernelGGL(pmx::k_spheres_grid, dim3((unsigned)std::min<int64_t>(n_spheres, 65536)), dim3(256), 0, stream,
                       spheres, n_spheres, g, (uint8_t)(1 - outside), img);
    return pmx::check_launch("spheres on grid");
}

}  // extern "C"
