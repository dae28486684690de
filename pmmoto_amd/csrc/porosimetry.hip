// porosimetry.hip -- element-wise steps of the porosimetry / drainage sweeps.
//
// pore_size_distribution (src/pmmoto/filters/porosimetry.py:163-227) keeps,
// per voxel, the first (largest) radius whose porosimetry image covers it:
//   img_results = where(pm.img == 1 & img_results == 0 & img_temp == 1, r, img_results)
// Here the map holds radius codes (index + 1 into the descending radius
// list, 0 = never invaded) so the sweep stays integer; pmx_codes_to_values
// writes the float64 radius map at the end, and pmx_code_histogram counts
// own pore voxels per code for saturation curves.
#include "common.hpp"

namespace pmx {
namespace {

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 65536); }

__global__ __launch_bounds__(256) void k_psd_update(const uint8_t *__restrict__ pm,
                                                    const uint8_t *__restrict__ temp, int64_t n,
                                                    int32_t code, int32_t *__restrict__ map)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (pm[i] == 1 && temp[i] == 1 && map[i] == 0) map[i] = code;
}

__global__ __launch_bounds__(256) void k_codes_to_values(const int32_t *__restrict__ map, int64_t n,
                                                         const double *__restrict__ values,
                                                         int32_t n_values, double *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t c = map[i];
        out[i] = (c > 0 && c <= n_values) ? values[c - 1] : 0.0;
    }
}

// Wetting saturation after each code from the code histogram, in float64
// in the reference's order: S_i = (wet_i / V) / (pore / V) with wet_i the
// pore voxels whose code is 0 or above i + 1 (Multiphase.get_saturation =
// get_volume_fraction / porosity, multiphase.py:56-95, porousmedia.py:56-66).
// One lane walks the codes (at most 4096); a domain with no pore voxel gives
// 0 / 0 = NaN, as the reference's numpy division does.
__global__ void k_saturation(const unsigned long long *__restrict__ counts, int32_t n_codes, double volume,
                             double *__restrict__ curve)
{
    if (threadIdx.x != 0) return;
    unsigned long long pore = 0;
    for (int c = 0; c <= n_codes; ++c) pore += counts[c];
    const double porosity = (double)pore / volume;
    unsigned long long wet = pore;
    for (int i = 0; i < n_codes; ++i) {
        wet -= counts[i + 1];
        curve[i] = ((double)wet / volume) / porosity;
    }
}

// counts[c] += number of own voxels with pm == 1 and map == c
__global__ __launch_bounds__(256) void k_code_hist(const uint8_t *__restrict__ pm,
                                                   const int32_t *__restrict__ map, int64_t n1,
                                                   int64_t n2, int64_t lo0, int64_t lo1, int64_t lo2,
                                                   int64_t e0, int64_t e1, int64_t e2, int32_t n_codes,
                                                   unsigned long long *__restrict__ counts)
{
    extern __shared__ unsigned long long sh[];
    for (int c = threadIdx.x; c <= n_codes; c += blockDim.x) sh[c] = 0;
    __syncthreads();
    const int64_t n = e0 * e1 * e2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = i % e2, t = i / e2, y = t % e1, x = t / e1;
        const int64_t v = ((lo0 + x) * n1 + lo1 + y) * n2 + lo2 + z;
        if (pm[v] == 1) {
            int32_t c = map[v];
            if (c < 0 || c > n_codes) c = 0;
            atomicAdd(&sh[c], 1ull);
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c <= n_codes; c += blockDim.x)
        if (sh[c]) atomicAdd(&counts[c], sh[c]);
}

__global__ __launch_bounds__(256) void k_both_one(const uint8_t *__restrict__ a,
                                                  const uint8_t *__restrict__ b, int64_t n,
                                                  uint8_t *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (a[i] == 1 && b[i] == 1) ? 1 : 0;
}

// drainage step (equilibrium_distribution.py:98-101): img = 1 where the
// morphology image is 1 and the voxel holds outlet-connected wetting phase
__global__ __launch_bounds__(256) void k_drain_update(const uint8_t *__restrict__ morph,
                                                     const uint8_t *__restrict__ wet, int64_t n,
                                                     uint8_t *__restrict__ img)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        if (morph[i] == 1 && wet[i] == 2) img[i] = 1;
}

__global__ __launch_bounds__(256) void k_box_op(uint8_t *__restrict__ a, int64_t n1, int64_t n2, int64_t lo0,
                                                int64_t lo1, int64_t lo2, int64_t e0, int64_t e1, int64_t e2,
                                                int op, uint8_t value, unsigned long long *__restrict__ count)
{
    const int64_t n = e0 * e1 * e2;
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = i % e2, t = i / e2, y = t % e1, x = t / e1;
        uint8_t *p = a + ((lo0 + x) * n1 + lo1 + y) * n2 + lo2 + z;
        if (op == 0)
            *p = value;
        else
            c += *p == value;
    }
    if (op != 0) {
        for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
        if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
    }
}

// max of non-negative float32 values (IEEE order equals integer order)
__global__ __launch_bounds__(256) void k_max_f32(const float *__restrict__ a, int64_t n, unsigned int *__restrict__ out)
{
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = fmaxf(m, a[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// max of n bytes: 16 per lane per step (the tail byte by byte), wave max, one atomic per wave
__global__ __launch_bounds__(256) void k_max_u8(const uint8_t *__restrict__ a, int64_t n, unsigned int *__restrict__ out)
{
    unsigned int m = 0u;
    const int64_t nv = n / 16;
    const uint4 *v = reinterpret_cast<const uint4 *>(a);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
        const uint4 q = v[i];
        const unsigned int w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            m = max(m, max(max(w[k] & 0xffu, (w[k] >> 8) & 0xffu), max((w[k] >> 16) & 0xffu, w[k] >> 24)));
    }
    for (int64_t i = 16 * nv + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, (unsigned int)a[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(out, m);
}

__global__ __launch_bounds__(64) void k_fill_u32(unsigned int *a, int64_t n)
{
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) a[i] = 0;
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long *a, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_psd_update(const uint8_t *pm, const uint8_t *temp, int64_t n, int32_t code, int32_t *map,
                   hipStream_t s)
{
    pmx::clear_error();
    if (!pm || !temp || !map || n < 0 || code <= 0) return pmx::fail(PMX_ERR_INVALID, "bad argument");
    if (n == 0) return PMX_OK;
    hipLaunchKernelGGL(pmx::k_psd_update, dim3(pmx::grid_for(n)), dim3(256), 0, s, pm, temp, n, code, map);
    return pmx::check_launch("psd update");
}

int pmx_codes_to_values(const int32_t *map, int64_t n, const double *values, int32_t n_values,
                        double *out, hipStream_t s)
{
    pmx::clear_error();
    if (!map || !values || !out || n < 0 || n_values < 0) return pmx::fail(PMX_ERR_INVALID, "bad argument");
    if (n == 0) return PMX_OK;
    hipLaunchKernelGGL(pmx::k_codes_to_values, dim3(pmx::grid_for(n)), dim3(256), 0, s, map, n, values,
                       n_values, out);
    return pmx::check_launch("codes to values");
}

int pmx_code_histogram(const uint8_t *pm, const int32_t *map, const int64_t *shape, const int64_t *lo,
                       const int64_t *extent, int32_t n_codes, uint64_t *counts, hipStream_t s)
{
    pmx::clear_error();
    if (!pm || !map || !shape || !lo || !extent || !counts || n_codes < 0 || n_codes > 4096)
        return pmx::fail(PMX_ERR_INVALID, "bad argument");
    for (int a = 0; a < 3; ++a)
        if (lo[a] < 0 || extent[a] < 0 || lo[a] + extent[a] > shape[a])
            return pmx::fail(PMX_ERR_INVALID, "sub-box outside the array on axis %d", a);
    auto *c = reinterpret_cast<unsigned long long *>(counts);
    hipLaunchKernelGGL(pmx::k_fill_u64, dim3(1), dim3(256), 0, s, c, (int64_t)n_codes + 1);
    const int64_t n = extent[0] * extent[1] * extent[2];
    if (n == 0) return pmx::check_launch("histogram clear");
    const size_t shm = sizeof(unsigned long long) * (size_t)(n_codes + 1);
    hipLaunchKernelGGL(pmx::k_code_hist, dim3(std::min<unsigned>(pmx::grid_for(n), 2048)), dim3(256), shm, s,
                       pm, map, shape[1], shape[2], lo[0], lo[1], lo[2], extent[0], extent[1], extent[2],
                       n_codes, c);
    return pmx::check_launch("code histogram");
}

int pmx_saturation_curve(const uint8_t *pm, const int32_t *map, const int64_t *shape, const int64_t *lo,
                         const int64_t *extent, int32_t n_codes, double volume, uint64_t *counts, double *curve,
                         hipStream_t s)
{
    pmx::clear_error();
    if (!curve || !(volume > 0.0)) return pmx::fail(PMX_ERR_INVALID, "curve is NULL or volume is not positive");
    if (int e = pmx_code_histogram(pm, map, shape, lo, extent, n_codes, counts, s)) return e;
    hipLaunchKernelGGL(pmx::k_saturation, dim3(1), dim3(64), 0, s,
                       reinterpret_cast<const unsigned long long *>(counts), n_codes, volume, curve);
    return pmx::check_launch("saturation curve");
}

int pmx_drain_update(const uint8_t *morph, const uint8_t *wet, int64_t n, uint8_t *img, hipStream_t s)
{
    pmx::clear_error();
    if (!morph || !wet || !img || n < 0) return pmx::fail(PMX_ERR_INVALID, "bad argument");
    if (n == 0) return PMX_OK;
    hipLaunchKernelGGL(pmx::k_drain_update, dim3(pmx::grid_for(n)), dim3(256), 0, s, morph, wet, n, img);
    return pmx::check_launch("drainage update");
}

// op 0: fill the sub-box with value; op 1: *count = number of voxels == value
int pmx_box_op(uint8_t *arr, const int64_t *shape, const int64_t *lo, const int64_t *extent, int op,
               uint8_t value, uint64_t *count, hipStream_t s)
{
    pmx::clear_error();
    if (!arr || !shape || !lo || !extent || (op != 0 && op != 1) || (op == 1 && !count))
        return pmx::fail(PMX_ERR_INVALID, "bad argument");
    for (int a = 0; a < 3; ++a)
        if (lo[a] < 0 || extent[a] < 0 || lo[a] + extent[a] > shape[a])
            return pmx::fail(PMX_ERR_INVALID, "sub-box outside the array on axis %d", a);
    auto *c = reinterpret_cast<unsigned long long *>(count);
    if (op == 1) {
        hipLaunchKernelGGL(pmx::k_fill_u64, dim3(1), dim3(64), 0, s, c, (int64_t)1);
        if (int e = pmx::check_launch("count clear")) return e;
    }
    const int64_t n = extent[0] * extent[1] * extent[2];
    if (n == 0) return PMX_OK;
    hipLaunchKernelGGL(pmx::k_box_op, dim3(std::min<unsigned>(pmx::grid_for(n), 4096)), dim3(256), 0, s, arr,
                       shape[1], shape[2], lo[0], lo[1], lo[2], extent[0], extent[1], extent[2], op, value, c);
    return pmx::check_launch("box op");
}

int pmx_max_f32(const float *a, int64_t n, float *out, hipStream_t s)
{
    pmx::clear_error();
    if (!a || !out || n < 1) return pmx::fail(PMX_ERR_INVALID, "bad argument");
    auto *o = reinterpret_cast<unsigned int *>(out);
    hipLaunchKernelGGL(pmx::k_fill_u32, dim3(1), dim3(64), 0, s, o, (int64_t)1);
    if (int e = pmx::check_launch("max clear")) return e;
    hipLaunchKernelGGL(pmx::k_max_f32, dim3(std::min<unsigned>(pmx::grid_for(n), 4096)), dim3(256), 0, s, a, n, o);
    return pmx::check_launch("max");
}

int pmx_max_u8(const uint8_t *a, int64_t n, uint32_t *out, hipStream_t s)
{
    pmx::clear_error();
    if (!a || !out || n < 1) return pmx::fail(PMX_ERR_INVALID, "bad argument");
    if ((uintptr_t)a % 16) return pmx::fail(PMX_ERR_INVALID, "a must be 16-byte aligned");
    hipLaunchKernelGGL(pmx::k_fill_u32, dim3(1), dim3(64), 0, s, out, (int64_t)1);
    if (int e = pmx::check_launch("max clear")) return e;
    hipLaunchKernelGGL(pmx::k_max_u8, dim3(std::min<unsigned>(pmx::grid_for(n / 16 + 1), 2048)), dim3(256), 0, s, a, n,
                       out);
    return pmx::check_launch("max u8");
}

int pmx_both_one(const uint8_t *a, const uint8_t *b, int64_t n, uint8_t *out, hipStream_t s)
{
    pmx::clear_error();
    if (!a || !b || !out || n < 0) return pmx::fail(PMX_ERR_INVALID, "bad argument");
    if (n == 0) return PMX_OK;
    hipLaunchKernelGGL(pmx::k_both_one, dim3(pmx::grid_for(n)), dim3(256), 0, s, a, b, n, out);
    return pmx::check_launch("both one");
}

}  // extern "C"
