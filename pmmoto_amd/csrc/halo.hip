// halo.hip -- the reference's halo-image morphology, for non-unit resolution.
//
// At resolution 1 the fused integer kernels (pmx_morph) reproduce the
// reference exactly.  At other resolutions the reference's float32 EDT
// values depend on the extent of the image it is run on, so this path
// rebuilds that image: the padded array extended by ceil(r/res) voxels of
// periodic copies on periodic faces (update_extended_buffer,
// communication.py:121-157; no halo beyond END or 1-wide WALL faces,
// subdomain_padded.py:116-167), then pmx_edt_ref on it (non-periodic, as
// morphological_operators.py:137-141,220-222 call edt3d), then the
// float comparison with the radius, cropping and set_wall_bcs.
#include "common.hpp"

namespace pmx {
namespace {

struct Ext {
    int64_t src[3];  // padded source extent
    int64_t lo[3];   // own range start in the source, per axis
    int64_t own[3];  // own extent (period) per axis
    int64_t ext[3];  // halo width added on each side
    int64_t dst[3];  // extended extent
};

__device__ __forceinline__ int64_t map_axis(int64_t i, int a, const Ext &e)
{
    const int64_t j = i - e.ext[a];  // coordinate in the source frame
    if (j >= 0 && j < e.src[a]) return j;
    // periodic copy of the own region
    int64_t k = (j - e.lo[a]) % e.own[a];
    if (k < 0) k += e.own[a];
    return e.lo[a] + k;
}

__global__ __launch_bounds__(256) void k_extend(const uint8_t *__restrict__ src, Ext e, uint8_t *__restrict__ dst)
{
    const int64_t total = e.dst[0] * e.dst[1] * e.dst[2];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = i % e.dst[2], t = i / e.dst[2], y = t % e.dst[1], x = t / e.dst[1];
        const int64_t sx = map_axis(x, 0, e), sy = map_axis(y, 1, e), sz = map_axis(z, 2, e);
        dst[i] = src[(sx * e.src[1] + sy) * e.src[2] + sz];
    }
}

// out (shape n, contiguous) = cmp(d[ext + i], r) ? v_le : v_gt, WALL layers 0
__global__ __launch_bounds__(256) void k_crop_threshold(const float *__restrict__ d, Ext e, double r64,
                                                        float r32, int use_f64, uint8_t v_le, uint8_t v_gt,
                                                        uint32_t wall_mask, uint8_t *__restrict__ out)
{
    const int64_t total = e.src[0] * e.src[1] * e.src[2];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t z = i % e.src[2], t = i / e.src[2], y = t % e.src[1], x = t / e.src[1];
        const float v = d[((x + e.ext[0]) * e.dst[1] + y + e.ext[1]) * e.dst[2] + z + e.ext[2]];
        const bool le = use_f64 ? ((double)v <= r64) : (v <= r32);
        uint8_t r = le ? v_le : v_gt;
        if (((wall_mask & 1u) && x == 0) || ((wall_mask & 2u) && x == e.src[0] - 1) ||
            ((wall_mask & 4u) && y == 0) || ((wall_mask & 8u) && y == e.src[1] - 1) ||
            ((wall_mask & 16u) && z == 0) || ((wall_mask & 32u) && z == e.src[2] - 1))
            r = 0;
        out[i] = r;
    }
}

__global__ __launch_bounds__(256) void k_invert(const uint8_t *a, int64_t n, uint8_t *o)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        o[i] = a[i] == 0 ? 1 : 0;
}

int make_ext(const int64_t *shape, const int64_t *pad, const int64_t *ext, Ext *e)
{
    for (int a = 0; a < 3; ++a) {
        if (shape[a] < 1 || ext[a] < 0 || pad[2 * a] < 0 || pad[2 * a + 1] < 0)
            return fail(PMX_ERR_INVALID, "axis %d: bad shape/pad/halo", a);
        e->src[a] = shape[a];
        e->lo[a] = pad[2 * a];
        e->own[a] = shape[a] - pad[2 * a] - pad[2 * a + 1];
        if (e->own[a] < 1) return fail(PMX_ERR_INVALID, "axis %d: no own voxels", a);
        e->ext[a] = ext[a];
        e->dst[a] = shape[a] + 2 * ext[a];
        if (e->dst[a] > 16384) return fail(PMX_ERR_INVALID, "axis %d: extended extent above 16384", a);
    }
    return PMX_OK;
}

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_halo_morph_workspace_size(const int64_t *shape, const int64_t *ext, size_t *bytes)
{
    pmx::clear_error();
    if (!shape || !ext || !bytes) return pmx::fail(PMX_ERR_INVALID, "NULL argument");
    int64_t n[3];
    for (int a = 0; a < 3; ++a) n[a] = shape[a] + 2 * ext[a];
    size_t ref = 0;
    if (int e = pmx_edt_ref_workspace_size(n, &ref)) return e;
    const size_t vox = (size_t)(n[0] * n[1] * n[2]);
    pmx::Carver c(nullptr, SIZE_MAX);
    c.take<uint8_t>(vox);
    c.take<float>(vox);
    c.take<char>(ref);
    *bytes = c.need();
    return PMX_OK;
}

int pmx_halo_morph(const uint8_t *img, const int64_t *shape, const int64_t *pad, const int64_t *halo_width,
                   const float *resolution, int invert, double radius, int compare_f64, uint8_t v_le,
                   uint8_t v_gt, uint32_t wall_mask, uint8_t *out, void *workspace, size_t workspace_bytes,
                   hipStream_t s)
{
    using namespace pmx;
    clear_error();
    if (!img || !shape || !pad || !halo_width || !resolution || !out)
        return fail(PMX_ERR_INVALID, "NULL argument");
    const int64_t *ext = halo_width;
    Ext e;
    if (int rc = make_ext(shape, pad, ext, &e)) return rc;
    for (int a = 0; a < 3; ++a)
        if (ext[a] > e.own[a])
            return fail(PMX_ERR_INVALID, "axis %d: halo %lld exceeds own voxels %lld", a, (long long)ext[a],
                        (long long)e.own[a]);
    size_t need = 0;
    if (int rc = pmx_halo_morph_workspace_size(shape, ext, &need)) return rc;
    if (!workspace || workspace_bytes < need)
        return fail(PMX_ERR_WORKSPACE, "workspace %zu bytes < required %zu", workspace_bytes, need);
    const size_t vox = (size_t)(e.dst[0] * e.dst[1] * e.dst[2]);
    Carver c(workspace, workspace_bytes);
    uint8_t *halo = c.take<uint8_t>(vox);
    float *d = c.take<float>(vox);
    size_t ref = 0;
    pmx_edt_ref_workspace_size(e.dst, &ref);
    char *rws = c.take<char>(ref);
    hipLaunchKernelGGL(k_extend, dim3(4096), dim3(256), 0, s, img, e, halo);
    if (int rc = check_launch("halo extend")) return rc;
    if (invert) {
        hipLaunchKernelGGL(k_invert, dim3(4096), dim3(256), 0, s, halo, (int64_t)vox, halo);
        if (int rc = check_launch("halo invert")) return rc;
    }
    const int64_t st[2] = {e.dst[1] * e.dst[2], e.dst[2]};
    const int32_t none[3] = {0, 0, 0};
    if (int rc = pmx_edt_ref(halo, e.dst, st, d, st, 0, none, nullptr, resolution, 1, rws, ref, s)) return rc;
    hipLaunchKernelGGL(k_crop_threshold, dim3(4096), dim3(256), 0, s, d, e, radius, (float)radius,
                       compare_f64, v_le, v_gt, wall_mask, out);
    return check_launch("halo threshold");
}

}  // extern "C"
