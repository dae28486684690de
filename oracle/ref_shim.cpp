// ref_shim.cpp -- exposes the reference's own per-line C++ kernels
// (src/pmmoto/filters/distance/_distance.hpp, src/pmmoto/core/_voxels.hpp),
// compiled from where they lie under the reference tree, through a C ABI
// matching the oracle's per-line function table (orc_line_ops in
// pmx_oracle.c).  TEST INFRASTRUCTURE ONLY: built into oracle/_ref/ by
// oracle/Makefile when the reference tree is present; used by tests/ and
// by the CPU-baseline leg of bench.py, never by the product.
#include <cstdint>
#include <vector>

#include "_distance.hpp"
#include "_voxels.hpp"

extern "C" {

struct ref_hull {
    int vertex;
    float height;
    float range;
};

// _distance.hpp:135-182, instantiated for uint8_t as _distance.pyx:502 does.
__attribute__((visibility("default"))) void ref_sq_edt_1d(const uint8_t *img, float *d, int n,
                                                          long stride, float res, float lo,
                                                          float hi)
{
    squared_edt_1d<uint8_t>(const_cast<uint8_t *>(img), d, n, stride, res, lo, hi);
}

// _distance.hpp:184-272; the Hull vectors are built as the Cython layer
// converts its lists (_distance.pyx:169-191).
__attribute__((visibility("default"))) void ref_parabolic(float *img, int n, float res,
                                                          long stride, const ref_hull *lo,
                                                          int n_lo, const ref_hull *hi, int n_hi)
{
    std::vector<Hull> lower, upper;
    for (int i = 0; i < n_lo; i++) lower.push_back({lo[i].vertex, lo[i].height, lo[i].range});
    for (int i = 0; i < n_hi; i++) upper.push_back({hi[i].vertex, hi[i].height, hi[i].range});
    squared_edt_1d_parabolic(img, n, res, stride, lower, upper);
}

// _distance.hpp:296-356
__attribute__((visibility("default"))) int ref_boundary_hull(float *img, int n, float res,
                                                             long stride, int num_hull, int ic,
                                                             int forward, ref_hull *out)
{
    std::vector<Hull> h = return_boundary_hull(img, n, res, stride, num_hull, ic, forward != 0);
    for (size_t i = 0; i < h.size(); i++) out[i] = {h[i].vertex, h[i].height, h[i].range};
    return (int)h.size();
}

// _voxels.hpp:60-79 with the skip window of _voxels.pyx:126-155
__attribute__((visibility("default"))) long ref_nearest_index(const uint8_t *line, int n,
                                                              long stride, uint8_t label,
                                                              int lower_skip, int upper_skip,
                                                              int forward)
{
    return (long)_get_nearest_boundary_index(const_cast<uint8_t *>(line) + (long)lower_skip * stride,
                                             label, n - lower_skip - upper_skip, stride,
                                             lower_skip, forward != 0);
}
}
