// psd_internal.hpp -- the pore-size-distribution sweep shared by the single
// and batched entry points (internal).
#pragma once

#include "edt_internal.hpp"

namespace pmx {

size_t psd_workspace_bytes(const pmx_geom *g);

int check_psd(const uint8_t *img, const pmx_geom *g, const int64_t *inlet_lo, const int64_t *inlet_ext, int n_levels,
              const int64_t *t_erode, const int64_t *t_dilate, const int32_t *code_map);

// The whole sweep on stream s, with no host synchronisation; tile_rows as
// for run_edt.
int run_psd(const uint8_t *img, const pmx_geom *g, const int64_t *inlet_lo, const int64_t *inlet_ext, int n_levels,
            const int64_t *t_erode, const int64_t *t_dilate, uint32_t wall_mask, int32_t *code_map, void *workspace,
            size_t workspace_bytes, hipStream_t s, int tile_rows);

}  // namespace pmx
