// edt_internal.hpp -- the exact EDT pipeline shared by libpmx's entry points.
#pragma once

#include "common.hpp"

namespace pmx {

// Outputs of the last EDT pass; any subset may be set.  Element (x, y, z)
// of the box is written at x*sx + y*sy + z.
struct PassXOut {
    int32_t *d2 = nullptr;       // squared distance
    float *f32sq = nullptr;      // float(d2), FLT_MAX for "no solid"
    float *f32 = nullptr;        // sqrtf of f32sq
    double *f64 = nullptr;       // sqrt(double(d2))
    uint8_t *u8 = nullptr;       // d2 <= t ? v_le : v_gt, WALL layers 0
    int64_t sx = 0, sy = 0;
    int32_t t = 0;
    uint8_t v_le = 0, v_gt = 0;
    uint32_t wall_mask = 0;
    // porosimetry (first invasion): where d2 <= t and psd_state == 1 (a pore
    // voxel not yet invaded), psd_state = 2 and psd_map = psd_code.  psd_map
    // has the strides sx, sy; psd_state its own, ssx and ssy.
    uint8_t *psd_state = nullptr;
    int64_t ssx = 0, ssy = 0;
    int32_t *psd_map = nullptr;
    int32_t psd_code = 0;
};

struct EdtPlan {
    uint16_t *dz = nullptr;   // pass-Z distances, [x][y][z], row pitch `pitch` voxels
    int32_t *b = nullptr;     // pass-Y squared distances, stored [y][x][z], same pitch
    int64_t pitch = 0;
    uint32_t *counters = nullptr, *ovf_list = nullptr;
    uint32_t *bits_y = nullptr, *bits_x = nullptr;  // deferred-line bitmaps (one allocation)
    size_t n_bits = 0;                              // words in both bitmaps
    int *stk = nullptr;
    int cap = 0;            // fallback stack entries per lane
    int fallback_lanes = 0;
};

// Fallback lanes of the exact EDT at most (one per deferred line).
constexpr int kMaxFallbackLanes = 32768;

// Workspace layout of run_edt for geometry g (sizes only when ws is null);
// max_lanes bounds the persistent fallback lanes (and their stacks).
size_t plan_edt(const pmx_geom *g, void *ws, EdtPlan *p, int max_lanes = kMaxFallbackLanes);

// Exact squared EDT of img over the box of g (distance to voxels == 0, or
// != 0 with invert), clamped to min(d2, cap) (cap >= 2^30: unclamped),
// delivered through `o`; stream-ordered on s.  A caller that only tests
// d2 <= t passes cap = t + 1.  With `proceed` (device memory) set, every
// kernel does nothing when *proceed == 0 at run time.  tile_rows: rows per
// tile of the windowed passes, 0 (default 64), 32 or 64.  match >= 0: the
// distance is to the voxels whose byte equals match (invert is ignored).
// max_lanes: as for plan_edt (the workspace must follow the same plan).
int run_edt(const uint8_t *img, const pmx_geom *g, bool invert, const PassXOut &o, int64_t cap,
            void *ws, size_t ws_bytes, hipStream_t s, const unsigned int *proceed = nullptr,
            int tile_rows = 0, int match = -1, int max_lanes = kMaxFallbackLanes);

}  // namespace pmx
