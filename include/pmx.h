/*
 * pmx.h -- C ABI of libpmx, the MI355X voxel-morphology library.
 *
 * All pointers are device pointers unless stated otherwise.  The caller
 * owns every buffer; scratch memory comes from a caller-provided workspace
 * whose size the *_workspace_size functions report.  Every call is
 * stream-ordered on the given hipStream_t (0 = the null stream) and
 * returns PMX_OK or a negative status; pmx_last_error() then describes
 * the failure (thread-local).
 *
 * Arrays are C-order [x][y][z] with z contiguous, as PMMoTo's numpy
 * arrays are (src/pmmoto/core/_voxels.pyx:62-74 derives element strides
 * from numpy strides the same way).
 *
 * Entry points and the reference interface each one replaces:
 *
 *   pmx_edt           the three per-line passes of edt3d (src/pmmoto/filters/distance/
 *                     euclidean_distance_transform.py:208-303) driven line by line through
 *                     _distance.get_initial_envelope (_distance.pyx:308-353) and
 *                     _distance.get_parabolic_envelope (_distance.pyx:86-127), whose C++
 *                     kernels are squared_edt_1d / squared_edt_1d_parabolic
 *                     (_distance.hpp:135-272): the exact integer EDT, equal to the
 *                     reference's float32 output at unit resolution on 0/1 images without
 *                     periodic axes.
 *   pmx_edt_ref       the same passes in the reference's float32 arithmetic at any
 *                     resolution, with its label-aware first pass, its five-member
 *                     periodic boundary hulls (edt3d) and corrected_edt's padded-subdomain
 *                     correctors and hulls (euclidean_distance_transform.py:56-101,
 *                     :304-528): the drop-in for periodic axes, multi-label images and
 *                     non-unit resolution.
 *   pmx_morph         the EDT-mode dilation / erosion bodies of
 *                     addition / subtraction (src/pmmoto/filters/morphological_operators.py:103-228):
 *                     EDT of the (inverted) halo image, threshold, set_wall_bcs.
 *   pmx_ccl26         cc3d.connected_components(img, return_N=True) at
 *                     src/pmmoto/filters/connected_components.py:52-54 plus the periodic
 *                     merge of connect_subdomain_boundaries (:65-101).
 *   pmx_mark_labels   gen_inlet_label_map / gen_outlet_label_map
 *                     (src/pmmoto/core/voxels.py:400-454).
 *   pmx_keep_labels   the label->phase map + renumber_img of inlet_connected_img and
 *                     its siblings (connected_components.py:254-392, _voxels.pyx:453-494).
 *   pmx_fill_periodic_pads  update_buffer's self-neighbour copy for periodic faces
 *                     (src/pmmoto/core/communication.py:101-118, :278-285).
 *   pmx_spheres_on_grid, pmx_sphere_first_cover, pmx_sphere_image
 *                     gen_pm_spheres_domain's centroid-in-sphere rasterisation
 *                     (src/pmmoto/domain_generation/domain_generation.py:125-155,
 *                     particles/spheres.hpp:42-48), and seeded packs grown to a porosity.
 *   pmx_psd_sweep     the radius loop of pore_size_distribution (porosimetry.py:163-227)
 *                     as one call.
 * Each declaration below cites the reference code it replaces, or says that it
 * has none.
 */
#ifndef PMX_H
#define PMX_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PMX_OK 0
#define PMX_ERR_INVALID -1   /* bad argument (maps to ValueError)   */
#define PMX_ERR_HIP -2       /* a HIP runtime call failed           */
#define PMX_ERR_WORKSPACE -3 /* workspace smaller than required      */

/* Boundary type of one face (src/pmmoto/core/boundary_types.py:11-17). */
#define PMX_BC_END 0      /* nothing beyond the face                  */
#define PMX_BC_WALL 1     /* a solid layer just beyond the face       */
#define PMX_BC_PERIODIC 2 /* wrap to the opposite face (both faces)   */

/* Squared distance where no solid voxel is reachable.  The reference's
 * float32 EDT holds FLT_MAX there (to_finite, _distance.hpp:22-32). */
#define PMX_D2_INF INT32_MAX

/* The box the kernels compute on, and how the caller's arrays map onto it.
 * Element (x, y, z) of the box lives at x*sx + y*sy + z of a view. */
typedef struct {
    int64_t n[3];    /* box extents (x, y, z); each in [1, 16384]          */
    int64_t in_s[2]; /* element strides (x, y) of the input view           */
    int64_t out_s[2];/* element strides (x, y) of every output view        */
    int32_t bc[6];   /* PMX_BC_* for faces x-, x+, y-, y+, z-, z+           */
} pmx_geom;

/* Library version (major*10000 + minor*100 + patch). */
int pmx_version(void);

/* Description of the last failure on this thread ("" if none). */
const char *pmx_last_error(void);

/* Per-kernel timing for benchmarks: after pmx_profile_begin every library
 * kernel launch is bracketed by HIP events on its own stream.
 * pmx_profile_end waits for them, stops recording and reports, per kernel
 * name (newline-separated in names), the summed milliseconds and launch
 * count (host buffers; up to max_kernels entries; *n_kernels set). */
int pmx_profile_begin(void);
int pmx_profile_end(char *names, int names_len, double *ms, int64_t *calls, int max_kernels,
                    int *n_kernels);

/* Workspace bytes pmx_edt / pmx_morph need for geometry g. */
int pmx_edt_workspace_size(const pmx_geom *g, size_t *bytes);

/* Exact squared Euclidean distance from every voxel to the nearest voxel
 * whose value is 0, with per-face boundary handling.  Outputs (any may be
 * NULL, at least one must not be): d2 (int32, PMX_D2_INF where no solid is
 * reachable), f32sq (float(d2), FLT_MAX for PMX_D2_INF: the reference's
 * squared float32 output), f32 (correctly rounded sqrtf of f32sq: the
 * reference's np.sqrt of it), f64 (correctly rounded sqrt of double(d2),
 * sqrt(double(FLT_MAX)) for PMX_D2_INF). */
int pmx_edt(const uint8_t *img, const pmx_geom *g, int32_t *d2, float *f32sq, float *f32,
            double *f64, void *workspace, size_t workspace_bytes, hipStream_t stream);

/* No counterpart in the reference.  pmx_edt for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc): the same kernels and outputs, with the workspace
 * in a static device array of PMX_EDT_NOALLOC_SCRATCH_BYTES that every GPU
 * the library runs on holds until the process ends.  A box whose workspace
 * exceeds it (beyond about 340^3 voxels) returns PMX_ERR_INVALID; use pmx_edt.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams or host threads) race, so serialise them.  Allocates
 * nothing. */
#define PMX_EDT_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int pmx_edt_noalloc(const uint8_t *img, const pmx_geom *g, int32_t *d2, float *f32sq, float *f32,
                    double *f64, hipStream_t stream);

/* The reference's float32 EDT arithmetic, bit for bit, at any resolution
 * (edt3d / edt2d / corrected_edt, euclidean_distance_transform.py:56-303,
 * per-line kernels _distance.hpp:135-356).  img: uint8 view (x, y strides
 * in_s[2]) of the box n[3]; d: float32 view (strides out_s[2]) receiving the
 * squared distance, or its float32 sqrt with take_sqrt.  first_axis gets the
 * 1-D label-aware sweep, then the other axes in ascending order get the
 * parabolic envelope (3-D: first_axis 0; 2-D image as a (1, X, Y) box:
 * first_axis 2).  periodic[3] selects the reference's corrector and
 * five-member boundary-hull handling on that axis; resolution[3] are the
 * float32 voxel sizes.  pads: NULL for edt3d / edt2d (correctors and hulls
 * from whole lines); or the six pad widths (x-, x+, y-, y+, z-, z+) of a
 * PMMoTo padded subdomain, whose whole padded array is the box: then the
 * periodic axes follow corrected_edt for one rank (pad-skipping face
 * searches, :304-528), first_axis must be 0.  Voxels with no reachable
 * solid hold FLT_MAX.  Replaces the per-line loops of
 * _distance.pyx:86-127,308-353,544-618. */
int pmx_edt_ref_workspace_size(const int64_t *n, size_t *bytes);
int pmx_edt_ref(const uint8_t *img, const int64_t *n, const int64_t *in_s, float *d,
                const int64_t *out_s, int first_axis, const int32_t *periodic, const int64_t *pads,
                const float *resolution, int take_sqrt, void *workspace, size_t workspace_bytes,
                hipStream_t stream);

/* No counterpart in the reference: introspection of pmx_edt_ref.  For the
 * 3-D modes (first_axis 0: edt3d, or corrected_edt with pads) at one
 * power-of-two resolution on every axis and extents up to 1024, pmx_edt_ref
 * first runs the same arithmetic in integer form (every float the reference forms there is an exact multiple of
 * res^2; only the boundary hulls depend on rounding, and they are computed
 * in float), and falls back to the float kernels, on the device, when the
 * image or its distances leave that form's range.  pmx_edt_ref_int_form
 * switches the integer form on or off for the process (for comparisons) and
 * returns the previous setting.  pmx_edt_ref_last_form waits for `stream` and
 * reports which path produced the last result written with this workspace
 * (box n): 0 the integer form, 1 the float kernels after the integer form
 * set its result aside, 2 the float kernels alone. */
int pmx_edt_ref_int_form(int enable);
int pmx_edt_ref_last_form(const int64_t *n, const void *workspace, size_t workspace_bytes,
                          hipStream_t stream, int *form);

/* The reference's halo-image morphology at any resolution
 * (morphological_operators.py:103-228): the padded uint8 array (shape[3],
 * pad widths pad[6]) is extended by halo_width[3] voxels per side (periodic
 * copies of the own region; 0 on axes without a halo), its float32 EDT is
 * taken with pmx_edt_ref (invert: of np.logical_not), and out (shape[3],
 * contiguous) = (d <= radius ? v_le : v_gt) compared in float64 when
 * compare_f64, else against (float)radius; WALL layers (wall_mask) -> 0. */
int pmx_halo_morph_workspace_size(const int64_t *shape, const int64_t *halo_width, size_t *bytes);
int pmx_halo_morph(const uint8_t *img, const int64_t *shape, const int64_t *pad,
                   const int64_t *halo_width, const float *resolution, int invert, double radius,
                   int compare_f64, uint8_t v_le, uint8_t v_gt, uint32_t wall_mask, uint8_t *out,
                   void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Morphology by EDT threshold.  With invert == 0 the distance is taken to
 * voxels equal to 0 (erosion: subtraction), with invert != 0 to voxels not
 * equal to 0 (dilation: addition of np.logical_not(img)).  out = d2 <= t ?
 * v_le : v_gt, then the outermost layer of every face whose bit is set in
 * wall_mask (bit 2*axis+side) is forced to 0 (set_wall_bcs). */
int pmx_morph(const uint8_t *img, const pmx_geom *g, int invert, int64_t t, uint8_t v_le,
              uint8_t v_gt, uint32_t wall_mask, uint8_t *out, void *workspace,
              size_t workspace_bytes, hipStream_t stream);

/* Copy own voxels into the pads of periodic axes of a padded C-order array
 * of `elem` bytes per element (1, 4 or 8): shape[3] is the padded extent,
 * pad[6] the pad widths (x-, x+, y-, y+, z-, z+), periodic[3] flags. */
int pmx_fill_periodic_pads(void *arr, int elem, const int64_t *shape, const int64_t *pad,
                           const int32_t *periodic, hipStream_t stream);

/* Workspace bytes pmx_ccl26 needs for geometry g. */
int pmx_ccl_workspace_size(const pmx_geom *g, size_t *bytes);

/* 26-connected labelling of voxels != 0; neighbours join when their values
 * are equal (multi-label, 0 is background), wrapping on PERIODIC faces
 * (WALL is rejected: wall pads are ordinary zero voxels, label them as END).
 * Components are numbered 1..N in C-order of their first voxel within the
 * box; exactly one of labels32 / labels64 receives them (output view);
 * *count (device int64) receives N. */
int pmx_ccl26(const uint8_t *img, const pmx_geom *g, int32_t *labels32, uint64_t *labels64,
              int64_t *count, void *workspace, size_t workspace_bytes, hipStream_t stream);

/* flags[l] = 1 for every label l in [1, n_flags) in the sub-box lo[3] +
 * [0, extent[3]) of a C-order label array (label_bytes 4 or 8, element
 * strides[2] for x and y); flags holds n_flags bytes (count + 1).  A label
 * >= n_flags (a stale count) is not written but counted into *bad (device
 * uint32, may be NULL), which the caller checks.  Replaces the label
 * collection of gen_inlet_label_map / gen_outlet_label_map
 * (src/pmmoto/core/voxels.py:400-454). */
int pmx_mark_labels(const void *labels, int label_bytes, const int64_t *strides, const int64_t *lo,
                    const int64_t *extent, uint8_t *flags, uint64_t n_flags, unsigned int *bad,
                    hipStream_t stream);

/* out[i] = keep ? img[i] : 0 over n contiguous voxels, keep decided from
 * the voxel's label l in [1, n_flags): mode 0 flags_a[l]; mode 1
 * flags_a[l] && flags_b[l]; mode 2 !flags_a[l] && !flags_b[l]; and, if
 * phase >= 0, img[i] == phase.  A nonzero label >= n_flags gives 0 and is
 * counted into *bad (may be NULL).  Replaces the label->phase map and
 * renumber_img of inlet_connected_img and its siblings
 * (connected_components.py:254-392, _voxels.pyx:453-494). */
int pmx_keep_labels(const uint8_t *img, const void *labels, int label_bytes, int64_t n,
                    const uint8_t *flags_a, const uint8_t *flags_b, uint64_t n_flags, int mode,
                    int phase, uint8_t *out, unsigned int *bad, hipStream_t stream);

/* Porosimetry steps over n contiguous voxels.  pmx_psd_update: map[i] =
 * code where pm[i] == 1 && temp[i] == 1 && map[i] == 0 (the first-invasion
 * rule of pore_size_distribution, porosimetry.py:220-225, with radius codes
 * index+1 instead of radius values).  pmx_codes_to_values: out[i] =
 * values[map[i]-1] (0.0 for code 0).  pmx_code_histogram: counts[c] =
 * number of voxels with pm == 1 and map == c inside the sub-box lo +
 * [0, extent) of a C-order array of `shape` (c in 0..n_codes <= 4096).
 * pmx_both_one: out = (a == 1 && b == 1). */
int pmx_psd_update(const uint8_t *pm, const uint8_t *temp, int64_t n, int32_t code, int32_t *map,
                   hipStream_t stream);
int pmx_codes_to_values(const int32_t *map, int64_t n, const double *values, int32_t n_values,
                        double *out, hipStream_t stream);
int pmx_code_histogram(const uint8_t *pm, const int32_t *map, const int64_t *shape, const int64_t *lo,
                       const int64_t *extent, int32_t n_codes, uint64_t *counts, hipStream_t stream);
int pmx_both_one(const uint8_t *a, const uint8_t *b, int64_t n, uint8_t *out, hipStream_t stream);

/* Wetting-saturation curve of a radius-code map: the code histogram of
 * pmx_code_histogram (counts: n_codes + 1 device uint64, also returned),
 * then curve[i] (device float64, n_codes values) = (wet_i / volume) /
 * (pore / volume), wet_i = pore voxels of the sub-box with code 0 or code >
 * i + 1, in the reference's order of float64 operations (Multiphase.
 * get_saturation = get_volume_fraction / porosity, multiphase.py:56-95,
 * porousmedia.py:56-66); volume = the domain's voxel count. */
int pmx_saturation_curve(const uint8_t *pm, const int32_t *map, const int64_t *shape, const int64_t *lo,
                         const int64_t *extent, int32_t n_codes, double volume, uint64_t *counts,
                         double *curve, hipStream_t stream);

/* Drainage (equilibrium_distribution.py:92-123, multiphase.py:43-95).
 * pmx_drain_update: img[i] = 1 where morph[i] == 1 && wet[i] == 2 (n voxels).
 * pmx_box_op on the sub-box lo + [0, extent) of a uint8 array of `shape`:
 * op 0 fills it with value (update_reservoir); op 1 writes the number of
 * voxels equal to value to *count (device uint64; volume fractions).
 * pmx_max_f32: *out (device) = max of n non-negative float32 values
 * (determine_maximum, utils.py:148-162).
 * pmx_max_u8: *out (device uint32) = max of n bytes (a 16-byte aligned); the
 * EDT drop-in uses it to find multi-label images, whose first pass resets
 * at label changes (_distance.hpp:155-165). */
int pmx_drain_update(const uint8_t *morph, const uint8_t *wet, int64_t n, uint8_t *img,
                     hipStream_t stream);
int pmx_box_op(uint8_t *arr, const int64_t *shape, const int64_t *lo, const int64_t *extent, int op,
               uint8_t value, uint64_t *count, hipStream_t stream);
int pmx_max_f32(const float *a, int64_t n, float *out, hipStream_t stream);
int pmx_max_u8(const uint8_t *a, int64_t n, uint32_t *out, hipStream_t stream);

/* The pore-size-distribution sweep of pore_size_distribution (porosimetry.py:
 * 163-227) in one call, on a box of img and code_map that share their
 * strides (a view into a larger array).  A padded PMMoTo image with END
 * faces, the same box with an all-solid wall pad layer cut off and that
 * face marked WALL, or the own region with WALL / PERIODIC faces (the torus
 * for the erosions, dilations and the inlet connectivity) give the same
 * codes on the own voxels.  Level i (radii
 * in descending order) erodes to d2 > t_erode[i] (non-increasing), keeps
 * the part 26-connected to the inlet slab inlet_lo + [0, inlet_ext) (all of
 * it when inlet_lo is NULL), dilates it to d2 <= t_dilate[i] and writes
 * code i+1 into code_map where img == 1 and code_map is still 0.  One EDT
 * of img, one union-find grown level by level, one EDT per level (of the
 * centres that joined at that level only).  No host synchronisation: the
 * launches can be captured in a graph.  n_levels <= 250. */
int pmx_psd_workspace_size(const pmx_geom *g, size_t *bytes);
int pmx_psd_sweep(const uint8_t *img, const pmx_geom *g, const int64_t *inlet_lo,
                  const int64_t *inlet_ext, int n_levels, const int64_t *t_erode,
                  const int64_t *t_dilate, uint32_t wall_mask, int32_t *code_map,
                  void *workspace, size_t workspace_bytes, hipStream_t stream);

/* Batched sweeps (no counterpart in the reference, which runs one domain
 * per process: for callers that sweep hundreds of unevenly sized domains).
 * Every domain runs pmx_psd_sweep's algorithm with the shared thresholds;
 * domains are spread over the caller's streams `lanes` (longest first, to
 * the least loaded lane), each lane's domains run in turn in the lane's
 * slice of the workspace, and nothing synchronises with the host, so the
 * whole batch can be captured in one graph.  Outputs are bit-equal to
 * per-domain pmx_psd_sweep calls. */
typedef struct {
    const uint8_t *img;    /* padded image, contiguous box with END/PERIODIC faces  */
    pmx_geom g;            /* its box (in and out strides equal the box)            */
    int64_t inlet_lo[3];   /* inlet slab, used when has_inlet != 0                  */
    int64_t inlet_ext[3];
    int32_t has_inlet;
    uint32_t wall_mask;    /* faces whose layer is zeroed in the dilations          */
    int32_t *code_map;     /* output: int32, the box                                */
    int32_t tile_rows;     /* rows per EDT tile: 0 (default), 32 or 64              */
} pmx_psd_domain;
int pmx_psd_batch_workspace_size(int n_domains, const pmx_psd_domain *domains, int n_lanes,
                                 size_t *bytes);
int pmx_psd_sweep_batch(int n_domains, const pmx_psd_domain *domains, int n_levels,
                        const int64_t *t_erode, const int64_t *t_dilate, int n_lanes,
                        const hipStream_t *lanes, void *workspace, size_t workspace_bytes);

/* The fused batch (no counterpart in the reference; for callers that sweep
 * hundreds of domains at once): every pass of every level is one launch
 * over all domains, each block finding its domain in a table of per-domain
 * arguments, and the erosion EDT is each domain's full exact EDT (written
 * to edt_d2[i], an int32 C-contiguous box, when edt_d2 and edt_d2[i] are
 * given; pmx_edt's squared distances).  Codes and distances are bit-equal to
 * per-domain pmx_psd_sweep / pmx_edt calls.  Domains without an inlet
 * requirement only, thresholds t_dilate below 128^2 (no dilation defers a
 * line).  pmx_psd_batch_create checks the domains, lays out the workspace and
 * uploads the launch tables (synchronising with the stream);
 * pmx_psd_batch_run then only launches, with no host synchronisation (it
 * can be captured in a graph; the plan, the workspace and the domains'
 * arrays must outlive every replay). */
typedef struct pmx_psd_batch pmx_psd_batch;
int pmx_psd_batch_plan_size(int n_domains, const pmx_psd_domain *domains, int32_t *const *edt_d2,
                            int n_levels, const int64_t *t_erode, const int64_t *t_dilate,
                            size_t *bytes);
int pmx_psd_batch_create(int n_domains, const pmx_psd_domain *domains, int32_t *const *edt_d2,
                         int n_levels, const int64_t *t_erode, const int64_t *t_dilate,
                         void *workspace, size_t workspace_bytes, hipStream_t stream,
                         pmx_psd_batch **out);
int pmx_psd_batch_run(const pmx_psd_batch *batch, hipStream_t stream);
void pmx_psd_batch_destroy(pmx_psd_batch *batch);

/* Tile autotuner (no counterpart in the reference): times the windowed EDT
 * passes of img with 32- and 64-row tiles (a threshold EDT clamped at cap;
 * invert as for pmx_morph) and returns the faster in *tile_rows.  The
 * choice is kept per (shape, boundary, clamp class) for the life of the
 * process, so later calls for the same key return at once.  Synchronises
 * with the stream; not for graph capture. */
int pmx_edt_autotune_workspace_size(const pmx_geom *g, size_t *bytes);
int pmx_edt_autotune(const uint8_t *img, const pmx_geom *g, int64_t cap, int invert,
                     int32_t *tile_rows, void *workspace, size_t workspace_bytes,
                     hipStream_t stream);
/* The autotuner's choices on disk (no counterpart in the reference): save
 * writes every choice of this process to a text file (one line per key:
 * shape, faces, invert, clamp class, tile rows); load adds a file's choices,
 * so pmx_edt_autotune returns them without timing.  A missing file loads
 * nothing; a malformed line is an error. */
int pmx_tune_cache_save(const char *path);
int pmx_tune_cache_load(const char *path);

/* Learned-surrogate kernels (no counterpart in the reference, which has no
 * learned component; for the porosimetry surrogate).  BF16 operands are
 * raw bfloat16 bits (uint16), accumulation is FP32.
 * pmx_gemm_bf16: C[M x N] (+)= op(A) . op(B) on the matrix cores, op(A)(m,k)
 * = A[m*lda+k] or A[k*lda+m] when a_trans, likewise B; k_splits > 1 splits
 * K over workgroups that add into C atomically (accumulate must be 1).
 * pmx_im2col3d / pmx_col2im3d: kernel-3, padding-1 patches of a
 * channels-last [N][D][H][W][C] tensor (in_shape) with stride 1 or 2,
 * patch rows of ldk >= 27C, k = ((kz*3+ky)*3+kx)*C + c; col2im is the
 * adjoint (a gather).  pmx_bias_act: y + bias, optional ReLU, to BF16 and/or
 * FP32.  pmx_relu_backward: dy where act > 0.  pmx_col_sum: out[j] += sum_i
 * a[i][j].  pmx_avgpool(_backward): mean over the p positions of each
 * sample.  pmx_gnll_loss: per-sample Gaussian negative log-likelihood of
 * [curve(q), log variance] and the gradient of its batch mean.  pmx_adam:
 * one Adam step on FP32 weights, refreshing their BF16 copy.
 * pmx_conv3d_first / pmx_conv3d_bf16: inference convolutions (kernel 3,
 * padding 1, stride 1 or 2) with the bias, the ReLU and the BF16 rounding
 * fused, writing channels-last BF16 [N][od][oh][ow][co] without a patch
 * matrix; w is BF16 [co][3][3][3][C] and bias FP32 [co], both on the device.
 * in_shape is the tensor's own shape; with in_pad / out_pad = 1 the buffer
 * is [N][D+2][H+2][W+2][C] holding the tensor in its interior (the caller
 * keeps the one-voxel border zero; the kernels write only the interior).
 * _first reads a one-channel uint8 image (co a multiple of 4 up to 32; ReLU
 * always; out 8-byte aligned); _bf16 reads a BF16 tensor whose C is a
 * multiple of 8 (x, w, out 8-byte aligned) on the matrix cores.  One sample
 * of the input must be below 2 GiB.  pmx_conv3d_first_pair: _first with
 * co = 8 followed by _bf16 with co = 16 (both stride 2, ReLU, the first
 * unpadded in between) in one pass that keeps the 8-channel activations on
 * chip; w0 [8][3][3][3][1], w1 [16][3][3][3][8]; the image's width must be
 * 128, its height a multiple of 32 and its depth a multiple of 4; the
 * output [N][D/4][H/4][32][16] (out_pad as above) equals the two calls'
 * bit for bit. */
int pmx_gemm_bf16(int64_t m, int n, int k, const uint16_t *a, int64_t lda, int a_trans,
                  const uint16_t *b, int64_t ldb, int b_trans, float *c, int64_t ldc,
                  int accumulate, int k_splits, hipStream_t stream);
int pmx_conv3d_first(const uint8_t *x, const int32_t *in_shape, int stride, int co, const uint16_t *w_bf16,
                     const float *bias, uint16_t *out, int out_pad, hipStream_t stream);
int pmx_conv3d_bf16(const uint16_t *x, const int32_t *in_shape, int in_pad, int stride, int co,
                    const uint16_t *w_bf16, const float *bias, int relu, uint16_t *out, int out_pad,
                    hipStream_t stream);
int pmx_conv3d_first_pair(const uint8_t *x, const int32_t *in_shape, const uint16_t *w0_bf16, const float *b0,
                          const uint16_t *w1_bf16, const float *b1, uint16_t *out, int out_pad,
                          hipStream_t stream);
int pmx_im2col3d(const uint16_t *x, const int32_t *in_shape, int stride, int64_t ldk,
                 uint16_t *cols, hipStream_t stream);
int pmx_col2im3d(const float *dcols, const int32_t *in_shape, int stride, int64_t ldk, float *dx,
                 hipStream_t stream);
int pmx_bias_act(const float *y, const float *bias, int64_t m, int c, int relu, uint16_t *out_bf16,
                 float *out_f32, hipStream_t stream);
int pmx_relu_backward(const float *dy, const uint16_t *act, int64_t n, float *dx, uint16_t *dx_bf16,
                      hipStream_t stream);
int pmx_to_bf16(const float *a, int64_t n, uint16_t *out, hipStream_t stream);
int pmx_u8_to_bf16(const uint8_t *a, int64_t n, uint16_t *out, hipStream_t stream);
int pmx_col_sum(const float *a, int64_t m, int n, float *out, hipStream_t stream);
int pmx_avgpool(const uint16_t *x, int n, int64_t p, int c, float *out_f32, uint16_t *out_bf16,
                hipStream_t stream);
int pmx_avgpool_backward(const float *dy, int n, int64_t p, int c, float *dx, hipStream_t stream);
int pmx_gnll_loss(const float *pred, const float *target, int n, int q, float *loss, float *dpred,
                  hipStream_t stream);
int pmx_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
             float eps, int step, uint16_t *w_bf16, hipStream_t stream);

/* Sphere packs.  spheres: n_spheres x (cx, cy, cz, r) float64 in voxel
 * units (voxel (i,j,k) has its centroid at (i+.5, j+.5, k+.5)); a voxel is
 * covered by a sphere when dx*dx+dy*dy (+) dz*dz <= r*r in float64, periodic
 * axes using the minimum image when a sphere spans the whole period.
 * pmx_sphere_first_cover writes first[v] = index of the first sphere that
 * covers voxel v (INT32_MAX if none) and counts[s] = number of voxels whose
 * first cover is s (n_spheres <= 2^24).  After k spheres the pore voxels
 * are those with first >= k; pmx_sphere_image writes img = first >= k. */
int pmx_sphere_first_cover(const double *spheres, int64_t n_spheres, const int64_t *shape,
                           const int32_t *periodic, int32_t *first, int64_t *counts,
                           hipStream_t stream);
/* pmx_spheres_on_grid: the reference's gen_pm_shape (_domain_generation.pyx
 * :30-70 with Sphere::contains, particles/spheres.hpp:42-48) for spheres
 * (n x (cx, cy, cz, r) float64, device) in the domain's own coordinates:
 * img (shape[0] x shape[1] x shape[2] uint8) = 1, and 0 at every voxel whose
 * centroid (x[i], y[j], z[k]) (device float64 coordinate arrays) satisfies
 * (cx-x)^2 + (cy-y)^2 + (cz-z)^2 <= r*r; invert swaps 0 and 1.  grid_lo and
 * spacing (host, 3 each) describe the nominal uniform grid (centroid i near
 * grid_lo + (i + 0.5) spacing) used only to bound each sphere's voxels. */
int pmx_spheres_on_grid(const double *spheres, int64_t n_spheres, const double *x, const double *y,
                        const double *z, const int64_t *shape, const double *grid_lo, const double *spacing,
                        int invert, uint8_t *img, hipStream_t stream);
int pmx_sphere_image(const int32_t *first, const int64_t *shape, int32_t k, uint8_t *img,
                     hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PMX_H */
