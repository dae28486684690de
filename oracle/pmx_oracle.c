/*
 * pmx_oracle.c -- CPU oracle for the voxel-morphology path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (pmmoto_amd/) links,
 * loads or calls this file; only tests/, the smoke check and the CPU
 * baseline leg of bench.py do, and only as the checker.
 *
 * Two families of functions live here:
 *
 *   orc_ref_*   A restatement of the reference CPU algorithm in float32,
 *               operation for operation, so its outputs are bit-identical
 *               to PMMoTo's (pass order, float rounding, FLT_MAX sentinel,
 *               5-member periodic boundary hulls).  Each function cites the
 *               reference file:line it follows.
 *
 *   orc_exact_* Exact integer algorithms (squared EDT with END / WALL /
 *               PERIODIC faces, 26-connected labelling) used to check the
 *               GPU kernels on inputs where the reference's float path is
 *               not integer-exact, and as the baseline "port".
 *
 * How each function is pinned is stated in tests/test_oracle_*.py.
 */
#include <float.h>
#include <pthread.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Reference float32 restatement                                            */
/* ------------------------------------------------------------------------ */

/* One member of a lower envelope: src/pmmoto/filters/distance/_distance.hpp:15-20 */
typedef struct {
    int vertex;
    float height;
    float range;
} orc_hull;

#define ORC_FMAX_M1 (FLT_MAX - 1.0f) /* equals FLT_MAX in float32 */

/* _distance.hpp:22-32 */
static void to_finite(float *f, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (f[i] == INFINITY) f[i] = ORC_FMAX_M1;
}

/* _distance.hpp:60-72.  The numerator is evaluated in float; the literal
 * 2.0 is a double, so the division happens in double and the quotient is
 * rounded back to float on assignment. */
static float intersect_parabolas(float f_r, float r, float f_q, float q, float aniso)
{
    float factor1 = (r - q) * aniso;
    float factor2 = r + q;
    float num = f_r - f_q + factor1 * factor2;
    float s = (float)((double)num / (2.0 * (double)factor1));
    return s;
}

/* _distance.hpp:89-112 */
static void update_hull(int *k, int i, float ff, int *hv, float *hh, float *ranges, float aniso)
{
    float s = intersect_parabolas(ff, (float)i, hh[*k], (float)hv[*k], aniso);
    while (*k > 0 && s <= ranges[*k]) {
        (*k)--;
        s = intersect_parabolas(ff, (float)i, hh[*k], (float)hv[*k], aniso);
    }
    (*k)++;
    hv[*k] = i;
    hh[*k] = ff;
    ranges[*k] = s;
    ranges[*k + 1] = INFINITY;
}

/* squared_edt_1d<uint8_t>: _distance.hpp:135-182 (multi-label forward sweep,
 * corrector seeding/capping, backward min, in-place square). */
ORC_EXPORT void orc_ref_sq_edt_1d(const uint8_t *img, float *d, int n, long stride,
                                  float res, float lower_corrector, float upper_corrector)
{
    long i;
    uint8_t seg = img[0];
    d[0] = seg == 0 ? 0.0f : lower_corrector * res;
    for (i = stride; i < (long)n * stride; i += stride) {
        if (img[i] == 0) {
            d[i] = 0.0f;
        } else if (img[i] == seg) {
            d[i] = d[i - stride] + res;
        } else {
            d[i] = res;
            d[i - stride] = (float)(img[i - stride] != 0) * res;
            seg = img[i];
        }
    }
    if (d[(long)(n - 1) * stride] > upper_corrector * res)
        d[(long)(n - 1) * stride] = upper_corrector * res;
    for (i = (long)(n - 2) * stride; i >= 0; i -= stride)
        d[i] = fminf(d[i], d[i + stride] + res);
    for (i = 0; i < (long)n * stride; i += stride)
        d[i] *= d[i];
}

/* squared_edt_1d_parabolic: _distance.hpp:184-272.  lower/upper are the
 * injected boundary hulls in the order the reference receives them. */
ORC_EXPORT void orc_ref_sq_edt_1d_parabolic(float *img, int n, float res, long stride,
                                            const orc_hull *lower, int n_lower,
                                            const orc_hull *upper, int n_upper)
{
    if (n == 0) return;
    const float aniso = res * res;
    int cap = n + n_lower + n_upper;
    float *ff = (float *)calloc((size_t)n, sizeof(float));
    float *hh = (float *)calloc((size_t)cap, sizeof(float));
    int *hv = (int *)calloc((size_t)cap, sizeof(int));
    float *ranges = (float *)calloc((size_t)cap + 1, sizeof(float));
    for (long i = 0; i < n; i++) ff[i] = img[i * stride];

    long loop_start = 1;
    ranges[0] = -INFINITY;
    ranges[1] = INFINITY;
    int k = 0;
    if (n_lower > 0) {
        loop_start = 0;
        for (int m = n_lower - 1; m >= 0; m--) { /* pushed in reverse (:220-228) */
            hv[k] = lower[m].vertex;
            hh[k] = lower[m].height;
            ranges[k] = lower[m].range;
            ranges[k + 1] = INFINITY;
            k++;
        }
        k--;
    } else {
        hv[0] = 0;
        hh[0] = ff[0];
    }
    for (long i = loop_start; i < n; i++)
        update_hull(&k, (int)i, ff[i], hv, hh, ranges, aniso);
    for (int m = 0; m < n_upper; m++) /* vertices shifted by n (:245-257) */
        update_hull(&k, upper[m].vertex + n, upper[m].height, hv, hh, ranges, aniso);

    k = 0;
    for (long i = 0; i < n; i++) {
        while (ranges[k + 1] < (float)i) k++;
        float di = (float)(i - hv[k]);
        img[i * stride] = aniso * (di * di) + hh[k];
    }
    free(ff);
    free(hh);
    free(hv);
    free(ranges);
}

/* return_boundary_hull: _distance.hpp:296-356.  `out` must hold
 * num_hull + 1 entries; returns the number written.  `fin_limit` bounds the
 * contiguous to_finite sweep of :305 to the caller's array. */
static int ref_boundary_hull(float *img, int n, float res, long stride, int num_hull,
                             int index_corrector, int forward, orc_hull *out, size_t fin_limit)
{
    to_finite(img, (size_t)n < fin_limit ? (size_t)n : fin_limit);
    int count = 0;
    if (n == 1 && img[0] < ORC_FMAX_M1) {
        out[0].vertex = index_corrector;
        out[0].height = img[0];
        out[0].range = INFINITY;
        return 1;
    }
    const float aniso = res * res;
    float *ff = (float *)calloc((size_t)n, sizeof(float));
    float *hh = (float *)calloc((size_t)n + 1, sizeof(float));
    int *hv = (int *)calloc((size_t)n + 1, sizeof(int));
    float *ranges = (float *)calloc((size_t)n + 2, sizeof(float));
    for (long i = 0; i < n; i++) ff[i] = img[i * stride];
    hh[0] = ff[0];
    ranges[0] = -INFINITY;
    ranges[1] = INFINITY;
    int k = 1; /* the reference starts the stack top at 1 (:329) */
    for (long i = 1; i < n; i++) update_hull(&k, (int)i, ff[i], hv, hh, ranges, aniso);
    if (k < num_hull) num_hull = k;
    int kk = forward ? 0 : k;
    int step = forward ? 1 : -1;
    while (count <= num_hull && kk >= 0 && kk <= k) {
        if (hh[kk] < ORC_FMAX_M1) {
            out[count].vertex = hv[kk] + index_corrector;
            out[count].height = hh[kk];
            out[count].range = ranges[kk] + (float)index_corrector;
            count++;
            if (hh[kk] == 0.9f) break;
        }
        kk += step;
    }
    free(ff);
    free(hh);
    free(hv);
    free(ranges);
    return count;
}

ORC_EXPORT int orc_ref_boundary_hull_1d(float *img, long start, int n, float res, int num_hull,
                                        int forward, orc_hull *out)
{
    /* get_boundary_hull_1d: _distance.pyx:679-702 (index_corrector = start) */
    return ref_boundary_hull(img + start, n, res, 1, num_hull, (int)start, forward, out,
                             (size_t)n);
}

/* _voxels.hpp:60-79 through _voxels.pyx:231-283: first (forward) or last
 * index equal to `label` within [lower_skip, n - upper_skip), or -1. */
static long ref_nearest_index(const uint8_t *line, int n, long stride, uint8_t label,
                              int lower_skip, int upper_skip, int forward)
{
    int m = n - lower_skip - upper_skip;
    const uint8_t *p = line + (long)lower_skip * stride;
    if (forward) {
        for (int c = 0; c < m; c++)
            if (p[(long)c * stride] == label) return c + lower_skip;
    } else {
        for (int c = 0; c < m; c++)
            if (p[(long)(m - 1 - c) * stride] == label) return (m - 1) - c + lower_skip;
    }
    return -1;
}

ORC_EXPORT long orc_ref_nearest_index_1d(const uint8_t *img, int n, uint8_t label,
                                         int lower_skip, int upper_skip, int forward)
{
    return ref_nearest_index(img, n, 1, label, lower_skip, upper_skip, forward);
}

/* Per-line kernels used by the whole-volume drivers below.  The drivers run
 * either with the restated kernels above (ops == NULL) or with the
 * reference's own kernels compiled from _distance.hpp / _voxels.hpp into
 * oracle/_ref (oracle/ref_shim.cpp), which pins the restatement line by
 * line and volume by volume. */
typedef struct {
    void (*sq_edt_1d)(const uint8_t *, float *, int, long, float, float, float);
    void (*parabolic)(float *, int, float, long, const orc_hull *, int, const orc_hull *, int);
    int (*boundary_hull)(float *, int, float, long, int, int, int, orc_hull *);
    long (*nearest_index)(const uint8_t *, int, long, uint8_t, int, int, int);
} orc_line_ops;

static int restated_boundary_hull(float *img, int n, float res, long stride, int num_hull,
                                  int index_corrector, int forward, orc_hull *out)
{
    return ref_boundary_hull(img, n, res, stride, num_hull, index_corrector, forward, out,
                             (size_t)n);
}

static const orc_line_ops restated_ops = {orc_ref_sq_edt_1d, orc_ref_sq_edt_1d_parabolic,
                                          restated_boundary_hull, ref_nearest_index};

/* Geometry helper for a C-order array of `nd` dims: the line along `dim`
 * through the multi-index whose other coordinates are (a, b) in ascending
 * dimension order, as _distance.pyx:get_dimensions (:64-83) orders them. */
typedef struct {
    int nd;
    long shape[3];
    long stride[3];
} geom;

static geom make_geom(int nd, const long *shape)
{
    geom g;
    g.nd = nd;
    for (int i = 0; i < 3; i++) g.shape[i] = 1;
    for (int i = 0; i < nd; i++) g.shape[i] = shape[i];
    long s = 1;
    for (int i = nd - 1; i >= 0; i--) {
        g.stride[i] = s;
        s *= g.shape[i];
    }
    return g;
}

static void other_dims(int nd, int dim, int *d1, int *d2)
{
    int o[2] = {0, 0}, c = 0;
    for (int i = 0; i < nd; i++)
        if (i != dim) o[c++] = i;
    *d1 = o[0];
    *d2 = nd == 3 ? o[1] : -1;
}

/* get_initial_envelope(_2d) with optional periodic correctors:
 * euclidean_distance_transform.py:226-244 (3-D), :120-149 (2-D);
 * correctors from _distance.pyx:240-279 / edt2d :126-140. */
static void ref_initial_envelope(const uint8_t *img, float *out, geom g, int dim, float res,
                                 int periodic, const orc_line_ops *ops)
{
    int d1, d2;
    other_dims(g.nd, dim, &d1, &d2);
    long s1 = g.shape[d1], s2 = d2 >= 0 ? g.shape[d2] : 1;
    long st1 = g.stride[d1], st2 = d2 >= 0 ? g.stride[d2] : 0;
    int n = (int)g.shape[dim];
    long st = g.stride[dim];
    for (long a = 0; a < s1; a++)
        for (long b = 0; b < s2; b++) {
            long base = a * st1 + b * st2;
            float lo = INFINITY, hi = INFINITY;
            if (periodic) {
                long first = ops->nearest_index(img + base, n, st, 0, 0, 0, 1);
                long last = ops->nearest_index(img + base, n, st, 0, 0, 0, 0);
                /* float64 corrector arrays converted to float at the call */
                lo = last != -1 ? (float)((double)n - (double)last) : INFINITY;
                hi = first != -1 ? (float)((double)first + 1.0) : INFINITY;
            }
            ops->sq_edt_1d(img + base, out + base, n, st, res, lo, hi);
        }
    long total = g.shape[0] * g.shape[1] * g.shape[2];
    to_finite(out, (size_t)total);
}

/* One parabolic pass with the periodic boundary-hull exchange of
 * euclidean_distance_transform.py:246-298 (3-D) / :151-203 (2-D):
 * hulls from _distance.pyx:get_boundary_hull (:544-618), swapped, then
 * get_parabolic_envelope (:86-127). */
static void ref_parabolic_pass(const uint8_t *og, float *out, geom g, int dim, float res,
                               int periodic, const orc_line_ops *ops)
{
    int d1, d2;
    other_dims(g.nd, dim, &d1, &d2);
    long s1 = g.shape[d1], s2 = d2 >= 0 ? g.shape[d2] : 1;
    long st1 = g.stride[d1], st2 = d2 >= 0 ? g.stride[d2] : 0;
    int n = (int)g.shape[dim];
    long st = g.stride[dim];
    long total = g.shape[0] * g.shape[1] * g.shape[2];
    const int num_hull = 4;
    orc_hull lower[8], upper[8];
    if (!periodic) {
        for (long a = 0; a < s1; a++)
            for (long b = 0; b < s2; b++)
                ops->parabolic(out + a * st1 + b * st2, n, res, st, NULL, 0, NULL, 0);
        return;
    }
    /* All hulls are taken from the pass input before any line is updated. */
    long lines = s1 * s2;
    orc_hull *hl = (orc_hull *)malloc(sizeof(orc_hull) * (size_t)lines * (num_hull + 1));
    orc_hull *hu = (orc_hull *)malloc(sizeof(orc_hull) * (size_t)lines * (num_hull + 1));
    int *nl = (int *)malloc(sizeof(int) * (size_t)lines);
    int *nu = (int *)malloc(sizeof(int) * (size_t)lines);
    for (long a = 0; a < s1; a++)
        for (long b = 0; b < s2; b++) {
            long base = a * st1 + b * st2;
            long li = a * s2 + b;
            long lv = ops->nearest_index(og + base, n, st, 0, 0, 0, 1);
            long uv = ops->nearest_index(og + base, n, st, 0, 0, 0, 0);
            for (int fw = 1; fw >= 0; fw--) {
                long bound = fw ? lv : uv;
                long start;
                int nv, ic;
                if (bound == -1) {
                    start = 0;
                    nv = n;
                    ic = fw ? 0 : -nv;
                } else if (fw) {
                    start = 0;
                    nv = (int)bound + 1;
                    ic = 0;
                } else {
                    start = bound;
                    nv = n - (int)bound;
                    ic = -nv;
                }
                long off = base + start * st;
                orc_hull *dst = fw ? &hl[li * (num_hull + 1)] : &hu[li * (num_hull + 1)];
                (void)total;
                int cnt = ops->boundary_hull(out + off, nv, res, st, num_hull, ic, fw, dst);
                if (fw)
                    nl[li] = cnt;
                else
                    nu[li] = cnt;
            }
        }
    for (long a = 0; a < s1; a++)
        for (long b = 0; b < s2; b++) {
            long li = a * s2 + b;
            int cl = nu[li], cu = nl[li]; /* swap: lower <- upper face hull */
            memcpy(lower, &hu[li * (num_hull + 1)], sizeof(orc_hull) * (size_t)cl);
            memcpy(upper, &hl[li * (num_hull + 1)], sizeof(orc_hull) * (size_t)cu);
            ops->parabolic(out + a * st1 + b * st2, n, res, st, lower, cl, upper, cu);
        }
    free(hl);
    free(hu);
    free(nl);
    free(nu);
}

/* edt3d(img, periodic, resolution, squared=True):
 * euclidean_distance_transform.py:208-303.  Output: squared float32. */
ORC_EXPORT void orc_ref_edt3d_sq(const uint8_t *img, long nx, long ny, long nz,
                                 const int *periodic, const float *res, float *out,
                                 const orc_line_ops *ops)
{
    if (!ops) ops = &restated_ops;
    long shape[3] = {nx, ny, nz};
    geom g = make_geom(3, shape);
    ref_initial_envelope(img, out, g, 0, res[0], periodic[0], ops);
    for (int dim = 1; dim <= 2; dim++)
        ref_parabolic_pass(img, out, g, dim, res[dim], periodic[dim], ops);
}

/* The same non-periodic 3-D EDT with each pass's lines split over nthreads
 * POSIX threads (ranges of the first other axis; a join between passes),
 * standing in for the reference's MPI ranks when it is timed as a CPU
 * baseline.  Line arithmetic and pass order are those of orc_ref_edt3d_sq,
 * so the output is identical. */
typedef struct {
    const uint8_t *img;
    float *out;
    geom g;
    int dim;
    float res;
    const orc_line_ops *ops;
    long a0, a1;      /* lines with first-other-axis index in [a0, a1) */
    long e0, e1;      /* elements [e0, e1) for to_finite */
    int stage;        /* 0: initial envelope, 1: to_finite, 2: parabolic */
} orc_job;

static void *orc_worker(void *arg)
{
    orc_job *j = (orc_job *)arg;
    if (j->stage == 1) {
        to_finite(j->out + j->e0, (size_t)(j->e1 - j->e0));
        return NULL;
    }
    int d1, d2;
    other_dims(j->g.nd, j->dim, &d1, &d2);
    long s2 = d2 >= 0 ? j->g.shape[d2] : 1;
    long st1 = j->g.stride[d1], st2 = d2 >= 0 ? j->g.stride[d2] : 0;
    int n = (int)j->g.shape[j->dim];
    long st = j->g.stride[j->dim];
    for (long a = j->a0; a < j->a1; a++)
        for (long b = 0; b < s2; b++) {
            long base = a * st1 + b * st2;
            if (j->stage == 0)
                j->ops->sq_edt_1d(j->img + base, j->out + base, n, st, j->res, INFINITY, INFINITY);
            else
                j->ops->parabolic(j->out + base, n, j->res, st, NULL, 0, NULL, 0);
        }
    return NULL;
}

static void orc_run_stage(orc_job proto, int nthreads)
{
    pthread_t th[256];
    orc_job jobs[256];
    int d1, d2;
    other_dims(proto.g.nd, proto.dim, &d1, &d2);
    long s1 = proto.g.shape[d1];
    long total = proto.g.shape[0] * proto.g.shape[1] * proto.g.shape[2];
    int created[256];
    for (int t = 0; t < nthreads; t++) {
        jobs[t] = proto;
        jobs[t].a0 = s1 * t / nthreads;
        jobs[t].a1 = s1 * (t + 1) / nthreads;
        jobs[t].e0 = total * t / nthreads;
        jobs[t].e1 = total * (t + 1) / nthreads;
        created[t] = pthread_create(&th[t], NULL, orc_worker, &jobs[t]) == 0;
        if (!created[t]) orc_worker(&jobs[t]); /* no thread: run the share here */
    }
    for (int t = 0; t < nthreads; t++)
        if (created[t]) pthread_join(th[t], NULL);
}

ORC_EXPORT void orc_ref_edt3d_sq_mt(const uint8_t *img, long nx, long ny, long nz, const float *res,
                                    float *out, const orc_line_ops *ops, int nthreads)
{
    if (!ops) ops = &restated_ops;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    long shape[3] = {nx, ny, nz};
    orc_job j;
    memset(&j, 0, sizeof(j));
    j.img = img;
    j.out = out;
    j.g = make_geom(3, shape);
    j.ops = ops;
    j.dim = 0;
    j.res = res[0];
    j.stage = 0;
    orc_run_stage(j, nthreads);
    j.stage = 1;
    orc_run_stage(j, nthreads);
    for (int dim = 1; dim <= 2; dim++) {
        j.dim = dim;
        j.res = res[dim];
        j.stage = 2;
        orc_run_stage(j, nthreads);
    }
}

/* corrected_edt(img, subdomain) for one rank of a padded subdomain
 * (euclidean_distance_transform.py:56-101): the passes run over the whole
 * padded array; on a periodic axis the rank exchanges features with itself
 * (communication.py:278-285), so the correctors and hulls it receives are
 * the ones it computed for the opposite face.
 *
 * Pass 0 correctors (get_initial_correctors, :395-447, through
 * get_nearest_boundary_distance, :304-392, with which_voxels="pad" and
 * distance_to="own").  The faces are visited lower face first
 * (orientation.py:79-114) and that function's skip variables persist
 * across faces, so the upper face's search also skips the lower pad:
 *   lower face: first zero in [2 p_lo, S)          -> d_lo = idx - p_lo
 *   upper face: last zero in [2 p_lo, S - 2 p_hi)  -> d_hi = S - idx - p_hi - 1
 * swapped: the line start is seeded with d_hi, the line end capped at d_lo.
 *
 * Passes 1, 2 hulls (get_boundary_hull, :450-528): the lower face's
 * forward hull of the segment from 2 p_lo to the first zero at or after it
 * (vertices relative to 2 p_lo), the upper face's backward hull of the
 * segment from the last zero before S - 2 p_hi to S - 2 p_hi (vertices
 * relative to that end, negative); swapped into the line's lower and
 * upper hulls.  Output: squared float32 over the whole padded array. */
ORC_EXPORT void orc_ref_corrected_edt_sq(const uint8_t *img, long nx, long ny, long nz,
                                         const int *periodic, const long *pads, const float *res,
                                         float *out, const orc_line_ops *ops)
{
    if (!ops) ops = &restated_ops;
    long shape[3] = {nx, ny, nz};
    geom g = make_geom(3, shape);
    long total = nx * ny * nz;
    const int num_hull = 4;
    /* pass 0 */
    {
        int d1, d2;
        other_dims(3, 0, &d1, &d2);
        int S = (int)g.shape[0];
        long st = g.stride[0];
        int plo = (int)pads[0], phi = (int)pads[1];
        for (long a = 0; a < g.shape[d1]; a++)
            for (long b = 0; b < g.shape[d2]; b++) {
                long base = a * g.stride[d1] + b * g.stride[d2];
                float lo = INFINITY, hi = INFINITY;
                if (periodic[0]) {
                    long f = ops->nearest_index(img + base, S, st, 0, 2 * plo, 0, 1);
                    long l = ops->nearest_index(img + base, S, st, 0, 2 * plo, 2 * phi, 0);
                    float d_lo = f != -1 ? (float)(f - plo) : INFINITY;
                    float d_hi = l != -1 ? (float)(S - l - phi - 1) : INFINITY;
                    lo = d_hi;
                    hi = d_lo;
                }
                ops->sq_edt_1d(img + base, out + base, S, st, res[0], lo, hi);
            }
        to_finite(out, (size_t)total);
    }
    for (int dim = 1; dim <= 2; dim++) {
        int d1, d2;
        other_dims(3, dim, &d1, &d2);
        long s1 = g.shape[d1], s2 = g.shape[d2];
        int S = (int)g.shape[dim];
        long st = g.stride[dim];
        if (!periodic[dim]) {
            for (long a = 0; a < s1; a++)
                for (long b = 0; b < s2; b++)
                    ops->parabolic(out + a * g.stride[d1] + b * g.stride[d2], S, res[dim], st, NULL, 0, NULL, 0);
            continue;
        }
        int plo = (int)pads[2 * dim], phi = (int)pads[2 * dim + 1];
        long lines = s1 * s2;
        orc_hull *hf = (orc_hull *)malloc(sizeof(orc_hull) * (size_t)lines * (num_hull + 1));
        orc_hull *hb = (orc_hull *)malloc(sizeof(orc_hull) * (size_t)lines * (num_hull + 1));
        int *nf = (int *)malloc(sizeof(int) * (size_t)lines);
        int *nb = (int *)malloc(sizeof(int) * (size_t)lines);
        for (long a = 0; a < s1; a++)
            for (long b = 0; b < s2; b++) {
                long base = a * g.stride[d1] + b * g.stride[d2];
                long li = a * s2 + b;
                for (int fw = 1; fw >= 0; fw--) {
                    int lsk = fw ? 2 * plo : 0, usk = fw ? 0 : 2 * phi;
                    long bound = ops->nearest_index(img + base, S, st, 0, lsk, usk, fw);
                    long start;
                    int nv, ic;
                    if (bound == -1) {
                        start = lsk;
                        nv = S - usk - lsk;
                        ic = fw ? 0 : -nv;
                    } else if (fw) {
                        start = lsk;
                        nv = (int)bound - lsk + 1;
                        ic = 0;
                    } else {
                        start = bound;
                        nv = S - (int)bound - usk;
                        ic = -nv;
                    }
                    orc_hull *dst = fw ? &hf[li * (num_hull + 1)] : &hb[li * (num_hull + 1)];
                    int cnt = nv > 0 ? ops->boundary_hull(out + base + start * st, nv, res[dim], st, num_hull, ic, fw,
                                                          dst)
                                     : 0;
                    if (fw)
                        nf[li] = cnt;
                    else
                        nb[li] = cnt;
                }
            }
        for (long a = 0; a < s1; a++)
            for (long b = 0; b < s2; b++) {
                long li = a * s2 + b;
                /* lower <- the upper face's backward hull, upper <- the lower face's forward hull */
                ops->parabolic(out + a * g.stride[d1] + b * g.stride[d2], S, res[dim], st, &hb[li * (num_hull + 1)],
                               nb[li], &hf[li * (num_hull + 1)], nf[li]);
            }
        free(hf);
        free(hb);
        free(nf);
        free(nb);
    }
}

/* edt2d(img, periodic, resolution): euclidean_distance_transform.py:104-205
 * (pass 0 along axis 1, parabolic along axis 0).  Output: squared float32. */
ORC_EXPORT void orc_ref_edt2d_sq(const uint8_t *img, long nx, long ny, const int *periodic,
                                 const float *res, float *out, const orc_line_ops *ops)
{
    if (!ops) ops = &restated_ops;
    long shape[2] = {nx, ny};
    geom g = make_geom(2, shape);
    ref_initial_envelope(img, out, g, 1, res[1], periodic[1], ops);
    ref_parabolic_pass(img, out, g, 0, res[0], periodic[0], ops);
}

/* ------------------------------------------------------------------------ */
/* Exact integer squared EDT                                                */
/* ------------------------------------------------------------------------ */

#define ORC_BC_END 0
#define ORC_BC_WALL 1
#define ORC_BC_PERIODIC 2
#define ORC_D2_INF INT32_MAX

/* Exact lower envelope of {(v, f(v))} evaluated at 0..n-1, with virtual
 * zero-height parabolas at -1 / n for WALL faces and the neighbouring
 * periods for PERIODIC lines.  Integer arithmetic throughout; infinite
 * inputs are never pushed. */
static void exact_line(const int64_t *f, int64_t *outv, int n, int bc_lo, int bc_hi, int64_t *v,
                       int64_t *h, int64_t *z_num, int64_t *z_den)
{
    int k = -1;
    int periodic = bc_lo == ORC_BC_PERIODIC;
    int lo_i = periodic ? -n : 0, hi_i = periodic ? 2 * n : n;
    /* push helper: maintains the envelope; intersection of (q, hq) and
     * (r, hr) with q < r is ((hr - hq) + r*r - q*q) / (2 (r - q)). */
#define PUSH(vert, hgt)                                                                  \
    do {                                                                                 \
        int64_t r_ = (vert), hr_ = (hgt);                                                \
        while (k >= 0) {                                                                 \
            int64_t q_ = v[k], hq_ = h[k];                                               \
            int64_t num_ = (hr_ - hq_) + r_ * r_ - q_ * q_, den_ = 2 * (r_ - q_);        \
            /* pop while intersection <= start of top's range */                        \
            if (k > 0 && num_ * z_den[k] <= z_num[k] * den_) {                           \
                k--;                                                                     \
                continue;                                                                \
            }                                                                            \
            k++;                                                                         \
            v[k] = r_;                                                                   \
            h[k] = hr_;                                                                  \
            z_num[k] = num_;                                                             \
            z_den[k] = den_;                                                             \
            break;                                                                       \
        }                                                                                \
        if (k < 0) {                                                                     \
            k = 0;                                                                       \
            v[0] = r_;                                                                   \
            h[0] = hr_;                                                                  \
            z_num[0] = -1;                                                               \
            z_den[0] = 0; /* -infinity */                                                \
        }                                                                                \
    } while (0)
    if (bc_lo == ORC_BC_WALL) PUSH(-1, 0);
    for (int i = lo_i; i < hi_i; i++) {
        int src = ((i % n) + n) % n;
        if (f[src] == ORC_D2_INF) continue;
        PUSH(i, f[src]);
    }
    if (bc_hi == ORC_BC_WALL) PUSH(n, 0);
#undef PUSH
    if (k < 0) {
        for (int i = 0; i < n; i++) outv[i] = ORC_D2_INF;
        return;
    }
    int j = 0;
    for (int i = 0; i < n; i++) {
        /* advance while the next member's range starts strictly below i:
         * z_{j+1} < i  <=>  num < i * den (den > 0) */
        while (j < k && z_num[j + 1] < (int64_t)i * z_den[j + 1]) j++;
        int64_t dd = (int64_t)i - v[j];
        outv[i] = dd * dd + h[j];
    }
}

/* One pass of the exact squared EDT over the lines along `dim` whose first
 * other-axis index lies in [a0, a1): each line is read from `out`, run
 * through exact_line and written back. */
typedef struct {
    int32_t *out;
    geom g;
    int nd, dim, bc_lo, bc_hi;
    long a0, a1;
} exact_job;

static void *exact_worker(void *arg)
{
    exact_job *j = (exact_job *)arg;
    int d1, d2;
    other_dims(j->nd, j->dim, &d1, &d2);
    long s2 = d2 >= 0 ? j->g.shape[d2] : 1;
    long st1 = j->g.stride[d1], st2 = d2 >= 0 ? j->g.stride[d2] : 0;
    int n = (int)j->g.shape[j->dim];
    long st = j->g.stride[j->dim];
    size_t cap = (size_t)(3 * n + 4);
    int64_t *f = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *o = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int64_t *v = (int64_t *)malloc(sizeof(int64_t) * cap), *h = (int64_t *)malloc(sizeof(int64_t) * cap);
    int64_t *zn = (int64_t *)malloc(sizeof(int64_t) * cap), *zd = (int64_t *)malloc(sizeof(int64_t) * cap);
    for (long a = j->a0; a < j->a1; a++)
        for (long b = 0; b < s2; b++) {
            long base = a * st1 + b * st2;
            for (int i = 0; i < n; i++) f[i] = j->out[base + i * st];
            exact_line(f, o, n, j->bc_lo, j->bc_hi, v, h, zn, zd);
            for (int i = 0; i < n; i++)
                j->out[base + i * st] = o[i] >= ORC_D2_INF ? ORC_D2_INF : (int32_t)o[i];
        }
    free(f);
    free(o);
    free(v);
    free(h);
    free(zn);
    free(zd);
    return NULL;
}

/* Exact squared EDT of a C-order 3-D (nd=3) or 2-D (nd=2) binary image,
 * solid = (img == 0), with per-face boundary types bc[2*axis + side].
 * Output int32; ORC_D2_INF where no solid is reachable.  The lines of each
 * pass are split over `nthreads` POSIX threads by their first other-axis
 * index (a join between passes); every line is computed the same way
 * whatever the split, so the output does not depend on nthreads. */
ORC_EXPORT void orc_exact_edt_sq_mt(const uint8_t *img, int nd, const long *shape, const int *bc,
                                    int32_t *out, int nthreads)
{
    geom g = make_geom(nd, shape);
    long total = g.shape[0] * g.shape[1] * g.shape[2];
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    for (long i = 0; i < total; i++) out[i] = img[i] == 0 ? 0 : ORC_D2_INF;
    pthread_t th[256];
    exact_job jobs[256];
    int created[256];
    for (int dim = 0; dim < nd; dim++) {
        int d1, d2;
        other_dims(nd, dim, &d1, &d2);
        long s1 = g.shape[d1];
        for (int t = 0; t < nthreads; t++) {
            exact_job j = {out, g, nd, dim, bc[2 * dim], bc[2 * dim + 1], s1 * t / nthreads,
                           s1 * (t + 1) / nthreads};
            jobs[t] = j;
            created[t] = nthreads > 1 && pthread_create(&th[t], NULL, exact_worker, &jobs[t]) == 0;
            if (!created[t]) exact_worker(&jobs[t]); /* no thread: run the share here */
        }
        for (int t = 0; t < nthreads; t++)
            if (created[t]) pthread_join(th[t], NULL);
    }
}

ORC_EXPORT void orc_exact_edt_sq(const uint8_t *img, int nd, const long *shape, const int *bc,
                                 int32_t *out)
{
    orc_exact_edt_sq_mt(img, nd, shape, bc, out, 1);
}

/* ------------------------------------------------------------------------ */
/* 26-connected multi-label connected components                            */
/* ------------------------------------------------------------------------ */

static int64_t uf_find(int64_t *p, int64_t x)
{
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

static void uf_union(int64_t *p, int64_t a, int64_t b)
{
    a = uf_find(p, a);
    b = uf_find(p, b);
    if (a == b) return;
    if (a < b)
        p[b] = a;
    else
        p[a] = b;
}

/* Labels every non-zero voxel so that 26-neighbours (with wrap on periodic
 * axes) holding the same value share a label.  Labels are numbered 1..N in
 * C-order of each component's first voxel; background is 0.  Returns N. */
ORC_EXPORT int64_t orc_exact_ccl26(const uint8_t *img, long nx, long ny, long nz,
                                   const int *periodic, int64_t *labels)
{
    long n = nx * ny * nz;
    int64_t *p = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    for (long i = 0; i < n; i++) p[i] = i;
    long dims[3] = {nx, ny, nz};
    for (long x = 0; x < nx; x++)
        for (long y = 0; y < ny; y++)
            for (long z = 0; z < nz; z++) {
                long i = (x * ny + y) * nz + z;
                uint8_t val = img[i];
                if (!val) continue;
                for (int dx = -1; dx <= 1; dx++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dz = -1; dz <= 1; dz++) {
                            if (!dx && !dy && !dz) continue;
                            long c[3] = {x + dx, y + dy, z + dz};
                            int ok = 1;
                            for (int a = 0; a < 3; a++) {
                                if (c[a] < 0 || c[a] >= dims[a]) {
                                    if (!periodic[a]) {
                                        ok = 0;
                                        break;
                                    }
                                    c[a] = (c[a] + dims[a]) % dims[a];
                                }
                            }
                            if (!ok) continue;
                            long j = (c[0] * ny + c[1]) * nz + c[2];
                            if (img[j] == val) uf_union(p, i, j);
                        }
            }
    int64_t count = 0;
    for (long i = 0; i < n; i++) labels[i] = 0;
    int64_t *root_label = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    for (long i = 0; i < n; i++) {
        if (!img[i]) continue;
        int64_t r = uf_find(p, i);
        if (!root_label[r]) root_label[r] = ++count;
        labels[i] = root_label[r];
    }
    free(root_label);
    free(p);
    return count;
}

/* ------------------------------------------------------------------------ */
/* Seeded overlapping-sphere rasteriser (input synthesis for tests/bench)   */
/* ------------------------------------------------------------------------ */

/* Voxel (i,j,k) becomes solid (0) when its centroid (i+.5, j+.5, k+.5) lies
 * within a sphere: dx^2+dy^2+dz^2 <= r^2 (particles/spheres.hpp:42-48 on
 * voxel centroids, domain_discretization.py:70-90, box (0,N)).  Periodic
 * axes use the minimum image.  Spheres are added in order; the function
 * stops after the first sphere that brings the porosity (fraction of
 * voxels equal to 1) to <= target, and returns the number of spheres used
 * (or -1 if the list ran out first).  img must be all ones on entry, or
 * the state a previous call left (pore_count carries the count across
 * calls, so the list can be supplied in chunks). */
ORC_EXPORT long orc_rasterize_spheres(uint8_t *img, long nx, long ny, long nz, const int *periodic,
                                      const double *spheres, long n_spheres, double target,
                                      long *pore_count)
{
    long dims[3] = {nx, ny, nz};
    double total = (double)nx * (double)ny * (double)nz;
    long pores = *pore_count;
    for (long s = 0; s < n_spheres; s++) {
        const double *sp = spheres + 4 * s;
        double r = sp[3], r2 = r * r;
        long lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = (long)floor(sp[a] - r - 0.5) - 1;
            hi[a] = (long)ceil(sp[a] + r - 0.5) + 1;
            if (!periodic[a]) {
                if (lo[a] < 0) lo[a] = 0;
                if (hi[a] > dims[a] - 1) hi[a] = dims[a] - 1;
            } else if (hi[a] - lo[a] + 1 > dims[a]) {
                lo[a] = 0;
                hi[a] = dims[a] - 1;
            }
        }
        for (long x = lo[0]; x <= hi[0]; x++) {
            long xi = ((x % nx) + nx) % nx;
            double dx = (double)x + 0.5 - sp[0];
            if (periodic[0] && hi[0] - lo[0] + 1 == nx) {
                dx = (double)xi + 0.5 - sp[0];
                dx -= (double)nx * nearbyint(dx / (double)nx);
            }
            for (long y = lo[1]; y <= hi[1]; y++) {
                long yi = ((y % ny) + ny) % ny;
                double dy = (double)y + 0.5 - sp[1];
                if (periodic[1] && hi[1] - lo[1] + 1 == ny) {
                    dy = (double)yi + 0.5 - sp[1];
                    dy -= (double)ny * nearbyint(dy / (double)ny);
                }
                double dxy = dx * dx + dy * dy;
                if (dxy > r2) continue;
                uint8_t *row = img + (xi * ny + yi) * nz;
                for (long z = lo[2]; z <= hi[2]; z++) {
                    long zi = ((z % nz) + nz) % nz;
                    double dz = (double)z + 0.5 - sp[2];
                    if (periodic[2] && hi[2] - lo[2] + 1 == nz) {
                        dz = (double)zi + 0.5 - sp[2];
                        dz -= (double)nz * nearbyint(dz / (double)nz);
                    }
                    if (dxy + dz * dz <= r2 && row[zi]) {
                        row[zi] = 0;
                        pores--;
                    }
                }
            }
        }
        if ((double)pores / total <= target) {
            *pore_count = pores;
            return s + 1;
        }
    }
    *pore_count = pores;
    return -1;
}
