"""Pins the CPU oracle's EDT functions before they are trusted as checkers.

Each oracle function is checked against something it does not share code
with: the reference's own known-answer vectors (tests/filters/test_edt.py),
the reference's own per-line kernels compiled into oracle/_ref, brute force
on tiny inputs, scipy's EDT (exact integer distances recovered from its
feature indices), and the tiled-image identity for periodic domains.
"""

import itertools

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O

needs_ref = pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built (reference tree absent)")


def rand_binary(shape, p_zero, seed):
    # gen_img_random_binary, domain_generation.py:40-60
    rng = np.random.default_rng(seed)
    return rng.choice([0, 1], size=shape, p=[p_zero, 1 - p_zero]).astype(np.uint8)


def brute_d2(img, bc):
    """Exact squared EDT by enumeration: solids plus virtual wall layers,
    periodic axes by all images within one period."""
    nd = img.ndim
    shape = img.shape
    pts = [np.array(p) for p in zip(*np.nonzero(img == 0))]
    # virtual walls: every position of the face layer just outside
    for ax in range(nd):
        for side in (0, 1):
            if bc[2 * ax + side] == O.BC_WALL:
                ranges = [range(-1, s + 1) for s in shape]
                ranges[ax] = [-1] if side == 0 else [shape[ax]]
                pts += [np.array(p) for p in itertools.product(*ranges)]
    out = np.full(shape, O.D2_INF, dtype=np.int64)
    shifts = [[0] if bc[2 * ax] != O.BC_PERIODIC else [-shape[ax], 0, shape[ax]] for ax in range(nd)]
    for idx in itertools.product(*[range(s) for s in shape]):
        best = O.D2_INF
        p = np.array(idx)
        for q in pts:
            for sh in itertools.product(*shifts):
                d = int(np.sum((p - q - np.array(sh)) ** 2))
                best = min(best, d)
        out[idx] = best
    return out


# ---------------------------------------------------------------------------
# Reference known-answer vectors (tests/filters/test_edt.py)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("use_ref", [False, pytest.param(True, marks=needs_ref)])
def test_initial_envelope_kats(use_ref):
    # test_edt.py:31-61
    img = np.array([1, 1, 0, 0, 1, 1, 0, 1, 1, 1, 1, 1, 1, 0, 1], dtype=np.uint8)
    np.testing.assert_array_equal(
        O.sq_edt_1d(img, use_ref=use_ref),
        [4.0, 1.0, 0.0, 0.0, 1.0, 1.0, 0.0, 1.0, 4.0, 9.0, 9.0, 4.0, 1.0, 0.0, 1.0])
    img = np.array([1, 1, 0, 1, 1, 1, 1, 1, 1], dtype=np.uint8)
    np.testing.assert_array_equal(O.sq_edt_1d(img, 1, 1, use_ref=use_ref),
                                  [1.0, 1.0, 0.0, 1.0, 4.0, 9.0, 9.0, 4.0, 1.0])
    np.testing.assert_array_equal(O.sq_edt_1d(img, 2, 4, use_ref=use_ref),
                                  [4.0, 1.0, 0.0, 1.0, 4.0, 9.0, 16.0, 25.0, 16.0])


def label_sources_pass0(v, lower, upper):
    """The first pass as pmx_edt_ref's integer form computes it on a
    multi-label line: sources are the zeros (value 0), the correctors'
    virtual zeros at -lower and n - 1 + upper, and the voxels beside a
    change between two different non-zero labels (value 1); squared."""
    n = len(v)
    chg = [x > 0 and v[x - 1] != 0 and v[x] != 0 and v[x - 1] != v[x] for x in range(n)] + [False]
    zeros = [x for x in range(n) if v[x] == 0]
    edges = [x for x in range(n) if chg[x] or chg[x + 1]]
    out = np.empty(n, np.float32)
    for x in range(n):
        c = [abs(x - z) for z in zeros] + [1 + abs(x - e) for e in edges]
        if lower is not None:
            c.append(x + lower)
        if upper is not None:
            c.append(n - 1 + upper - x)
        out[x] = np.inf if not c else float(min(c) ** 2)
    return out


@pytest.mark.parametrize("use_ref", [False, pytest.param(True, marks=needs_ref)])
def test_label_pass0_sources_match_the_first_pass(use_ref):
    """The label-aware integer pass 0 (edge voxels of a label change as
    distance-1 sources) equals squared_edt_1d on random multi-label lines
    with and without correctors (_distance.hpp:136-182)."""
    rng = np.random.default_rng(0)
    for _ in range(1500):
        n = int(rng.integers(1, 40))
        p0 = rng.random() * 0.3
        v = rng.choice([0, 1, 2, 3], size=n, p=[p0, (1 - p0) / 2, (1 - p0) / 4, (1 - p0) / 4]).astype(np.uint8)
        lo = None if rng.random() < 0.5 else int(rng.integers(1, 10))
        up = None if rng.random() < 0.5 else int(rng.integers(1, 10))
        want = O.sq_edt_1d(v, np.inf if lo is None else float(lo), np.inf if up is None else float(up), 1.0,
                           use_ref=use_ref)
        np.testing.assert_array_equal(label_sources_pass0(v, lo, up), want, err_msg=f"{v} {lo} {up}")


def test_boundary_hull_kats():
    # test_edt.py:146-218
    img = np.array([4, np.inf, 0, np.inf, np.inf, np.inf, 4, 6, 16], dtype=np.float32)
    h = O.boundary_hull_1d(img, 0, 9, 1.0, 2, True)
    assert (h[0]["height"], h[0]["vertex"], h[0]["range"]) == (4.0, 0, -np.inf)
    assert (h[1]["height"], h[1]["vertex"], h[1]["range"]) == (0.0, 2, 0.0)
    h = O.boundary_hull_1d(img, 1, 9, 1.0, 3, True)
    assert (h[0]["height"], h[0]["vertex"]) == (0.0, 2)
    h = O.boundary_hull_1d(img, 0, 9, 1.0, 1, False)
    assert (h[0]["height"], h[0]["vertex"], h[0]["range"]) == (16.0, 8, 12.5)
    h = O.boundary_hull_1d(img, 0, 9, 1.0, 4, False)
    want = [(16.0, 8, 12.5), (6.0, 7, 7.5), (4.0, 6, 4.5), (0.0, 2, 0.0)]
    assert [(x["height"], x["vertex"], x["range"]) for x in h[:4]] == want
    h = O.boundary_hull_1d(img, 0, 8, 1.0, 5, False)
    assert [(x["height"], x["vertex"], x["range"]) for x in h[:3]] == want[1:]


def test_nearest_index_kats():
    # tests/core/test_voxels.py nearest-boundary semantics: first/last index
    # of label within [lower_skip, n - upper_skip), else -1
    img = np.array([1, 1, 0, 1, 0, 1], dtype=np.uint8)
    assert O.nearest_index_1d(img, 0) == 2
    assert O.nearest_index_1d(img, 0, forward=False) == 4
    assert O.nearest_index_1d(img, 0, lower_skip=3) == 4
    assert O.nearest_index_1d(img, 0, upper_skip=2, forward=False) == 2
    assert O.nearest_index_1d(np.ones(5, np.uint8), 0) == -1


# ---------------------------------------------------------------------------
# Restated per-line kernels == the reference's compiled kernels
# ---------------------------------------------------------------------------

@needs_ref
def test_lines_match_reference_kernels():
    rng = np.random.default_rng(7)
    for trial in range(400):
        n = int(rng.integers(1, 40))
        img = rng.choice([0, 1, 2, 3], size=n, p=[0.2, 0.6, 0.1, 0.1]).astype(np.uint8)
        res = float(rng.choice([1.0, 0.5, 0.01, 0.32325245, 233.0]))
        lo = float(rng.choice([np.inf, 0, 1, 3, 7]))
        hi = float(rng.choice([np.inf, 0, 2, 5]))
        a = O.sq_edt_1d(img, lo, hi, res)
        b = O.sq_edt_1d(img, lo, hi, res, use_ref=True)
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        f = O.sq_edt_1d(img, lo, hi, res)
        f[f == np.inf] = O.FLT_MAX
        lower = [(int(-rng.integers(1, 6)), float(rng.integers(0, 30)), float(rng.uniform(-20, 0)))
                 for _ in range(int(rng.integers(0, 4)))]
        upper = [(int(rng.integers(0, 6)), float(rng.integers(0, 30)), float(rng.uniform(0, 20)))
                 for _ in range(int(rng.integers(0, 4)))]
        a = O.parabolic_1d(f, lower, upper, res)
        b = O.parabolic_1d(f, lower, upper, res, use_ref=True)
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        for fw in (True, False):
            nh = int(rng.integers(1, 6))
            if n == 1 and f[0] >= O.FLT_MAX:
                # return_boundary_hull reads hull_height[1] of a 1-element
                # vector here (_distance.hpp:316,329,342): undefined
                # behaviour, so there is nothing to match.
                continue
            got = O.boundary_hull_1d(f, 0, n, res, nh, fw)
            got = [(g["vertex"], g["height"], g["range"]) for g in got]
            want = O.ref_boundary_hull_line(f, res, nh, 0, fw)
            assert got == want


@needs_ref
@pytest.mark.parametrize("shape,periodic,res,p0,seed", [
    ((10, 50), (False, False), (1.0, 1.0), 0.1, 1),
    ((60, 60), (True, True), (1.0, 1.0), 0.1, 4),
    ((6, 6), (True, True), (1.0, 1.0), 0.1, 5),
    ((17, 23, 19), (False, False, False), (1.0, 1.0, 1.0), 0.1, 3),
    ((20, 20, 20), (True, True, True), (1.0, 1.0, 1.0), 0.1, 1),
    ((15, 12, 18), (True, False, True), (1.0, 1.0, 1.0), 0.05, 9),
    ((14, 14, 14), (True, True, True), (1, 0.32325245, 233), 0.2, 1246),
    ((12, 16, 10), (False, True, False), (0.01, 0.01, 0.01), 0.3, 11),
])
def test_volume_matches_reference_kernels(shape, periodic, res, p0, seed):
    img = rand_binary(shape, p0, seed)
    a = O.ref_edt_sq(img, periodic, res)
    b = O.ref_edt_sq(img, periodic, res, use_ref=True)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


# ---------------------------------------------------------------------------
# Exact integer EDT: brute force, scipy, tiled periodic images
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(6))
def test_exact_vs_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    for nd in (2, 3):
        shape = tuple(int(s) for s in rng.integers(1, 6, size=nd))
        img = rand_binary(shape, float(rng.uniform(0.05, 0.4)), seed)
        bc = []
        for ax in range(nd):
            t = int(rng.integers(0, 3))
            bc += [t, t] if t == O.BC_PERIODIC else [int(rng.integers(0, 2)), int(rng.integers(0, 2))]
        got = O.exact_edt_sq(img, bc).astype(np.int64)
        np.testing.assert_array_equal(got, brute_d2(img, bc))


def scipy_exact_d2(img, sampling=None):
    _, ind = ndimage.distance_transform_edt(img, return_indices=True)
    grid = np.indices(img.shape)
    return np.sum((ind - grid) ** 2, axis=0)


@pytest.mark.parametrize("shape,p0,seed", [((40, 33, 29), 0.1, 1), ((64, 64), 0.02, 2), ((30, 30, 30), 0.005, 3)])
def test_exact_vs_scipy(shape, p0, seed):
    img = rand_binary(shape, p0, seed)
    np.testing.assert_array_equal(O.exact_edt_sq(img), scipy_exact_d2(img))


def test_exact_wall_equals_zero_padding():
    img = rand_binary((21, 18, 25), 0.02, 8)
    bc = [O.BC_WALL] * 6
    padded = np.pad(img, 1, constant_values=0)
    want = scipy_exact_d2(padded)[1:-1, 1:-1, 1:-1]
    np.testing.assert_array_equal(O.exact_edt_sq(img, bc), want)


@pytest.mark.parametrize("shape,p0,seed", [((20, 20, 20), 0.1, 1), ((16, 22, 12), 0.01, 5), ((60, 60), 0.1, 4)])
def test_exact_periodic_vs_tiled(shape, p0, seed):
    # test_edt.py:64-143,221-263: periodic EDT == centre block of the tiled EDT
    img = rand_binary(shape, p0, seed)
    nd = img.ndim
    tiled = np.tile(img, (3,) * nd)
    centre = tuple(slice(s, 2 * s) for s in shape)
    want = scipy_exact_d2(tiled)[centre]
    np.testing.assert_array_equal(O.exact_edt_sq(img, [O.BC_PERIODIC] * (2 * nd)), want)


def test_exact_no_solid_is_inf():
    img = np.ones((5, 6, 7), np.uint8)
    assert np.all(O.exact_edt_sq(img) == O.D2_INF)
    assert np.all(O.ref_edt_sq(img) == O.FLT_MAX)


# ---------------------------------------------------------------------------
# Reference float path == exact integers at unit resolution
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("shape,p0,seed", [((40, 40, 40), 0.1, 1), ((33, 47, 21), 0.01, 2), ((64, 64, 64), 0.001, 3)])
def test_reference_float_is_integer_exact(shape, p0, seed):
    img = rand_binary(shape, p0, seed)
    ref = O.ref_edt_sq(img)
    np.testing.assert_array_equal(ref, O.d2_to_f32(O.exact_edt_sq(img)))


def test_reference_edt_matches_scipy_kats():
    # test_edt.py:10-28
    img = rand_binary((10, 50), 0.1, 1)
    np.testing.assert_array_almost_equal(O.ref_edt(img), ndimage.distance_transform_edt(img))
    img = np.ones([4, 4, 4], dtype=np.uint8)
    img[3, 3, 3] = 0
    np.testing.assert_array_almost_equal(O.ref_edt(img), ndimage.distance_transform_edt(img))


def test_reference_periodic_vs_tiled():
    # test_edt.py:221-263 (100^3 there; 40^3 here keeps the CPU suite short)
    img = rand_binary((40, 40, 40), 0.1, 1)
    tiled = np.tile(img, (3, 3, 3))
    want = ndimage.distance_transform_edt(tiled)[40:80, 40:80, 40:80]
    got = np.sqrt(O.ref_edt_sq(img, (True, True, True)))
    np.testing.assert_array_almost_equal(got, want)


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("res", [(1.0, 1.0, 1.0), (0.5, 0.25, 0.125)])
def test_threaded_driver_is_identical(threads, res):
    """The threaded CPU-baseline driver splits each pass's lines over host
    threads; its output equals the single-threaded driver bit for bit."""
    rng = np.random.default_rng(threads)
    img = (rng.random((23, 19, 17)) < 0.75).astype(np.uint8)
    use_ref = O.ref_lib() is not None
    want = O.ref_edt_sq(img, resolution=res, use_ref=use_ref)
    got = O.ref_edt_sq_threads(img, threads, resolution=res, use_ref=use_ref)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
