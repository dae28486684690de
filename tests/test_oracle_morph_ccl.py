"""Pins the oracle's morphology and labelling before they are trusted:
scipy's binary morphology with the reference's ball element
(tests/filters/test_morphological_operators.py:195-225 uses the same
check), scipy's 26-connected labelling, brute-force BFS on tiny tori, and
the reference's connected-components known answers."""

import itertools

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O
from tests.helpers import rand_binary


def ball(r):
    R = int(np.ceil(r))
    g = np.arange(-R, R + 1, dtype=np.float64)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return (x * x + y * y + z * z <= r * r).astype(np.uint8)


@pytest.mark.parametrize("r", [1.0, 2.5, 3, 4.2])
def test_erode_dilate_vs_scipy(r):
    img = rand_binary((30, 26, 22), 0.08, 3)
    box = O.Box(bc=((O.BC_END, O.BC_END),) * 3, pad=((0, 0),) * 3)
    want_e = ndimage.binary_erosion(img, ball(r), border_value=1).astype(np.uint8)
    np.testing.assert_array_equal(O.erode(box, img, r), want_e)
    want_d = ndimage.binary_dilation(img, ball(r)).astype(np.uint8)
    np.testing.assert_array_equal(O.dilate(box, img, r), want_d)


def test_periodic_dilate_vs_tiled_scipy():
    own = rand_binary((20, 18, 16), 0.97, 5)
    r = 3.5
    box = O.Box(bc=((O.BC_PERIODIC, O.BC_PERIODIC),) * 3, pad=((1, 1),) * 3)
    padded = np.pad(own, 1, mode="wrap")
    got = O.dilate(box, padded, r)[1:-1, 1:-1, 1:-1]
    tiled = np.tile(own, (3, 3, 3))
    want = ndimage.binary_dilation(tiled, ball(r)).astype(np.uint8)[20:40, 18:36, 16:32]
    np.testing.assert_array_equal(got, want)


def test_ccl_vs_scipy_label():
    for seed, p in ((1, 0.5), (2, 0.7), (3, 0.3)):
        img = rand_binary((25, 30, 20), p, seed)
        lab, n = O.ccl26(img)
        ref, m = ndimage.label(img, structure=np.ones((3, 3, 3)))
        assert n == m
        np.testing.assert_array_equal(lab, O.canonical_relabel(ref))


def bfs_labels(img, periodic):
    shape = img.shape
    lab = np.zeros(shape, np.int64)
    n = 0
    for idx in itertools.product(*[range(s) for s in shape]):
        if img[idx] == 0 or lab[idx]:
            continue
        n += 1
        stack = [idx]
        lab[idx] = n
        while stack:
            c = stack.pop()
            for d in itertools.product((-1, 0, 1), repeat=3):
                q = []
                ok = True
                for a in range(3):
                    v = c[a] + d[a]
                    if v < 0 or v >= shape[a]:
                        if not periodic[a]:
                            ok = False
                            break
                        v %= shape[a]
                    q.append(v)
                q = tuple(q)
                if ok and img[q] == img[idx] and not lab[q]:
                    lab[q] = n
                    stack.append(q)
    return lab, n


@pytest.mark.parametrize("seed", range(5))
def test_ccl_periodic_multilabel_vs_bfs(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in rng.integers(1, 7, 3))
    img = rng.choice([0, 1, 2], size=shape, p=[0.4, 0.4, 0.2]).astype(np.uint8)
    per = tuple(bool(b) for b in rng.integers(0, 2, 3))
    lab, n = O.ccl26(img, per)
    want, m = bfs_labels(img, per)
    assert n == m
    np.testing.assert_array_equal(lab, want)


def test_ccl_reference_kats():
    # test_connect_components.py:81-136: every voxel its own value
    img = (np.arange(120) % 256).astype(np.uint8).reshape(6, 5, 4)
    assert O.ccl26(img)[1] == 119
    assert O.ccl26(img + 1)[1] == 120


@pytest.mark.parametrize("seed,inlet,radii", [
    (1, True, [np.float64(r) for r in range(6, 0, -1)]),
    (2, False, [4, 3.5, 2, 1]),
    (3, True, [5.5, 3, 1.5]),
])
def test_psd_codes_exact_equals_reference_chain(seed, inlet, radii):
    """The fast pore-size-distribution restatement (exact EDT computed once,
    scipy labelling) gives the radius codes of the reference chain
    (pore_size_distribution: float32 EDT per radius, the oracle's
    labelling), on a padded x END / y,z WALL box with inlet x = 0."""
    from tests.helpers import padded_case
    own = rand_binary((26, 22, 20), 0.25, seed)
    img, _, box = padded_case(own, (("end", "end"), ("wall", "wall"), ("wall", "wall")),
                              inlet=((True, False), (False, False), (False, False)))
    got = O.psd_codes_exact(box, img, radii, inlet=inlet, threads=2)
    ref = O.pore_size_distribution(box, img, radii, inlet=inlet, mode="distance")
    want = np.zeros(img.shape, np.uint8)
    for i, r in enumerate(radii):
        want[ref == r] = i + 1
    assert got.max() > 0
    np.testing.assert_array_equal(got, want)
    # the saturation arithmetic against a direct float64 evaluation
    sw = O.saturation_from_codes(box, img, got, len(radii), own.shape)
    own_sl = (slice(None), slice(1, -1), slice(1, -1))
    pore = int((img[own_sl] == 1).sum())
    for i in range(len(radii)):
        wet = int(((img[own_sl] == 1) & ((got[own_sl] == 0) | (got[own_sl] > i + 1))).sum())
        assert sw[i] == (wet / own.size) / (pore / own.size)


def test_radius_threshold_closed_form():
    """T(r) = r^2 for integer radii (Python int compares in float32, numpy
    float64 in float64); for other radii the largest d2 whose float32 root is
    within r, checked against a direct scan."""
    for r in range(1, 40):
        assert O.radius_threshold(r) == r * r
        assert O.radius_threshold(np.float64(r)) == r * r
    for r in (0.5, 1.5, 2.25, 7.1, np.float64(3.3)):
        want = max(d for d in range(200) if np.sqrt(np.float32(d)) <= (np.float64(r) if isinstance(r, np.float64) else np.float32(r)))
        assert O.radius_threshold(r) == want


def test_threaded_halo_edt_gives_the_same_morphology():
    """erode/dilate with the halo EDT's lines split over threads equal the
    single-threaded chain (the threaded driver runs the same line code)."""
    img = rand_binary((40, 36, 30), 0.1, 9)
    box = O.Box(bc=((O.BC_PERIODIC, O.BC_PERIODIC), (O.BC_END, O.BC_END), (O.BC_WALL, O.BC_WALL)),
                pad=((1, 1), (0, 0), (1, 1)))
    for r in (2, 3.5):
        np.testing.assert_array_equal(O.erode(box, img, r, threads=4), O.erode(box, img, r))
        np.testing.assert_array_equal(O.dilate(box, img, r, threads=4), O.dilate(box, img, r))
