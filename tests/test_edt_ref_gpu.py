"""GPU parity of the reference-arithmetic EDT (pmx_edt_ref) and the halo
morphology (pmx_halo_morph) at non-unit resolution: bitwise float32
against the oracle's restatement of the reference (itself pinned against
the reference's compiled kernels in tests/test_oracle_edt.py)."""

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import E, P, W, padded_case, rand_binary

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,periodic,res,p0,seed", [
    ((17, 23, 19), (False, False, False), (1.0, 1.0, 1.0), 0.1, 3),
    ((20, 20, 20), (True, True, True), (1.0, 1.0, 1.0), 0.1, 1),
    ((15, 12, 18), (True, False, True), (1.0, 1.0, 1.0), 0.05, 9),
    ((14, 14, 14), (True, True, True), (1, 0.32325245, 233), 0.2, 1246),
    ((12, 16, 10), (False, True, False), (0.01, 0.01, 0.01), 0.3, 11),
    ((30, 31, 29), (False, False, False), (0.01, 0.01, 0.01), 0.05, 12),
    ((25, 9, 40), (True, True, False), (0.1, 0.1, 0.1), 0.02, 13),
])
def test_edt3d_reference_mode_bitwise(shape, periodic, res, p0, seed):
    from pmmoto_amd.filters import distance
    img = rand_binary(shape, p0, seed)
    print(f"[pmx_edt_ref] {shape} periodic={periodic} res={res}", flush=True)
    got = distance.edt3d(img, periodic=periodic, resolution=res, squared=True, mode="reference")
    want = O.ref_edt_sq(img, periodic, res)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    got = distance.edt3d(img, periodic=periodic, resolution=res, mode="reference")
    np.testing.assert_array_equal(got.view(np.uint32), np.sqrt(want).view(np.uint32))


def test_multilabel_pass0():
    from pmmoto_amd.filters import distance
    rng = np.random.default_rng(4)
    img = rng.choice([0, 1, 2, 3], size=(21, 17, 15), p=[0.1, 0.5, 0.2, 0.2]).astype(np.uint8)
    got = distance.edt3d(img, squared=True, mode="reference")
    np.testing.assert_array_equal(got, O.ref_edt_sq(img))


def test_multilabel_auto_mode_follows_reference():
    """A multi-label image at unit resolution (the default path) gets the
    reference's label-aware first pass, not the binary exact EDT."""
    from pmmoto_amd.filters import distance
    rng = np.random.default_rng(7)
    img = rng.choice([0, 1, 2], size=(19, 16, 13), p=[0.1, 0.6, 0.3]).astype(np.uint8)
    want = O.ref_edt_sq(img)
    got = distance.edt3d(img, squared=True)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(distance.edt(img).view(np.uint32), np.sqrt(want).view(np.uint32))
    # a 0/1 image keeps the exact kernels, which agree with the reference there
    b = (img > 0).astype(np.uint8)
    np.testing.assert_array_equal(distance.edt3d(b, squared=True), O.ref_edt_sq(b))


@pytest.mark.parametrize("n", [1, 15, 16, 17, 4099, 1 << 20])
def test_max_u8(n):
    import torch

    from pmmoto_amd import _device, _lib
    rng = np.random.default_rng(n)
    a = rng.integers(0, 3, size=n, dtype=np.uint8)
    a[rng.integers(0, n)] = 200 if n > 16 else 9
    t = torch.from_numpy(a).cuda()
    out = torch.empty(1, dtype=torch.int32, device="cuda")
    print(f"[max_u8] n={n}", flush=True)
    _lib.check(_lib.lib().pmx_max_u8(_device.ptr(t), n, _device.ptr(out), _device.stream()))
    assert int(out.item()) == int(a.max())


@pytest.mark.parametrize("periodic,res", [((False, False), (1.0, 1.0)), ((True, True), (1.0, 1.0)),
                                          ((True, False), (0.5, 0.25))])
def test_edt2d_reference_mode(periodic, res):
    from pmmoto_amd.filters import distance
    img = rand_binary((60, 60), 0.1, 4)
    got = distance.edt2d(img, periodic=periodic, resolution=res, mode="reference")
    want = np.sqrt(O.ref_edt_sq(img, periodic, res))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_edt_subdomain_nonunit_resolution():
    import pmmoto_amd
    from pmmoto_amd.filters import distance
    # PMMoTo's default unit box: resolution 1/N (test_edt.py:353-366)
    img = rand_binary((10, 10, 10), 0.1, 1)
    sd = pmmoto_amd.initialize(voxels=(10, 10, 10), pad=(1, 1, 1))
    out = distance.edt(img, sd)
    np.testing.assert_array_equal(out, O.ref_edt(img, resolution=sd.domain.resolution))
    assert np.max(out) < 1.0


@pytest.mark.parametrize("bts", [((E, E), (E, E), (E, E)), ((W, W), (W, W), (W, W)),
                                 ((P, P), (P, P), (P, P)), ((E, E), (P, P), (W, W))])
@pytest.mark.parametrize("radius", [0.03, 0.05, np.float64(0.0412)])
def test_halo_morphology_nonunit_resolution(bts, radius):
    import pmmoto_amd
    from pmmoto_amd.filters import morphological_operators as M
    own = rand_binary((24, 22, 20), 0.05, 21)
    img, sd, box = padded_case(own, bts)
    # rescale to a 0.01 voxel size, as the reference's tests use
    BT = pmmoto_amd.BoundaryType
    sd = pmmoto_amd.initialize(own.shape, box=tuple((0, 0.01 * v) for v in own.shape),
                               boundary_types=tuple((BT(a), BT(b)) for a, b in bts))
    box.resolution = sd.domain.resolution
    print(f"[pmx_halo_morph] {bts} r={radius!r}", flush=True)
    np.testing.assert_array_equal(M.erode(sd, img, radius), O.erode(box, img, radius))
    sparse = rand_binary(own.shape, 0.995, 22)
    img2, _, _ = padded_case(sparse, bts)
    np.testing.assert_array_equal(M.dilate(sd, img2, radius), O.dilate(box, img2, radius))
