"""GPU parity of morphology, labelling, inlet connectivity and porosimetry
(through the C ABI) against the oracle's restatement of the reference
drivers.  Bar: bit-exact uint8 images and labels."""

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import E, P, W, padded_case, rand_binary

pytestmark = pytest.mark.gpu

BT_CASES = {
    "end": ((E, E), (E, E), (E, E)),
    "wall": ((W, W), (W, W), (W, W)),
    "periodic": ((P, P), (P, P), (P, P)),
    "mixed": ((E, E), (P, P), (W, W)),
    "cfg3": ((E, E), (W, W), (W, W)),
}


@pytest.mark.parametrize("bts", list(BT_CASES))
@pytest.mark.parametrize("radius", [1, 2.5, np.float64(3.0000001), np.float32(4.5), 5])
def test_erode_dilate_match_oracle(bts, radius):
    from pmmoto_amd.filters import morphological_operators as M
    own = rand_binary((22, 25, 19), 0.06, 7)
    img, sd, box = padded_case(own, BT_CASES[bts])
    print(f"[pmx_morph] {bts} r={radius!r} shape={img.shape}", flush=True)
    np.testing.assert_array_equal(M.erode(sd, img, radius), O.erode(box, img, radius))
    sparse = rand_binary((22, 25, 19), 0.995, 8)
    img2, _, _ = padded_case(sparse, BT_CASES[bts])
    np.testing.assert_array_equal(M.dilate(sd, img2, radius), O.dilate(box, img2, radius))


@pytest.mark.parametrize("bts", ["wall", "periodic", "mixed"])
@pytest.mark.parametrize("radius", [12, 20])
def test_erode_dilate_large_radius_match_oracle(bts, radius):
    """Radii whose threshold needs the 32-row halo of the windowed passes
    (20) or fills the 16-row one (12), against the oracle."""
    from pmmoto_amd.filters import morphological_operators as M
    density = {12: 3e-4, 20: 6e-5}[radius]  # about 10 % eroded-kept and 70-90 % dilated: mixed results
    own = rand_binary((44, 38, 70), density, 17)
    img, sd, box = padded_case(own, BT_CASES[bts])
    print(f"[pmx_morph] {bts} r={radius} shape={img.shape}", flush=True)
    np.testing.assert_array_equal(M.erode(sd, img, radius), O.erode(box, img, radius))
    sparse = rand_binary((44, 38, 70), 1 - density, 18)
    img2, _, _ = padded_case(sparse, BT_CASES[bts])
    np.testing.assert_array_equal(M.dilate(sd, img2, radius), O.dilate(box, img2, radius))


def test_opening_closing():
    from pmmoto_amd.filters import morphological_operators as M
    own = rand_binary((30, 30, 30), 0.1, 9)
    img, sd, box = padded_case(own, BT_CASES["mixed"])
    want_open = O.dilate(box, O.erode(box, img, 2.0), 2.0)
    np.testing.assert_array_equal(M.opening(sd, img, 2.0), want_open)
    want_close = O.erode(box, O.dilate(box, img, 2.0), 2.0)
    np.testing.assert_array_equal(M.closing(sd, img, 2.0), want_close)


@pytest.mark.parametrize("bts", ["end", "periodic", "mixed"])
def test_connect_components_match_oracle(bts):
    from pmmoto_amd.filters import connected_components as CC
    rng = np.random.default_rng(3)
    own = rng.choice([0, 1, 2], size=(30, 28, 33), p=[0.45, 0.45, 0.1]).astype(np.uint8)
    img, sd, box = padded_case(own, BT_CASES[bts])
    print(f"[pmx_ccl26] {bts} shape={img.shape}", flush=True)
    labels, count = CC.connect_components(img, sd)
    want, n = O.label_padded(box, img)
    assert labels.dtype == np.uint64
    np.testing.assert_array_equal(labels.astype(np.int64), want)
    assert count == (n + 1 if bts in ("periodic", "mixed") else n)


def test_connect_components_reference_kats():
    from pmmoto_amd.filters import connected_components as CC
    # tests/filters/test_connect_components.py:81-186 on a 6x5x4 domain
    own = (np.arange(120) % 256).astype(np.uint8).reshape(6, 5, 4)
    img, sd, _ = padded_case(own, BT_CASES["end"])
    assert CC.connect_components(img, sd)[1] == 119
    img, sd, _ = padded_case(own + 1, BT_CASES["end"])
    assert CC.connect_components(img, sd)[1] == 120
    for bts in (((P, P), (E, E), (E, E)), ((E, E), (P, P), (E, E)), ((E, E), (E, E), (P, P)),
                ((P, P), (P, P), (E, E)), ((P, P), (E, E), (P, P))):
        img, sd, _ = padded_case(own, bts)
        assert CC.connect_components(img, sd)[1] == 120


def test_inlet_outlet_isolated_reference_kats():
    import pmmoto_amd
    from pmmoto_amd.filters import connected_components as CC
    # test_connect_components.py:250-279
    sd = pmmoto_amd.initialize((20, 20, 20), box=((0, 20),) * 3, inlet=((1, 0), (0, 0), (0, 0)))
    img = np.zeros(sd.voxels, dtype=np.uint8)
    img[0:10, 5:10, 5:10] = 1
    img[12:15, 12:15, 12:15] = 1
    out = CC.inlet_connected_img(sd, img)
    assert np.all(out[12:15, 12:15, 12:15] == 0) and np.all(out[0:10, 5:10, 5:10] == 1)
    sd = pmmoto_amd.initialize((20, 20, 20), box=((0, 20),) * 3, outlet=((1, 0), (0, 0), (0, 0)))
    out = CC.outlet_connected_img(sd, img)
    assert np.all(out[12:15, 12:15, 12:15] == 0) and np.all(out[0:10, 5:10, 5:10] == 1)
    # :314-344
    sd = pmmoto_amd.initialize((20, 20, 20), box=((0, 20),) * 3, inlet=((1, 0), (0, 0), (0, 0)),
                               outlet=((0, 1), (0, 0), (0, 0)))
    img = np.zeros(sd.voxels, dtype=np.uint8)
    img[:, 2:4, 5:10] = 1
    img[:, 6:8, 1:2] = 2
    out = CC.inlet_outlet_connected_img(sd, img)
    assert np.all(out[:, 2:4, 5:10] == 1) and np.all(out[:, 6:8, 1:2] == 2)
    assert sorted(CC.inlet_outlet_connections(sd, out)) == [1, 2]
    out = CC.inlet_outlet_connected_img(sd, img, phase=2)
    assert np.all(out[:, 2:4, 5:10] == 0) and np.all(out[:, 6:8, 1:2] == 2)
    # :347-377
    sd = pmmoto_amd.initialize((30, 30, 30), box=((0, 30),) * 3, inlet=((1, 0), (0, 0), (0, 0)),
                               outlet=((0, 1), (0, 0), (0, 0)))
    img = np.zeros(sd.voxels, dtype=np.uint8)
    img[1:3, 2:4, 5:10] = 1
    img[6:9, 6:8, 1:2] = 2
    img[:, 6:8, 3:4] = 3
    out = CC.isolated_img(sd, img)
    assert np.all(out[1:3, 2:4, 5:10] == 1) and np.all(out[6:9, 6:8, 1:2] == 2)
    assert np.all(out[:, 6:8, 3:4] == 0)
    out = CC.isolated_img(sd, img, phase=2)
    assert np.all(out[1:3, 2:4, 5:10] == 0) and np.all(out[6:9, 6:8, 1:2] == 2)


@pytest.mark.parametrize("bts", ["cfg3", "mixed", "periodic"])
def test_inlet_connected_match_oracle(bts):
    from pmmoto_amd.filters import connected_components as CC
    own = rand_binary((26, 24, 22), 0.35, 12)
    inlet = ((True, False), (False, False), (False, False))
    if bts == "periodic":
        inlet = ((False, False),) * 3
    img, sd, box = padded_case(own, BT_CASES[bts], inlet=inlet)
    np.testing.assert_array_equal(CC.inlet_connected_img(sd, img), O.inlet_connected(box, img))


@pytest.mark.parametrize("inlet", [False, True])
def test_porosimetry_match_oracle(inlet):
    from pmmoto_amd.domain_generation import gen_pm, pack_radii, sphere_pack
    from pmmoto_amd.filters import porosimetry as PO
    img_t, _ = sphere_pack((40, 36, 44), 0.45, pack_radii("uniform", lo=3, hi=7), seed=21)
    own = img_t.cpu().numpy()
    img, sd, box = padded_case(own, BT_CASES["cfg3"], inlet=((True, False), (False, False), (False, False)),
                               outlet=((False, True), (False, False), (False, False)))
    pm = gen_pm(sd, img)
    for r in (1, 2.0, 3.5, np.float64(4.0)):
        print(f"[porosimetry] inlet={inlet} r={r!r}", flush=True)
        got = PO.porosimetry(sd, pm, r, inlet=inlet, mode="distance")
        np.testing.assert_array_equal(got, O.porosimetry(box, img, r, inlet))


def test_pore_size_distribution_match_oracle():
    from pmmoto_amd.domain_generation import gen_pm, pack_radii, sphere_pack
    from pmmoto_amd.filters import porosimetry as PO
    img_t, _ = sphere_pack((36, 40, 32), 0.4, pack_radii("lognormal", mean=5.0, sigma=0.3), seed=31)
    img, sd, box = padded_case(img_t.cpu().numpy(), BT_CASES["cfg3"],
                               inlet=((True, False), (False, False), (False, False)))
    pm = gen_pm(sd, img)
    radii = np.arange(8, 0, -1).astype(np.float64)
    got = PO.pore_size_distribution(sd, pm, radii, inlet=True, mode="distance")
    np.testing.assert_array_equal(got, O.pore_size_distribution(box, img, radii, inlet=True))
    got_all = PO.pore_size_distribution(sd, pm, [1, 3, 2], inlet=False)
    np.testing.assert_array_equal(got_all, O.pore_size_distribution(box, img, [3, 2, 1], inlet=False,
                                                                     mode="hybrid"))


@pytest.mark.parametrize("mode", ["distance", "hybrid", "morph"])
def test_fused_sweep_equals_per_radius_sweep_and_oracle(mode):
    """pmx_psd_sweep (one EDT + incremental union-find + one EDT per level)
    against the per-radius porosimetry path and the oracle."""
    from pmmoto_amd.domain_generation import pack_radii, sphere_pack
    from pmmoto_amd.filters import porosimetry as PO
    img_t, _ = sphere_pack((48, 40, 44), 0.38, pack_radii("lognormal", mean=5.0, sigma=0.3), seed=51)
    img, sd, box = padded_case(img_t.cpu().numpy(), BT_CASES["cfg3"],
                               inlet=((True, False), (False, False), (False, False)))
    import torch
    pm_t = torch.from_numpy(img).cuda()
    radii = [np.float64(r) for r in (7.5, 6.0, 5.2, 4.0, 3.3, 2.5, 2.0, 1.5, 1.0)]
    fused = PO.psd_sweep_fused(sd, pm_t, radii, inlet=True, mode=mode).cpu().numpy()
    levels = PO.psd_sweep_levels(sd, pm_t, radii, inlet=True, mode=mode).cpu().numpy()
    np.testing.assert_array_equal(fused, levels)
    want = O.pore_size_distribution(box, img, np.array(radii), inlet=True, mode=mode)
    got = np.where(fused > 0, np.array([0.0] + [float(r) for r in radii])[fused], 0.0)
    np.testing.assert_array_equal(got, want)


def test_fused_sweep_pad_with_pore_voxel():
    """A wall pad layer holding a pore voxel is not cut off (only all-solid
    pad layers are): the fused sweep still equals the per-radius path."""
    from pmmoto_amd.domain_generation import pack_radii, sphere_pack
    from pmmoto_amd.filters import porosimetry as PO
    img_t, _ = sphere_pack((30, 26, 28), 0.4, pack_radii("lognormal", mean=4.0, sigma=0.3), seed=53)
    img, sd, _ = padded_case(img_t.cpu().numpy(), BT_CASES["cfg3"],
                             inlet=((True, False), (False, False), (False, False)))
    img = img.copy()
    img[5:9, 0, 6:12] = 1  # pore voxels in the y = 0 pad, next to the own region
    import torch
    pm_t = torch.from_numpy(img).cuda()
    origin, n, bc = PO._sweep_box(sd, pm_t)
    assert origin[1] == 0 and origin[2] == 1 and n[2] == img.shape[2] - 2
    radii = [np.float64(r) for r in (5.0, 3.5, 2.0, 1.0)]
    fused = PO.psd_sweep_fused(sd, pm_t, radii, inlet=True, mode="distance").cpu().numpy()
    levels = PO.psd_sweep_levels(sd, pm_t, radii, inlet=True, mode="distance").cpu().numpy()
    np.testing.assert_array_equal(fused, levels)


def test_porosimetry_modes_and_inlet_kat():
    import pmmoto_amd
    from pmmoto_amd.domain_generation import gen_pm
    from pmmoto_amd.filters import porosimetry as PO
    # test_porosimetry.py:81-103 in voxel units (radius 2 voxels)
    sd = pmmoto_amd.initialize((100, 100, 100), box=((0, 100),) * 3, inlet=((1, 0), (0, 0), (0, 0)))
    img = np.zeros(sd.voxels, dtype=np.uint8)
    img[0:10, 5:10, 5:10] = 1
    img[12:35, 12:35, 12:35] = 1
    pm = gen_pm(sd, img)
    no_inlet = PO.porosimetry(sd, pm, 2, mode="morph", inlet=False)
    with_inlet = PO.porosimetry(sd, pm, 2, mode="morph", inlet=True)
    assert not np.array_equal(no_inlet, with_inlet)
    assert np.array_equal(PO.porosimetry(sd, pm, 2, mode="hybrid"), PO.porosimetry(sd, pm, [2, 2], mode="distance"))
    with pytest.raises(ValueError):
        PO.porosimetry(sd, pm, [2, 2, 2])


def test_ccl26_many_rank_tiles_match_oracle():
    """More than 8192 rank tiles of 4096 voxels: the tile-count scan runs
    several rounds, and every later round's offsets reach the labels."""
    rng = np.random.default_rng(5)
    img = rng.choice([0, 1], size=(344, 336, 300), p=[0.9, 0.1]).astype(np.uint8)
    labels, count = _ccl26_gpu(img, (True, False, True))
    want, n = O.ccl26(img, (True, False, True))
    assert count == n
    np.testing.assert_array_equal(labels, want)


def _ccl26_gpu(img, per):
    import ctypes as C

    import torch

    from pmmoto_amd import _device, _lib
    from pmmoto_amd.filters.distance import euclidean_distance_transform as EDT
    shape = img.shape
    bc = [_lib.BC_PERIODIC if per[a // 2] else _lib.BC_END for a in range(6)]
    g = EDT._geom(shape, (0, 0, 0), shape, bc)
    L = _lib.lib()
    nb = C.c_size_t()
    _lib.check(L.pmx_ccl_workspace_size(C.byref(g), C.byref(nb)))
    ws = _device.workspace(nb.value)
    dimg = torch.from_numpy(img).cuda()
    labels = torch.empty(shape, dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    print(f"[pmx_ccl26] shape={shape} periodic={per}", flush=True)
    _lib.check(L.pmx_ccl26(_device.ptr(dimg), C.byref(g), _device.ptr(labels), None, _device.ptr(count),
                           _device.ptr(ws), nb.value, _device.stream()))
    torch.cuda.synchronize()
    return labels.cpu().numpy().astype(np.int64), int(count.item())


@pytest.mark.parametrize("shape,per", [
    ((20, 19, 150), (True, True, True)),     # three z tiles, the last partial; wrap on every axis
    ((17, 9, 130), (False, False, False)),
    ((9, 70, 65), (False, True, True)),      # a one-voxel z tile
    ((3, 2, 200), (True, False, True)),      # one x/y tile, wrap inside it
    ((1, 1, 64), (False, False, True)),      # one full z tile that wraps onto itself
    ((26, 17, 2), (True, True, True)),
])
def test_ccl26_tiles_match_oracle(shape, per):
    """pmx_ccl26 on boxes of several (and partial) 8 x 8 x 64 labelling
    tiles per axis, multi-label values, periodic and END faces."""
    import ctypes as C

    import torch

    from pmmoto_amd import _device, _lib
    from pmmoto_amd.filters.distance import euclidean_distance_transform as EDT
    rng = np.random.default_rng(sum(shape))
    img = rng.choice([0, 1, 2], size=shape, p=[0.42, 0.46, 0.12]).astype(np.uint8)
    bc = [_lib.BC_PERIODIC if per[a // 2] else _lib.BC_END for a in range(6)]
    g = EDT._geom(shape, (0, 0, 0), shape, bc)
    L = _lib.lib()
    nb = C.c_size_t()
    _lib.check(L.pmx_ccl_workspace_size(C.byref(g), C.byref(nb)))
    ws = _device.workspace(nb.value)
    dimg = torch.from_numpy(img).cuda()
    labels = torch.empty(shape, dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int64, device="cuda")
    print(f"[pmx_ccl26] shape={shape} periodic={per}", flush=True)
    _lib.check(L.pmx_ccl26(_device.ptr(dimg), C.byref(g), _device.ptr(labels), None, _device.ptr(count),
                           _device.ptr(ws), nb.value, _device.stream()))
    torch.cuda.synchronize()
    want, n = O.ccl26(img, per)
    assert int(count.item()) == n
    np.testing.assert_array_equal(labels.cpu().numpy().astype(np.int64), want)


@pytest.mark.parametrize("inlet,outlet", [
    (((True, False), (True, False), (False, False)), ((False, True), (False, False), (False, True))),
    (((True, True), (False, False), (False, True)), ((False, False), (True, True), (False, False))),
])
def test_several_inlet_faces_follow_the_reference(inlet, outlet):
    """With several inlet (outlet) faces the reference keeps the labels of the
    last flagged face in its face order (gen_inlet_label_map overwrites per
    face); the oracle restates that and is pinned to the reference's own
    output in tests/test_oracle_golden.py."""
    from pmmoto_amd.domain_generation import gen_pm, pack_radii, sphere_pack
    from pmmoto_amd.filters import connected_components as CC
    from pmmoto_amd.filters import porosimetry as PO
    img_t, _ = sphere_pack((34, 30, 28), 0.45, pack_radii("uniform", lo=3, hi=6), seed=77)
    img, sd, box = padded_case(img_t.cpu().numpy(), BT_CASES["end"], inlet=inlet, outlet=outlet)
    print(f"[inlet faces] inlet={inlet} outlet={outlet}", flush=True)
    np.testing.assert_array_equal(CC.inlet_connected_img(sd, img), O.inlet_connected(box, img))
    np.testing.assert_array_equal(CC.outlet_connected_img(sd, img),
                                  O.inlet_connected(box, img, which="outlet"))
    pm = gen_pm(sd, img)
    radii = np.arange(6, 0, -1).astype(np.float64)
    got = PO.pore_size_distribution(sd, pm, radii, inlet=True, mode="distance")
    np.testing.assert_array_equal(got, O.pore_size_distribution(box, img, radii, inlet=True))


def test_stale_label_count_raises():
    """Label kernels meeting labels beyond the component count (a stale
    count) write nothing out of range and make the host raise."""
    import torch

    from pmmoto_amd import _lib
    from pmmoto_amd.filters import connected_components as CC
    own = rand_binary((20, 18, 16), 0.85, 5)  # sparse pores: many components
    img, sd, _ = padded_case(own, BT_CASES["end"], inlet=((True, False), (False, False), (False, False)))
    t = torch.from_numpy(img).cuda()
    labels, count = CC.label_device(t, sd)
    n = int(count.item())
    assert n > 20
    with pytest.raises(_lib.PmxError):
        CC._face_flags(sd, labels, 2, "inlet")
    fa = torch.ones(3, dtype=torch.uint8, device="cuda")
    with pytest.raises(_lib.PmxError):
        CC.keep_device(t, labels, fa, None, 0, None)
    # the matching count passes
    flags = CC._face_flags(sd, labels, n, "inlet")
    CC.keep_device(t, labels, flags, None, 0, None)


def test_images_beyond_uint8_are_rejected():
    """The kernels take uint8 phase / label values: wider or fractional
    images raise ValueError instead of wrapping (256 -> 0 would become
    solid); integral values in 0..255 of any dtype are accepted."""
    import torch

    from pmmoto_amd.filters import connected_components as CC
    from pmmoto_amd.filters import distance
    img = rand_binary((12, 10, 8), 0.2, 1).astype(np.int64)
    img2 = img.copy()
    img2[3, 3, 3] = 256
    with pytest.raises(ValueError, match="0..255"):
        distance.edt(img2)
    with pytest.raises(ValueError, match="0..255"):
        CC.connect_components(img2, padded_case(img2.astype(np.uint8), BT_CASES["end"])[1])
    with pytest.raises(ValueError, match="non-integer"):
        distance.edt(img.astype(np.float64) * 0.5)
    with pytest.raises(ValueError):
        distance.edt(torch.from_numpy(img2).cuda())
    with pytest.raises(ValueError):
        distance.edt(-img)
    np.testing.assert_array_equal(distance.edt(img), distance.edt(img.astype(np.uint8)))
    np.testing.assert_array_equal(distance.edt(img.astype(np.float32)), distance.edt(img.astype(np.uint8)))


@pytest.mark.parametrize("bts,inlet", [("periodic", False), ("mixed", True), ("mixed", False)])
def test_fused_sweep_periodic_axes(bts, inlet):
    """Periodic axes in the fused sweep (own voxels with wrap, code-map pads
    copied): equal to the per-radius path and to the oracle."""
    import torch

    from pmmoto_amd.domain_generation import pack_radii, sphere_pack
    from pmmoto_amd.filters import porosimetry as PO
    per = tuple(BT_CASES[bts][a][0] == P for a in range(3))
    img_t, _ = sphere_pack((40, 36, 44), 0.42, pack_radii("uniform", lo=2, hi=6), seed=61, periodic=per)
    inl = ((True, False), (False, False), (False, False)) if inlet else ((False, False),) * 3
    img, sd, box = padded_case(img_t.cpu().numpy(), BT_CASES[bts], inlet=inl)
    assert PO.fused_sweep_ok(sd, "distance")
    pm_t = torch.from_numpy(img).cuda()
    radii = [np.float64(r) for r in (6.0, 4.5, 3.0, 2.0, 1.0)]
    print(f"[fused sweep] {bts} inlet={inlet}", flush=True)
    fused = PO.psd_sweep_fused(sd, pm_t, radii, inlet=inlet, mode="distance").cpu().numpy()
    levels = PO.psd_sweep_levels(sd, pm_t, radii, inlet=inlet, mode="distance").cpu().numpy()
    np.testing.assert_array_equal(fused, levels)
    want = O.pore_size_distribution(box, img, np.array(radii), inlet=inlet, mode="distance")
    got = np.where(fused > 0, np.array([0.0] + [float(r) for r in radii])[fused], 0.0)
    assert np.count_nonzero(got) > 0
    np.testing.assert_array_equal(got, want)


def test_saturation_curve_kernel_matches_oracle_arithmetic():
    """pmx_saturation_curve: the code histogram of the own box and
    (wet / V) / (pore / V) in float64, bit for bit against the oracle's
    evaluation of the reference's arithmetic, on a padded subdomain and on a
    batched domain (no pads)."""
    import torch

    from pmmoto_amd.filters import porosimetry as PO
    from pmmoto_amd.filters.batch import batch_saturation
    rng = np.random.default_rng(8)
    own = rand_binary((30, 28, 26), 0.4, 8)
    img, sd, box = padded_case(own, BT_CASES["mixed"])
    codes = np.where(img == 1, rng.integers(0, 13, size=img.shape), rng.integers(0, 13, size=img.shape))
    codes = codes.astype(np.int32)
    got = PO.saturation_curve(sd, torch.from_numpy(img).cuda(), torch.from_numpy(codes).cuda(), 12)
    want = O.saturation_from_codes(box, img, codes.astype(np.uint8), 12, own.shape)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    bbox = O.Box(bc=((O.BC_END, O.BC_END),) * 3, pad=((0, 0),) * 3)
    c2 = codes[:20, :20, :20].copy()
    i2 = img[:20, :20, :20].copy()
    got = batch_saturation(torch.from_numpy(c2).cuda(), torch.from_numpy(i2).cuda(), 12)
    want = O.saturation_from_codes(bbox, i2, c2.astype(np.uint8), 12, i2.shape)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    none = batch_saturation(torch.zeros((4, 4, 4), dtype=torch.int32, device="cuda"),
                            torch.zeros((4, 4, 4), dtype=torch.uint8, device="cuda"), 3)
    assert np.all(np.isnan(none))  # 0 / 0, as the reference's numpy division gives
