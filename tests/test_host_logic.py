"""CPU tests of the host-side logic of the drop-in (no GPU): argument
checks and known answers the reference's tests pin, and the integer
radius threshold against brute force."""

import numpy as np
import pytest

from pmmoto_amd.filters import morphological_operators as M
from pmmoto_amd.filters import porosimetry as PO


def test_gen_struct_ratio_kats():
    # tests/filters/test_morphological_operators.py:10-26
    assert M.gen_struct_ratio(resolution=[0.5, 0.5, 0.5], radius=1) == (2, 2, 2)
    with pytest.raises(ValueError):
        M.gen_struct_ratio(resolution=[0.5], radius=1)
    with pytest.raises(ValueError):
        M.gen_struct_ratio(resolution=[0.5, 0.5, 0.5], radius=0.1)


def test_gen_struct_element_kats():
    # :29-50
    _, se = M.gen_struct_element(resolution=[0.01, 0.01, 0.01], radius=0.03)
    assert np.sum(se) == 123
    for res, r in (([0.5], 1), ([0.5, 0.5, 0.5], 0.1), ([0.5, 0.25, 0.5], 0.1)):
        with pytest.raises(ValueError):
            M.gen_struct_element(resolution=res, radius=r)


def test_get_sizes_kats():
    # tests/filters/test_porosimetry.py:8-41
    np.testing.assert_array_almost_equal(PO.get_sizes(0, 10, 4, "linear"), [10.0, 6.666667, 3.333333, 0.0])
    np.testing.assert_array_almost_equal(PO.get_sizes(1, 10, 4, "log"), [10.0, 4.64158883, 2.15443469, 1.0])
    for args in ((5, 1, 5), (0, 10, 0), (0, 10, 5, "log"), (0, 10, 5, "undefined")):
        with pytest.raises(ValueError):
            PO.get_sizes(*args)


@pytest.mark.parametrize("radius", [0, 1, 2, 2.5, 3.0000001, 5, 7.999999, 31.5, np.float64(4.0000001),
                                    np.float32(4.0000001), np.float64(2.9999999999), np.int64(6)])
def test_radius_threshold_vs_brute_force(radius):
    t = M.radius_threshold(radius)
    d = np.arange(0, 4000, dtype=np.int64)
    dist = np.sqrt(d.astype(np.float32))
    # numpy's own comparison with the reference's operand types
    ok = dist <= radius
    want = int(d[ok].max()) if ok.any() else -1
    assert t == want


def test_radius_threshold_negative():
    assert M.radius_threshold(-1.0) == -1


def test_sweep_box_cuts_only_all_solid_wall_pads():
    """The fused sweep's box: a WALL face's pad layer is cut off (and the face
    marked WALL) only when it holds no pore voxel; END faces and pads with
    pore voxels stay in the box as END faces; periodic axes keep their own
    voxels with wrap."""
    import torch

    import pmmoto_amd
    from pmmoto_amd import _lib
    BT = pmmoto_amd.BoundaryType
    sd = pmmoto_amd.initialize((10, 12, 14), box=((0, 10), (0, 12), (0, 14)),
                               boundary_types=((BT.END, BT.END), (BT.WALL, BT.WALL), (BT.WALL, BT.WALL)))
    pm = torch.zeros(sd.voxels, dtype=torch.uint8)
    pm[:, 1:-1, 1:-1] = 1
    origin, n, bc = PO._sweep_box(sd, pm)
    shape = tuple(pm.shape)
    assert origin == [0, 1, 1] and n == [shape[0], shape[1] - 2, shape[2] - 2]
    assert bc == [_lib.BC_END, _lib.BC_END] + [_lib.BC_WALL] * 4
    pm[3, 0, 5] = 1  # a pore voxel in the y = 0 pad keeps that layer
    origin, n, bc = PO._sweep_box(sd, pm)
    assert origin == [0, 0, 1] and n == [shape[0], shape[1] - 1, shape[2] - 2]
    assert bc == [_lib.BC_END, _lib.BC_END, _lib.BC_END, _lib.BC_WALL, _lib.BC_WALL, _lib.BC_WALL]
    per = pmmoto_amd.initialize((10, 12, 14), box=((0, 10), (0, 12), (0, 14)),
                                boundary_types=((BT.PERIODIC, BT.PERIODIC),) * 3)
    origin, n, bc = PO._sweep_box(per, torch.zeros(per.voxels, dtype=torch.uint8))
    # periodic axes: the own voxels with wrap (pads are copies, filled afterwards)
    assert origin == [1, 1, 1] and n == [10, 12, 14] and bc == [_lib.BC_PERIODIC] * 6


def test_probe_radius_and_pc_known_answers():
    """tests/domain_generation/test_multiphase.py:60-120."""
    from pmmoto_amd.domain_generation.multiphase import Multiphase as M
    assert M.get_probe_radius(1, 1) == 2.0
    assert M.get_probe_radius(2, 1) == 1.0
    assert M.get_probe_radius(3, 0.5) == 1 / 3
    assert M.get_probe_radius(0, 0.5) == 0
    np.testing.assert_approx_equal(M.get_probe_radius(1, 1, 20), 1.87938524)
    np.testing.assert_approx_equal(M.get_probe_radius(2, 1, 20), 0.93969262)
    np.testing.assert_approx_equal(M.get_probe_radius(3, 0.5, 85), 0.029051914)
    assert M.get_pc(M.get_probe_radius(1, 1), 1) == 1


def test_plain_subdomain_has_no_pads_or_reservoir():
    import pmmoto_amd
    sd = pmmoto_amd.initialize((5, 6, 7), reservoir_voxels=5, return_subdomain=True,
                               inlet=((True, False), (False, False), (False, False)))
    assert not sd.padded and sd.voxels == (5, 6, 7)
    assert sd.reservoir_pad == ((0, 0),) * 3


def test_tile_choice_cache_round_trip(tmp_path):
    """pmx_tune_cache_load / pmx_tune_cache_save (host only): a saved file
    loads back to the same choices, a missing file loads nothing, and a
    malformed line is an error."""
    import os

    from pmmoto_amd import _lib
    L = _lib.lib()
    src = tmp_path / "in.txt"
    src.write_text("# choices\n40 33 50 1 1 1 1 2 2 0 1 32\n7 8 9 0 0 2 2 1 1 1 0 64\n")
    _lib.check(L.pmx_tune_cache_load(os.fsencode(src)))
    out = tmp_path / "out.txt"
    _lib.check(L.pmx_tune_cache_save(os.fsencode(out)))
    lines = set(out.read_text().splitlines()[1:])
    assert {"40 33 50 1 1 1 1 2 2 0 1 32", "7 8 9 0 0 2 2 1 1 1 0 64"} <= lines
    _lib.check(L.pmx_tune_cache_load(os.fsencode(tmp_path / "missing.txt")))
    bad = tmp_path / "bad.txt"
    bad.write_text("40 33 50 1 1 1 1 2 2 0 1 48\n")
    with pytest.raises(ValueError, match="not a tile choice"):
        _lib.check(L.pmx_tune_cache_load(os.fsencode(bad)))
