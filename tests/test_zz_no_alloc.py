"""pmx_edt_noalloc (the exact EDT with its workspace in the library's static
device scratch) against pmx_edt with a caller workspace, through the C ABI:
bit-exact on every output, on two calls in a row with the second smaller
(the scratch keeps the first call's contents), and against the oracle; a box
beyond the scratch is rejected on the host with a clear error."""

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rand_binary

pytestmark = pytest.mark.gpu

BC = [O.BC_WALL, O.BC_WALL, O.BC_PERIODIC, O.BC_PERIODIC, O.BC_END, O.BC_END]


def run_both(img, bc):
    import torch

    from pmmoto_amd import _device, _lib
    from pmmoto_amd.filters.distance.euclidean_distance_transform import _geom
    L = _lib.lib()
    t = torch.from_numpy(img).cuda()
    shape = img.shape
    g = _geom(shape, (0, 0, 0), shape, bc)
    outs = {}
    for name in ("noalloc", "ws"):
        d2 = torch.full(shape, -5, dtype=torch.int32, device="cuda")
        f32 = torch.full(shape, -5.0, dtype=torch.float32, device="cuda")
        f64 = torch.full(shape, -5.0, dtype=torch.float64, device="cuda")
        print(f"[{name}] shape={shape} bc={bc}", flush=True)
        if name == "noalloc":
            _lib.check(L.pmx_edt_noalloc(_device.ptr(t), C.byref(g), _device.ptr(d2), None, _device.ptr(f32),
                                         _device.ptr(f64), _device.stream()))
        else:
            nb = C.c_size_t()
            _lib.check(L.pmx_edt_workspace_size(C.byref(g), C.byref(nb)))
            ws = _device.workspace(nb.value)
            _lib.check(L.pmx_edt(_device.ptr(t), C.byref(g), _device.ptr(d2), None, _device.ptr(f32),
                                 _device.ptr(f64), _device.ptr(ws), nb.value, _device.stream()))
        torch.cuda.synchronize()
        outs[name] = (d2.cpu().numpy(), f32.cpu().numpy(), f64.cpu().numpy())
    return outs


def test_noalloc_equals_pmx_edt_on_two_calls_second_smaller():
    cases = [((150, 140, 170), 0.02, 1), ((40, 33, 90), 0.05, 2)]
    for shape, p0, seed in cases:
        img = rand_binary(shape, p0, seed)
        outs = run_both(img, BC)
        for a, b in zip(outs["noalloc"], outs["ws"]):
            np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))
        np.testing.assert_array_equal(outs["noalloc"][0], O.exact_edt_sq(img, BC))


def test_noalloc_sparse_image_uses_the_fallback():
    """One solid voxel: every line defers to the fallback lanes, whose stacks
    live in the scratch too."""
    img = np.ones((64, 72, 80), np.uint8)
    img[30, 2, 77] = 0
    outs = run_both(img, BC)
    for a, b in zip(outs["noalloc"], outs["ws"]):
        np.testing.assert_array_equal(a.view(np.uint8), b.view(np.uint8))
    np.testing.assert_array_equal(outs["noalloc"][0], O.exact_edt_sq(img, BC))


def test_noalloc_rejects_a_box_beyond_the_scratch():
    import torch

    from pmmoto_amd import _device, _lib
    from pmmoto_amd.filters.distance.euclidean_distance_transform import _geom
    L = _lib.lib()
    small = torch.ones((4, 4, 4), dtype=torch.uint8, device="cuda")
    out = torch.empty((4, 4, 4), dtype=torch.int32, device="cuda")
    g = _geom((512, 512, 512), (0, 0, 0), (512, 512, 512), BC)  # checked on the host before any launch
    rc = L.pmx_edt_noalloc(_device.ptr(small), C.byref(g), _device.ptr(out), None, None, None, _device.stream())
    assert rc == _lib.ERR_INVALID
    assert b"pmx_edt_noalloc holds" in L.pmx_last_error()
