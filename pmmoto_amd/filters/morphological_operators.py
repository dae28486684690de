"""Morphological dilation / erosion by EDT threshold on the GPU.

Same call surface as src/pmmoto/filters/morphological_operators.py:35-329.
The reference extends the padded image by a halo of ceil(r/res) voxels
(periodic copies, none beyond END or 1-wide WALL faces), takes the
non-periodic EDT of it, thresholds ``sqrt(d2) <= r``, drops the halo and
zeroes WALL faces.  On one GPU the halo is unnecessary: periodic faces wrap
inside the kernel (the nearest solid within r of a voxel is the same
either way), so one fused pmx_morph launch computes the EDT and writes the
thresholded uint8 image with WALL faces zeroed.

The threshold is an integer: the comparison ``np.sqrt(float32 d2) <= r``
is monotone in d2, so it holds exactly for d2 <= T(r) with
T(r) = max{d : cmp(sqrt_f32(d), r)} -- computed on the host with the
reference's dtype rules (NumPy 2 / NEP 50: a Python int/float radius is
compared in float32, a numpy float64 radius in float64).
``fft=True`` (scipy.signal.fftconvolve with the ball element, :131-133,
:203-217) is an integer threshold too: the ball holds the offsets with
|o|^2 <= (r/res)^2 in float64, so the FFT result is the exact EDT
thresholded at floor((r/res)^2) -- at any isotropic resolution.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from .. import _device, _lib
from ..core.boundary_types import BoundaryType
from .distance.euclidean_distance_transform import _geom, _unit, compute_box, fill_periodic_pads

__all__ = [
    "gen_struct_ratio",
    "gen_struct_element",
    "addition",
    "dilate",
    "subtraction",
    "erode",
    "opening",
    "closing",
    "check_radii",
    "radius_threshold",
]

_D2_MAX = 3 * 16384 * 16384


def gen_struct_ratio(resolution, radius):
    """morphological_operators.py:35-63."""
    if len(resolution) not in {2, 3}:
        raise ValueError("Resolution must be a list of length 2 or 3")
    if radius < np.min(resolution):
        raise ValueError(f"The chosen radius {radius} is too small! The resolutions is {resolution}")
    return tuple(int(math.ceil(radius / res)) for res in resolution)


def gen_struct_element(resolution, radius):
    """morphological_operators.py:66-100 (ball of radius r/res[0])."""
    if len(resolution) not in {2, 3}:
        raise ValueError("Resolution must be a list or tuple of length 2 or 3.")
    if resolution[0] != resolution[1] != resolution[2]:
        raise ValueError("Resolution must be isotropic for morphological methods.")
    ratio = gen_struct_ratio(resolution, radius)
    grids = [np.linspace(-r, r, r * 2 + 1) for r in ratio]
    xg, yg, zg = np.meshgrid(*grids, indexing="ij")
    s = xg**2 + yg**2 + zg**2
    _radius = radius / resolution[0]
    return ratio, np.array(s <= _radius * _radius, dtype=np.uint8)


def radius_threshold(radius) -> int:
    """Largest integer d2 with ``np.sqrt(np.float32(d2)) <= radius`` under
    NumPy's promotion of the radius' type (-1 if none)."""
    if isinstance(radius, (np.floating, np.integer)) and not isinstance(radius, (np.float32, np.float16)):
        r = np.float64(radius)
        ok = lambda d: np.float64(np.sqrt(np.float32(d))) <= r  # noqa: E731
    elif isinstance(radius, (np.float32, np.float16)):
        r = np.float32(radius)
        ok = lambda d: np.sqrt(np.float32(d)) <= r  # noqa: E731
    else:
        r = np.float32(float(radius))
        ok = lambda d: np.sqrt(np.float32(d)) <= r  # noqa: E731
    if not ok(0):
        return -1
    lo, hi = 0, _D2_MAX
    if ok(hi):
        return hi
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if ok(mid):
            lo = mid
        else:
            hi = mid
    return lo


def fft_threshold(resolution, radius) -> int:
    """Integer form of the FFT modes (morphological_operators.py:131-133,
    :203-217): fftconvolve with the ball element {o : |o|^2 <= (r/res)^2}
    (float64, gen_struct_element :96-98) marks a voxel when some (dilation)
    or every (erosion) voxel at such an offset qualifies, i.e. when the
    exact squared voxel distance to the nearest such voxel is <= floor of
    (r/res)^2."""
    _radius = radius / resolution[0]
    x = _radius * _radius
    return int(math.floor(x)) if x >= 0 else -1


def check_radii(subdomain, radii) -> None:
    """morphological_operators.py:300-329: the halo may not exceed the own
    voxels (the reference aborts all ranks; here ValueError)."""
    for radius in radii:
        ratio, _ = gen_struct_element(subdomain.domain.resolution, radius)
        if np.any(subdomain.own_voxels < np.array(ratio)):
            raise ValueError(f"The radius ({radius:.2f}) exceeds at least one dimension of the subdomain "
                             f"({tuple(subdomain.own_voxels)}).")


def _wall_mask(subdomain) -> int:
    m = 0
    for a in range(3):
        for s in (0, 1):
            if subdomain.boundary_types[a][s] == BoundaryType.WALL:
                m |= 1 << (2 * a + s)
    return m


def morph_device(img_t: torch.Tensor, subdomain, t: int, invert: bool) -> torch.Tensor:
    """One fused pmx_morph launch on a padded uint8 device tensor with the
    integer squared-distance threshold t."""
    shape = tuple(img_t.shape)
    origin, n, bc, pads = compute_box(shape, subdomain=subdomain)
    g = _geom(shape, origin, n, bc)
    off = (origin[0] * shape[1] + origin[1]) * shape[2] + origin[2]
    out = torch.empty(shape, dtype=torch.uint8, device=img_t.device)
    L = _lib.lib()
    nbytes = C.c_size_t()
    _lib.check(L.pmx_edt_workspace_size(C.byref(g), C.byref(nbytes)))
    ws = _device.workspace(nbytes.value)
    v_le, v_gt = (1, 0) if invert else (0, 1)
    _lib.check(L.pmx_morph(_device.ptr(img_t, off), C.byref(g), 1 if invert else 0, t, v_le, v_gt,
                           _wall_mask(subdomain), _device.ptr(out, off), _device.ptr(ws), nbytes.value,
                           _device.stream()))
    fill_periodic_pads([out], shape, pads)
    return out


def _compare_f64(radius) -> bool:
    """NEP 50: float32 array <= radius promotes to float64 for numpy
    float64 / integer scalars, stays float32 for Python numbers."""
    return isinstance(radius, (np.floating, np.integer)) and not isinstance(radius, (np.float32, np.float16))


def halo_morph_device(img_t: torch.Tensor, subdomain, radius, invert: bool) -> torch.Tensor:
    """The reference's halo-image morphology in its float32 arithmetic
    (pmx_halo_morph), for resolutions other than 1."""
    shape = tuple(img_t.shape)
    ratio = gen_struct_ratio(subdomain.domain.resolution, radius)
    halo = []
    for a in range(3):
        lo_bt, hi_bt = subdomain.boundary_types[a]
        # subdomain_padded.py:116-167: no halo beyond END or 1-wide WALL faces
        periodic = lo_bt == BoundaryType.PERIODIC
        halo.append(int(ratio[a]) if periodic else 0)
    pads = [int(p) for pair in subdomain.pad for p in pair]
    out = torch.empty(shape, dtype=torch.uint8, device=img_t.device)
    L = _lib.lib()
    sh = (C.c_int64 * 3)(*shape)
    hw = (C.c_int64 * 3)(*halo)
    nbytes = C.c_size_t()
    _lib.check(L.pmx_halo_morph_workspace_size(sh, hw, C.byref(nbytes)))
    ws = _device.workspace(nbytes.value)
    v_le, v_gt = (1, 0) if invert else (0, 1)
    res = (C.c_float * 3)(*[float(r) for r in subdomain.domain.resolution])
    _lib.check(L.pmx_halo_morph(_device.ptr(img_t), sh, (C.c_int64 * 6)(*pads), hw, res, 1 if invert else 0,
                                float(radius), 1 if _compare_f64(radius) else 0, v_le, v_gt, _wall_mask(subdomain),
                                _device.ptr(out), _device.ptr(ws), nbytes.value, _device.stream()))
    return out


def morph_any_device(img_t, subdomain, radius, invert, fft=False):
    """Erosion (invert False) or dilation (invert True) with the reference's
    semantics for the mode: FFT mode and unit-resolution EDT mode are
    integer thresholds of the exact EDT (pmx_morph); EDT mode at other
    resolutions runs the reference's float32 halo path (pmx_halo_morph)."""
    if fft:
        return morph_device(img_t, subdomain, fft_threshold(subdomain.domain.resolution, radius), invert)
    if _unit(subdomain.domain.resolution):
        return morph_device(img_t, subdomain, radius_threshold(radius), invert)
    return halo_morph_device(img_t, subdomain, radius, invert)


def _run(subdomain, img, radius, invert, fft):
    gen_struct_element(subdomain.domain.resolution, radius)  # the reference's argument checks
    t, as_np = _device.as_device_u8(img)
    if t.dim() != 3:
        raise ValueError(f"Wrong img dimension {tuple(t.shape)}")
    return _device.to_host(morph_any_device(t, subdomain, radius, invert, fft), as_np)


def addition(subdomain, img, radius, fft=False):
    """Dilation: voxels within ``radius`` of a non-zero voxel become 1
    (morphological_operators.py:103-149)."""
    return _run(subdomain, img, radius, invert=True, fft=fft)


def dilate(subdomain, img, radius, fft=False):
    return addition(subdomain, img, radius, fft)


def subtraction(subdomain, img, radius, fft=False):
    """Erosion: voxels farther than ``radius`` from every zero voxel stay 1
    (morphological_operators.py:175-228)."""
    return _run(subdomain, img, radius, invert=False, fft=fft)


def erode(subdomain, img, radius, fft=False):
    return subtraction(subdomain, img, radius, fft)


def opening(subdomain, img, radius, fft=False):
    """morphological_operators.py:254-274."""
    return addition(subdomain, subtraction(subdomain, img, radius, fft), radius, fft)


def closing(subdomain, img, radius, fft=False):
    """morphological_operators.py:277-297."""
    return subtraction(subdomain, addition(subdomain, img, radius, fft), radius, fft)
