"""Morphological porosimetry on the GPU.

Same call surface as src/pmmoto/filters/porosimetry.py:33-227:
``get_sizes``, ``porosimetry``, ``pore_size_distribution``.  Every step of
one porosimetry call is a libpmx launch on the padded device image:

  erosion       pmx_morph (EDT threshold, invert = 0)
  inlet         pmx_ccl26 + pmx_mark_labels on inlet face slabs + pmx_keep_labels
  multiphase    the same on the multiphase image, then pmx_both_one
  dilation      pmx_morph (invert = 1)

``mode`` follows the reference: "morph" uses the FFT semantics for both
steps, "hybrid" for the dilation only, "distance" none (the FFT semantics
are an integer threshold of the exact EDT, see morphological_operators).  The sweep keeps integer
radius codes on the device and converts to the float64 radius map once.
"""

from __future__ import annotations

import ctypes as C
import logging

import numpy as np
import torch

from .. import _device, _lib
from . import connected_components as cc
from . import morphological_operators as morph
from .distance.euclidean_distance_transform import _geom, fill_periodic_pads

__all__ = ["get_sizes", "porosimetry", "pore_size_distribution", "psd_sweep", "saturation_curve"]


def get_sizes(min_value, max_value, num_values, spacing="linear"):
    """porosimetry.py:33-81."""
    if min_value >= max_value:
        raise ValueError(f"Error: min_value {min_value} must be greater than max value {max_value}")
    if num_values <= 0:
        raise ValueError(f"Error: num_values {num_values} must be greater than 0")
    if spacing == "linear":
        return np.linspace(min_value, max_value, num_values)[::-1]
    if spacing == "log":
        if min_value < 1:
            raise ValueError(f"Error: min_value {min_value} must be greater than or equal to 1")
        return np.logspace(np.log10(min_value), np.log10(max_value), num_values)[::-1]
    raise ValueError(f"spacing {spacing} can only be 'linear' or 'log'")


def _split_radius(radius):
    if isinstance(radius, (int, float)):
        return radius, radius
    if isinstance(radius, list) and len(radius) == 2:
        return radius[0], radius[1]
    raise ValueError(f"Radius {radius} must either be a number or a list of length 2")


def porosimetry_device(subdomain, pm_t: torch.Tensor, erosion_radius, dilation_radius, inlet=False,
                       mp_t: torch.Tensor | None = None, mode="hybrid") -> torch.Tensor:
    """One porosimetry step (porosimetry.py:117-160) on device tensors.
    mode: "morph" FFT erosion and dilation, "hybrid" EDT erosion and FFT
    dilation, "distance" EDT both (:120-158)."""
    res = morph.morph_any_device(pm_t, subdomain, erosion_radius, invert=False, fft=(mode == "morph"))
    if inlet:
        labels, count = cc.label_device(res, subdomain)
        n = int(count.item())
        fin = cc._face_flags(subdomain, labels, n, "inlet")
        res = cc.keep_device(res, labels, fin, None, 0, None)
    if mp_t is not None:
        # porosimetry.py:137-148: keep erosion centres inside inlet-connected phase 1
        labels, count = cc.label_device(mp_t, subdomain)
        n = int(count.item())
        fin = cc._face_flags(subdomain, labels, n, "inlet")
        n_connected = cc.keep_device(mp_t, labels, fin, None, 0, 1)
        both = torch.empty_like(res)
        _lib.check(_lib.lib().pmx_both_one(_device.ptr(res), _device.ptr(n_connected), res.numel(),
                                           _device.ptr(both), _device.stream()))
        res = both
    # phase_exists(img_results, 1) gates the dilation in the reference; the
    # dilation of an empty set is empty, so it runs unconditionally here.
    return morph.morph_any_device(res, subdomain, dilation_radius, invert=True, fft=(mode in ("morph", "hybrid")))


def porosimetry(subdomain, porous_media, radius, inlet=False, multiphase=None, mode="hybrid"):
    """porosimetry.py:84-160 (numpy in -> numpy out, tensor -> tensor)."""
    erosion_radius, dilation_radius = _split_radius(radius)
    morph.check_radii(subdomain, [erosion_radius, dilation_radius])
    if mode not in ("hybrid", "distance", "morph"):
        raise ValueError(f"mode {mode} not recognised")
    pm_t, as_np = _device.as_device_u8(porous_media.img)
    mp_t = None
    if multiphase is not None and hasattr(multiphase, "img"):
        mp_t, _ = _device.as_device_u8(multiphase.img)
        if int((mp_t == 1).sum().item()) == 0:  # utils.phase_exists(multiphase.img, 1)
            mp_t = None
    out = porosimetry_device(subdomain, pm_t, erosion_radius, dilation_radius, inlet, mp_t, mode)
    return _device.to_host(out, as_np)


def _sorted_radii(subdomain, porous_media, radii, num_radii):
    if radii is not None:
        if isinstance(radii, (int, float)):
            return [radii]
        if isinstance(radii, list):
            return sorted(radii, reverse=True)
        return np.sort(np.asarray(radii))[::-1]
    # porosimetry.py:193-201: the global maximum of the distance, less 1e-6
    # of the same maximum (porous_media.max_distance), in float64
    max_d = float(porous_media.max_distance)
    global_max = max_d - 1e-6 * max_d
    r_min = np.nextafter(np.min(subdomain.domain.resolution), np.inf)
    return get_sizes(r_min, global_max, num_radii, "linear")


def _thresholds(subdomain, radius, fft):
    if fft:
        return morph.fft_threshold(subdomain.domain.resolution, radius)
    return morph.radius_threshold(radius)


def fused_sweep_ok(subdomain, mode) -> bool:
    """pmx_psd_sweep covers domains whose erosion and dilation are integer
    thresholds (unit resolution, or FFT semantics throughout).  Periodic
    axes run on the own voxels with wrap: the reference's ceil(r) halo of
    periodic copies (morphological_operators.py:125-129) gives the same
    erosion and dilation, and its labelling wraps those axes."""
    from .distance.euclidean_distance_transform import _unit
    return _unit(subdomain.domain.resolution) or mode == "morph"


def _sweep_box(subdomain, pm_t: torch.Tensor):
    """(origin, extent, faces) of the box the fused sweep runs on.  The
    padded image with END faces gives the reference's codes; a WALL face's
    pad layer that holds no pore voxel is cut off and the face marked WALL
    (a solid layer just outside the box), which gives the same codes: the
    erosion sees the same solid, the layer holds no dilation centre and is
    never invaded (its codes stay 0).  The dilations and the internal arrays
    then run on the own extent (1024 rather than 1026 voxels per row for a
    1024^3 domain with y/z walls)."""
    from ..core import BoundaryType
    shape = list(pm_t.shape)
    origin, n, bc = [0, 0, 0], list(shape), [_lib.BC_END] * 6
    for a in range(3):
        if subdomain.boundary_types[a][0] == BoundaryType.PERIODIC:
            # the own voxels with wrap; the pads are periodic copies, filled afterwards
            lo, hi = subdomain.pad[a]
            origin[a], n[a] = lo, shape[a] - lo - hi
            bc[2 * a] = bc[2 * a + 1] = _lib.BC_PERIODIC
            continue
        for s in (0, 1):
            if subdomain.boundary_types[a][s] != BoundaryType.WALL or not subdomain.pad[a][s] or n[a] < 3:
                continue
            layer = origin[a] if s == 0 else origin[a] + n[a] - 1
            if bool(pm_t.select(a, layer).any()):
                continue
            if s == 0:
                origin[a] += 1
            n[a] -= 1
            bc[2 * a + s] = _lib.BC_WALL
    return origin, n, bc


def psd_sweep_fused(subdomain, pm_t: torch.Tensor, radii, inlet=False, mode="hybrid") -> torch.Tensor:
    """The whole sweep in one pmx_psd_sweep call (same codes as psd_sweep)."""
    shape = tuple(pm_t.shape)
    te = [_thresholds(subdomain, r, mode == "morph") for r in radii]
    td = [_thresholds(subdomain, r, mode in ("morph", "hybrid")) for r in radii]
    origin, n, bc = _sweep_box(subdomain, pm_t)
    g = _geom(shape, origin, n, bc)
    off = (origin[0] * shape[1] + origin[1]) * shape[2] + origin[2]
    code_map = torch.zeros(shape, dtype=torch.int32, device=pm_t.device)
    L = _lib.lib()
    nbytes = C.c_size_t()
    _lib.check(L.pmx_psd_workspace_size(C.byref(g), C.byref(nbytes)))
    ws = _device.workspace(nbytes.value)
    lo = ext = None
    if inlet:
        # the one face whose labels the reference's inlet map keeps
        face = cc.selected_face(subdomain, "inlet")
        if face is not None:
            sl = subdomain.face_own_slab(*face)
            # in box coordinates, clipped to the box (a cut pad layer holds no pore voxel)
            b0 = [min(max(sl[a].start - origin[a], 0), n[a]) for a in range(3)]
            b1 = [min(max(sl[a].stop - origin[a], 0), n[a]) for a in range(3)]
            lo = (C.c_int64 * 3)(*b0)
            ext = (C.c_int64 * 3)(*[b1[a] - b0[a] for a in range(3)])
        else:
            lo = (C.c_int64 * 3)(0, 0, 0)  # inlet requested but no inlet face: nothing connects
            ext = (C.c_int64 * 3)(0, 0, 0)
    _lib.check(L.pmx_psd_sweep(_device.ptr(pm_t, off), C.byref(g), lo, ext, len(radii), (C.c_int64 * len(te))(*te),
                               (C.c_int64 * len(td))(*td), morph._wall_mask(subdomain), _device.ptr(code_map, off),
                               _device.ptr(ws), nbytes.value, _device.stream()))
    pads = [0] * 6
    for a in range(3):
        if bc[2 * a] == _lib.BC_PERIODIC:
            pads[2 * a], pads[2 * a + 1] = subdomain.pad[a]
    fill_periodic_pads([code_map], shape, pads)
    return code_map


def psd_sweep(subdomain, pm_t: torch.Tensor, radii, inlet=False, mode="hybrid") -> torch.Tensor:
    """Radius-code map: code i+1 where radii[i] (descending) first invades
    a pore voxel, 0 where none does."""
    if fused_sweep_ok(subdomain, mode):
        return psd_sweep_fused(subdomain, pm_t, radii, inlet, mode)
    return psd_sweep_levels(subdomain, pm_t, radii, inlet, mode)


def psd_sweep_levels(subdomain, pm_t: torch.Tensor, radii, inlet=False, mode="hybrid") -> torch.Tensor:
    """The sweep as one porosimetry call per radius (any domain)."""
    code_map = torch.zeros(tuple(pm_t.shape), dtype=torch.int32, device=pm_t.device)
    for i, r in enumerate(radii):
        temp = porosimetry_device(subdomain, pm_t, r, r, inlet, None, mode)
        _lib.check(_lib.lib().pmx_psd_update(_device.ptr(pm_t), _device.ptr(temp), pm_t.numel(), i + 1,
                                             _device.ptr(code_map), _device.stream()))
    return code_map


def pore_size_distribution(subdomain, porous_media, radii=None, num_radii=25, inlet=False, mode="hybrid"):
    """float64 map of the largest radius whose porosimetry image covers each
    pore voxel (porosimetry.py:163-227)."""
    radii = _sorted_radii(subdomain, porous_media, radii, num_radii)
    morph.check_radii(subdomain, radii)
    pm_t, as_np = _device.as_device_u8(porous_media.img)
    for n, r in enumerate(radii):
        logging.debug("Processing radius %i of %i with value: %f", n + 1, len(radii), r)
    code_map = psd_sweep(subdomain, pm_t, radii, inlet, mode)
    values = torch.tensor([float(r) for r in radii], dtype=torch.float64, device=pm_t.device)
    out = torch.empty(tuple(pm_t.shape), dtype=torch.float64, device=pm_t.device)
    _lib.check(_lib.lib().pmx_codes_to_values(_device.ptr(code_map), code_map.numel(), _device.ptr(values),
                                              len(radii), _device.ptr(out), _device.stream()))
    return _device.to_host(out, as_np)


def saturation_curve_box(pm_t: torch.Tensor, code_map: torch.Tensor, lo, ext, n_radii: int, volume: float):
    """pmx_saturation_curve on the sub-box lo + [0, ext): the float64 wetting
    saturation after each radius, (wet / V) / (pore / V) on the device."""
    shape = tuple(pm_t.shape)
    counts = torch.empty(n_radii + 1, dtype=torch.int64, device=pm_t.device)
    curve = torch.empty(n_radii, dtype=torch.float64, device=pm_t.device)
    _lib.check(_lib.lib().pmx_saturation_curve(_device.ptr(pm_t), _device.ptr(code_map), (C.c_int64 * 3)(*shape),
                                               (C.c_int64 * 3)(*lo), (C.c_int64 * 3)(*ext), int(n_radii),
                                               float(volume), _device.ptr(counts), _device.ptr(curve),
                                               _device.stream()))
    return curve.cpu().numpy()


def saturation_curve(subdomain, pm_t: torch.Tensor, code_map: torch.Tensor, n_radii: int) -> np.ndarray:
    """Wetting saturation after each radius of a sweep:
    S_w(r_i) = (count(own pore & not invaded at radii[0..i]) / V) / (count(own pore) / V),
    in float64 in that order (multiphase.py:56-95, porousmedia.py:56-66),
    V = the domain's voxel count."""
    own = subdomain.get_own_voxels()
    lo = (int(own[0]), int(own[2]), int(own[4]))
    ext = (int(own[1] - own[0]), int(own[3] - own[2]), int(own[5] - own[4]))
    return saturation_curve_box(pm_t, code_map, lo, ext, n_radii, float(np.prod(subdomain.domain.voxels)))
