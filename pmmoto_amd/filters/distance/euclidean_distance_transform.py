"""Euclidean distance transform on the GPU.

Same call surface as pmmoto.filters.distance
(src/pmmoto/filters/distance/euclidean_distance_transform.py:25-303):
``edt``, ``edt2d``, ``edt3d``.  numpy in -> numpy out (the reference's
contract); a torch tensor on the GPU in -> tensor out (device-resident use).

Two GPU paths.  pmx_edt is the exact integer EDT: at unit resolution the
reference's float32 EDT holds exactly these integers wherever a solid is
reachable and every axis is at most 1024 voxels, so non-periodic domains
match the reference bit for bit on the fast path.  pmx_edt_ref runs the
reference's float32 arithmetic line by line: it is the default at other
resolutions and on periodic axes, where the reference keeps at most five
boundary parabolas per line from across the face and so can exceed the
exact torus distance near it (``mode="exact"`` gives the latter).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ... import _device, _lib
from ...core.boundary_types import BoundaryType

__all__ = ["edt", "edt2d", "edt3d", "edt_squared_int", "compute_box"]

_OUT_DT = {"d2": torch.int32, "f32sq": torch.float32, "f32": torch.float32, "f64": torch.float64}


def _unit(resolution) -> bool:
    return all(float(r) == 1.0 for r in resolution)


def run_edt_ref(img_t: torch.Tensor, origin, n, first_axis, periodic, resolution, take_sqrt, out=None, pads=None):
    """pmx_edt_ref: the reference's float32 arithmetic (any resolution) on
    the box `origin`+`n` of a contiguous 3-D uint8 tensor; returns a
    float32 tensor of the full shape (box region written).  ``pads`` (six
    widths) selects corrected_edt's padded-subdomain handling of periodic
    axes, with the box the whole padded array."""
    shape = tuple(img_t.shape)
    off = (origin[0] * shape[1] + origin[1]) * shape[2] + origin[2]
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=img_t.device)
    nn = (C.c_int64 * 3)(*[int(v) for v in n])
    st = (C.c_int64 * 2)(shape[1] * shape[2], shape[2])
    per = (C.c_int32 * 3)(*[1 if p else 0 for p in periodic])
    res = (C.c_float * 3)(*[float(r) for r in resolution])
    pd = None if pads is None else (C.c_int64 * 6)(*[int(v) for v in pads])
    L = _lib.lib()
    nbytes = C.c_size_t()
    _lib.check(L.pmx_edt_ref_workspace_size(nn, C.byref(nbytes)))
    ws = _device.workspace(nbytes.value)
    _lib.check(L.pmx_edt_ref(_device.ptr(img_t, off), nn, st, _device.ptr(out, off), st, int(first_axis), per, pd,
                             res, 1 if take_sqrt else 0, _device.ptr(ws), nbytes.value, _device.stream()))
    return out


def compute_box(shape, subdomain=None, periodic=None):
    """Box the kernels compute on inside an array of `shape`, and its face
    codes.  Periodic axes of a padded subdomain compute on the own voxels
    (pads are periodic copies filled afterwards); other axes compute on the
    full extent with END faces, their pads being ordinary data (WALL pads
    hold the zeros set_wall_bcs writes), as edt3d on the padded image does.
    Returns (origin[3], n[3], bc[6], pads[6] for the periodic fill)."""
    origin, n, bc, pads = [0, 0, 0], list(shape), [], [0] * 6
    for a in range(3):
        per = False
        if subdomain is not None:
            per = subdomain.boundary_types[a][0] == BoundaryType.PERIODIC
        elif periodic is not None:
            per = bool(periodic[a])
        if per:
            lo, hi = (subdomain.pad[a] if subdomain is not None else (0, 0))
            origin[a], n[a] = lo, shape[a] - lo - hi
            pads[2 * a], pads[2 * a + 1] = lo, hi
            bc += [_lib.BC_PERIODIC, _lib.BC_PERIODIC]
        else:
            bc += [_lib.BC_END, _lib.BC_END]
    return origin, n, bc, pads


def _geom(shape, origin, n, bc):
    g = _lib.Geom()
    for a in range(3):
        g.n[a] = int(n[a])
    s1, s0 = int(shape[2]), int(shape[1]) * int(shape[2])
    g.in_s[0], g.in_s[1] = s0, s1
    g.out_s[0], g.out_s[1] = s0, s1
    for f in range(6):
        g.bc[f] = int(bc[f])
    return g


def run_edt(img_t: torch.Tensor, origin, n, bc, pads, outputs=("f32",)):
    """Run pmx_edt on a contiguous 3-D uint8 device tensor; returns the
    requested full-shape outputs with periodic pads filled."""
    shape = tuple(img_t.shape)
    g = _geom(shape, origin, n, bc)
    off = (origin[0] * shape[1] + origin[1]) * shape[2] + origin[2]
    outs = {k: torch.empty(shape, dtype=_OUT_DT[k], device=img_t.device) for k in outputs}
    L = _lib.lib()
    nbytes = C.c_size_t()
    _lib.check(L.pmx_edt_workspace_size(C.byref(g), C.byref(nbytes)))
    ws = _device.workspace(nbytes.value)
    p = lambda k: _device.ptr(outs[k], off) if k in outs else None  # noqa: E731
    _lib.check(L.pmx_edt(_device.ptr(img_t, off), C.byref(g), p("d2"), p("f32sq"), p("f32"), p("f64"),
                         _device.ptr(ws), nbytes.value, _device.stream()))
    fill_periodic_pads(list(outs.values()), shape, pads)
    return outs


def fill_periodic_pads(tensors, shape, pads):
    if not any(pads):
        return
    L = _lib.lib()
    sh = (C.c_int64 * 3)(*shape)
    pd = (C.c_int64 * 6)(*pads)
    per = (C.c_int32 * 3)(*[1 if (pads[2 * a] or pads[2 * a + 1]) else 0 for a in range(3)])
    for t in tensors:
        _lib.check(L.pmx_fill_periodic_pads(_device.ptr(t), t.element_size(), sh, pd, per,
                                            _device.stream()))


def _as3d(img_t):
    if img_t.dim() == 3:
        return img_t, False
    if img_t.dim() == 2:
        return img_t.unsqueeze(0).contiguous(), True
    raise ValueError(f"Wrong img dimension {tuple(img_t.shape)}")


def _finish(t, squeeze, as_numpy):
    if squeeze:
        t = t[0]
    return _device.to_host(t, as_numpy)


def _multilabel(t: torch.Tensor) -> bool:
    """True when the image holds a value above 1 (one device reduction)."""
    if t.data_ptr() % 16 or not t.is_contiguous():
        t = t.contiguous().clone()
    out = torch.empty(1, dtype=torch.int32, device=t.device)
    _lib.check(_lib.lib().pmx_max_u8(_device.ptr(t), t.numel(), _device.ptr(out), _device.stream()))
    return int(out.item()) > 1


def _use_reference(mode, resolution, periodic, img_t=None) -> bool:
    """mode "auto": the exact integer kernels where they equal the reference
    bit for bit (unit resolution, no periodic axis, a 0/1 image); the
    reference's float32 arithmetic otherwise -- at periodic faces the
    reference keeps only five boundary parabolas per line, which differs from
    the exact torus distance near the face on some images, and on a
    multi-label image its first pass restarts at every label change
    (_distance.hpp:155-165) while the later passes ignore labels.  "exact":
    the integer kernels (exact torus distance on periodic axes; every nonzero
    voxel is pore); "reference": always the reference arithmetic."""
    if mode == "reference":
        return True
    if mode == "exact":
        if not _unit(resolution):
            raise ValueError("mode='exact' needs unit resolution")
        return False
    if mode != "auto":
        raise ValueError(f"mode {mode!r} must be 'auto', 'exact' or 'reference'")
    if not _unit(resolution) or any(bool(p) for p in periodic):
        return True
    return img_t is not None and img_t.numel() > 0 and _multilabel(img_t)


def edt3d(img, periodic=(False, False, False), resolution=(1.0, 1.0, 1.0), squared=False, mode="auto"):
    """edt3d (euclidean_distance_transform.py:208-303): float32 distance (or
    squared distance with ``squared=True``) to the nearest zero voxel."""
    t, as_np = _device.as_device_u8(img)
    if t.dim() != 3:
        raise ValueError(f"edt3d needs a 3-D image, got {tuple(t.shape)}")
    if _use_reference(mode, resolution, periodic, t):
        out = run_edt_ref(t, (0, 0, 0), t.shape, 0, periodic, resolution, not squared)
        return _device.to_host(out, as_np)
    origin, n, bc, pads = compute_box(tuple(t.shape), periodic=periodic)
    key = "f32sq" if squared else "f32"
    return _finish(run_edt(t, origin, n, bc, pads, (key,))[key], False, as_np)


def edt2d(img, periodic=(False, False), resolution=(1.0, 1.0), mode="auto"):
    """edt2d (euclidean_distance_transform.py:104-205)."""
    t, as_np = _device.as_device_u8(img)
    if t.dim() != 2:
        raise ValueError(f"edt2d needs a 2-D image, got {tuple(t.shape)}")
    t3, _ = _as3d(t)
    if _use_reference(mode, resolution, periodic, t):
        # the reference sweeps axis 1 first, then axis 0: box (1, X, Y), axis 2 first
        out = run_edt_ref(t3, (0, 0, 0), t3.shape, 2, (False,) + tuple(periodic),
                          (1.0,) + tuple(resolution), True)
        return _finish(out, True, as_np)
    origin, n, bc, pads = compute_box(tuple(t3.shape), periodic=(False,) + tuple(periodic))
    return _finish(run_edt(t3, origin, n, bc, pads, ("f32",))["f32"], True, as_np)


def edt(img, subdomain=None, mode="auto"):
    """edt (euclidean_distance_transform.py:25-53): with a subdomain, the
    domain's resolution and periodic faces apply to the padded image and the
    periodic pads of the result hold periodic copies; without one, edt3d or
    edt2d by dimension (1-D raises ValueError)."""
    if subdomain is None:
        nd = img.dim() if isinstance(img, torch.Tensor) else np.ndim(img)
        if nd == 3:
            return edt3d(img, mode=mode)
        if nd == 2:
            return edt2d(img, mode=mode)
        raise ValueError(f"Wrong img dimension {np.shape(img)}")
    t, as_np = _device.as_device_u8(img)
    if t.dim() != 3:
        raise ValueError(f"Wrong img dimension {tuple(t.shape)}")
    origin, n, bc, pads = compute_box(tuple(t.shape), subdomain=subdomain)
    res = subdomain.domain.resolution
    per = [bc[2 * a] == _lib.BC_PERIODIC for a in range(3)]
    if _use_reference(mode, res, per, t):
        # non-periodic: edt3d of the padded image; periodic: corrected_edt,
        # whose passes run over the whole padded array with correctors and
        # boundary hulls measured past the pads (the pad voxels are computed,
        # not copied), as the reference does for one rank
        if any(per):
            sd_pads = [int(p) for a in range(3) for p in subdomain.pad[a]]
            out = run_edt_ref(t, (0, 0, 0), t.shape, 0, per, res, True, pads=sd_pads)
        else:
            out = run_edt_ref(t, (0, 0, 0), t.shape, 0, per, res, True)
        return _device.to_host(out, as_np)
    return _device.to_host(run_edt(t, origin, n, bc, pads, ("f32",))["f32"], as_np)


def edt_squared_int(img, subdomain=None, periodic=None, bc=None, outputs=("d2",)):
    """Device-resident exact EDT with every output libpmx offers: ``d2``
    (int32 squared distance, INT32_MAX where no solid is reachable),
    ``f32sq``, ``f32`` and ``f64``.  ``bc`` (six PMX_BC_* codes) overrides
    the faces of a plain array, e.g. WALL faces without a padded copy."""
    t, as_np = _device.as_device_u8(img)
    t3, squeeze = _as3d(t)
    if bc is not None:
        if squeeze:
            bc = [_lib.BC_END, _lib.BC_END] + list(bc)
        origin, n, pads = [0, 0, 0], list(t3.shape), [0] * 6
    else:
        per = periodic
        if squeeze and per is not None:
            per = (False,) + tuple(per)
        origin, n, bc, pads = compute_box(tuple(t3.shape), subdomain=subdomain, periodic=per)
    outs = run_edt(t3, origin, n, bc, pads, tuple(outputs))
    return {k: _finish(v, squeeze, as_np) for k, v in outs.items()}
