"""26-connected labelling and inlet/outlet connectivity on the GPU.

Same call surface as src/pmmoto/filters/connected_components.py:32-392.
Labels come from pmx_ccl26: components numbered 1..N in C-order of their
first voxel (cc3d's numbering is not part of the contract; the partition
and the counts are).  Periodic faces are joined inside the kernel instead
of by matching pad voxels across ranks, and the pads of periodic axes get
the labels of the voxels they copy.  The returned count follows the
reference's convention: N for a domain with no periodic face, N + 1 when
a face is periodic (the global relabelling of
core/voxels.py:344-397 counts one more there;
tests/filters/test_connect_components.py:77,187,247).
"""

from __future__ import annotations

import ctypes as C
from collections import defaultdict

import numpy as np
import torch

from .. import _device, _lib
from .distance.euclidean_distance_transform import _geom, compute_box, fill_periodic_pads

__all__ = [
    "connect_components",
    "gen_img_to_label_map",
    "phase_count",
    "inlet_outlet_labels",
    "inlet_outlet_connections",
    "inlet_connected_img",
    "outlet_connected_img",
    "inlet_outlet_connected_img",
    "isolated_img",
]


def label_device(img_t: torch.Tensor, subdomain, dtype=torch.int64):
    """Labels (padded shape, int32 or uint64 stored as int64) and N."""
    shape = tuple(img_t.shape)
    origin, n, bc, pads = compute_box(shape, subdomain=subdomain)
    g = _geom(shape, origin, n, bc)
    off = (origin[0] * shape[1] + origin[1]) * shape[2] + origin[2]
    out = torch.empty(shape, dtype=dtype, device=img_t.device)
    count = torch.zeros(1, dtype=torch.int64, device=img_t.device)
    L = _lib.lib()
    nbytes = C.c_size_t()
    _lib.check(L.pmx_ccl_workspace_size(C.byref(g), C.byref(nbytes)))
    ws = _device.workspace(nbytes.value)
    p32 = _device.ptr(out, off) if dtype == torch.int32 else None
    p64 = _device.ptr(out, off) if dtype == torch.int64 else None
    _lib.check(L.pmx_ccl26(_device.ptr(img_t, off), C.byref(g), p32, p64, _device.ptr(count),
                           _device.ptr(ws), nbytes.value, _device.stream()))
    fill_periodic_pads([out], shape, pads)
    return out, count


def _check_bad(bad: torch.Tensor, what: str) -> None:
    """Raise if a label kernel met labels beyond the flag array (a label
    image that does not match its count)."""
    n = int(bad.item())
    if n:
        raise _lib.PmxError(f"{what}: {n} voxels hold labels beyond the component count (stale labels?)")


# (axis, side) of the six faces in the reference's feature order: the unit
# directions of itertools.product([-1, 0, 1], repeat=3), i.e. x-, y-, z-, z+,
# y+, x+ (core/orientation.py:79-114)
FACE_ORDER = ((0, 0), (1, 0), (2, 0), (2, 1), (1, 1), (0, 1))


def selected_face(subdomain, which: str):
    """The face whose labels gen_inlet_label_map / gen_outlet_label_map
    return: each flagged face overwrites the previous one's labels
    (core/voxels.py:400-454, "currently restricted to a single face"), so the
    last flagged face in the reference's face order; None if no face is
    flagged."""
    faces = subdomain.inlet if which == "inlet" else subdomain.outlet
    last = None
    for a, s in FACE_ORDER:
        if faces[a][s]:
            last = (a, s)
    return last


def _face_flags(subdomain, labels: torch.Tensor, n_labels: int, which: str) -> torch.Tensor:
    """flags[l] = 1 for labels on the own slab of the selected inlet (or
    outlet) face (gen_inlet_label_map / gen_outlet_label_map,
    core/voxels.py:400-454)."""
    flags = torch.zeros(n_labels + 1, dtype=torch.uint8, device=labels.device)
    face = selected_face(subdomain, which)
    if face is None:
        return flags
    bad = torch.zeros(1, dtype=torch.int32, device=labels.device)
    shape = tuple(labels.shape)
    strides = (C.c_int64 * 2)(shape[1] * shape[2], shape[2])
    sl = subdomain.face_own_slab(*face)
    lo = (C.c_int64 * 3)(*[x.start for x in sl])
    ext = (C.c_int64 * 3)(*[x.stop - x.start for x in sl])
    _lib.check(_lib.lib().pmx_mark_labels(_device.ptr(labels), labels.element_size(), strides, lo, ext,
                                          _device.ptr(flags), flags.numel(), _device.ptr(bad), _device.stream()))
    _check_bad(bad, "face labels")
    return flags


def keep_device(img_t, labels, fa, fb, mode, phase):
    out = torch.empty_like(img_t)
    if fb is not None and fb.numel() != fa.numel():
        raise ValueError("flag arrays differ in size")
    bad = torch.zeros(1, dtype=torch.int32, device=img_t.device)
    _lib.check(_lib.lib().pmx_keep_labels(_device.ptr(img_t), _device.ptr(labels), labels.element_size(),
                                          img_t.numel(), _device.ptr(fa), _device.ptr(fb) if fb is not None else None,
                                          fa.numel(), mode, -1 if phase is None else int(phase), _device.ptr(out),
                                          _device.ptr(bad), _device.stream()))
    _check_bad(bad, "keep labels")
    return out


def _reported_count(subdomain, n: int) -> int:
    return n + 1 if subdomain.domain.periodic else n


def connect_components(img, subdomain):
    """(uint64 labels, count) -- connected_components.py:32-62."""
    t, as_np = _device.as_device_u8(img)
    labels, count = label_device(t, subdomain)
    n = int(count.item())
    if as_np:
        return labels.cpu().numpy().astype(np.uint64), _reported_count(subdomain, n)
    return labels, _reported_count(subdomain, n)


def gen_img_to_label_map(img, labeled_img):
    """label -> phase (connected_components.py:104-120, _voxels.pyx:453-475)."""
    img = np.asarray(img)
    lab = np.asarray(labeled_img)
    vals, idx = np.unique(lab.ravel(), return_index=True)
    return {np.uint64(v): np.uint64(img.ravel()[i]) for v, i in zip(vals, idx)}


def phase_count(phase_map):
    """connected_components.py:123-141."""
    out = {}
    for label in phase_map:
        out[phase_map[label]] = out.get(phase_map[label], 0) + 1
    return out


def inlet_outlet_labels(subdomain, labeled_img):
    """{label: {"inlet": bool, "outlet": bool}} (connected_components.py:144-177)."""
    lab = np.asarray(labeled_img.cpu() if isinstance(labeled_img, torch.Tensor) else labeled_img)
    connected = defaultdict(lambda: {"inlet": False, "outlet": False})
    for which in ("inlet", "outlet"):
        face = selected_face(subdomain, which)
        if face is not None:
            for l in np.unique(lab[subdomain.face_own_slab(*face)]):
                if l > 0:
                    connected[np.uint64(l)][which] = True
    return connected


def inlet_outlet_connections(subdomain, labeled_img):
    """connected_components.py:180-200."""
    return [k for k, v in inlet_outlet_labels(subdomain, labeled_img).items() if v["inlet"] and v["outlet"]]


def _select(subdomain, img, phase, mode_name):
    t, as_np = _device.as_device_u8(img)
    labels, count = label_device(t, subdomain)
    n = int(count.item())
    fin = _face_flags(subdomain, labels, n, "inlet")
    fout = _face_flags(subdomain, labels, n, "outlet")
    if mode_name == "inlet":
        out = keep_device(t, labels, fin, None, 0, phase)
    elif mode_name == "outlet":
        out = keep_device(t, labels, fout, None, 0, phase)
    elif mode_name == "both":
        out = keep_device(t, labels, fin, fout, 1, phase)
    else:
        out = keep_device(t, labels, fin, fout, 2, phase)
    return _device.to_host(out, as_np)


def inlet_connected_img(subdomain, img, phase=None):
    """Voxels whose component touches an inlet face keep their value
    (connected_components.py:254-296)."""
    return _select(subdomain, img, phase, "inlet")


def outlet_connected_img(subdomain, img, phase=None):
    """connected_components.py:299-341."""
    return _select(subdomain, img, phase, "outlet")


def inlet_outlet_connected_img(subdomain, img, phase=None):
    """connected_components.py:203-251."""
    return _select(subdomain, img, phase, "both")


def isolated_img(subdomain, img, phase=None):
    """connected_components.py:344-392."""
    return _select(subdomain, img, phase, "neither")
