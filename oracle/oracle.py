"""CPU oracle for the voxel-morphology path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The product
(``pmmoto_amd``) never imports it.

Layers:

* ctypes wrappers over ``liboracle.so`` (``pmx_oracle.c``): the float32
  restatement of the reference per-line kernels and drivers, the exact
  integer EDT, the 26-connected labelling and the sphere rasteriser;
* optional access to ``_ref/libpmmoto_ref.so``, the reference's own
  ``_distance.hpp`` / ``_voxels.hpp`` compiled by ``oracle/Makefile``;
* numpy restatements of the Python-level reference drivers:
  morphology (``filters/morphological_operators.py:103-228``), inlet
  connectivity (``filters/connected_components.py:254-296``,
  ``core/voxels.py:400-425``) and porosimetry
  (``filters/porosimetry.py:84-227``).
"""

from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libpmmoto_ref.so")

BC_END, BC_WALL, BC_PERIODIC = 0, 1, 2
D2_INF = np.iinfo(np.int32).max
FLT_MAX = np.float32(np.finfo(np.float32).max)


class Hull(C.Structure):
    _fields_ = [("vertex", C.c_int), ("height", C.c_float), ("range", C.c_float)]


class LineOps(C.Structure):
    _fields_ = [
        ("sq_edt_1d", C.c_void_p),
        ("parabolic", C.c_void_p),
        ("boundary_hull", C.c_void_p),
        ("nearest_index", C.c_void_p),
    ]


_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_ref_sq_edt_1d.argtypes = [P, P, C.c_int, C.c_long, C.c_float, C.c_float, C.c_float]
        L.orc_ref_sq_edt_1d_parabolic.argtypes = [P, C.c_int, C.c_float, C.c_long, P, C.c_int, P, C.c_int]
        L.orc_ref_boundary_hull_1d.argtypes = [P, C.c_long, C.c_int, C.c_float, C.c_int, C.c_int, P]
        L.orc_ref_boundary_hull_1d.restype = C.c_int
        L.orc_ref_nearest_index_1d.argtypes = [P, C.c_int, C.c_uint8, C.c_int, C.c_int, C.c_int]
        L.orc_ref_nearest_index_1d.restype = C.c_long
        L.orc_ref_edt3d_sq.argtypes = [P, C.c_long, C.c_long, C.c_long, P, P, P, P]
        L.orc_ref_edt2d_sq.argtypes = [P, C.c_long, C.c_long, P, P, P, P]
        L.orc_ref_corrected_edt_sq.argtypes = [P, C.c_long, C.c_long, C.c_long, P, P, P, P, P]
        L.orc_ref_edt3d_sq_mt.argtypes = [P, C.c_long, C.c_long, C.c_long, P, P, P, C.c_int]
        L.orc_exact_edt_sq.argtypes = [P, C.c_int, P, P, P]
        L.orc_exact_edt_sq_mt.argtypes = [P, C.c_int, P, P, P, C.c_int]
        L.orc_exact_ccl26.argtypes = [P, C.c_long, C.c_long, C.c_long, P, P]
        L.orc_exact_ccl26.restype = C.c_int64
        L.orc_rasterize_spheres.argtypes = [P, C.c_long, C.c_long, C.c_long, P, P, C.c_long, C.c_double, P]
        L.orc_rasterize_spheres.restype = C.c_long
        _lib = L
    return _lib


def ref_lib():
    """The reference's own per-line kernels, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = C.CDLL(REF_PATH)
        P = C.c_void_p
        R.ref_sq_edt_1d.argtypes = [P, P, C.c_int, C.c_long, C.c_float, C.c_float, C.c_float]
        R.ref_parabolic.argtypes = [P, C.c_int, C.c_float, C.c_long, P, C.c_int, P, C.c_int]
        R.ref_boundary_hull.argtypes = [P, C.c_int, C.c_float, C.c_long, C.c_int, C.c_int, C.c_int, P]
        R.ref_boundary_hull.restype = C.c_int
        R.ref_nearest_index.argtypes = [P, C.c_int, C.c_long, C.c_uint8, C.c_int, C.c_int, C.c_int]
        R.ref_nearest_index.restype = C.c_long
        _ref = R
    return _ref


def ref_ops():
    """An orc_line_ops table pointing at the reference's kernels (or None)."""
    R = ref_lib()
    if R is None:
        return None
    ops = LineOps()
    for name, fn in (("sq_edt_1d", R.ref_sq_edt_1d), ("parabolic", R.ref_parabolic),
                     ("boundary_hull", R.ref_boundary_hull), ("nearest_index", R.ref_nearest_index)):
        setattr(ops, name, C.cast(fn, C.c_void_p).value)
    return ops


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------
# Per-line kernels (restated), used by the known-answer tests
# --------------------------------------------------------------------------

def sq_edt_1d(img, lower_corrector=np.inf, upper_corrector=np.inf, resolution=1.0, use_ref=False):
    """squared_edt_1d on a contiguous uint8 line (_distance.hpp:135-182)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    out = np.zeros(img.shape[0], dtype=np.float32)
    fn = ref_lib().ref_sq_edt_1d if use_ref else lib().orc_ref_sq_edt_1d
    fn(_p(img), _p(out), img.shape[0], 1, resolution, lower_corrector, upper_corrector)
    return out


def parabolic_1d(f, lower=(), upper=(), resolution=1.0, use_ref=False):
    """squared_edt_1d_parabolic on a float32 line (_distance.hpp:184-272)."""
    f = np.array(f, dtype=np.float32)
    lo = (Hull * max(1, len(lower)))(*[Hull(*h) for h in lower])
    hi = (Hull * max(1, len(upper)))(*[Hull(*h) for h in upper])
    fn = ref_lib().ref_parabolic if use_ref else lib().orc_ref_sq_edt_1d_parabolic
    fn(_p(f), f.shape[0], resolution, 1, C.cast(lo, C.c_void_p), len(lower), C.cast(hi, C.c_void_p), len(upper))
    return f


def boundary_hull_1d(img, start, end, resolution, num_hull, forward):
    """get_boundary_hull_1d (_distance.pyx:679-702) -> list of dicts.
    Like the reference, reads img[start:start+end]; the buffer is zero-padded
    so a window running past the array (as one reference test does) reads
    zeros instead of foreign memory."""
    img = np.concatenate([np.array(img, dtype=np.float32), np.zeros(start + end, np.float32)])
    out = (Hull * (num_hull + 2))()
    n = lib().orc_ref_boundary_hull_1d(_p(img), start, end, resolution, num_hull, 1 if forward else 0,
                                       C.cast(out, C.c_void_p))
    return [dict(vertex=out[i].vertex, height=out[i].height, range=out[i].range) for i in range(n)]


def ref_boundary_hull_line(f, resolution, num_hull, index_corrector, forward):
    """The reference's return_boundary_hull on a contiguous float32 line."""
    f = np.array(f, dtype=np.float32)
    out = (Hull * (num_hull + 2))()
    n = ref_lib().ref_boundary_hull(_p(f), f.shape[0], resolution, 1, num_hull, index_corrector,
                                    1 if forward else 0, C.cast(out, C.c_void_p))
    return [(out[i].vertex, out[i].height, out[i].range) for i in range(n)]


def nearest_index_1d(img, label, lower_skip=0, upper_skip=0, forward=True):
    img = np.ascontiguousarray(img, dtype=np.uint8)
    return lib().orc_ref_nearest_index_1d(_p(img), img.shape[0], label, lower_skip, upper_skip,
                                          1 if forward else 0)


# --------------------------------------------------------------------------
# Whole-volume reference EDT (float32, reference pass order)
# --------------------------------------------------------------------------

def ref_edt_sq(img, periodic=None, resolution=None, use_ref=False):
    """Squared float32 EDT as edt3d/edt2d compute it before the sqrt
    (euclidean_distance_transform.py:104-303)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    nd = img.ndim
    periodic = periodic or (False,) * nd
    resolution = resolution or (1.0,) * nd
    per = np.array([1 if p else 0 for p in periodic], dtype=np.int32)
    res = np.array(resolution, dtype=np.float32)
    out = np.empty(img.shape, dtype=np.float32)
    ops = ref_ops() if use_ref else None
    if use_ref and ops is None:
        raise RuntimeError("reference kernels not built (make -C oracle ref)")
    opp = C.byref(ops) if ops is not None else None
    if nd == 3:
        lib().orc_ref_edt3d_sq(_p(img), *img.shape, _p(per), _p(res), _p(out), opp)
    elif nd == 2:
        lib().orc_ref_edt2d_sq(_p(img), *img.shape, _p(per), _p(res), _p(out), opp)
    else:
        raise ValueError(f"Wrong img dimension {img.shape}")
    return out


def ref_corrected_edt_sq(img, periodic, pads, resolution, use_ref=False):
    """Squared float32 output of corrected_edt(img, subdomain) for one rank
    of a padded subdomain (euclidean_distance_transform.py:56-101, before
    its np.sqrt): periodic[a] per axis, pads = six pad widths (lo, hi per
    axis), over the whole padded array."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    per = np.array([1 if p else 0 for p in periodic], dtype=np.int32)
    pd = np.array([int(v) for v in pads], dtype=np.int64)
    res = np.array(resolution, dtype=np.float32)
    out = np.empty(img.shape, dtype=np.float32)
    ops = ref_ops() if use_ref else None
    if use_ref and ops is None:
        raise RuntimeError("reference kernels not built (make -C oracle ref)")
    lib().orc_ref_corrected_edt_sq(_p(img), *img.shape, _p(per), _p(pd), _p(res), _p(out),
                                   C.byref(ops) if ops is not None else None)
    return out


def ref_edt_sq_threads(img, threads, resolution=None, use_ref=False):
    """ref_edt_sq of a non-periodic 3-D image with each pass's lines split
    over `threads` host threads (orc_ref_edt3d_sq_mt; identical output)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    if img.ndim != 3:
        raise ValueError("3-D images only")
    res = np.array(resolution or (1.0, 1.0, 1.0), dtype=np.float32)
    out = np.empty(img.shape, dtype=np.float32)
    ops = ref_ops() if use_ref else None
    if use_ref and ops is None:
        raise RuntimeError("reference kernels not built (make -C oracle ref)")
    lib().orc_ref_edt3d_sq_mt(_p(img), *img.shape, _p(res), _p(out), C.byref(ops) if ops is not None else None,
                              int(threads))
    return out


def ref_edt(img, periodic=None, resolution=None, use_ref=False):
    """edt3d/edt2d output: np.sqrt of the float32 squared EDT."""
    return np.sqrt(ref_edt_sq(img, periodic, resolution, use_ref))


# --------------------------------------------------------------------------
# Exact integer EDT and labelling
# --------------------------------------------------------------------------

def exact_edt_sq(img, bc=None, threads=1):
    """Exact squared EDT (int32; D2_INF where no solid is reachable).
    bc: per-face boundary types, sequence of 2*ndim ints (END/WALL/PERIODIC).
    threads > 1 splits each pass's lines over host threads (same output)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    nd = img.ndim
    bc = np.array(bc if bc is not None else [BC_END] * (2 * nd), dtype=np.int32)
    shape = np.array(img.shape, dtype=np.int64)
    out = np.empty(img.shape, dtype=np.int32)
    lib().orc_exact_edt_sq_mt(_p(img), nd, _p(shape), _p(bc), _p(out), int(threads))
    return out


def d2_to_f32(d2):
    """Map exact integer squared distances to the reference's float32
    squared values (D2_INF -> FLT_MAX, the to_finite sentinel)."""
    out = d2.astype(np.float32)
    out[d2 == D2_INF] = FLT_MAX
    return out


def ccl26(img, periodic=(False, False, False)):
    """26-connected multi-label labelling, canonical numbering (C-order of
    first voxel).  Returns (int64 labels, count)."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    per = np.array([1 if p else 0 for p in periodic], dtype=np.int32)
    labels = np.empty(img.shape, dtype=np.int64)
    n = lib().orc_exact_ccl26(_p(img), *img.shape, _p(per), _p(labels))
    return labels, int(n)


def canonical_relabel(labels):
    """Renumber labels 1..N by C-order of first occurrence (0 stays 0)."""
    flat = labels.ravel()
    nz = flat != 0
    vals, first = np.unique(flat[nz], return_index=True)
    order = np.argsort(first)
    remap = np.zeros(vals.shape[0], dtype=np.int64)
    remap[order] = np.arange(1, vals.shape[0] + 1)
    out = np.zeros(flat.shape, dtype=np.int64)
    out[nz] = remap[np.searchsorted(vals, flat[nz])]
    return out.reshape(labels.shape)


def rasterize_spheres(shape, periodic, spheres, target):
    """Sequential overlapping-sphere rasteriser: returns (img, n_used)."""
    img = np.ones(shape, dtype=np.uint8)
    per = np.array([1 if p else 0 for p in periodic], dtype=np.int32)
    sp = np.ascontiguousarray(spheres, dtype=np.float64)
    pores = np.array([img.size], dtype=np.int64)
    n = lib().orc_rasterize_spheres(_p(img), *shape, _p(per), _p(sp), sp.shape[0], float(target), _p(pores))
    return img, int(n)


def spheres_on_grid(spheres, coords, invert=False):
    """gen_pm_shape (_domain_generation.pyx:30-70) with Sphere::contains
    (particles/spheres.hpp:42-48): 1 everywhere, 0 where a voxel centroid
    (coords[0][i], coords[1][j], coords[2][k]) satisfies
    (cx-x)^2 + (cy-y)^2 + (cz-z)^2 <= r*r, in float64 and in that order;
    ``invert`` swaps 0 and 1 (domain_generation.py:148-149)."""
    x, y, z = (np.asarray(c, dtype=np.float64) for c in coords)
    img = np.ones((x.size, y.size, z.size), dtype=np.uint8)
    for cx, cy, cz, r in np.asarray(spheres, dtype=np.float64):
        dx = (cx - x)[:, None, None]
        dy = (cy - y)[None, :, None]
        dz = (cz - z)[None, None, :]
        img[(dx * dx + dy * dy) + dz * dz <= r * r] = 0
    return (1 - img).astype(np.uint8) if invert else img


def periodic_pads(img, pads):
    """update_buffer (communication.py:101-118) for one rank: each periodic
    pad holds the own voxels at the opposite side, axis by axis."""
    img = img.copy()
    for ax in range(3):
        lo, hi = pads[ax]
        n = img.shape[ax]
        if lo:
            src = np.take(img, np.arange(n - hi - lo, n - hi), axis=ax)
            idx = [slice(None)] * 3
            idx[ax] = slice(0, lo)
            img[tuple(idx)] = src
        if hi:
            src = np.take(img, np.arange(lo, lo + hi), axis=ax)
            idx = [slice(None)] * 3
            idx[ax] = slice(n - hi, n)
            img[tuple(idx)] = src
    return img


# --------------------------------------------------------------------------
# Padded-subdomain model and the Python-level reference drivers
# --------------------------------------------------------------------------

@dataclass
class Box:
    """What the reference drivers read from a PaddedSubdomain: per-face
    boundary types, pad widths, inlet/outlet flags and resolution."""

    bc: tuple  # ((lo, hi),) * 3 of BC_*
    pad: tuple  # ((lo, hi),) * 3
    inlet: tuple = ((False, False),) * 3
    outlet: tuple = ((False, False),) * 3
    resolution: tuple = (1.0, 1.0, 1.0)
    reservoir: tuple = ((0, 0),) * 3


def _own_slices(box: Box, shape):
    return tuple(slice(p[0] + r[0], s - p[1] - r[1]) for p, r, s in zip(box.pad, box.reservoir, shape))


def halo_image(box: Box, img, halo):
    """update_extended_buffer (communication.py:121-157) for one rank:
    periodic faces receive the periodic extension of the own region, END
    faces and 1-wide WALL pads receive no halo (subdomain_padded.py:116-167)."""
    out = img
    pads = []
    for ax in range(3):
        lo_bc, hi_bc = box.bc[ax]
        plo, phi = box.pad[ax]
        elo = 0 if lo_bc in (BC_END,) or (lo_bc == BC_WALL and plo == 1) else halo[ax]
        ehi = 0 if hi_bc in (BC_END,) or (hi_bc == BC_WALL and phi == 1) else halo[ax]
        if elo or ehi:
            n = out.shape[ax]
            own = np.take(out, np.arange(plo, n - phi), axis=ax)
            ext = np.pad(own, [(0, 0)] * ax + [(plo + elo, phi + ehi)] + [(0, 0)] * (2 - ax), mode="wrap")
            # keep the caller's pad voxels, as buffer_unpack writes only the halo
            idx = [slice(None)] * 3
            idx[ax] = slice(elo, elo + n)
            ext[tuple(idx)] = out
            out = ext
        pads.append((elo, ehi))
    return out, tuple(pads)


def unpad(img, pads):
    return img[tuple(slice(p[0], img.shape[i] - p[1]) for i, p in enumerate(pads))]


def set_wall_bcs(box: Box, img):
    """subdomain.py:273-287: outermost layer of WALL faces forced to 0."""
    img = img.copy()
    for ax in range(3):
        for side in (0, 1):
            if box.bc[ax][side] == BC_WALL:
                idx = [slice(None)] * 3
                idx[ax] = 0 if side == 0 else -1
                img[tuple(idx)] = 0
    return img


def struct_ratio(resolution, radius):
    """gen_struct_ratio (morphological_operators.py:35-63)."""
    if len(resolution) not in (2, 3):
        raise ValueError("Resolution must be a list of length 2 or 3")
    if radius < np.min(resolution):
        raise ValueError("radius too small")
    return tuple(math.ceil(radius / r) for r in resolution)


def _halo_edt(halo, resolution, threads):
    """The non-periodic float32 EDT of a halo image (np.sqrt of edt3d's
    squared output); threads > 1 splits each pass's lines (same output)."""
    if threads > 1:
        return np.sqrt(ref_edt_sq_threads(halo, threads, resolution=resolution))
    return ref_edt(halo, resolution=resolution)


def erode(box: Box, img, radius, threads=1):
    """subtraction, EDT mode (morphological_operators.py:175-228)."""
    halo, pads = halo_image(box, img, struct_ratio(box.resolution, radius))
    d = _halo_edt(halo.astype(np.uint8), box.resolution, threads)
    out = np.where(d <= radius, 0, 1).astype(np.uint8)
    return set_wall_bcs(box, unpad(out, pads))


def dilate(box: Box, img, radius, threads=1):
    """addition, EDT mode (morphological_operators.py:103-149)."""
    halo, pads = halo_image(box, img, struct_ratio(box.resolution, radius))
    d = _halo_edt(np.logical_not(halo).astype(np.uint8), box.resolution, threads)
    out = np.where(d <= radius, 1, 0).astype(np.uint8)
    return set_wall_bcs(box, unpad(out, pads))


def label_padded(box: Box, img):
    """connect_components for one rank (connected_components.py:32-101):
    periodic axes label the own region with wrap and copy labels into the
    pads; other axes label the padded extent as data."""
    per = tuple(box.bc[a][0] == BC_PERIODIC for a in range(3))
    sl = tuple(slice(box.pad[a][0], img.shape[a] - box.pad[a][1]) if per[a] else slice(None) for a in range(3))
    core = np.ascontiguousarray(img[sl])
    lab, n = ccl26(core, per)
    pads = [(box.pad[a] if per[a] else (0, 0)) for a in range(3)]
    full = np.pad(lab, pads, mode="wrap") if any(per) else lab
    return full, n


# faces in the reference's feature order (core/orientation.py:79-114): the
# unit directions of itertools.product([-1, 0, 1], repeat=3)
FACE_ORDER = ((0, 0), (1, 0), (2, 0), (2, 1), (1, 1), (0, 1))


def face_slab(box: Box, shape, ax, side):
    """The voxels a face feature owns (core/features.py:76-139): on the
    face's axis 1 voxel thick on END faces (the outermost layer of a
    reservoir); on the other axes the array without its pads, reservoir
    layers included (subdomain_features.py:46-50 passes the padded voxel
    counts and the pads)."""
    idx = [slice(box.pad[a][0], shape[a] - box.pad[a][1]) for a in range(3)]
    p = box.pad[ax][side]
    n = shape[ax]
    if side == 0:
        idx[ax] = slice(p, 2 * p) if p else slice(0, 1)
    else:
        idx[ax] = slice(n - 2 * p, n - p) if p else slice(n - 1, n)
    return tuple(idx)


def inlet_labels(box: Box, labels, which="inlet"):
    """gen_inlet_label_map / gen_outlet_label_map (core/voxels.py:400-454):
    each flagged END face overwrites the labels of the faces before it in
    the reference's face order, so the last one's labels are returned."""
    faces = box.inlet if which == "inlet" else box.outlet
    found = set()
    for ax, side in FACE_ORDER:
        if faces[ax][side] and box.bc[ax][side] == BC_END:
            found = set(np.unique(labels[face_slab(box, labels.shape, ax, side)]).tolist())
    found.discard(0)
    return found


def inlet_connected(box: Box, img, phase=None, which="inlet"):
    """inlet_connected_img / outlet_connected_img
    (connected_components.py:254-341): voxels whose component touches the
    face keep their phase value, others become 0."""
    lab, _ = label_padded(box, img)
    keep = np.zeros(int(lab.max()) + 1, dtype=bool)
    for l in inlet_labels(box, lab, which):
        keep[l] = True
    out = np.where(keep[lab], img, 0).astype(np.uint8)
    if phase is not None:
        out = np.where(out == phase, out, 0).astype(np.uint8)
    return out


def struct_element(resolution, radius):
    """gen_struct_element (morphological_operators.py:66-100)."""
    ratio = struct_ratio(resolution, radius)
    grids = [np.linspace(-r, r, r * 2 + 1) for r in ratio]
    xg, yg, zg = np.meshgrid(*grids, indexing="ij")
    s = xg**2 + yg**2 + zg**2
    _radius = radius / resolution[0]
    return ratio, np.array(s <= _radius * _radius, dtype=np.uint8)


def erode_fft(box: Box, img, radius):
    """subtraction, FFT mode (morphological_operators.py:203-217)."""
    from scipy.signal import fftconvolve

    ratio, se = struct_element(box.resolution, radius)
    halo, pads = halo_image(box, img, ratio)
    _pad = max(ratio) if any(b == BC_END for pair in box.bc for b in pair) else 1
    _img = np.pad(halo, _pad, mode="constant", constant_values=1)
    _img = fftconvolve(_img, se, mode="same") > (se.sum() - 0.1)
    out = unpad(_img, ((_pad, _pad),) * 3).astype(np.uint8)
    return set_wall_bcs(box, unpad(out, pads))


def dilate_fft(box: Box, img, radius):
    """addition, FFT mode (morphological_operators.py:131-133)."""
    from scipy.signal import fftconvolve

    ratio, se = struct_element(box.resolution, radius)
    halo, pads = halo_image(box, img, ratio)
    out = (fftconvolve(halo, se, mode="same") > 0.1).astype(np.uint8)
    return set_wall_bcs(box, unpad(out, pads))


def porosimetry(box: Box, pm_img, radius, inlet=False, mp_img=None, mode="distance"):
    """porosimetry (porosimetry.py:84-160)."""
    er, dr = (radius, radius) if not isinstance(radius, list) else radius
    res = erode_fft(box, pm_img, er) if mode == "morph" else erode(box, pm_img, er)
    if inlet:
        res = inlet_connected(box, res)
    if np.count_nonzero(res == 1) > 0:  # utils.phase_exists (utils.py:127-145)
        if mp_img is not None and np.count_nonzero(mp_img == 1) > 0:
            n_connected = inlet_connected(box, mp_img, 1)
            res = np.where((res != 1) | (n_connected != 1), 0, res)
        res = dilate_fft(box, res, dr) if mode in ("morph", "hybrid") else dilate(box, res, dr)
    return res.astype(np.uint8)


def update_reservoir(box: Box, img, value):
    """subdomain_padded.py:285-307."""
    img = img.copy()
    for a, (lo, hi) in enumerate(box.reservoir):
        idx = [slice(None)] * 3
        if lo:
            idx[a] = slice(0, lo)
            img[tuple(idx)] = value
        if hi:
            idx[a] = slice(img.shape[a] - hi, None)
            img[tuple(idx)] = value
    return img


def drainage(box: Box, pm_img, mp_img, capillary_pressures, domain_voxels, max_distance, mode="hybrid"):
    """drainage, standard method (equilibrium_distribution.py:25-179);
    returns (wetting saturations, final phase image)."""
    own = _own_slices(box, pm_img.shape)
    vol = np.prod(domain_voxels)
    porosity = np.count_nonzero(pm_img[own] == 1) / vol
    out = np.zeros(len(capillary_pressures))
    for n, pc in enumerate(capillary_pressures):
        radius = 2.0 * 1 / pc * np.cos(np.deg2rad(0))
        if radius < min(box.resolution) or radius > max_distance:
            morph = np.zeros_like(pm_img, dtype=np.uint8)
        else:
            morph = porosimetry(box, pm_img, radius, inlet=True, mp_img=mp_img, mode=mode)
        w = inlet_connected(box, mp_img, 2, which="outlet")
        mp_img = update_reservoir(box, np.where((morph == 1) & (w == 2), 1, mp_img).astype(np.uint8), 1)
        out[n] = float(np.count_nonzero(mp_img[own] == 2) / vol) / porosity
    return out, mp_img


def pore_size_distribution(box: Box, pm_img, radii, inlet=False, mode="distance"):
    """pore_size_distribution (porosimetry.py:163-227); a list keeps its
    Python scalars (sorted descending), an array becomes numpy scalars."""
    radii = sorted(radii, reverse=True) if isinstance(radii, list) else np.sort(np.asarray(radii))[::-1]
    out = np.zeros(pm_img.shape, dtype=np.float64)
    for r in radii:
        t = porosimetry(box, pm_img, r, inlet, mode=mode)
        if np.any(t):
            out = np.where((pm_img == 1) & (out == 0) & (t == 1), r, out)
    return out


def psd_codes_exact(box: Box, pm_img, radii, inlet=False, threads=1):
    """pore_size_distribution (porosimetry.py:163-227) in mode "distance"
    at unit resolution on a box with no periodic face, as radius codes:
    code i+1 where radii[i] (descending) first covers a pore voxel, 0 where
    none does.  The same steps as ``pore_size_distribution``, with the
    float32 EDT replaced by the exact integer EDT it equals at unit
    resolution, the erosion EDT computed once (it does not depend on the
    radius), the threshold sqrt_f32(d2) <= r written d2 <= T(r) and the
    labelling done by scipy.ndimage.label with a 3x3x3 structure; for sizes
    where that chain is too slow.  tests/test_oracle_morph_ccl.py checks it
    against ``pore_size_distribution``."""
    from scipy import ndimage

    if any(b == BC_PERIODIC for pair in box.bc for b in pair):
        raise ValueError("psd_codes_exact covers boxes without periodic faces")
    if any(float(r) != 1.0 for r in box.resolution):
        raise ValueError("psd_codes_exact covers unit resolution")
    pm_img = np.ascontiguousarray(pm_img, dtype=np.uint8)
    own = _own_slices(box, pm_img.shape)
    d2_pm = exact_edt_sq(pm_img, [BC_END] * 6, threads=threads)
    codes = np.zeros(pm_img.shape, dtype=np.uint8)
    pore = pm_img == 1
    for i, r in enumerate(radii):
        t = radius_threshold(r)
        eroded = d2_pm > t
        if inlet:
            lab, _ = ndimage.label(eroded, structure=np.ones((3, 3, 3), dtype=bool))
            keep = np.zeros(int(lab.max()) + 1, dtype=bool)
            keep[list(inlet_labels(box, lab))] = True
            keep[0] = False
            eroded = keep[lab]
            del lab
        if not eroded.any():
            continue
        d2 = exact_edt_sq(np.logical_not(eroded).astype(np.uint8), [BC_END] * 6, threads=threads)
        cover = d2 <= t
        del d2
        cover = set_wall_bcs(box, cover.astype(np.uint8)).astype(bool)
        codes[pore & (codes == 0) & cover] = i + 1
    return codes


def radius_threshold(r):
    """Largest integer d2 with np.sqrt(np.float32(d2)) <= r under NumPy 2
    promotion (a Python float compares in float32, a numpy float64 in
    float64): the threshold the EDT-mode erosion and dilation apply
    (morphological_operators.py:147,226)."""
    lo, hi = 0, 1
    dt = np.float64 if isinstance(r, np.floating) and np.asarray(r).dtype == np.float64 else np.float32
    rr = dt(r)

    def ok(d):
        return bool(np.sqrt(np.float32(d)) <= rr) if dt is np.float32 else bool(np.float64(np.sqrt(np.float32(d))) <= rr)

    if not ok(0):
        return -1
    while ok(hi):
        hi *= 2
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if ok(mid):
            lo = mid
        else:
            hi = mid
    return lo


def saturation_from_codes(box: Box, pm_img, codes, n_radii, domain_voxels):
    """Wetting saturation after each radius of a sweep, in float64 as the
    reference forms it: S_w = (count(own pore not yet invaded) / V) /
    (count(own pore) / V) (multiphase.py:56-95, porousmedia.py:56-66)."""
    own = _own_slices(box, pm_img.shape)
    pm_own = pm_img[own] == 1
    c = np.bincount(codes[own][pm_own].ravel(), minlength=n_radii + 1)
    vol = float(np.prod(domain_voxels))
    pore = int(pm_own.sum())
    porosity = pore / vol
    out = np.zeros(n_radii, dtype=np.float64)
    invaded = 0
    for i in range(n_radii):
        invaded += int(c[i + 1])
        out[i] = ((pore - invaded) / vol) / porosity
    return out
