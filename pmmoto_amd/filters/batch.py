"""Pore-size-distribution sweeps of many unevenly sized domains at once.

The reference runs one domain per (MPI) job; here a batch of domains runs
as one captured graph.  Without inlet requirements it is the fused batch
(``pmx_psd_batch_create`` / ``pmx_psd_batch_run``): every pass of every
level is one launch over all domains, and each domain's full exact EDT is
computed on the way (``edts``).  Domains with an inlet requirement run the
per-domain sweeps spread over a few HIP streams (``pmx_psd_sweep_batch``).
The EDT tile size of each domain shape is chosen by timing
(``pmx_edt_autotune``), and the choices can be kept in a file across
processes (``tune_cache``, or the ``PMX_TUNE_CACHE`` environment variable).
Each domain is its own region (no pads) with per-axis PERIODIC, WALL or END
faces and gives exactly the codes of the single-domain sweep
(``pmx_psd_sweep``) and the distances of ``pmx_edt`` on the same input.

    sweep = BatchSweep(domains, radii=range(24, 0, -1))
    maps = sweep.run()          # first call tunes and captures, later calls replay
    curves = [saturation(m, d) for ...]
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np
import torch

from .. import _device, _lib
from . import morphological_operators as morph

__all__ = ["BatchDomain", "BatchSweep", "batch_saturation"]


_BC = {"end": _lib.BC_END, "wall": _lib.BC_WALL, "periodic": _lib.BC_PERIODIC}


@dataclass
class BatchDomain:
    """img: CUDA uint8 (X, Y, Z), 1 = pore (own voxels, no pads); bc: per
    axis "periodic", "wall" or "end" (both faces); inlet: (axis, side) of
    an END inlet face (its own slab seeds the inlet connectivity), or None
    for no inlet requirement."""

    img: torch.Tensor
    bc: tuple = ("wall", "wall", "wall")
    inlet: tuple | None = None


def _geom(shape, bc):
    g = _lib.Geom()
    g.n[:] = shape
    g.in_s[:] = (shape[1] * shape[2], shape[2])
    g.out_s[:] = (shape[1] * shape[2], shape[2])
    codes = []
    for a in range(3):
        if bc[a] not in _BC:
            raise ValueError(f"boundary {bc[a]!r}: use 'periodic', 'wall' or 'end'")
        codes += [_BC[bc[a]]] * 2
    g.bc[:] = codes
    return g


class BatchSweep:
    """radii: descending radii in voxels (unit resolution, "distance" mode:
    erosion d <= r by float32 sqrt, as porosimetry's EDT mode).  graph:
    capture the batch once and replay it.  fused: one launch per pass over
    all domains (only without inlet requirements; otherwise, or with
    fused=False, the per-domain sweeps spread over `lanes` HIP streams).
    edt: keep each domain's full squared EDT (int32, ``edts``; fused batch
    only).  tune_cache: a file of tile choices read before and written after
    autotuning (default: the PMX_TUNE_CACHE environment variable, if set)."""

    def __init__(self, domains, radii, lanes: int = 4, autotune: bool = True, graph: bool = True,
                 fused: bool = True, edt: bool = False, tune_cache: str | None = None):
        _device.require_gpu()
        radii = [float(r) for r in radii]
        if any(radii[i] < radii[i + 1] for i in range(len(radii) - 1)):
            raise ValueError("radii must be in descending order")
        self.domains = list(domains)
        self.radii = radii
        t = [morph.radius_threshold(np.float64(r)) for r in radii]
        self.te = (C.c_int64 * len(t))(*t)
        self.td = (C.c_int64 * len(t))(*t)
        self.maps = []
        arr = (_lib.PsdDomain * len(self.domains))()
        for i, d in enumerate(self.domains):
            if d.img.dtype != torch.uint8 or not d.img.is_cuda or d.img.dim() != 3:
                raise ValueError(f"domain {i}: img must be a CUDA uint8 (X, Y, Z) tensor")
            img = d.img.contiguous()
            self.domains[i] = BatchDomain(img, tuple(d.bc), d.inlet)
            shape = tuple(img.shape)
            m = torch.empty(shape, dtype=torch.int32, device=img.device)
            self.maps.append(m)
            e = arr[i]
            e.img = _device.ptr(img)
            e.g = _geom(shape, self.domains[i].bc)
            if d.inlet is not None:
                a, s = d.inlet
                if d.bc[a] != "end":
                    raise ValueError(f"domain {i}: the inlet face must be an END face")
                lo = [0, 0, 0]
                ext = list(shape)
                lo[a] = 0 if s == 0 else shape[a] - 1
                ext[a] = 1
                e.inlet_lo[:] = lo
                e.inlet_ext[:] = ext
                e.has_inlet = 1
            e.code_map = _device.ptr(m)
            e.tile_rows = 0
        self.arr = arr
        self.n_lanes = max(1, min(int(lanes), len(self.domains)))
        self.fused = bool(fused) and all(d.inlet is None for d in self.domains)
        if edt and not self.fused:
            raise ValueError("edt=True needs the fused batch (no inlet requirements, fused=True)")
        L = _lib.lib()
        self.tune_cache = tune_cache if tune_cache is not None else os.environ.get("PMX_TUNE_CACHE")
        if autotune:
            self._autotune()
        self.graph = graph
        self._g = None
        self._plan = None
        self.edts = []
        if self.fused:
            if edt:
                self.edts = [torch.empty(tuple(d.img.shape), dtype=torch.int32, device=d.img.device)
                             for d in self.domains]
            self._edt_ptrs = (C.c_void_p * len(self.domains))(*[_device.ptr(t) for t in self.edts]) if edt else None
            nbytes = C.c_size_t()
            _lib.check(L.pmx_psd_batch_plan_size(len(self.domains), arr, self._edt_ptrs, len(self.radii), self.te,
                                                 self.td, C.byref(nbytes)))
            self.ws_bytes = nbytes.value
            self.ws = _device.workspace(self.ws_bytes)
            plan = C.c_void_p()
            _lib.check(L.pmx_psd_batch_create(len(self.domains), arr, self._edt_ptrs, len(self.radii), self.te,
                                              self.td, _device.ptr(self.ws), self.ws_bytes, _device.stream(),
                                              C.byref(plan)))
            self._plan = plan
            self.streams = []
        else:
            nbytes = C.c_size_t()
            _lib.check(L.pmx_psd_batch_workspace_size(len(self.domains), arr, self.n_lanes, C.byref(nbytes)))
            self.ws_bytes = nbytes.value
            self.ws = _device.workspace(self.ws_bytes)
            self.streams = [torch.cuda.Stream() for _ in range(self.n_lanes)]

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            _lib.lib().pmx_psd_batch_destroy(plan)
            self._plan = None

    def _autotune(self):
        """Tile rows per domain shape: the faster of 32 and 64 on the
        domain's own erosion EDT (clamped at the largest threshold)."""
        L = _lib.lib()
        if self.tune_cache:
            _lib.check(L.pmx_tune_cache_load(os.fsencode(self.tune_cache)))
        need = C.c_size_t()
        big = 0
        for e in self.arr:
            _lib.check(L.pmx_edt_autotune_workspace_size(C.byref(e.g), C.byref(need)))
            big = max(big, need.value)
        ws = _device.workspace(big)
        tr = C.c_int32()
        for e in self.arr:
            _lib.check(L.pmx_edt_autotune(e.img, C.byref(e.g), int(self.te[0]) + 1, 0, C.byref(tr), _device.ptr(ws),
                                          big, _device.stream()))
            e.tile_rows = tr.value
        del ws
        if self.tune_cache:
            d = os.path.dirname(os.path.abspath(self.tune_cache))
            os.makedirs(d, exist_ok=True)
            _lib.check(L.pmx_tune_cache_save(os.fsencode(self.tune_cache)))

    def tile_rows(self):
        return [e.tile_rows for e in self.arr]

    def _launch(self):
        if self.fused:
            _lib.check(_lib.lib().pmx_psd_batch_run(self._plan, _device.stream()))
            return
        cur = torch.cuda.current_stream()
        for st in self.streams:
            st.wait_stream(cur)
        handles = (C.c_void_p * self.n_lanes)(*[st.cuda_stream for st in self.streams])
        _lib.check(_lib.lib().pmx_psd_sweep_batch(len(self.domains), self.arr, len(self.radii), self.te, self.td,
                                                  self.n_lanes, handles, _device.ptr(self.ws), self.ws_bytes))
        for st in self.streams:
            cur.wait_stream(st)

    def run(self):
        """Codes per domain: i+1 where radii[i] first invades a pore voxel."""
        if not self.graph:
            self._launch()
            return self.maps
        if self._g is None:
            self._launch()  # eager first run (also warms the kernels up)
            torch.cuda.synchronize()
            self._g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self._g):
                self._launch()
        self._g.replay()
        return self.maps


def batch_saturation(code_map: torch.Tensor, img: torch.Tensor, n_levels: int) -> np.ndarray:
    """Wetting saturation after each level of one batched domain (no pads:
    V = its voxel count): (wet / V) / (pore / V) in float64, on the device
    (pmx_saturation_curve, the arithmetic of Multiphase.get_saturation)."""
    from .porosimetry import saturation_curve_box
    shape = tuple(img.shape)
    return saturation_curve_box(img, code_map, (0, 0, 0), shape, n_levels, float(img.numel()))
