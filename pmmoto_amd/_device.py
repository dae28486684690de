"""Tensor plumbing between callers' arrays and the HIP kernels.

PyTorch-ROCm provides device memory, the current HIP stream and the
host<->device copies; the computation itself is libpmx's.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise _lib.PmxError("pmmoto_amd needs a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _u8_values_error(lo, hi, integral) -> str | None:
    if not integral:
        return "image holds non-integer values; the kernels take uint8 phase/label values 0..255"
    if lo < 0 or hi > 255:
        return (f"image values span [{lo}, {hi}]; the kernels take uint8 phase/label values 0..255 "
                "(relabel wider label images before calling)")
    return None


def as_device_u8(img):
    """Return (contiguous uint8 cuda tensor, came_from_numpy).  Images of
    another dtype are accepted when every value is an integer in 0..255;
    otherwise ValueError (a cast would silently merge or drop labels)."""
    dev = require_gpu()
    if isinstance(img, torch.Tensor):
        t = img
        if t.dtype == torch.bool:
            t = t.to(torch.uint8)
        if t.dtype != torch.uint8:
            if t.numel():
                integral = True
                if t.is_floating_point():
                    integral = bool(torch.all(t == torch.round(t)).item())
                lo, hi = (float(v) for v in torch.aminmax(t.to(torch.float64) if t.is_floating_point() else t))
                err = _u8_values_error(lo, hi, integral)
                if err:
                    raise ValueError(err)
            t = t.to(torch.uint8)
        return t.to(dev).contiguous(), False
    a = np.asarray(img)
    if a.dtype != np.uint8 and a.dtype != np.bool_:
        if a.size:
            if a.dtype.kind not in "biuf":
                raise ValueError(f"image dtype {a.dtype} is not numeric")
            integral = a.dtype.kind != "f" or bool(np.all(np.floor(a) == a))
            err = _u8_values_error(float(np.min(a)), float(np.max(a)), integral)
            if err:
                raise ValueError(err)
    if a.dtype != np.uint8:
        a = a.astype(np.uint8)
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev), True


def ptr(t: torch.Tensor, offset_elems: int = 0) -> C.c_void_p:
    return C.c_void_p(t.data_ptr() + offset_elems * t.element_size())


def workspace(nbytes: int) -> torch.Tensor:
    """Scratch for one call, from torch's stream-aware caching allocator."""
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=require_gpu())


def to_host(t: torch.Tensor, as_numpy: bool):
    return t.cpu().numpy() if as_numpy else t
