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


def as_device_u8(img):
    """Return (contiguous uint8 cuda tensor, came_from_numpy)."""
    dev = require_gpu()
    if isinstance(img, torch.Tensor):
        t = img
        if t.dtype == torch.bool:
            t = t.to(torch.uint8)
        if t.dtype != torch.uint8:
            t = t.to(torch.uint8)
        return t.to(dev).contiguous(), False
    a = np.asarray(img)
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
