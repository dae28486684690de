"""Multiphase (src/pmmoto/domain_generation/multiphase.py:22-130): a phase
image over a PorousMedia (0 solid, 1..num_phases fluids)."""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _device, _lib


def own_box(subdomain):
    o = subdomain.get_own_voxels()
    lo = (C.c_int64 * 3)(int(o[0]), int(o[2]), int(o[4]))
    ext = (C.c_int64 * 3)(int(o[1] - o[0]), int(o[3] - o[2]), int(o[5] - o[4]))
    return lo, ext


def count_value_device(t: torch.Tensor, subdomain, value: int) -> int:
    """Number of own voxels equal to value (pmx_box_op count)."""
    lo, ext = own_box(subdomain)
    cnt = torch.empty(1, dtype=torch.int64, device=t.device)
    _lib.check(_lib.lib().pmx_box_op(_device.ptr(t), (C.c_int64 * 3)(*t.shape), lo, ext, 1, int(value),
                                     _device.ptr(cnt), _device.stream()))
    return int(cnt.item())


def update_reservoir_device(t: torch.Tensor, subdomain, value: int) -> None:
    """subdomain_padded.py:285-307 on a device image (in place)."""
    shape = tuple(t.shape)
    for a, (lo_w, hi_w) in enumerate(subdomain.reservoir_pad):
        for start, width in ((0, lo_w), (shape[a] - hi_w, hi_w)):
            if width <= 0:
                continue
            lo = [0, 0, 0]
            ext = list(shape)
            lo[a], ext[a] = start, width
            _lib.check(_lib.lib().pmx_box_op(_device.ptr(t), (C.c_int64 * 3)(*shape), (C.c_int64 * 3)(*lo),
                                             (C.c_int64 * 3)(*ext), 0, int(value), None, _device.stream()))


class Multiphase:
    """Phase image, saturations and the capillary-pressure / radius maps."""

    def __init__(self, porous_media, img, num_phases: int):
        self.porous_media = porous_media
        self.pm_img = porous_media.img
        self.subdomain = porous_media.subdomain
        self.num_phases = num_phases
        self.img = img
        self.fluids = list(range(1, num_phases + 1))

    def update_img(self, img) -> None:
        """multiphase.py:43-54: reservoir voxels are forced to phase 1; a
        subdomain without pads (initialize(return_subdomain=True)) raises
        TypeError, as the reference's plain Subdomain does."""
        if not getattr(self.subdomain, "padded", True):
            raise TypeError("subdomain does not support update_reservoir")
        if isinstance(img, torch.Tensor):
            update_reservoir_device(img, self.subdomain, 1)
        else:
            img = self.subdomain.update_reservoir(np.asarray(img), np.asarray(img).dtype.type(1))
        self.img = img

    def get_volume_fraction(self, phase: int, img=None) -> float:
        """multiphase.py:56-82: count(own == phase) / prod(domain.voxels)."""
        if img is None:
            img = self.img
        t, _ = _device.as_device_u8(img)
        return float(count_value_device(t, self.subdomain, phase) / np.prod(self.subdomain.domain.voxels))

    def get_saturation(self, phase: int, img=None) -> float:
        """multiphase.py:84-95."""
        return self.get_volume_fraction(phase, img) / self.porous_media.porosity

    @staticmethod
    def get_probe_radius(pc: float, gamma: float = 1, contact_angle: float = 0) -> float:
        """multiphase.py:97-116: r = 2 gamma / pc cos(theta) (0 if pc == 0)."""
        if pc == 0:
            return 0
        return 2.0 * gamma / pc * np.cos(np.deg2rad(contact_angle))

    @staticmethod
    def get_pc(radius: float, gamma: float = 1) -> float:
        """multiphase.py:118-130."""
        return 2.0 * gamma / radius
