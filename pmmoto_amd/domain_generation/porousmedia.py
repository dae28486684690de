"""PorousMedia (src/pmmoto/domain_generation/porousmedia.py:21-118)."""

from __future__ import annotations

import numpy as np
import torch

from .. import _device, _lib
from ..filters import distance


class PorousMedia:
    """Image (pore = 1) plus cached porosity and distance transform."""

    def __init__(self, subdomain, img):
        self.subdomain = subdomain
        self.img = img
        self._porosity = None
        self._distance = None
        self._max_distance = None

    @property
    def porosity(self) -> float:
        """porousmedia.py:43-66: count(own == 1) / prod(domain.voxels)."""
        if self._porosity is None:
            from .multiphase import count_value_device

            t, _ = _device.as_device_u8(self.img)
            self._porosity = count_value_device(t, self.subdomain, 1) / np.prod(self.subdomain.domain.voxels)
        return self._porosity

    @property
    def distance(self):
        """porousmedia.py:68-101 (edt with the subdomain; periodic pads copied)."""
        if self._distance is None:
            self._distance = distance.edt(img=self.img, subdomain=self.subdomain)
        return self._distance

    @property
    def max_distance(self) -> float:
        """utils.determine_maximum of the distance (utils.py:148-162)."""
        if self._max_distance is None:
            d = self.distance
            t = d if isinstance(d, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(d)).to(_device.require_gpu())
            out = torch.empty(1, dtype=torch.float32, device=t.device)
            _lib.check(_lib.lib().pmx_max_f32(_device.ptr(t), t.numel(), _device.ptr(out), _device.stream()))
            self._max_distance = float(out.item())
        return self._max_distance


def gen_pm(subdomain, img) -> PorousMedia:
    """porousmedia.py:104-118."""
    return PorousMedia(subdomain=subdomain, img=img)
