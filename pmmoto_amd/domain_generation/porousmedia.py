"""PorousMedia (src/pmmoto/domain_generation/porousmedia.py:21-118)."""

from __future__ import annotations

import numpy as np

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
        # porousmedia.py:43-66: count(own == 1) / prod(domain.voxels)
        if self._porosity is None:
            o = self.subdomain.get_own_voxels()
            own = np.asarray(self.img)[o[0]:o[1], o[2]:o[3], o[4]:o[5]]
            self._porosity = np.count_nonzero(own == 1) / np.prod(self.subdomain.domain.voxels)
        return self._porosity

    @property
    def distance(self):
        # porousmedia.py:68-101 (edt + update_buffer: periodic pads copied)
        if self._distance is None:
            self._distance = distance.edt(img=self.img, subdomain=self.subdomain)
        return self._distance

    @property
    def max_distance(self) -> float:
        if self._max_distance is None:
            self._max_distance = float(np.max(self.distance))
        return self._max_distance


def gen_pm(subdomain, img) -> PorousMedia:
    return PorousMedia(subdomain=subdomain, img=img)
