"""The slice of PMMoTo's domain model the filters read.

One GPU holds the whole domain, so there is exactly one subdomain (rank 0)
and no INTERNAL faces.  What is kept is what the filters consume: box,
voxels, resolution, per-face boundary types, inlet/outlet flags, pad
widths and own-voxel bounds, reservoir pads, and set_wall_bcs.

Follows src/pmmoto/core/pmmoto.py:20-80 (initialize),
domain_discretization.py:56-68 (resolution), subdomain.py:186-287
(inlet/outlet on END faces, set_wall_bcs), subdomain_padded.py:71-114
(pads: END -> 0, WALL -> 1, PERIODIC -> pad), :218-307 (reservoir),
features.py:76-139 (face own slabs).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .boundary_types import BoundaryType

_END2 = ((BoundaryType.END, BoundaryType.END),) * 3
_FALSE2 = ((False, False),) * 3


@dataclass
class Domain:
    box: tuple
    boundary_types: tuple
    inlet: tuple
    outlet: tuple
    voxels: tuple
    subdomains: tuple = (1, 1, 1)
    num_subdomains: int = 1
    dims: int = 3

    @property
    def length(self):
        return tuple(b[1] - b[0] for b in self.box)

    @property
    def resolution(self) -> tuple:
        # domain_discretization.py:56-68
        return tuple(float(l / v) for l, v in zip(self.length, self.voxels))

    @property
    def periodic(self) -> bool:
        return any(bt == BoundaryType.PERIODIC for pair in self.boundary_types for bt in pair)

    @property
    def periodic_axes(self) -> tuple:
        return tuple(self.boundary_types[a][0] == BoundaryType.PERIODIC for a in range(3))


@dataclass
class PaddedSubdomain:
    domain: Domain
    pad: tuple  # ((lo, hi),) * 3
    reservoir_pad: tuple
    rank: int = 0
    index: tuple = (0, 0, 0)
    # False for initialize(return_subdomain=True): the reference's plain
    # Subdomain, which has no pads, no reservoir and no update_reservoir
    padded: bool = True
    boundary_types: tuple = field(init=False)
    inlet: tuple = field(init=False)
    outlet: tuple = field(init=False)

    def __post_init__(self):
        self.boundary_types = self.domain.boundary_types
        # subdomain.py:186-248: inlet/outlet only on END faces
        self.inlet = tuple(
            tuple(bool(self.domain.inlet[a][s]) and self.boundary_types[a][s] == BoundaryType.END
                  for s in (0, 1)) for a in range(3))
        self.outlet = tuple(
            tuple(bool(self.domain.outlet[a][s]) and self.boundary_types[a][s] == BoundaryType.END
                  for s in (0, 1)) for a in range(3))

    @property
    def voxels(self) -> tuple:
        """Padded extent of this subdomain's arrays."""
        return tuple(v + p[0] + p[1] + r[0] + r[1]
                     for v, p, r in zip(self.domain.voxels, self.pad, self.reservoir_pad))

    @property
    def start(self) -> tuple:
        """Global index of the first (padded) voxel: minus the low pads and
        reservoir (subdomain_padded.py:242-269, the image extent in VTK files)."""
        return tuple(-(p[0] + r[0]) for p, r in zip(self.pad, self.reservoir_pad))

    @property
    def own_voxels(self) -> np.ndarray:
        return np.array(self.domain.voxels, dtype=np.int64)

    def get_own_voxels(self) -> np.ndarray:
        """[x0, x1, y0, y1, z0, z1] bounds of owned voxels (subdomain_padded.py:271-283)."""
        out = np.zeros(6, dtype=np.int64)
        for a, (p, r) in enumerate(zip(self.pad, self.reservoir_pad)):
            out[2 * a] = p[0] + r[0]
            out[2 * a + 1] = self.voxels[a] - p[1] - r[1]
        return out

    @property
    def box(self) -> tuple:
        """Physical extent including pads and reservoirs (subdomain_padded.py:204-216)."""
        res = self.domain.resolution
        return tuple((b[0] - (p[0] + r[0]) * res[a], b[1] + (p[1] + r[1]) * res[a])
                     for a, (b, p, r) in enumerate(zip(self.domain.box, self.pad, self.reservoir_pad)))

    @property
    def coords(self) -> list:
        """Voxel-centroid coordinates per axis (domain_discretization.py:70-90)."""
        out = []
        for (lo, hi), n, res in zip(self.box, self.voxels, self.domain.resolution):
            half = 0.5 * res
            out.append(np.linspace(lo + half, hi - half, n))
        return out

    def check_boundary_type(self, bt: BoundaryType) -> bool:
        return any(b == bt for pair in self.boundary_types for b in pair)

    def set_wall_bcs(self, img):
        """subdomain.py:273-287: outermost layer of WALL faces forced to 0."""
        for a in range(3):
            for s in (0, 1):
                if self.boundary_types[a][s] == BoundaryType.WALL:
                    idx = [slice(None)] * 3
                    idx[a] = 0 if s == 0 else -1
                    img[tuple(idx)] = 0
        return img

    def update_reservoir(self, img, value):
        """subdomain_padded.py:285-307."""
        for a, (lo, hi) in enumerate(self.reservoir_pad):
            if lo > 0:
                idx = [slice(None)] * img.ndim
                idx[a] = slice(0, lo)
                img[tuple(idx)] = value
            if hi > 0:
                idx = [slice(None)] * img.ndim
                idx[a] = slice(-hi, None)
                img[tuple(idx)] = value
        return img

    def face_own_slab(self, axis: int, side: int) -> tuple:
        """Index of the voxels a face feature owns (features.py:76-139): on the
        face's axis the layer next to the face; on the other axes the padded
        extent without the pads (reservoir layers included), as
        SubdomainFeatures builds them from the padded voxel counts and the
        pads (subdomain_features.py:46-50)."""
        idx = [slice(int(self.pad[a][0]), int(self.voxels[a] - self.pad[a][1])) for a in range(3)]
        p = self.pad[axis][side]
        n = self.voxels[axis]
        wall = self.boundary_types[axis][side] == BoundaryType.WALL
        if side == 0:
            idx[axis] = slice(1, 2) if (p and wall) else (slice(p, 2 * p) if p else slice(0, 1))
        else:
            idx[axis] = slice(n - 2, n - 1) if (p and wall) else (slice(n - 2 * p, n - p) if p else slice(n - 1, n))
        return tuple(idx)


def initialize(
    voxels,
    box=((0, 1.0), (0, 1.0), (0, 1.0)),
    subdomains=(1, 1, 1),
    boundary_types=_END2,
    inlet=_FALSE2,
    outlet=_FALSE2,
    reservoir_voxels: int = 0,
    rank: int = 0,
    pad=(1, 1, 1),
    verlet_domains=(1, 1, 1),
    return_subdomain: bool = False,
) -> PaddedSubdomain:
    """pmmoto.initialize (core/pmmoto.py:20-80) for one GPU holding the
    whole domain.  ``subdomains`` other than (1, 1, 1) are rejected: the
    single-GPU build replaces spatial decomposition."""
    if tuple(subdomains) != (1, 1, 1):
        raise ValueError("one GPU holds the whole domain: subdomains must be (1, 1, 1)")
    if rank != 0:
        raise ValueError("rank must be 0")
    bts = tuple(tuple(BoundaryType(b) for b in pair) for pair in boundary_types)
    for a in range(3):
        if (bts[a][0] == BoundaryType.PERIODIC) != (bts[a][1] == BoundaryType.PERIODIC):
            raise ValueError(f"axis {a}: periodic boundaries must be periodic on both faces")
    dom = Domain(box=tuple(tuple(b) for b in box), boundary_types=bts, inlet=tuple(map(tuple, inlet)),
                 outlet=tuple(map(tuple, outlet)), voxels=tuple(int(v) for v in voxels))
    if return_subdomain:
        pads = ((0, 0),) * 3
    else:
        # subdomain_padded.py:71-114
        pads = tuple(tuple(0 if bts[a][s] == BoundaryType.END else (1 if bts[a][s] == BoundaryType.WALL
                                                                    else int(pad[a])) for s in (0, 1))
                     for a in range(3))
    sd = PaddedSubdomain(domain=dom, pad=pads, reservoir_pad=((0, 0),) * 3, padded=not return_subdomain)
    if reservoir_voxels and not return_subdomain:
        # subdomain_padded.py:218-240: reservoir pads on inlet faces
        sd.reservoir_pad = tuple(tuple(reservoir_voxels if sd.inlet[a][s] else 0 for s in (0, 1))
                                 for a in range(3))
    return sd
