"""Training data for the surrogate: seeded sphere packs labelled by the
exact GPU porosimetry.

Sample i: a size^3 pack drawn with numpy.random.default_rng(seed0 + i):
porosity uniform in 0.25..0.6, sphere radii uniform in 4..12 voxels, each
axis periodic or wall with probability 1/2 (the batched-domain generator
at a fixed edge).  Label: the wetting saturation after each radius of
`radii` (descending) of the exact pore-size-distribution sweep without an
inlet requirement -- the fraction of pore voxels not yet invaded.
"""

from __future__ import annotations

import numpy as np
import torch

from ..domain_generation import pack_radii, sphere_pack
from ..filters.batch import BatchDomain, BatchSweep, batch_saturation

RADII = tuple(float(r) for r in range(32, 0, -1))


def draw_domain(seed: int, size: int):
    rng = np.random.default_rng(seed)
    por = float(rng.uniform(0.25, 0.6))
    bc = tuple("periodic" if rng.random() < 0.5 else "wall" for _ in range(3))
    img, _ = sphere_pack((size, size, size), por, pack_radii("uniform", lo=4, hi=12), seed=seed,
                         periodic=tuple(b == "periodic" for b in bc))
    return img, bc


def exact_curves(images, bcs, radii=RADII, chunk=256, lanes=4):
    """Exact saturation curves (float32 [n, len(radii)]) of CUDA uint8 images."""
    out = []
    for i in range(0, len(images), chunk):
        doms = [BatchDomain(images[j], bcs[j], None) for j in range(i, min(len(images), i + chunk))]
        maps = BatchSweep(doms, radii, lanes=lanes, autotune=False, graph=False).run()
        for d, m in zip(doms, maps):
            out.append(batch_saturation(m, d.img, len(radii)).astype(np.float32))
        del maps, doms
    return np.stack(out)


def make_dataset(n: int, seed0: int, size: int = 128, radii=RADII):
    """(images CUDA uint8 [n, size, size, size], curves float32 [n, q], bcs)."""
    images = torch.empty((n, size, size, size), dtype=torch.uint8, device="cuda")
    bcs = []
    for i in range(n):
        img, bc = draw_domain(seed0 + i, size)
        images[i].copy_(img)
        bcs.append(bc)
    curves = exact_curves([images[i] for i in range(n)], bcs, radii)
    return images, curves, bcs
