"""Shared helpers: padded PMMoTo-style cases for the product and the
matching oracle Box."""

import numpy as np

from oracle import oracle as O

E, W, P = "end", "wall", "periodic"
CODE = {E: O.BC_END, W: O.BC_WALL, P: O.BC_PERIODIC}


def rand_binary(shape, p_zero, seed):
    rng = np.random.default_rng(seed)
    return rng.choice([0, 1], size=shape, p=[p_zero, 1 - p_zero]).astype(np.uint8)


def padded_case(own, bts, inlet=((False, False),) * 3, outlet=((False, False),) * 3, pad=1):
    """PMMoTo padded image of an own-voxel image: periodic axes get `pad`
    periodic copies, WALL faces a zero layer, END faces nothing.  Returns
    (padded image, subdomain, oracle Box)."""
    import pmmoto_amd
    BT = pmmoto_amd.BoundaryType
    img = own
    pads = []
    for a in range(3):
        lo, hi = bts[a]
        if lo == P:
            img = np.pad(img, [(0, 0)] * a + [(pad, pad)] + [(0, 0)] * (2 - a), mode="wrap")
            pads.append((pad, pad))
        else:
            pl = 1 if lo == W else 0
            ph = 1 if hi == W else 0
            img = np.pad(img, [(0, 0)] * a + [(pl, ph)] + [(0, 0)] * (2 - a), constant_values=0)
            pads.append((pl, ph))
    n = own.shape
    sd = pmmoto_amd.initialize(n, box=tuple((0, float(v)) for v in n),
                               boundary_types=tuple((BT(lo), BT(hi)) for lo, hi in bts),
                               inlet=inlet, outlet=outlet, pad=(pad, pad, pad))
    assert sd.voxels == img.shape, (sd.voxels, img.shape)
    box = O.Box(bc=tuple((CODE[lo], CODE[hi]) for lo, hi in bts), pad=tuple(pads),
                inlet=sd.inlet, outlet=sd.outlet)
    return np.ascontiguousarray(img), sd, box


def cpu_sphere_pack(shape, porosity, kind, seed, periodic=(False, False, False), **kw):
    """The product's seeded draw sequence (domain_generation.sphere_pack),
    rasterised by the oracle's sequential rule on the CPU."""
    import math
    rng = np.random.default_rng(seed)
    if kind == "fixed":
        sampler, mean_r3 = (lambda g, n: np.full(n, float(kw["r"]))), float(kw["r"]) ** 3
    elif kind == "uniform":
        lo, hi = float(kw["lo"]), float(kw["hi"])
        sampler, mean_r3 = (lambda g, n: g.uniform(lo, hi, n)), (hi ** 4 - lo ** 4) / (4 * (hi - lo))
    else:
        mean, sigma = float(kw["mean"]), float(kw["sigma"])
        mu = math.log(mean) - sigma * sigma / 2
        sampler, mean_r3 = (lambda g, n: g.lognormal(mu, sigma, n)), math.exp(3 * mu + 4.5 * sigma * sigma)
    vol = float(np.prod(shape))
    est = int(math.ceil(-math.log(porosity) * vol / (4.0 / 3.0 * math.pi * mean_r3) * 1.25)) + 64
    batches = []
    while True:
        m = est if not batches else max(64, est // 4)
        centres = rng.uniform(0.0, 1.0, (m, 3)) * np.array(shape, dtype=np.float64)
        batches.append(np.column_stack([centres, np.asarray(sampler(rng, m), dtype=np.float64)]))
        spheres = np.concatenate(batches)
        img, n = O.rasterize_spheres(shape, periodic, spheres, porosity)
        if n > 0:
            return img, spheres[:n]
