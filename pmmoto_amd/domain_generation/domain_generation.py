"""Seeded synthetic porous media, rasterised on the GPU.

``gen_img_random_binary`` mirrors the reference generator
(src/pmmoto/domain_generation/domain_generation.py:40-60) exactly (same
numpy calls, so the same image for the same seed).  ``sphere_pack`` grows
an overlapping-sphere pack in voxel units until the porosity reaches a
target: centres are uniform in the box, radii come from a sampler, voxel
(i,j,k) is solid when its centroid lies in a sphere (the reference's
Sphere::contains, particles/spheres.hpp:42-48, on voxel centroids,
domain_discretization.py:70-90).  The stopping rule -- keep the first
sphere after which porosity <= target -- is evaluated exactly from
libpmx's first-cover counts (pmx_sphere_first_cover).
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from .. import _device, _lib

__all__ = ["gen_img_random_binary", "sphere_pack", "pack_radii", "gen_pm_spheres_domain", "gen_pm_inkbottle",
           "gen_mp_constant"]


def gen_img_random_binary(shape, p_zero=0.5, seed=None):
    """domain_generation.py:40-60 (numpy; pore = 1)."""
    if not (0 <= p_zero <= 1):
        raise ValueError("Probability p_zero must be between 0 and 1.")
    rng = np.random.default_rng(seed)
    return rng.choice([0, 1], size=shape, p=[p_zero, 1 - p_zero]).astype(np.uint8)


def pack_radii(kind: str, **kw):
    """Radius samplers: ("fixed", r), ("uniform", lo, hi),
    ("lognormal", mean, sigma) with E[r] = mean."""
    if kind == "fixed":
        r = float(kw["r"])
        return (lambda rng, n: np.full(n, r)), r ** 3
    if kind == "uniform":
        lo, hi = float(kw["lo"]), float(kw["hi"])
        return (lambda rng, n: rng.uniform(lo, hi, n)), (hi ** 4 - lo ** 4) / (4 * (hi - lo))
    if kind == "lognormal":
        mean, sigma = float(kw["mean"]), float(kw["sigma"])
        mu = math.log(mean) - sigma * sigma / 2
        return (lambda rng, n: rng.lognormal(mu, sigma, n)), math.exp(3 * mu + 4.5 * sigma * sigma)
    raise ValueError(f"unknown radius distribution {kind}")


def sphere_pack(shape, porosity, radii, seed, periodic=(False, False, False), device_out=True):
    """Overlapping-sphere pack: returns (uint8 image, pore = 1, as a device
    tensor or numpy array; float64 (n, 4) array of the spheres used).

    Spheres are drawn in batches from ``numpy.random.default_rng(seed)``:
    each batch draws centres ``rng.uniform(0, shape, (m, 3))`` then radii;
    batches are appended until the target is reached."""
    sampler, mean_r3 = radii
    shape = tuple(int(s) for s in shape)
    vol = float(np.prod(shape))
    if not (0.0 < porosity < 1.0):
        raise ValueError("porosity must be in (0, 1)")
    rng = np.random.default_rng(seed)
    est = int(math.ceil(-math.log(porosity) * vol / (4.0 / 3.0 * math.pi * mean_r3) * 1.25)) + 64
    batches = []
    dev = _device.require_gpu()
    first = torch.empty(shape, dtype=torch.int32, device=dev)
    L = _lib.lib()
    sh = (C.c_int64 * 3)(*shape)
    per = (C.c_int32 * 3)(*[1 if p else 0 for p in periodic])
    while True:
        m = est if not batches else max(64, est // 4)
        centres = rng.uniform(0.0, 1.0, (m, 3)) * np.array(shape, dtype=np.float64)
        rad = np.asarray(sampler(rng, m), dtype=np.float64)
        batches.append(np.column_stack([centres, rad]))
        spheres = np.ascontiguousarray(np.concatenate(batches))
        if not np.all(np.isfinite(spheres)) or np.any(spheres[:, 3] < 0):
            raise ValueError("sphere parameters must be finite with r >= 0")
        sp = torch.from_numpy(spheres).to(dev)
        counts = torch.empty(spheres.shape[0], dtype=torch.int64, device=dev)
        _lib.check(L.pmx_sphere_first_cover(_device.ptr(sp), spheres.shape[0], sh, per,
                                            _device.ptr(first), _device.ptr(counts), _device.stream()))
        solid_after = np.cumsum(counts.cpu().numpy())  # solid voxels after k+1 spheres
        pores_after = vol - solid_after
        hit = np.nonzero(pores_after / vol <= porosity)[0]
        if hit.size:
            k = int(hit[0]) + 1
            break
        if spheres.shape[0] > (1 << 24) - est:
            raise ValueError("porosity target not reachable with this radius distribution")
    img = torch.empty(shape, dtype=torch.uint8, device=dev)
    _lib.check(L.pmx_sphere_image(_device.ptr(first), sh, k, _device.ptr(img), _device.stream()))
    del first
    return (img if device_out else img.cpu().numpy()), spheres[:k]


def gen_pm_spheres_domain(subdomain, spheres, kd: bool = False, invert: bool = False):
    """gen_pm_spheres_domain (domain_generation.py:125-155): spheres
    ``(n, 4)`` = (x, y, z, r) in the domain's coordinates are solid (0) on
    the subdomain's padded grid of voxel centroids (``subdomain.coords``),
    tested with the reference's float64 predicate (spheres.hpp:42-48); the
    periodic pads then hold periodic copies (update_buffer), ``invert``
    swaps pore and solid, and WALL faces get their solid layer
    (set_wall_bcs).  No periodic images of the spheres are added, as in the
    reference.  ``kd`` (a CPU search structure there) has no effect."""
    from ..filters.distance.euclidean_distance_transform import compute_box, fill_periodic_pads
    from .porousmedia import gen_pm

    sp = np.ascontiguousarray(np.asarray(spheres, dtype=np.float64))
    if sp.ndim != 2 or sp.shape[1] != 4:
        raise ValueError(f"spheres must be an (n, 4) array of x, y, z, r; got shape {sp.shape}")
    if not np.all(np.isfinite(sp)):
        raise ValueError("sphere data must be finite")
    dev = _device.require_gpu()
    shape = tuple(int(v) for v in subdomain.voxels)
    coords = [torch.from_numpy(np.ascontiguousarray(c, dtype=np.float64)).to(dev) for c in subdomain.coords]
    res = subdomain.domain.resolution
    lo = [float(b[0]) for b in subdomain.box]
    img = torch.empty(shape, dtype=torch.uint8, device=dev)
    sp_d = torch.from_numpy(sp).to(dev) if sp.shape[0] else None
    _lib.check(_lib.lib().pmx_spheres_on_grid(_device.ptr(sp_d) if sp_d is not None else None, sp.shape[0],
                                              _device.ptr(coords[0]), _device.ptr(coords[1]),
                                              _device.ptr(coords[2]), (C.c_int64 * 3)(*shape),
                                              (C.c_double * 3)(*lo), (C.c_double * 3)(*res), int(bool(invert)),
                                              _device.ptr(img), _device.stream()))
    _, _, _, pads = compute_box(shape, subdomain=subdomain)
    fill_periodic_pads([img], shape, pads)
    img = subdomain.set_wall_bcs(img)
    return gen_pm(subdomain, img)


def gen_pm_inkbottle(subdomain, r_y: float = 1.0, r_z: float = 1.0):
    """Ink-bottle pore (domain_generation.py:260-287, _domain_generation.pyx:
    107-133): pore where y^2/ry^2 + z^2/rz^2 <= 1 with
    r(x) = 0.01 cos(0.01 x) + 0.5 sin(x) + 0.75, pore for x < 0, and the
    reservoir pads set to pore.  r(x) uses the C library's sin/cos per x,
    the rest float64 numpy in the reference's operation order."""
    from .porousmedia import gen_pm

    x, y, z = subdomain.coords
    r = np.array([0.01 * math.cos(0.01 * xi) + 0.5 * math.sin(xi) + 0.75 for xi in x], dtype=np.float64)
    ry = r * r_y
    rz = r * r_z
    yy = (y * y)[None, :, None] / (ry * ry)[:, None, None]
    zz = (z * z)[None, None, :] / (rz * rz)[:, None, None]
    img = ((yy + zz) <= 1).astype(np.uint8)
    img[x < 0] = 1
    pm = gen_pm(subdomain, img)
    pm.img = subdomain.update_reservoir(pm.img, np.uint8(1))
    return pm


def gen_mp_constant(porous_media, fluid_phase: int = 1):
    """domain_generation.py:290-306: every pore voxel holds fluid_phase."""
    from .multiphase import Multiphase

    src = porous_media.img
    if isinstance(src, torch.Tensor):
        img = torch.where(src == 1, int(fluid_phase), 0).to(torch.uint8)
    else:
        img = np.where(np.asarray(src) == 1, fluid_phase, 0).astype(np.uint8)
    return Multiphase(porous_media, img, 2)
