"""Pins the oracle's drainage chain (float32 EDT at resolution 0.1, FFT
dilation, inlet/outlet connectivity with phases, the non-wetting
restriction, reservoir reset, saturation arithmetic) against the
reference's golden ink-bottle curve
(tests/applications/test_drainage_inkbottle.py:9-78, 115-124)."""

import math

import numpy as np

from oracle import oracle as O

CAPILLARY_PRESSURE = [
    1.58965, 1.5943, 1.60194, 1.61322, 1.62893, 1.65002, 1.67755, 1.7127, 1.75678, 1.81122, 1.87764,
    1.95783, 2.05388, 2.16814, 2.30332, 2.4625, 2.64914, 2.86704, 3.12024, 3.41274, 3.74806, 4.12854,
    4.55421, 5.02123, 5.52008, 6.03352, 6.53538, 6.9909, 7.36005, 7.60403, 7.69393, 7.69409,
]
W_SATURATION_EXPECTED = [
    0.914608928, 0.914608928, 0.914608928, 0.903539715, 0.903539715, 0.895876414, 0.885658679,
    0.877508819, 0.870696995, 0.865466488, 0.86047926, 0.856343511, 0.853545797, 0.840651989,
    0.838462474, 0.837124437, 0.830190974, 0.829461136, 0.827514901, 0.826055224, 0.822649313,
    0.821919475, 0.820824717, 0.818513563, 0.817418805, 0.816810607, 0.816810607, 0.816445688,
    0.81571585, 0.81571585, 0.81571585, 0.81571585,
]


def inkbottle_case():
    """140x30x30 voxels, box x (0, 14), y, z (-1.5, 1.5), inlet on x+ with a
    10-voxel reservoir, outlet on x-, END faces (no pads)."""
    res = 0.1
    nx = 150  # 140 + reservoir
    x = np.linspace(0.0 + 0.05, 15.0 - 0.05, nx)
    y = np.linspace(-1.5 + 0.05, 1.5 - 0.05, 30)
    z = y.copy()
    r = np.array([0.01 * math.cos(0.01 * xi) + 0.5 * math.sin(xi) + 0.75 for xi in x])
    yy = (y * y)[None, :, None] / (r * r)[:, None, None]
    zz = (z * z)[None, None, :] / (r * r)[:, None, None]
    img = ((yy + zz) <= 1).astype(np.uint8)
    img[140:] = 1  # reservoir
    box = O.Box(bc=((O.BC_END, O.BC_END),) * 3, pad=((0, 0),) * 3,
                inlet=((False, True), (False, False), (False, False)),
                outlet=((True, False), (False, False), (False, False)),
                resolution=(res, res, res), reservoir=((0, 10), (0, 0), (0, 0)))
    return img, box


def test_oracle_drainage_matches_golden_inkbottle_curve():
    img, box = inkbottle_case()
    mp = np.where(img == 1, 2, 0).astype(np.uint8)
    max_distance = float(np.max(O.ref_edt(img, resolution=box.resolution)))
    sw, _ = O.drainage(box, img, mp, CAPILLARY_PRESSURE, (140, 30, 30), max_distance)
    np.testing.assert_array_almost_equal(sw, W_SATURATION_EXPECTED)
