"""GPU drainage (a chain of libpmx launches) against the reference's golden
ink-bottle curve and the oracle's final phase image; FFT-mode morphology
against the oracle's scipy.signal.fftconvolve restatement."""

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import E, P, W, padded_case, rand_binary
from tests.test_oracle_drainage import CAPILLARY_PRESSURE, W_SATURATION_EXPECTED, inkbottle_case

pytestmark = pytest.mark.gpu


def test_drainage_inkbottle_golden_curve():
    import pmmoto_amd
    from pmmoto_amd.domain_generation import gen_mp_constant, gen_pm_inkbottle
    from pmmoto_amd.filters import equilibrium_distribution
    sd = pmmoto_amd.initialize((140, 30, 30), box=((0.0, 14.0), (-1.5, 1.5), (-1.5, 1.5)),
                               inlet=((0, 1), (0, 0), (0, 0)), outlet=((1, 0), (0, 0), (0, 0)),
                               reservoir_voxels=10)
    pm = gen_pm_inkbottle(sd)
    img, box = inkbottle_case()
    np.testing.assert_array_equal(pm.img, img)
    mp = gen_mp_constant(pm, 2)
    print("[drainage] ink bottle 150x30x30, 32 pressures", flush=True)
    sw = equilibrium_distribution.drainage(mp, CAPILLARY_PRESSURE)
    np.testing.assert_array_almost_equal(sw, W_SATURATION_EXPECTED)
    max_d = float(np.max(O.ref_edt(img, resolution=box.resolution)))
    want_sw, want_img = O.drainage(box, img, np.where(img == 1, 2, 0).astype(np.uint8), CAPILLARY_PRESSURE,
                                   (140, 30, 30), max_d)
    np.testing.assert_array_equal(sw, want_sw)
    np.testing.assert_array_equal(mp.img, want_img)


@pytest.mark.parametrize("bts", [((E, E), (E, E), (E, E)), ((P, P), (P, P), (P, P)), ((E, E), (P, P), (W, W))])
@pytest.mark.parametrize("radius", [1.5, 2.9999999, np.float64(3.2)])
def test_fft_mode_morphology_matches_oracle(bts, radius):
    from pmmoto_amd.filters import morphological_operators as M
    own = rand_binary((24, 26, 22), 0.08, 5)
    img, sd, box = padded_case(own, bts)
    np.testing.assert_array_equal(M.erode(sd, img, radius, fft=True), O.erode_fft(box, img, radius))
    sparse = rand_binary((24, 26, 22), 0.99, 6)
    img2, _, _ = padded_case(sparse, bts)
    np.testing.assert_array_equal(M.dilate(sd, img2, radius, fft=True), O.dilate_fft(box, img2, radius))


@pytest.mark.parametrize("mode", ["morph", "hybrid", "distance"])
def test_porosimetry_modes_match_oracle(mode):
    from pmmoto_amd.domain_generation import gen_pm, pack_radii, sphere_pack
    from pmmoto_amd.filters import porosimetry as PO
    img_t, _ = sphere_pack((36, 34, 40), 0.45, pack_radii("uniform", lo=3, hi=7), seed=41)
    img, sd, box = padded_case(img_t.cpu().numpy(), ((E, E), (W, W), (W, W)),
                               inlet=((True, False), (False, False), (False, False)))
    pm = gen_pm(sd, img)
    for r in (2.0, 3.5):
        got = PO.porosimetry(sd, pm, r, inlet=True, mode=mode)
        np.testing.assert_array_equal(got, O.porosimetry(box, img, r, inlet=True, mode=mode))
