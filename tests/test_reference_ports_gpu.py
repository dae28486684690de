"""The reference's own tests of drainage, porosimetry modes and the
PorousMedia / Multiphase objects, ported to the drop-in API, plus oracle
parity where the reference test only compares methods with each other:

* tests/filters/test_equilibrium_distribution.py:8-128 (drainage methods on
  bcc.out, save=True, argument errors, pressure sorting);
* tests/filters/test_porosimetry.py:43-165 (mode equality on bcc.in, the
  default-radii pore-size distribution, scalar / list / ndarray radii);
* tests/domain_generation/test_multiphase.py:8-57 and
  tests/domain_generation/test_porous_media.py:7-21 (known answers).
"""

import logging
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

PACKS = os.path.join(os.path.dirname(__file__), "golden", "sphere_packs")


def host(a):
    """numpy view of a numpy array or a device tensor."""
    return a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)


def bcc_case():
    import pmmoto_amd
    from pmmoto_amd.domain_generation import gen_mp_constant, gen_pm_spheres_domain
    from pmmoto_amd.io.data_read import read_sphere_pack_xyzr_domain
    spheres, box = read_sphere_pack_xyzr_domain(os.path.join(PACKS, "bcc.out"))
    sd = pmmoto_amd.initialize((50, 50, 50), box=box)
    pm = gen_pm_spheres_domain(sd, spheres)
    return sd, pm, gen_mp_constant(pm, 2)


def test_drainage_methods_agree_and_match_oracle():
    """test_equilibrium_distribution.py:8-53: with a zero contact angle the
    three methods give the same curve; the standard method also equals the
    oracle's drainage chain (its final phase image too)."""
    from pmmoto_amd.filters import equilibrium_distribution as ED
    for pc in (0.1, 20):
        sws = {}
        for method in ("standard", "contact_angle", "extended_contact_angle"):
            sd, pm, mp = bcc_case()
            print(f"[drainage] bcc 50^3 pc={pc} method={method}", flush=True)
            sws[method] = (ED.drainage(mp, pc, contact_angle=0, method=method), host(mp.img))
        np.testing.assert_array_equal(sws["standard"][0], sws["contact_angle"][0])
        np.testing.assert_array_equal(sws["standard"][0], sws["extended_contact_angle"][0])
        np.testing.assert_array_equal(sws["standard"][1], sws["extended_contact_angle"][1])
        sd, pm, mp = bcc_case()
        img = host(pm.img)
        box = O.Box(bc=((O.BC_END, O.BC_END),) * 3, pad=((0, 0),) * 3, resolution=sd.domain.resolution)
        max_d = float(np.max(O.ref_edt(img, resolution=sd.domain.resolution)))
        want_sw, want_img = O.drainage(box, img, np.where(img == 1, 2, 0).astype(np.uint8), [pc], (50, 50, 50),
                                       max_d)
        np.testing.assert_array_equal(sws["standard"][0], want_sw)
        np.testing.assert_array_equal(sws["standard"][1], want_img)


def test_drainage_save_img(tmp_path):
    """test_equilibrium_distribution.py:56-71: save=True writes one image per
    pressure (the porosimetry image alongside), read back here."""
    from pmmoto_amd.filters import equilibrium_distribution as ED
    sd, pm, mp = bcc_case()
    sw = ED.drainage(mp, [0.1, 20], contact_angle=0, method="standard", save=True, out_folder=str(tmp_path))
    files = sorted(os.listdir(tmp_path))
    assert files == ["capillary_pressure_0_100.vti", "capillary_pressure_20_000.vti"], files
    text = open(os.path.join(tmp_path, files[-1]), "rb").read()
    assert b'Name="img"' in text and b'Name="morph"' in text
    assert sw.shape == (2,)


def test_drainage_errors_and_sorting(caplog):
    """test_equilibrium_distribution.py:74-128."""
    from pmmoto_amd.filters import equilibrium_distribution as ED
    _, _, mp = bcc_case()
    with pytest.raises(ValueError):
        ED.drainage(mp, 0.1, contact_angle=10, method="standard")
    with pytest.raises(ValueError):
        ED.drainage(mp, 0.1, contact_angle=0, method="undefined")
    with caplog.at_level(logging.WARNING):
        ED.drainage(mp, [2, 1], contact_angle=0, method="standard")
    assert "must be monotonically increasing" in caplog.text


def test_porosimetry_modes_agree_on_bcc_in():
    """test_porosimetry.py:43-79: morph, distance, hybrid and [r, r] give the
    same image at radius 0.030001 on the 100^3 bcc domain, and a
    three-element radius raises; the distance mode also equals the oracle."""
    import pmmoto_amd
    from pmmoto_amd.domain_generation import gen_pm_spheres_domain
    from pmmoto_amd.filters import porosimetry as PO
    from pmmoto_amd.io.data_read import read_sphere_pack_xyzr_domain
    spheres, box = read_sphere_pack_xyzr_domain(os.path.join(PACKS, "bcc.in"))
    BT = pmmoto_amd.BoundaryType
    sd = pmmoto_amd.initialize((100, 100, 100), box=box, boundary_types=((BT.END, BT.END),) * 3)
    pm = gen_pm_spheres_domain(sd, spheres)
    radius = 0.030001
    print("[porosimetry modes] bcc.in 100^3", flush=True)
    morph = host(PO.porosimetry(sd, pm, radius, mode="morph"))
    dist = host(PO.porosimetry(sd, pm, radius, mode="distance"))
    hybrid = host(PO.porosimetry(sd, pm, radius, mode="hybrid"))
    pair = host(PO.porosimetry(sd, pm, [radius, radius], mode="morph"))
    np.testing.assert_array_equal(morph, dist)
    np.testing.assert_array_equal(morph, hybrid)
    np.testing.assert_array_equal(morph, pair)
    with pytest.raises(ValueError):
        PO.porosimetry(sd, pm, [radius, radius, radius], mode="morph")
    obox = O.Box(bc=((O.BC_END, O.BC_END),) * 3, pad=((0, 0),) * 3, resolution=sd.domain.resolution)
    np.testing.assert_array_equal(dist, O.porosimetry(obox, host(pm.img), radius, mode="distance"))


def test_pore_size_distribution_default_radii():
    """test_porosimetry.py:128-165: the default radii (25, linear from just
    above the resolution to just below the largest distance) give the same
    map in every mode; a scalar, list and ndarray radius give the same map;
    the distance mode equals the oracle's chain with the radii it derives
    from its own float32 EDT."""
    import pmmoto_amd
    from pmmoto_amd.domain_generation import gen_pm_spheres_domain
    from pmmoto_amd.filters import porosimetry as PO
    sd = pmmoto_amd.initialize((30, 30, 30))
    spheres = np.array([[0.5, 0.5, 0.5, 0.1]])
    pm = gen_pm_spheres_domain(sd, spheres, invert=True)
    print("[pore_size_distribution] default radii, 30^3", flush=True)
    hybrid = host(PO.pore_size_distribution(sd, pm, inlet=False, mode="hybrid"))
    dist = host(PO.pore_size_distribution(sd, pm, inlet=False, mode="distance"))
    morph = host(PO.pore_size_distribution(sd, pm, inlet=False, mode="morph"))
    np.testing.assert_array_equal(hybrid, dist)
    np.testing.assert_array_equal(hybrid, morph)
    r = 0.03333333333333334
    f = host(PO.pore_size_distribution(sd, pm, radii=r, inlet=False, mode="distance"))
    lst = host(PO.pore_size_distribution(sd, pm, radii=[r], inlet=False, mode="distance"))
    arr = host(PO.pore_size_distribution(sd, pm, radii=np.array([r]), inlet=False, mode="distance"))
    np.testing.assert_array_equal(f, lst)
    np.testing.assert_array_equal(f, arr)
    img = host(pm.img)
    res = sd.domain.resolution
    d = O.ref_edt(img, resolution=res)
    gmax = float(np.max(d)) - 1e-6 * float(np.max(d))
    radii = PO.get_sizes(np.nextafter(np.min(res), np.inf), gmax, 25, "linear")
    box = O.Box(bc=((O.BC_END, O.BC_END),) * 3, pad=((0, 0),) * 3, resolution=res)
    want = O.pore_size_distribution(box, img, radii, inlet=False, mode="distance")
    assert np.count_nonzero(want) > 0
    np.testing.assert_array_equal(dist, want)


def test_multiphase_and_porous_media_known_answers():
    """test_multiphase.py:8-57, test_porous_media.py:7-21."""
    import pmmoto_amd
    from pmmoto_amd.domain_generation import gen_mp_constant, gen_pm
    from pmmoto_amd.domain_generation.multiphase import Multiphase
    sd = pmmoto_amd.initialize((5, 5, 5))
    img = np.ones(sd.voxels, dtype=np.uint8)
    pm = gen_pm(sd, img)
    mp = gen_mp_constant(pm, 2)
    np.testing.assert_array_equal(mp.img, 2 * np.ones(sd.voxels))
    img = 3 * img
    mp.update_img(img)
    np.testing.assert_array_equal(mp.img, 3 * np.ones(sd.voxels))
    sd2 = pmmoto_amd.initialize((5, 5, 5), reservoir_voxels=5, return_subdomain=True)
    mp2 = gen_mp_constant(gen_pm(sd2, img), 2)
    with pytest.raises(TypeError):
        mp2.update_img(img)

    BT = pmmoto_amd.BoundaryType
    sd = pmmoto_amd.initialize(voxels=(10, 10, 10), boundary_types=((BT.WALL, BT.WALL),) * 3)
    img = np.zeros(sd.voxels)
    img[3:6, 3:6, 3:6] = 1
    pm = gen_pm(sd, img)
    assert pm.porosity == (27 / 1000)
    mp_img = np.copy(img)
    mp_img[4:6, 4:6, 4:6] = 2
    mp = Multiphase(porous_media=pm, img=mp_img, num_phases=2)
    assert mp.get_volume_fraction(1) == (19 / 1000)
    assert mp.get_volume_fraction(2) == (8 / 1000)
    assert mp.get_saturation(1) == (19 / 27)
    assert mp.get_saturation(2) == (8 / 27)
