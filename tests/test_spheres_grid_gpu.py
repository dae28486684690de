"""gen_pm_spheres_domain on the GPU (pmx_spheres_on_grid) against the
reference's known answer (tests/domain_generation/test_domain_generation.py:8-18:
one sphere of radius 0.25 at the centre of a 10^3 unit box leaves 944 pore
voxels) and against the oracle's restatement of the float64 centroid test,
the periodic pad copies and the WALL layer on packs read from the
reference's sphere-pack files (copies under tests/golden/sphere_packs)."""

import os

import numpy as np
import pytest

from oracle import oracle as O
from pmmoto_amd import initialize
from pmmoto_amd.core import BoundaryType
from pmmoto_amd.io import data_read

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sphere_packs")


def _host(t):
    return t.cpu().numpy() if hasattr(t, "cpu") else np.asarray(t)


def test_single_sphere_known_answer():
    from pmmoto_amd.domain_generation import gen_pm_spheres_domain
    sd = initialize(voxels=(10, 10, 10))
    print("[spheres_grid] known answer 944", flush=True)
    pm = gen_pm_spheres_domain(sd, np.array([[0.5, 0.5, 0.5, 0.25]]))
    assert int(_host(pm.img).sum()) == 944


@pytest.mark.parametrize("name,voxels", [("bcc.out", (40, 37, 33)), ("10pack.out", (48, 48, 48)),
                                         ("test_periodic_spheres.out", (20, 21, 22))])
@pytest.mark.parametrize("bt", [BoundaryType.END, BoundaryType.WALL, BoundaryType.PERIODIC])
@pytest.mark.parametrize("invert", [False, True])
def test_matches_oracle(name, voxels, bt, invert):
    from pmmoto_amd.domain_generation import gen_pm_spheres_domain
    spheres, box = data_read.read_sphere_pack_xyzr_domain(os.path.join(GOLDEN, name))
    sd = initialize(voxels=voxels, box=box, boundary_types=((bt, bt),) * 3)
    print(f"[spheres_grid] {name} {voxels} {bt} invert={invert}", flush=True)
    got = _host(gen_pm_spheres_domain(sd, spheres, invert=invert).img)
    want = O.spheres_on_grid(spheres, sd.coords, invert)
    if bt == BoundaryType.PERIODIC:
        want = O.periodic_pads(want, sd.pad)
    box_o = O.Box(bc=((O.BC_WALL, O.BC_WALL),) * 3 if bt == BoundaryType.WALL else ((O.BC_END, O.BC_END),) * 3,
                  pad=sd.pad)
    want = O.set_wall_bcs(box_o, want)
    assert got.shape == tuple(sd.voxels)
    np.testing.assert_array_equal(got, want)


def test_reservoir_and_empty_list():
    from pmmoto_amd.domain_generation import gen_pm_spheres_domain
    sd = initialize(voxels=(12, 10, 8), boundary_types=((BoundaryType.END, BoundaryType.END),) * 3,
                    inlet=((True, False), (False, False), (False, False)), reservoir_voxels=3)
    print("[spheres_grid] reservoir + empty", flush=True)
    sp = np.array([[0.1, 0.5, 0.5, 0.3], [-0.2, 0.5, 0.5, 0.3]])  # the second lies over the reservoir
    got = _host(gen_pm_spheres_domain(sd, sp).img)
    np.testing.assert_array_equal(got, O.spheres_on_grid(sp, sd.coords))
    empty = _host(gen_pm_spheres_domain(sd, np.zeros((0, 4))).img)
    assert empty.shape == tuple(sd.voxels) and empty.min() == 1
    with pytest.raises(ValueError):
        gen_pm_spheres_domain(sd, np.zeros((3, 3)))
