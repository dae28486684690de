"""GPU parity of the reference-arithmetic EDT (pmx_edt_ref) against vectors
written by the reference itself (tests/golden/make_ref_edt_vectors.py),
through the drop-in API:

* edt3d on periodic axes (the five-member boundary hulls) at unit,
  non-unit and anisotropic resolution, and the digest of a 128^3 periodic
  sphere pack whose reference output exceeds the exact torus EDT at 47
  voxels;
* edt(img, subdomain) on padded periodic subdomains (corrected_edt: the
  passes run over the padded array, correctors and hulls measured past the
  pads), at resolution 1/N and at unit resolution.

Then the same paths against the oracle's restatement on cases the vectors
leave open: lines whose stacks outgrow the LDS stack (sparse images), pads
of 2, and a padded periodic subdomain of a sphere pack at resolution 1/N.
Bar: float32 bitwise.
"""

import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rand_binary

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
VEC = np.load(os.path.join(GOLD, "ref_edt_vectors.npz"))
FACE = {0: "end", 1: "wall", 2: "periodic"}


def product_subdomain(own_shape, meta, unit):
    import pmmoto_amd
    BT = pmmoto_amd.BoundaryType
    bts = tuple((BT(FACE[int(meta[0][2 * a])]), BT(FACE[int(meta[0][2 * a + 1])])) for a in range(3))
    box = tuple((0.0, float(v)) for v in own_shape) if unit else ((0, 1.0), (0, 1.0), (0, 1.0))
    return pmmoto_amd.initialize(own_shape, box=box, boundary_types=bts)


@pytest.mark.parametrize("key,periodic,res", [
    ("edt3d_per_res", (True, True, True), (0.1, 0.1, 0.1)),
    ("edt3d_mixed_aniso", (False, True, True), (0.05, 1.0 / 16, 1.0 / 24)),
    ("edt3d_pack_res", (True, True, True), (1 / 40, 1 / 40, 1 / 40)),
])
def test_edt3d_matches_reference_vectors(key, periodic, res):
    from pmmoto_amd.filters import distance
    print(f"[edt3d golden] {key}", flush=True)
    got = distance.edt3d(VEC[key + "_in"], periodic=periodic, resolution=res, squared=True)
    np.testing.assert_array_equal(got.view(np.uint32), VEC[key + "_out"].view(np.uint32))


@pytest.mark.parametrize("key", ["sd_periodic", "sd_mixed", "sd_pack", "sd_per_end", "sd_pack_unit"])
def test_edt_padded_subdomain_matches_reference_vectors(key):
    from pmmoto_amd.filters import distance
    own = VEC[key + "_own"]
    sd = product_subdomain(own.shape, VEC[key + "_meta"], key.endswith("_unit"))
    assert tuple(sd.voxels) == VEC[key + "_in"].shape
    np.testing.assert_allclose(sd.domain.resolution, VEC[key + "_res"], rtol=0, atol=0)
    print(f"[edt subdomain golden] {key} {sd.voxels}", flush=True)
    got = distance.edt(VEC[key + "_in"], sd)
    np.testing.assert_array_equal(got.view(np.uint32), VEC[key + "_out"].view(np.uint32))


def test_periodic_pack_128_matches_reference_digest():
    from pmmoto_amd.domain_generation import pack_radii, sphere_pack
    from pmmoto_amd.filters import distance
    with open(os.path.join(GOLD, "ref_edt_periodic_128.json")) as f:
        doc = json.load(f)
    print("[sphere_pack] 128^3 periodic seed 2", flush=True)
    img_t, _ = sphere_pack((128, 128, 128), 0.5, pack_radii("uniform", lo=4, hi=12), seed=2,
                           periodic=(True, True, True))
    img = img_t.cpu().numpy()
    assert hashlib.sha256(img.tobytes()).hexdigest() == doc["image_sha256"]
    got = distance.edt3d(img_t, periodic=(True, True, True), squared=True).cpu().numpy()
    assert hashlib.sha256(got.tobytes()).hexdigest() == doc["output_sha256"]
    exact = distance.edt3d(img_t, periodic=(True, True, True), squared=True, mode="exact").cpu().numpy()
    diff = np.flatnonzero(got.ravel() != exact.ravel())
    assert diff.tolist() == [d[0] for d in doc["exact_differs_at"]]


@pytest.mark.parametrize("shape,p0,seed,periodic,res", [
    ((40, 36, 44), 0.002, 31, (True, True, True), (0.05, 0.05, 0.05)),
    ((50, 30, 20), 0.0005, 32, (False, True, True), (1.0, 1.0, 1.0)),
    ((33, 70, 65), 0.001, 33, (True, False, True), (0.02, 0.03, 0.01)),
    ((8, 9, 300), 0.0, 34, (True, True, True), (0.1, 0.1, 0.1)),
])
def test_edt3d_reference_mode_lds_overflow(shape, p0, seed, periodic, res):
    """Sparse images: most line stacks and boundary-hull segments outgrow
    the LDS stack and are recomputed by the workspace-stack fallback."""
    from pmmoto_amd.filters import distance
    img = rand_binary(shape, p0, seed)
    if p0 == 0.0:
        img[3, 4, 150] = 0
    print(f"[pmx_edt_ref sparse] {shape} p0={p0} periodic={periodic} res={res}", flush=True)
    got = distance.edt3d(img, periodic=periodic, resolution=res, squared=True, mode="reference")
    want = O.ref_edt_sq(img, periodic, res)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("bts,pad,p0,seed,unit", [
    ((("periodic", "periodic"),) * 3, (2, 2, 2), 0.05, 41, False),
    ((("periodic", "periodic"), ("end", "end"), ("wall", "wall")), (1, 1, 1), 0.003, 42, False),
    ((("wall", "wall"), ("periodic", "periodic"), ("periodic", "periodic")), (2, 1, 3), 0.1, 43, True),
])
def test_edt_padded_subdomain_vs_oracle(bts, pad, p0, seed, unit):
    import pmmoto_amd
    from pmmoto_amd.filters import distance
    own = rand_binary((26, 30, 22), p0, seed)
    BT = pmmoto_amd.BoundaryType
    box = tuple((0.0, float(v)) for v in own.shape) if unit else ((0, 1.0), (0, 1.0), (0, 1.0))
    sd = pmmoto_amd.initialize(own.shape, box=box, boundary_types=tuple((BT(a), BT(b)) for a, b in bts), pad=pad)
    # the padded image of this pad width: periodic copies, zero wall layers
    img = own
    for a in range(3):
        lo, hi = sd.pad[a]
        mode = "wrap" if bts[a][0] == "periodic" else "constant"
        img = np.pad(img, [(0, 0)] * a + [(lo, hi)] + [(0, 0)] * (2 - a), mode=mode)
    img = np.ascontiguousarray(img, dtype=np.uint8)
    assert img.shape == tuple(sd.voxels)
    periodic = [bts[a][0] == "periodic" for a in range(3)]
    pads = [int(p) for a in range(3) for p in sd.pad[a]]
    print(f"[edt subdomain] {bts} pads={pads} res={sd.domain.resolution}", flush=True)
    got = distance.edt(img, sd)
    want = np.sqrt(O.ref_corrected_edt_sq(img, periodic, pads, sd.domain.resolution))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_edt_padded_periodic_pack_unit_box():
    """A 96^3 periodic sphere pack in PMMoTo's unit box (resolution 1/96),
    padded: corrected_edt bitwise."""
    import pmmoto_amd
    from pmmoto_amd.filters import distance
    from tests.helpers import cpu_sphere_pack
    own, _ = cpu_sphere_pack((96, 96, 96), 0.5, "uniform", 44, periodic=(True, True, True), lo=4, hi=12)
    BT = pmmoto_amd.BoundaryType
    sd = pmmoto_amd.initialize((96, 96, 96), boundary_types=((BT.PERIODIC, BT.PERIODIC),) * 3)
    img = np.ascontiguousarray(np.pad(own, 1, mode="wrap"))
    print("[edt subdomain] 96^3 periodic pack, unit box", flush=True)
    got = distance.edt(img, sd)
    want = np.sqrt(O.ref_corrected_edt_sq(img, [True] * 3, [1] * 6, sd.domain.resolution))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
