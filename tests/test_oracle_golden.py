"""Pins the oracle's periodic EDT drivers to vectors written by the
reference itself (tests/golden/make_ref_edt_vectors.py runs the reference's
Cython drivers single-rank in the development container):

* edt3d with the five-member periodic boundary hulls
  (euclidean_distance_transform.py:208-303), on a 128^3 periodic sphere
  pack (digest and the voxels where it exceeds the exact torus EDT) and on
  small images at non-unit and anisotropic resolution;
* corrected_edt on padded single-rank subdomains
  (euclidean_distance_transform.py:56-101): periodic, periodic/wall/end and
  end/periodic faces of the unit box (resolution 1/N), and a 64^3 periodic
  pack at unit resolution.

Both with the oracle's restated per-line kernels and with the reference's
own kernels compiled into oracle/_ref (when built).
"""

import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import cpu_sphere_pack

GOLD = os.path.join(os.path.dirname(__file__), "golden")
VEC = np.load(os.path.join(GOLD, "ref_edt_vectors.npz"))
SD_CASES = ["sd_periodic", "sd_mixed", "sd_pack", "sd_per_end", "sd_pack_unit"]
EDT3D_CASES = {
    "edt3d_per_res": ((True, True, True), (0.1, 0.1, 0.1)),
    "edt3d_mixed_aniso": ((False, True, True), (0.05, 1.0 / 16, 1.0 / 24)),
    "edt3d_pack_res": ((True, True, True), (1 / 40, 1 / 40, 1 / 40)),
}
KERNELS = [False, pytest.param(True, marks=pytest.mark.skipif(O.ref_lib() is None, reason="oracle/_ref not built"))]


def sd_case(key):
    meta = VEC[key + "_meta"]
    periodic = [bool(meta[0][2 * a] == 2) for a in range(3)]
    return VEC[key + "_in"], periodic, [int(p) for p in meta[1]], tuple(float(r) for r in VEC[key + "_res"])


@pytest.mark.parametrize("use_ref", KERNELS)
def test_edt3d_periodic_pack_128_matches_reference(use_ref):
    with open(os.path.join(GOLD, "ref_edt_periodic_128.json")) as f:
        doc = json.load(f)
    img, _ = cpu_sphere_pack((128, 128, 128), 0.5, "uniform", 2, periodic=(True, True, True), lo=4, hi=12)
    assert hashlib.sha256(img.tobytes()).hexdigest() == doc["image_sha256"]
    got = O.ref_edt_sq(img, (True, True, True), use_ref=use_ref)
    assert hashlib.sha256(got.tobytes()).hexdigest() == doc["output_sha256"]
    # the reference exceeds the exact torus EDT exactly where the fixture says
    exact = O.d2_to_f32(O.exact_edt_sq(img, [O.BC_PERIODIC] * 6)).ravel()
    diff = np.flatnonzero(got.ravel() != exact)
    assert diff.tolist() == [d[0] for d in doc["exact_differs_at"]]
    assert np.all(got.ravel()[diff] > exact[diff])
    assert len(diff) == 47


@pytest.mark.parametrize("use_ref", KERNELS)
@pytest.mark.parametrize("key", list(EDT3D_CASES))
def test_edt3d_vectors_match_reference(key, use_ref):
    per, res = EDT3D_CASES[key]
    got = O.ref_edt_sq(VEC[key + "_in"], per, res, use_ref=use_ref)
    np.testing.assert_array_equal(got.view(np.uint32), VEC[key + "_out"].view(np.uint32))


@pytest.mark.parametrize("use_ref", KERNELS)
@pytest.mark.parametrize("key", SD_CASES)
def test_corrected_edt_padded_subdomain_matches_reference(key, use_ref):
    img, periodic, pads, res = sd_case(key)
    got = np.sqrt(O.ref_corrected_edt_sq(img, periodic, pads, res, use_ref=use_ref))
    np.testing.assert_array_equal(got.view(np.uint32), VEC[key + "_out"].view(np.uint32))


def test_own_box_restatement_is_not_the_reference_on_packs():
    """The reference computes a padded subdomain's passes on the padded
    array, so the own box with wrap (edt3d semantics) is not a substitute:
    on sphere packs it differs in the last ulp at non-unit resolution and,
    through the hull truncation, at unit resolution (there in a pad voxel,
    which the reference computes rather than copies)."""
    for key in ("sd_pack", "sd_pack_unit"):
        img, periodic, pads, res = sd_case(key)
        sl = tuple(slice(pads[2 * a], img.shape[a] - pads[2 * a + 1]) for a in range(3))
        own = np.sqrt(O.ref_edt_sq(np.ascontiguousarray(img[sl]), tuple(periodic), res))
        full = np.pad(own, [(pads[2 * a], pads[2 * a + 1]) for a in range(3)], mode="wrap")
        assert np.any(full.view(np.uint32) != VEC[key + "_out"].view(np.uint32))


def _label_cases():
    with open(os.path.join(GOLD, "ref_inlet_labels.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("case", _label_cases(), ids=lambda c: c["name"])
def test_inlet_outlet_label_maps_match_reference(case):
    """gen_inlet_label_map / gen_outlet_label_map with several flagged faces
    keep only the last flagged face's labels: the oracle's restatement and
    the product's face selection (on the host) against the reference's own
    output."""
    import pmmoto_amd
    from pmmoto_amd.filters.connected_components import selected_face
    BT = pmmoto_amd.BoundaryType
    bts = tuple((BT(a), BT(b)) for a, b in case["boundary_types"])
    inlet = tuple(tuple(p) for p in case["inlet"])
    outlet = tuple(tuple(p) for p in case["outlet"])
    sd = pmmoto_amd.initialize(tuple(case["voxels"]), boundary_types=bts, inlet=inlet, outlet=outlet,
                               reservoir_voxels=case["reservoir_voxels"])
    assert list(sd.voxels) == case["padded_shape"]
    assert [list(p) for p in sd.pad] == case["pad"]
    lab = np.random.default_rng(case["label_seed"]).integers(0, case["label_high"], size=tuple(sd.voxels))
    code = {"end": O.BC_END, "wall": O.BC_WALL, "periodic": O.BC_PERIODIC}
    box = O.Box(bc=tuple((code[a], code[b]) for a, b in case["boundary_types"]), pad=tuple(sd.pad),
                inlet=sd.inlet, outlet=sd.outlet, reservoir=tuple(sd.reservoir_pad))
    for which, key in (("inlet", "inlet_labels"), ("outlet", "outlet_labels")):
        want = set(case[key]) - {0}
        assert O.inlet_labels(box, lab, which) == want
        face = selected_face(sd, which)
        got = set(np.unique(lab[sd.face_own_slab(*face)]).tolist()) - {0}
        assert got == want
