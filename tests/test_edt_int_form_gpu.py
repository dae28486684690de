"""pmx_edt_ref's integer form (edt3d at one power-of-two resolution, extents
up to 1024): bit-identical float32 against the float kernels of the same
entry point (the integer form switched off) and, on small boxes, against the
oracle's restatement of the reference.  Each call's path is read back
(pmx_edt_ref_last_form), so a comparison is never between two runs of the
float kernels: dense images must take the integer form, and the cases it
does not cover (a byte above 1, a squared distance past 65534 or a plane
with no zero, a resolution that is not a power of two) must be handed to
the float kernels.  Lines along x with no zero (the reference's infinite
heights) stay in the integer form.  Bar: bit-exact float32 squared
distances and square roots."""

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rand_binary

pytestmark = pytest.mark.gpu


def ref_call(t, periodic, res, take_sqrt, int_form=True, pads=None):
    """pmx_edt_ref on a contiguous uint8 cuda tensor (pads: corrected_edt's
    padded-subdomain mode); (float32 tensor, form)."""
    import torch

    from pmmoto_amd import _device, _lib
    L = _lib.lib()
    shape = tuple(t.shape)
    nn = (C.c_int64 * 3)(*shape)
    st = (C.c_int64 * 2)(shape[1] * shape[2], shape[2])
    per = (C.c_int32 * 3)(*[1 if p else 0 for p in periodic])
    rr = (C.c_float * 3)(*[float(r) for r in res])
    nb = C.c_size_t()
    _lib.check(L.pmx_edt_ref_workspace_size(nn, C.byref(nb)))
    ws = _device.workspace(nb.value)
    out = torch.empty(shape, dtype=torch.float32, device="cuda")
    prev = L.pmx_edt_ref_int_form(1 if int_form else 0)
    try:
        pd = None if pads is None else (C.c_int64 * 6)(*pads)
        _lib.check(L.pmx_edt_ref(_device.ptr(t), nn, st, _device.ptr(out), st, 0, per, pd, rr,
                                 1 if take_sqrt else 0, _device.ptr(ws), nb.value, _device.stream()))
    finally:
        L.pmx_edt_ref_int_form(prev)
    form = C.c_int()
    _lib.check(L.pmx_edt_ref_last_form(nn, _device.ptr(ws), nb.value, _device.stream(), C.byref(form)))
    return out, form.value


def both_forms(img, periodic, res, take_sqrt=False, pads=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    a, fa = ref_call(t, periodic, res, take_sqrt, True, pads)
    b, fb = ref_call(t, periodic, res, take_sqrt, False, pads)
    assert fb == 2
    return a.cpu().numpy(), fa, b.cpu().numpy()


def pack(shape, porosity, lo, hi, seed, periodic):
    from pmmoto_amd.domain_generation import pack_radii, sphere_pack
    img, _ = sphere_pack(shape, porosity, pack_radii("uniform", lo=lo, hi=hi), seed=seed, periodic=periodic)
    return img.cpu().numpy()


PER = [(True, True, True), (False, True, True), (True, False, True), (True, True, False), (False, False, False),
       (False, True, False)]


@pytest.mark.parametrize("periodic", PER)
@pytest.mark.parametrize("res", [1.0, 2.0 ** -9, 0.5, 4.0])
def test_sphere_pack_forms_agree(periodic, res):
    img = pack((72, 60, 84), 0.5, 3, 9, 31, periodic)
    print(f"[int form] pack 72x60x84 periodic={periodic} res={res}", flush=True)
    for sq in (False, True):
        a, form, b = both_forms(img, periodic, (res,) * 3, take_sqrt=sq)
        assert form == 0, "a dense pack takes the integer form"
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("shape,p0,seed,periodic", [
    ((33, 17, 65), 0.3, 1, (True, True, True)),
    ((7, 130, 9), 0.25, 2, (True, True, True)),
    ((40, 3, 41), 0.5, 3, (True, True, True)),
    ((64, 64, 64), 0.08, 4, (True, True, True)),
    ((21, 50, 2), 0.2, 5, (True, False, True)),
    ((1, 40, 37), 0.3, 6, (True, True, True)),
    ((45, 1, 30), 0.3, 7, (False, True, True)),
    ((200, 24, 260), 0.05, 8, (True, True, True)),
])
def test_noise_forms_agree_and_match_oracle(shape, p0, seed, periodic):
    img = rand_binary(shape, p0, seed)
    print(f"[int form] noise {shape} p0={p0} periodic={periodic}", flush=True)
    res = (2.0 ** -9,) * 3
    a, form, b = both_forms(img, periodic, res)
    assert form == 0
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    if img.size <= 200_000:
        np.testing.assert_array_equal(a.view(np.uint32), O.ref_edt_sq(img, periodic, res).view(np.uint32))


def test_periodic_pack_512_forms_agree():
    """The 512^3 fully periodic sphere pack at resolution 1/512 (the unit
    box), squared and rooted: both paths, every voxel."""
    img = pack((512, 512, 512), 0.5, 4, 12, 2, (True, True, True))
    print("[int form] pack 512^3 periodic", flush=True)
    for sq in (False, True):
        a, form, b = both_forms(img, (True, True, True), (2.0 ** -9,) * 3, take_sqrt=sq)
        assert form == 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), \
            f"{int((a.view(np.uint32) != b.view(np.uint32)).sum())} voxels differ"


@pytest.mark.parametrize("case", ["label2", "inf_line", "inf_plane", "res_third", "far"])
def test_cases_at_the_edge_of_the_integer_form(case):
    """Lines along x with no zero and multi-label images keep the integer
    form; outside it the float kernels take over, on the device or by the
    arguments, and the result is still the reference's."""
    img = rand_binary((30, 26, 34), 0.3, 11)
    res = (2.0 ** -9,) * 3
    periodic = (True, True, True)
    want_form = 1
    if case == "label2":
        # a multi-label image takes the label-aware integer pass 0
        img[3, 4, 5] = 2
        want_form = 0
    elif case == "inf_line":
        img[:, 7, 9] = 1
        img[:, 20, 3] = 1
        want_form = 0
    elif case == "inf_plane":
        # no zero in the plane z = 9: its lines along y hold no finite height
        img[:, :, 9] = 1
    elif case == "res_third":
        res = (1.0 / 3.0,) * 3
        want_form = 2
    elif case == "far":
        # zeros only on the face x = 0 of a long box without wrap along x:
        # squared distances past 65534 leave the form
        img = np.ones((300, 20, 30), np.uint8)
        img[0] = 0
        periodic = (False, True, True)
    a, form, b = both_forms(img, periodic, res)
    assert form == want_form
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    if img.size <= 200_000:
        np.testing.assert_array_equal(a.view(np.uint32), O.ref_edt_sq(img, periodic, res).view(np.uint32))


@pytest.mark.parametrize("kind,own_shape,bts,pad,res", [
    ("pack", (64, 64, 64), ("periodic",) * 3, (1, 1, 1), 1 / 64),
    ("pack", (48, 40, 56), ("periodic", "wall", "periodic"), (2, 1, 3), 1.0),
    ("noise", (40, 33, 50), ("periodic", "periodic", "end"), (2, 2, 1), 0.5),
    ("noise", (26, 30, 22), ("wall", "periodic", "periodic"), (2, 1, 3), 1.0),
    ("noise", (30, 31, 29), ("periodic", "periodic", "periodic"), (3, 2, 1), 2.0 ** -5),
])
def test_padded_subdomain_forms_agree_and_match_oracle(kind, own_shape, bts, pad, res):
    """corrected_edt's padded-subdomain mode (periodic copies in the pads,
    zero wall layers): the integer form against the float kernels and the
    oracle's restatement of corrected_edt."""
    periodic = [b == "periodic" for b in bts]
    if kind == "pack":
        own = pack(own_shape, 0.5, 3, 9, 51, tuple(periodic))
    else:
        own = rand_binary(own_shape, 0.25, 52)
    img = own
    pads = []
    for a in range(3):
        p = pad[a]
        mode = "wrap" if periodic[a] else "constant"
        w = (p, p) if bts[a] != "end" else (0, 0)
        img = np.pad(img, [(0, 0)] * a + [w] + [(0, 0)] * (2 - a), mode=mode)
        pads += list(w)
    img = np.ascontiguousarray(img, dtype=np.uint8)
    print(f"[int form] padded {kind} {img.shape} {bts} pads={pads} res={res}", flush=True)
    a, form, b = both_forms(img, periodic, (res,) * 3, take_sqrt=True, pads=pads)
    assert form == 0
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    want = np.sqrt(O.ref_corrected_edt_sq(img, periodic, pads, (res,) * 3))
    np.testing.assert_array_equal(a.view(np.uint32), want.view(np.uint32))


def labelled(img, seed, n_labels):
    """The pore voxels of img split into labels 1 .. n_labels by random
    axis-aligned slabs (runs of one label with changes inside the pore
    space), solids kept at 0."""
    rng = np.random.default_rng(seed)
    lab = np.ones(img.shape, np.uint8)
    for a in range(3):
        cuts = np.sort(rng.integers(0, img.shape[a], size=3))
        idx = np.searchsorted(cuts, np.arange(img.shape[a]), side="right")
        sh = [1, 1, 1]
        sh[a] = img.shape[a]
        lab = ((lab.astype(np.int64) + idx.reshape(sh)) % n_labels + 1).astype(np.uint8)
    return np.where(img > 0, lab, 0).astype(np.uint8)


@pytest.mark.parametrize("periodic", [(True, True, True), (True, False, True), (False, False, False),
                                      (False, True, True)])
@pytest.mark.parametrize("kind", ["noise", "pack", "no_zero_lines"])
def test_multilabel_images_take_the_integer_form(periodic, kind):
    """Multi-label images (squared_edt_1d's reset at label changes in the
    first pass) in the integer form: bit-identical to the float kernels and
    to the oracle, squared and rooted, at a power-of-two and at unit
    resolution."""
    if kind == "noise":
        rng = np.random.default_rng(21)
        img = rng.choice([0, 1, 2, 3], size=(41, 33, 50), p=[0.1, 0.5, 0.2, 0.2]).astype(np.uint8)
    elif kind == "pack":
        img = labelled(pack((48, 40, 56), 0.5, 3, 9, 61, periodic), 62, 3)
    else:
        # lines along x with no zero but label changes, and lines with neither
        rng = np.random.default_rng(23)
        img = rng.choice([0, 1, 2], size=(30, 26, 34), p=[0.05, 0.6, 0.35]).astype(np.uint8)
        img[:, 5, 7] = 2
        img[:, 11, 3] = 1
        img[:15, 11, 3] = 3
    print(f"[int form] multi-label {kind} {img.shape} periodic={periodic}", flush=True)
    for res in (2.0 ** -9, 1.0):
        for sq in (False, True):
            a, form, b = both_forms(img, periodic, (res,) * 3, take_sqrt=sq)
            assert form == 0, "a multi-label image takes the integer form"
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
            want = O.ref_edt_sq(img, periodic, (res,) * 3)
            if sq:
                want = np.sqrt(want)
            np.testing.assert_array_equal(a.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("bts,pad,res", [(("periodic",) * 3, (2, 1, 1), 1.0 / 32),
                                          (("periodic", "wall", "periodic"), (2, 1, 3), 1.0),
                                          (("wall", "periodic", "end"), (1, 2, 1), 0.5)])
def test_multilabel_padded_subdomain_takes_the_integer_form(bts, pad, res):
    """corrected_edt's padded mode with labels (its first pass runs over the
    whole padded line): the integer form, equal to the float kernels and the
    oracle."""
    periodic = [b == "periodic" for b in bts]
    own = labelled(pack((30, 31, 29), 0.5, 3, 8, 72, tuple(periodic)), 73, 2)
    img = own
    pads = []
    for a in range(3):
        w = (pad[a], pad[a]) if bts[a] != "end" else (0, 0)
        img = np.pad(img, [(0, 0)] * a + [w] + [(0, 0)] * (2 - a), mode="wrap" if periodic[a] else "constant")
        pads += list(w)
    img = np.ascontiguousarray(img, dtype=np.uint8)
    print(f"[int form] multi-label padded {img.shape} {bts} pads={pads}", flush=True)
    a, form, b = both_forms(img, periodic, (res,) * 3, take_sqrt=True, pads=pads)
    assert form == 0
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    want = np.sqrt(O.ref_corrected_edt_sq(img, periodic, pads, (res,) * 3))
    np.testing.assert_array_equal(a.view(np.uint32), want.view(np.uint32))
