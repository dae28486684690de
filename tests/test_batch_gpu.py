"""Batched sweeps through BatchSweep (the fused batch, pmx_psd_batch_*, and
the per-domain sweeps on streams, pmx_psd_sweep_batch): every domain's codes
equal the single-domain sweep's (pmx_psd_sweep) and its full EDT equals
pmx_edt's on the same input, on every replay of the captured graph, and the
oracle's pore-size distribution on small domains.  Bar: bit-exact int32
codes and squared distances."""

import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import E, P, W, padded_case

pytestmark = pytest.mark.gpu

NAMES = {"end": E, "wall": W, "periodic": P}


def single_sweep(img_t, bc, radii, inlet):
    """The plain path: one pmx_psd_sweep call on the same own region."""
    import torch

    from pmmoto_amd import _device, _lib
    from pmmoto_amd.filters import morphological_operators as morph
    from pmmoto_amd.filters.batch import _geom
    shape = tuple(img_t.shape)
    g = _geom(shape, bc)
    t = [morph.radius_threshold(np.float64(r)) for r in radii]
    out = torch.empty(shape, dtype=torch.int32, device="cuda")
    L = _lib.lib()
    nb = C.c_size_t()
    _lib.check(L.pmx_psd_workspace_size(C.byref(g), C.byref(nb)))
    ws = _device.workspace(nb.value)
    lo = ext = None
    if inlet is not None:
        a, s = inlet
        lo_v, ext_v = [0, 0, 0], list(shape)
        lo_v[a], ext_v[a] = (0 if s == 0 else shape[a] - 1), 1
        lo, ext = (C.c_int64 * 3)(*lo_v), (C.c_int64 * 3)(*ext_v)
    _lib.check(L.pmx_psd_sweep(_device.ptr(img_t), C.byref(g), lo, ext, len(t), (C.c_int64 * len(t))(*t),
                               (C.c_int64 * len(t))(*t), 0, _device.ptr(out), _device.ptr(ws), nb.value,
                               _device.stream()))
    return out


def make_domains():
    from pmmoto_amd.domain_generation import pack_radii, sphere_pack
    from pmmoto_amd.filters.batch import BatchDomain
    specs = [((40, 33, 50), ("wall", "wall", "periodic"), None, 0.45),
             ((64, 20, 30), ("periodic", "periodic", "periodic"), None, 0.55),
             ((25, 60, 41), ("end", "wall", "wall"), (0, 0), 0.40),
             ((30, 30, 30), ("wall", "wall", "wall"), None, 0.30),
             ((17, 45, 23), ("periodic", "wall", "end"), (2, 1), 0.50),
             ((70, 16, 16), ("end", "periodic", "periodic"), (0, 1), 0.60)]
    out = []
    for i, (shape, bc, inlet, por) in enumerate(specs):
        img, _ = sphere_pack(shape, por, pack_radii("uniform", lo=2, hi=6), seed=700 + i,
                             periodic=tuple(b == "periodic" for b in bc))
        out.append(BatchDomain(img, bc, inlet))
    return out


def test_batch_equals_single_domain_sweeps_on_every_replay():
    import torch

    from pmmoto_amd.filters.batch import BatchSweep
    radii = list(range(10, 0, -1))
    doms = make_domains()
    want = [single_sweep(d.img, d.bc, radii, d.inlet).cpu().numpy() for d in doms]
    sweep = BatchSweep(doms, radii, lanes=3, autotune=True, graph=True)
    assert all(t in (32, 64) for t in sweep.tile_rows())
    for rep in range(3):
        print(f"[batch] replay {rep}", flush=True)
        maps = sweep.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            np.testing.assert_array_equal(m.cpu().numpy(), want[i], err_msg=f"domain {i} replay {rep}")
            m.fill_(-7)  # eager work between replays; the next replay must rewrite every code
        torch.cuda.synchronize()
    eager = BatchSweep(doms, radii, lanes=2, autotune=False, graph=False).run()
    for i, m in enumerate(eager):
        np.testing.assert_array_equal(m.cpu().numpy(), want[i], err_msg=f"domain {i} eager")


def test_batch_matches_oracle_pore_size_distribution():
    import torch

    from pmmoto_amd.filters.batch import BatchDomain, BatchSweep
    from tests.helpers import cpu_sphere_pack
    radii = np.arange(5, 0, -1).astype(np.float64)
    cases = [((20, 22, 24), ("wall", "wall", "wall"), None),
             ((21, 18, 26), ("periodic", "periodic", "periodic"), None),
             ((24, 20, 22), ("end", "wall", "periodic"), (0, 0))]
    doms, boxes = [], []
    for i, (shape, bc, inlet) in enumerate(cases):
        own, _ = cpu_sphere_pack(shape, 0.45, "uniform", 900 + i, periodic=tuple(b == "periodic" for b in bc),
                                 lo=2, hi=5)
        bts = tuple((NAMES[b], NAMES[b]) for b in bc)
        inl = ((False, False),) * 3 if inlet is None else tuple(
            (a == inlet[0] and inlet[1] == 0, a == inlet[0] and inlet[1] == 1) for a in range(3))
        padded, sd, box = padded_case(own, bts, inlet=inl)
        doms.append(BatchDomain(torch.from_numpy(own).cuda(), bc, inlet))
        boxes.append((padded, box, bts, inlet is not None))
    maps = BatchSweep(doms, radii, lanes=2).run()
    torch.cuda.synchronize()
    for i, (padded, box, bts, has_inlet) in enumerate(boxes):
        want = O.pore_size_distribution(box, padded, radii, inlet=has_inlet)
        sl = tuple(slice(p[0], padded.shape[a] - p[1]) for a, p in enumerate(box.pad))
        codes = maps[i].cpu().numpy()
        got = np.where(codes > 0, np.concatenate([[0.0], radii])[codes], 0.0)
        np.testing.assert_array_equal(got, want[sl], err_msg=f"case {i} {bts}")


def test_autotune_choice_is_kept_per_shape():
    from pmmoto_amd.filters.batch import BatchSweep
    doms = make_domains()[:2]
    a = BatchSweep(doms, [4.0, 2.0, 1.0], lanes=1).tile_rows()
    b = BatchSweep(doms, [4.0, 2.0, 1.0], lanes=1).tile_rows()
    assert a == b and all(t in (32, 64) for t in a)


def single_edt(img_t, bc):
    """The plain path: one pmx_edt call (int32 squared distances)."""
    import torch

    from pmmoto_amd import _device, _lib
    from pmmoto_amd.filters.batch import _geom
    shape = tuple(img_t.shape)
    g = _geom(shape, bc)
    L = _lib.lib()
    nb = C.c_size_t()
    _lib.check(L.pmx_edt_workspace_size(C.byref(g), C.byref(nb)))
    ws = _device.workspace(nb.value)
    out = torch.empty(shape, dtype=torch.int32, device="cuda")
    _lib.check(L.pmx_edt(_device.ptr(img_t), C.byref(g), _device.ptr(out), None, None, None, _device.ptr(ws),
                         nb.value, _device.stream()))
    return out


def fused_domains():
    """No inlet requirements: short rows (16 and 32 lanes per row), a long-row
    domain (64 lanes, nz > 512), one whose distances exceed 255 voxels (the
    int32 passes) and one whose lines defer to the envelope fallback."""
    import torch

    from pmmoto_amd.domain_generation import pack_radii, sphere_pack
    from pmmoto_amd.filters.batch import BatchDomain
    specs = [((40, 33, 50), ("wall", "wall", "periodic"), 0.45),
             ((64, 20, 30), ("periodic", "periodic", "periodic"), 0.55),
             ((30, 30, 300), ("wall", "wall", "wall"), 0.30),
             ((9, 7, 600), ("end", "periodic", "end"), 0.50),
             ((21, 45, 23), ("periodic", "wall", "end"), 0.50)]
    out = []
    for i, (shape, bc, por) in enumerate(specs):
        img, _ = sphere_pack(shape, por, pack_radii("uniform", lo=2, hi=6), seed=800 + i,
                             periodic=tuple(b == "periodic" for b in bc))
        out.append(BatchDomain(img, bc, None))
    # all pore but two solid voxels near one end: distances above 255 and lines past the
    # 128-row scan
    sparse = torch.ones((600, 6, 9), dtype=torch.uint8, device="cuda")
    sparse[3, 2, 4] = 0
    sparse[10, 5, 0] = 0
    out.append(BatchDomain(sparse, ("end", "wall", "periodic"), None))
    return out


def test_fused_batch_equals_single_domain_paths_on_every_replay(tmp_path):
    import torch

    from pmmoto_amd.filters.batch import BatchSweep
    radii = list(range(10, 0, -1))
    doms = fused_domains()
    want = [single_sweep(d.img, d.bc, radii, None).cpu().numpy() for d in doms]
    want_edt = [single_edt(d.img, d.bc).cpu().numpy() for d in doms]
    cache = tmp_path / "tiles.txt"
    sweep = BatchSweep(doms, radii, autotune=True, graph=True, edt=True, tune_cache=str(cache))
    assert sweep.fused and cache.exists()
    for rep in range(3):
        print(f"[fused batch] replay {rep}", flush=True)
        maps = sweep.run()
        torch.cuda.synchronize()
        for i, m in enumerate(maps):
            np.testing.assert_array_equal(m.cpu().numpy(), want[i], err_msg=f"domain {i} replay {rep}")
            np.testing.assert_array_equal(sweep.edts[i].cpu().numpy(), want_edt[i], err_msg=f"edt {i} replay {rep}")
            m.fill_(-7)  # eager work between replays: the next replay must rewrite every output
            sweep.edts[i].fill_(-7)
        torch.cuda.synchronize()
    eager = BatchSweep(doms, radii, autotune=False, graph=False)
    for i, m in enumerate(eager.run()):
        np.testing.assert_array_equal(m.cpu().numpy(), want[i], err_msg=f"domain {i} eager")


def test_fused_batch_rejects_inlets_and_large_thresholds():
    import torch

    from pmmoto_amd.filters.batch import BatchDomain, BatchSweep
    img = torch.ones((20, 20, 20), dtype=torch.uint8, device="cuda")
    with pytest.raises(ValueError, match="fused batch"):
        BatchSweep([BatchDomain(img, ("end", "wall", "wall"), (0, 0))], [3.0, 1.0], autotune=False, edt=True)
    # an inlet domain runs the per-domain sweeps instead
    assert not BatchSweep([BatchDomain(img, ("end", "wall", "wall"), (0, 0))], [3.0, 1.0], autotune=False).fused
    with pytest.raises(ValueError, match="below"):
        BatchSweep([BatchDomain(img)], [200.0, 1.0], autotune=False)
