"""Writes golden vectors of the reference's own periodic EDT drivers.

The float32 per-line kernels of the reference (_distance.hpp, _voxels.hpp)
are compiled into oracle/_ref and pin the oracle line by line.  The Python
drivers around them -- the periodic boundary-hull exchange of ``edt3d``
(src/pmmoto/filters/distance/euclidean_distance_transform.py:208-303) and
``corrected_edt`` on a padded subdomain (:56-101, with the correctors and
hulls of :306-528) -- are pinned here by running the reference itself, in
this development container only:

* ``_distance.pyx`` and ``_voxels.pyx`` are cythonized and compiled from
  the reference tree into tests/golden/_build/ (never into the reference
  tree, never committed, never shipped);
* ``pmmoto`` and ``pmmoto.filters(.distance)`` are empty namespace modules
  whose search paths point into the reference tree, so only ``pmmoto.core``
  (domains, subdomains, features, communication) and the EDT module run;
* ``mpi4py`` is replaced by a SINGLE-RANK STAND-IN written into the scratch
  directory (rank 0 of 1, collectives return their argument, point-to-point
  calls raise).  One rank with periodic faces exchanges features with itself
  inside ``communicate_features`` without MPI.

Outputs (committed, data only):

* tests/golden/ref_edt_periodic_128.json -- the squared float32 EDT of a
  128^3 periodic overlapping-sphere pack (radii U(4,12), porosity 0.5,
  seed 2), edt3d(periodic=(True,)*3, squared=True): SHA-256 of the input and
  of the output bytes, and every voxel where it differs from the exact torus
  EDT (flat index, reference value, exact value).
* tests/golden/ref_edt_vectors.npz -- small whole-array cases: periodic and
  mixed-face edt3d at non-unit resolution, and ``edt(img, subdomain)`` on
  padded periodic / mixed subdomains of the unit box (resolution 1/N) and
  on a 64^3 periodic pack at unit resolution.
* tests/golden/ref_inlet_labels.json -- the labels gen_inlet_label_map and
  gen_outlet_label_map (core/voxels.py:400-454) return on subdomains with
  several inlet / outlet faces, for seeded random label images.

Run:  python tests/golden/make_ref_edt_vectors.py   (REF_ROOT overrides
/root/reference)
"""

import hashlib
import importlib.abc
import importlib.machinery
import importlib.util
import json
import os
import sys
import types

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("REF_ROOT", "/root/reference")
SRC = os.path.join(REF, "src", "pmmoto")
SCRATCH = os.path.join(HERE, "_build", "refpy")

sys.dont_write_bytecode = True  # nothing is written next to the reference sources

FAKE_MPI = '''"""SINGLE-RANK STAND-IN for mpi4py (a fake backend for golden-vector runs):
rank 0 of 1; collectives return their argument; point-to-point raises."""


class _Op:
    def __init__(self, name):
        self.name = name


SUM, PROD, MAX, MIN = _Op("sum"), _Op("prod"), _Op("max"), _Op("min")
LAND, BAND, LOR, BOR = _Op("land"), _Op("band"), _Op("lor"), _Op("bor")
LXOR, BXOR, MINLOC, MAXLOC = _Op("lxor"), _Op("bxor"), _Op("minloc"), _Op("maxloc")
Op = _Op


class Intracomm:
    def Get_rank(self):
        return 0

    def Get_size(self):
        return 1

    def allgather(self, data):
        return [data]

    def gather(self, data, root=0):
        return [data]

    def allreduce(self, data, op=None):
        return data

    def bcast(self, data, root=0):
        return data

    def Barrier(self):
        return None

    barrier = Barrier

    def isend(self, *a, **k):
        raise RuntimeError("single-rank stand-in: no peer ranks")

    def recv(self, *a, **k):
        raise RuntimeError("single-rank stand-in: no peer ranks")


class Request:
    @staticmethod
    def waitall(requests):
        if requests:
            raise RuntimeError("single-rank stand-in: no peer ranks")


COMM_WORLD = Intracomm()
'''


def build_extensions():
    """Cythonize and compile the two reference extensions into SCRATCH."""
    import numpy
    from Cython.Build import cythonize
    from setuptools import Extension
    from setuptools.dist import Distribution

    lib_dir = os.path.join(SCRATCH, "lib")
    want = {"pmmoto.core._voxels": os.path.join(SRC, "core", "_voxels.pyx"),
            "pmmoto.filters.distance._distance": os.path.join(SRC, "filters", "distance", "_distance.pyx")}
    inc = {"pmmoto.core._voxels": [os.path.join(SRC, "core"), numpy.get_include()],
           "pmmoto.filters.distance._distance": [os.path.join(SRC, "filters", "distance"), numpy.get_include()]}
    exts = [Extension(name, [src], include_dirs=inc[name], language="c++",
                      extra_compile_args=["-std=c++17", "-O3", "-pthread"]) for name, src in want.items()]
    exts = cythonize(exts, build_dir=os.path.join(SCRATCH, "cython"), quiet=True,
                     compiler_directives={"language_level": "3"})
    dist = Distribution({"ext_modules": exts, "script_name": "make_ref_edt_vectors"})
    cmd = dist.get_command_obj("build_ext")
    cmd.build_lib = lib_dir
    cmd.build_temp = os.path.join(SCRATCH, "tmp")
    cmd.inplace = False
    cmd.ensure_finalized()
    cmd.run()
    found = {}
    for name in want:
        rel = os.path.join(lib_dir, *name.split("."))
        d, base = os.path.dirname(rel), os.path.basename(rel)
        for f in os.listdir(d):
            if f.startswith(base + ".") and f.endswith(".so"):
                found[name] = os.path.join(d, f)
    assert set(found) == set(want), found
    return found


class _ExtFinder(importlib.abc.MetaPathFinder):
    def __init__(self, paths):
        self.paths = paths

    def find_spec(self, fullname, path=None, target=None):
        if fullname in self.paths:
            loader = importlib.machinery.ExtensionFileLoader(fullname, self.paths[fullname])
            return importlib.util.spec_from_file_location(fullname, self.paths[fullname], loader=loader)
        return None


def load_reference():
    """Returns (pmmoto.core package, euclidean_distance_transform module)."""
    exts = build_extensions()
    shim = os.path.join(SCRATCH, "shim", "mpi4py")
    os.makedirs(shim, exist_ok=True)
    with open(os.path.join(shim, "__init__.py"), "w") as f:
        f.write('"""SINGLE-RANK STAND-IN package (fake backend)."""\nfrom . import MPI  # noqa: F401\n')
    with open(os.path.join(shim, "MPI.py"), "w") as f:
        f.write(FAKE_MPI)
    sys.path.insert(0, os.path.dirname(shim))
    sys.meta_path.insert(0, _ExtFinder(exts))
    for name, path in (("pmmoto", SRC), ("pmmoto.filters", os.path.join(SRC, "filters")),
                       ("pmmoto.filters.distance", os.path.join(SRC, "filters", "distance"))):
        m = types.ModuleType(name)
        m.__path__ = [path]
        sys.modules[name] = m
    os.makedirs(os.path.join(SCRATCH, "run"), exist_ok=True)
    os.chdir(os.path.join(SCRATCH, "run"))  # the reference logger writes logs/ into the cwd
    import pmmoto.core
    import pmmoto.core.pmmoto
    import pmmoto.filters.distance.euclidean_distance_transform as E
    return pmmoto.core, E


def main():
    import numpy as np

    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    from tests.helpers import cpu_sphere_pack, rand_binary

    core, E = load_reference()
    BT = core.boundary_types.BoundaryType
    initialize = core.pmmoto.initialize

    # 128^3 periodic pack at unit resolution: digests and the truncation diff
    img, spheres = cpu_sphere_pack((128, 128, 128), 0.5, "uniform", 2, periodic=(True, True, True), lo=4, hi=12)
    ref = np.ascontiguousarray(E.edt3d(img, periodic=(True, True, True), squared=True), dtype=np.float32)
    exact = O.d2_to_f32(O.exact_edt_sq(img, [O.BC_PERIODIC] * 6))
    diff = np.flatnonzero(ref.view(np.uint32).ravel() != exact.view(np.uint32).ravel())
    doc = {
        "workload": "edt3d(img, periodic=(True, True, True), squared=True) of a 128^3 periodic "
                    "overlapping-sphere pack (radii U(4,12), porosity 0.5, seed 2)",
        "shape": [128, 128, 128],
        "spheres": int(spheres.shape[0]),
        "image_sha256": hashlib.sha256(img.tobytes()).hexdigest(),
        "output_sha256": hashlib.sha256(ref.tobytes()).hexdigest(),
        "exact_differs_at": [[int(i), float(ref.ravel()[i]), float(exact.ravel()[i])] for i in diff],
    }
    with open(os.path.join(HERE, "ref_edt_periodic_128.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print(f"128^3 periodic pack: {diff.size} voxels differ from the exact torus EDT", flush=True)

    arrays = {}

    # edt3d at non-unit resolution: periodic and mixed faces
    img = rand_binary((18, 22, 26), 0.06, 7)
    arrays["edt3d_per_res_in"] = img
    arrays["edt3d_per_res_out"] = np.asarray(E.edt3d(img, periodic=(True, True, True), resolution=(0.1, 0.1, 0.1),
                                                     squared=True), dtype=np.float32)
    img = rand_binary((20, 16, 24), 0.04, 8)
    arrays["edt3d_mixed_aniso_in"] = img
    arrays["edt3d_mixed_aniso_out"] = np.asarray(E.edt3d(img, periodic=(False, True, True),
                                                         resolution=(0.05, 1.0 / 16, 1.0 / 24),
                                                         squared=True), dtype=np.float32)
    pack, _ = cpu_sphere_pack((40, 40, 40), 0.5, "uniform", 21, periodic=(True, True, True), lo=3, hi=7)
    arrays["edt3d_pack_res_in"] = pack
    arrays["edt3d_pack_res_out"] = np.asarray(E.edt3d(pack, periodic=(True, True, True),
                                                      resolution=(1 / 40, 1 / 40, 1 / 40), squared=True),
                                              dtype=np.float32)

    # edt(img, subdomain): padded single-rank subdomains of the unit box, and
    # one at unit resolution
    pack64, _ = cpu_sphere_pack((64, 64, 64), 0.5, "uniform", 23, periodic=(True, True, True), lo=4, hi=12)
    cases = {
        "sd_periodic": ((20, 24, 28), ((BT.PERIODIC, BT.PERIODIC),) * 3, 0.08, 9, None),
        "sd_mixed": ((22, 18, 26), ((BT.PERIODIC, BT.PERIODIC), (BT.WALL, BT.WALL), (BT.END, BT.END)), 0.06, 10,
                     None),
        "sd_pack": ((40, 40, 40), ((BT.PERIODIC, BT.PERIODIC),) * 3, 0.0, 22, pack),
        "sd_per_end": ((24, 20, 22), ((BT.END, BT.END), (BT.PERIODIC, BT.PERIODIC), (BT.PERIODIC, BT.PERIODIC)),
                       0.05, 11, None),
        "sd_pack_unit": ((64, 64, 64), ((BT.PERIODIC, BT.PERIODIC),) * 3, 0.0, 23, pack64),
    }
    for key, (n, bts, p0, seed, own) in cases.items():
        # the unit-resolution case spans box (0, n) per axis; the others the unit box
        box = tuple((0.0, float(v)) for v in n) if key.endswith("_unit") else ((0, 1.0), (0, 1.0), (0, 1))
        sd = initialize(n, box=box, boundary_types=bts)
        if own is None:
            own = rand_binary(n, p0, seed)
        padded = own
        for a in range(3):
            lo, hi = sd.pad[a]
            if bts[a][0] == BT.PERIODIC:
                padded = np.pad(padded, [(0, 0)] * a + [(lo, hi)] + [(0, 0)] * (2 - a), mode="wrap")
            else:
                padded = np.pad(padded, [(0, 0)] * a + [(lo, hi)] + [(0, 0)] * (2 - a), constant_values=0)
        padded = np.ascontiguousarray(padded, dtype=np.uint8)
        assert padded.shape == tuple(sd.voxels), (padded.shape, sd.voxels)
        arrays[f"{key}_own"] = own
        arrays[f"{key}_in"] = padded
        arrays[f"{key}_out"] = np.asarray(E.edt(padded, sd), dtype=np.float32)
        # row 0: face codes (0 END, 1 WALL, 2 PERIODIC) per face; row 1: pad widths
        arrays[f"{key}_meta"] = np.array(
            [[{"end": 0, "wall": 1, "periodic": 2}[b.value] for pair in bts for b in pair],
             [int(p) for pair in sd.pad for p in pair]], dtype=np.int64)
        arrays[f"{key}_res"] = np.asarray(sd.domain.resolution, dtype=np.float64)
        print(f"{key}: padded {padded.shape}, resolution {tuple(sd.domain.resolution)}", flush=True)
    np.savez_compressed(os.path.join(HERE, "ref_edt_vectors.npz"), **arrays)
    print("wrote ref_edt_periodic_128.json and ref_edt_vectors.npz", flush=True)
    inlet_label_vectors(core)


def inlet_label_vectors(core):
    """gen_inlet_label_map / gen_outlet_label_map (core/voxels.py:400-454)
    on subdomains with several inlet or outlet faces: the labels each
    returns for a seeded random label image of the padded shape."""
    import numpy as np

    BT = core.boundary_types.BoundaryType
    E_, W_, P_ = BT.END, BT.WALL, BT.PERIODIC
    F, T = False, True
    cases = [
        ("end_x_y", (12, 10, 14), ((E_, E_),) * 3, ((T, F), (T, F), (F, F)), ((F, T), (F, F), (F, T)), 0),
        ("end_xlo_xhi_zhi", (10, 12, 9), ((E_, E_),) * 3, ((T, T), (F, F), (F, T)), ((F, F), (T, T), (F, F)), 0),
        ("wall_y", (11, 9, 13), ((E_, E_), (W_, W_), (E_, E_)), ((T, F), (F, F), (F, T)), ((F, T), (F, F), (T, F)), 0),
        ("periodic_z_reservoir", (10, 11, 12), ((E_, E_), (E_, E_), (P_, P_)), ((T, F), (F, T), (F, F)),
         ((F, T), (T, F), (F, F)), 3),
    ]
    out = []
    for k, (name, n, bts, inlet, outlet, res_vox) in enumerate(cases):
        sd = core.pmmoto.initialize(n, boundary_types=bts, inlet=inlet, outlet=outlet, reservoir_voxels=res_vox)
        rng = np.random.default_rng(500 + k)
        lab = rng.integers(0, 2000, size=tuple(int(v) for v in sd.voxels)).astype(np.uint64)
        got_in = core.voxels.gen_inlet_label_map(sd, lab)
        got_out = core.voxels.gen_outlet_label_map(sd, lab)
        out.append({
            "name": name, "voxels": list(n), "boundary_types": [[b.value for b in pair] for pair in bts],
            "inlet": [list(p) for p in inlet], "outlet": [list(p) for p in outlet], "reservoir_voxels": res_vox,
            "padded_shape": [int(v) for v in sd.voxels], "pad": [[int(p) for p in pair] for pair in sd.pad],
            "label_seed": 500 + k, "label_high": 2000,
            "inlet_labels": sorted(int(v) for v in got_in), "outlet_labels": sorted(int(v) for v in got_out),
        })
        print(f"{name}: padded {tuple(sd.voxels)}, {len(got_in)} inlet / {len(got_out)} outlet labels", flush=True)
    with open(os.path.join(HERE, "ref_inlet_labels.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote ref_inlet_labels.json", flush=True)


if __name__ == "__main__":
    main()
