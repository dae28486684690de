"""CPU checks of the drop-in boundary: libpmx.so loads without a GPU and
exports every symbol include/pmx.h declares; the Python surface mirrors the
reference's names and argument errors."""

import os
import re

import numpy as np
import pytest

from pmmoto_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "pmx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pmx_\w+)\s*\(", src, flags=re.M)))


def test_header_lists_symbols():
    syms = declared_symbols()
    assert "pmx_edt" in syms and "pmx_morph" in syms and len(syms) >= 6


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_version_and_error_string():
    L = _lib.lib()
    assert L.pmx_version() >= 100
    assert isinstance(L.pmx_last_error(), bytes)


def test_invalid_geometry_is_rejected_without_gpu():
    import ctypes as C
    L = _lib.lib()
    g = _lib.Geom()
    g.n[0], g.n[1], g.n[2] = 0, 4, 4
    nbytes = C.c_size_t()
    rc = L.pmx_edt_workspace_size(C.byref(g), C.byref(nbytes))
    assert rc == _lib.ERR_INVALID
    assert b"extent" in L.pmx_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc)
    g.n[0] = 4
    g.in_s[0], g.in_s[1], g.out_s[0], g.out_s[1] = 16, 4, 16, 4
    g.bc[0], g.bc[1] = _lib.BC_PERIODIC, _lib.BC_END
    assert L.pmx_edt_workspace_size(C.byref(g), C.byref(nbytes)) == _lib.ERR_INVALID
    g.bc[1] = _lib.BC_PERIODIC
    assert L.pmx_edt_workspace_size(C.byref(g), C.byref(nbytes)) == _lib.OK
    assert nbytes.value > 4 * 64


def test_filters_fail_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from pmmoto_amd.filters import distance
    with pytest.raises(_lib.PmxError):
        distance.edt(np.ones((4, 4, 4), np.uint8))


def test_edt_dimension_errors():
    from pmmoto_amd.filters import distance
    with pytest.raises(ValueError):
        distance.edt(np.ones(4))  # test_edt.py:369-373


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors of pmx_geom and pmx_psd_domain have the C layout."""
    import ctypes as C
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "pmx.h"\n'
                   'int main(void) { printf("%zu %zu %zu %zu\\n", sizeof(pmx_geom), sizeof(pmx_psd_domain), '
                   'offsetof(pmx_psd_domain, code_map), offsetof(pmx_psd_domain, tile_rows)); return 0; }\n')
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("header does not compile with gcc here: " + r.stderr[-200:])
    sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert sizes == [C.sizeof(_lib.Geom), C.sizeof(_lib.PsdDomain), _lib.PsdDomain.code_map.offset,
                     _lib.PsdDomain.tile_rows.offset]


def test_documented_binding_signatures_match_the_library():
    """The ctypes binding documented for PMMoTo's maintainers declares the
    same argument counts as the package's own table (executed on the host:
    it only loads the library and defines functions)."""
    blocks = re.findall(r"```python\n(.*?)```", open(os.path.join(ROOT, "INTEGRATION.md")).read(), flags=re.S)
    os.environ["LIBPMX"] = _lib.LIB_PATH
    ns = {}
    exec(compile(blocks[0], "binding", "exec"), ns)
    for name in ("pmx_edt", "pmx_edt_workspace_size", "pmx_edt_ref", "pmx_edt_ref_workspace_size"):
        assert len(getattr(ns["_L"], name).argtypes) == len(_lib.SIGNATURES[name][1]), name
