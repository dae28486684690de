"""ctypes binding of libpmx.so (the HIP kernels behind include/pmx.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
pmmoto_amd/csrc``).  There is no CPU fallback: every filter in this package
runs on the GPU through this binding and raises if the library or a HIP
device is missing.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

# PMX_LIBRARY may name another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("PMX_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpmx.so")

OK, ERR_INVALID, ERR_HIP, ERR_WORKSPACE = 0, -1, -2, -3
BC_END, BC_WALL, BC_PERIODIC = 0, 1, 2
D2_INF = np.iinfo(np.int32).max


class Geom(C.Structure):
    """pmx_geom: the box the kernels compute on and the views' strides."""

    _fields_ = [
        ("n", C.c_int64 * 3),
        ("in_s", C.c_int64 * 2),
        ("out_s", C.c_int64 * 2),
        ("bc", C.c_int32 * 6),
    ]


class PsdDomain(C.Structure):
    """pmx_psd_domain: one domain of a batched sweep."""

    _fields_ = [
        ("img", C.c_void_p),
        ("g", Geom),
        ("inlet_lo", C.c_int64 * 3),
        ("inlet_ext", C.c_int64 * 3),
        ("has_inlet", C.c_int32),
        ("wall_mask", C.c_uint32),
        ("code_map", C.c_void_p),
        ("tile_rows", C.c_int32),
    ]


# name -> (restype, argtypes); every symbol include/pmx.h declares
P, I32, I64, U8, U32, SZ = C.c_void_p, C.c_int32, C.c_int64, C.c_uint8, C.c_uint32, C.c_size_t
SIGNATURES = {
    "pmx_version": (C.c_int, []),
    "pmx_last_error": (C.c_char_p, []),
    "pmx_profile_begin": (C.c_int, []),
    "pmx_profile_end": (C.c_int, [P, C.c_int, P, P, C.c_int, P]),
    "pmx_edt_workspace_size": (C.c_int, [P, P]),
    "pmx_edt": (C.c_int, [P, P, P, P, P, P, P, SZ, P]),
    "pmx_edt_noalloc": (C.c_int, [P, P, P, P, P, P, P]),
    "pmx_edt_ref_workspace_size": (C.c_int, [P, P]),
    "pmx_edt_ref": (C.c_int, [P, P, P, P, P, C.c_int, P, P, P, C.c_int, P, SZ, P]),
    "pmx_edt_ref_int_form": (C.c_int, [C.c_int]),
    "pmx_edt_ref_last_form": (C.c_int, [P, P, SZ, P, P]),
    "pmx_halo_morph_workspace_size": (C.c_int, [P, P, P]),
    "pmx_halo_morph": (C.c_int, [P, P, P, P, P, C.c_int, C.c_double, C.c_int, U8, U8, U32, P, P, SZ, P]),
    "pmx_morph": (C.c_int, [P, P, C.c_int, I64, U8, U8, U32, P, P, SZ, P]),
    "pmx_fill_periodic_pads": (C.c_int, [P, C.c_int, P, P, P, P]),
    "pmx_ccl_workspace_size": (C.c_int, [P, P]),
    "pmx_ccl26": (C.c_int, [P, P, P, P, P, P, SZ, P]),
    "pmx_mark_labels": (C.c_int, [P, C.c_int, P, P, P, P, C.c_uint64, P, P]),
    "pmx_keep_labels": (C.c_int, [P, P, C.c_int, I64, P, P, C.c_uint64, C.c_int, C.c_int, P, P, P]),
    "pmx_psd_workspace_size": (C.c_int, [P, P]),
    "pmx_psd_sweep": (C.c_int, [P, P, P, P, C.c_int, P, P, U32, P, P, SZ, P]),
    "pmx_psd_batch_workspace_size": (C.c_int, [C.c_int, P, C.c_int, P]),
    "pmx_psd_sweep_batch": (C.c_int, [C.c_int, P, C.c_int, P, P, C.c_int, P, P, SZ]),
    "pmx_edt_autotune_workspace_size": (C.c_int, [P, P]),
    "pmx_edt_autotune": (C.c_int, [P, P, I64, C.c_int, P, P, SZ, P]),
    "pmx_tune_cache_save": (C.c_int, [C.c_char_p]),
    "pmx_tune_cache_load": (C.c_int, [C.c_char_p]),
    "pmx_psd_batch_plan_size": (C.c_int, [C.c_int, P, P, C.c_int, P, P, P]),
    "pmx_psd_batch_create": (C.c_int, [C.c_int, P, P, C.c_int, P, P, P, SZ, P, P]),
    "pmx_psd_batch_run": (C.c_int, [P, P]),
    "pmx_psd_batch_destroy": (None, [P]),
    "pmx_psd_update": (C.c_int, [P, P, I64, I32, P, P]),
    "pmx_codes_to_values": (C.c_int, [P, I64, P, I32, P, P]),
    "pmx_code_histogram": (C.c_int, [P, P, P, P, P, I32, P, P]),
    "pmx_both_one": (C.c_int, [P, P, I64, P, P]),
    "pmx_saturation_curve": (C.c_int, [P, P, P, P, P, C.c_int32, C.c_double, P, P, P]),
    "pmx_drain_update": (C.c_int, [P, P, I64, P, P]),
    "pmx_box_op": (C.c_int, [P, P, P, P, C.c_int, U8, P, P]),
    "pmx_max_f32": (C.c_int, [P, I64, P, P]),
    "pmx_max_u8": (C.c_int, [P, I64, P, P]),
    "pmx_gemm_bf16": (C.c_int, [I64, C.c_int, C.c_int, P, I64, C.c_int, P, I64, C.c_int, P, I64, C.c_int, C.c_int, P]),
    "pmx_conv3d_first": (C.c_int, [P, P, C.c_int, C.c_int, P, P, P, C.c_int, P]),
    "pmx_conv3d_first_pair": (C.c_int, [P, P, P, P, P, P, P, C.c_int, P]),
    "pmx_conv3d_bf16": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, P, P, C.c_int, P, C.c_int, P]),
    "pmx_im2col3d": (C.c_int, [P, P, C.c_int, I64, P, P]),
    "pmx_col2im3d": (C.c_int, [P, P, C.c_int, I64, P, P]),
    "pmx_bias_act": (C.c_int, [P, P, I64, C.c_int, C.c_int, P, P, P]),
    "pmx_relu_backward": (C.c_int, [P, P, I64, P, P, P]),
    "pmx_to_bf16": (C.c_int, [P, I64, P, P]),
    "pmx_u8_to_bf16": (C.c_int, [P, I64, P, P]),
    "pmx_col_sum": (C.c_int, [P, I64, C.c_int, P, P]),
    "pmx_avgpool": (C.c_int, [P, C.c_int, I64, C.c_int, P, P, P]),
    "pmx_avgpool_backward": (C.c_int, [P, C.c_int, I64, C.c_int, P, P]),
    "pmx_gnll_loss": (C.c_int, [P, P, C.c_int, C.c_int, P, P, P]),
    "pmx_adam": (C.c_int, [P, P, P, P, I64, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, P, P]),
    "pmx_sphere_first_cover": (C.c_int, [P, I64, P, P, P, P, P]),
    "pmx_spheres_on_grid": (C.c_int, [P, I64, P, P, P, P, P, P, C.c_int, P, P]),
    "pmx_sphere_image": (C.c_int, [P, P, I32, P, P]),
}

_lib = None


class PmxError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libpmx.so once; raise loudly if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PmxError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "or `make -C pmmoto_amd/csrc` (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    """Map a pmx status to the exception type the reference raises."""
    if rc == OK:
        return
    msg = lib().pmx_last_error().decode()
    if rc == ERR_INVALID:
        raise ValueError(msg)
    raise PmxError(f"libpmx status {rc}: {msg}")


def profile_begin() -> None:
    check(lib().pmx_profile_begin())


def profile_end() -> dict:
    """{kernel name: (total ms, launches)} recorded since profile_begin()."""
    names = C.create_string_buffer(8192)
    ms = (C.c_double * 64)()
    calls = (C.c_int64 * 64)()
    n = C.c_int()
    check(lib().pmx_profile_end(names, 8192, ms, calls, 64, C.byref(n)))
    keys = names.value.decode().split("\n")[: n.value]
    return {k: (ms[i], int(calls[i])) for i, k in enumerate(keys)}
