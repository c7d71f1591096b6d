"""ctypes binding of libeuler_hip.so (the C ABI declared in include/euler_capi.h).

The library is loaded from this package's directory. There is no fallback: if the
shared object is missing, importing this module raises, so nothing can silently run
a non-HIP path.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libeuler_hip.so"

EB_OK = 0
EB_ERR_INVALID_ARGUMENT = 1
EB_ERR_INDEX_OVERFLOW = 2
EB_ERR_HIP = 3
EB_ERR_OUT_OF_MEMORY = 4
EB_ERR_UNSUPPORTED = 5


class EulerError(RuntimeError):
    """Raised when an eb_* entry point returns a non-zero status."""

    def __init__(self, status: int, func: str, message: str):
        super().__init__(f"{func}: status {status}: {message}")
        self.status = status


def _fields(names):
    return [(n, ctypes.c_void_p) for n in names]


class PrimPtrs(ctypes.Structure):
    """eb_prim_f64 / eb_prim_f32: {d, p, ux0, ux1, ux2} device pointers."""

    _fields_ = _fields(["d", "p", "ux0", "ux1", "ux2"])


class ConsPtrs(ctypes.Structure):
    """eb_cons_f64 / eb_cons_f32: {d, e, mx0, mx1, mx2} device pointers."""

    _fields_ = _fields(["d", "e", "mx0", "mx1", "mx2"])


def _instance_struct(real):
    class BatchInstance(ctypes.Structure):
        """eb_batch_instance_f64 / _f32."""

        _fields_ = [("prim_a", PrimPtrs), ("prim_b", PrimPtrs), ("cons", ConsPtrs), ("n0", ctypes.c_int64),
                    ("n1", ctypes.c_int64), ("n2", ctypes.c_int64), ("gamma", real), ("dx0", real), ("dx1", real),
                    ("dx2", real), ("cfl", real)]

    return BatchInstance


BatchInstanceF64 = _instance_struct(ctypes.c_double)
BatchInstanceF32 = _instance_struct(ctypes.c_float)

_P = ctypes.POINTER
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_u64 = ctypes.c_uint64
_f64 = ctypes.c_double
_f32 = ctypes.c_float


def _signatures():
    sig = {
        "eb_context_create": [_P(_vp)],
        "eb_context_destroy": [_vp],
        "eb_set_godunov_config": [_int, _int, _int],
        "eb_get_godunov_config": [_P(_int), _P(_int), _P(_int)],
        "eb_get_godunov_tile_height": [_int, _int, _P(_int)],
        "eb_batch_create_f64": [_P(_vp), _P(BatchInstanceF64), _int, _int],
        "eb_batch_create_f32": [_P(_vp), _P(BatchInstanceF32), _int, _int],
        "eb_batch_destroy": [_vp],
        "eb_batch_set_tiles": [_vp, _int, _int, _int],
        "eb_batch_get_tiles": [_vp, _P(_int), _P(_int), _P(_int), _P(_int), _P(_int)],
        "eb_batch_device_state": [_vp, _P(_vp), _P(_vp)],
        "eb_batch_copy_state": [_vp, _vp, _vp, _vp],
        "eb_batch_start": [_vp, _vp],
        "eb_batch_refresh_ghosts": [_vp, _vp, _int],
        "eb_batch_step": [_vp, _vp, _int],
        "eb_batch_step_per_instance": [_vp, _vp, _int],
        "eb_batch_autotune": [_vp, _vp, _int, _int, _P(_f32)],
        "eb_time_step_noalloc_f64": [_vp, _P(PrimPtrs), _i64, _i64, _i64, _int, _f64, _f64, _f64, _f64, _vp],
    }
    for suf, real in (("f64", _f64), ("f32", _f32)):
        ext = [_i64, _i64, _i64, _int]
        sig[f"eb_cons_to_prim_{suf}"] = [_vp, _P(ConsPtrs), _P(PrimPtrs)] + ext + [real]
        sig[f"eb_prim_to_cons_{suf}"] = [_vp, _P(PrimPtrs), _P(ConsPtrs)] + ext + [real]
        sig[f"eb_time_step_{suf}"] = [_vp, _vp, _P(PrimPtrs)] + ext + [real] * 4 + [_vp]
        sig[f"eb_time_step_sync_{suf}"] = [_vp, _vp, _P(PrimPtrs)] + ext + [real] * 4 + [_P(real)]
        sig[f"eb_godunov_{suf}"] = [_vp, _P(PrimPtrs), _P(ConsPtrs)] + ext + [real] * 5
        sig[f"eb_godunov_dev_{suf}"] = [_vp, _P(PrimPtrs), _P(ConsPtrs)] + ext + [real] * 4 + [_vp, real]
        sig[f"eb_godunov_reference_order_{suf}"] = [_vp, _P(PrimPtrs), _P(ConsPtrs)] + ext + [real] * 5
        sig[f"eb_ghost_refresh_{suf}"] = [_vp, _vp, _P(ConsPtrs), _P(PrimPtrs)] + ext + [real] * 4 + [_vp]
        sig[f"eb_step_fused_{suf}"] = (
            [_vp, _vp, _P(PrimPtrs), _P(ConsPtrs), _P(PrimPtrs)] + ext + [real] * 5 + [_vp, _vp, _vp, _vp]
        )
        sig[f"eb_step_fused_region_{suf}"] = (
            [_vp, _vp, _P(PrimPtrs), _P(ConsPtrs), _P(PrimPtrs)] + ext + [real] * 5 + [_vp, _vp, _vp, _vp, _vp]
        )
        sig[f"eb_init_implode_{suf}"] = [_vp, _P(PrimPtrs)] + ext + [real] * 3
        sig[f"eb_fill_{suf}"] = [_vp, _vp, ctypes.c_int64, real]
        sig[f"eb_stream_probe_{suf}"] = [_vp, _P(PrimPtrs), _P(ConsPtrs), ctypes.c_int64]
        sig[f"eb_init_diagonal_{suf}"] = [_vp, _P(PrimPtrs)] + ext + [real] * 8
        sig[f"eb_init_random_{suf}"] = [_vp, _P(PrimPtrs), _i64, _int, _u64] + [_f64] * 5
        sig[f"eb_godunov_vjp_{suf}"] = (
            [_vp, _P(PrimPtrs), _P(ConsPtrs), _P(ConsPtrs)] + ext + [real] * 4 + [_vp, real, _vp]
        )
        sig[f"eb_godunov_vjp_first_{suf}"] = (
            [_vp, _P(PrimPtrs), _P(ConsPtrs), _P(PrimPtrs)] + ext + [real] * 4 + [_vp, real, _vp]
        )
        sig[f"eb_cons_to_prim_vjp_{suf}"] = [_vp, _P(PrimPtrs), _P(PrimPtrs), _P(ConsPtrs)] + ext + [real]
        sig[f"eb_prim_to_cons_vjp_{suf}"] = [_vp, _P(PrimPtrs), _P(ConsPtrs), _P(PrimPtrs)] + ext + [real]
        sig[f"eb_conserved_totals_{suf}"] = [_vp, _vp, _P(ConsPtrs)] + ext + [_vp]
        sig[f"eb_boundary_flux_{suf}"] = [_vp, _vp, _P(PrimPtrs)] + ext + [real] * 4 + [_vp, real, _vp]
        sig[f"eb_hllc_flux_{suf}"] = [_vp, _int, _P(PrimPtrs), _P(PrimPtrs), _P(ConsPtrs), _i64, real]
        sig[f"eb_hllc_vjp_{suf}"] = (
            [_vp, _int, _P(PrimPtrs), _P(PrimPtrs), _P(ConsPtrs), _P(PrimPtrs), _P(PrimPtrs), _i64, real]
        )
    return sig


SIGNATURES = _signatures()


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C {LIB_PATH.parent}` "
            "(or __graft_entry__.build()); there is no non-HIP fallback"
        )
    lib = ctypes.CDLL(os.fspath(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.eb_status_string.argtypes = [ctypes.c_int]
    lib.eb_status_string.restype = ctypes.c_char_p
    lib.eb_last_error.argtypes = []
    lib.eb_last_error.restype = ctypes.c_char_p
    return lib


lib = _load()


def check(status: int, func: str) -> None:
    if status != EB_OK:
        msg = lib.eb_last_error().decode(errors="replace")
        raise EulerError(status, func, msg or lib.eb_status_string(status).decode())


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args), name)
