"""ctypes binding of libhymor_hip.so (declared in include/hymor.h).

The library is the only compute path: if it is missing or no GPU is present
the wrappers raise; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libhymor_hip.so"

F64, C128 = 0, 1
PREC_F64, PREC_F32 = 0, 1
FORM_NORMALISED, FORM_POWERS, FORM_HORNER = 0, 1, 2
OK, ERR_ARG, ERR_HIP, ERR_NOMEM, ERR_NODEV = 0, -1, -2, -3, -4

_i64 = C.c_int64
_p = C.c_void_p
_pp = C.POINTER(C.c_void_p)
_pi64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol include/hymor.h declares
SIGNATURES = {
    "hymor_device_available": (C.c_int, []),
    "hymor_last_error": (C.c_char_p, []),
    "hymor_op_create_csc": (C.c_int, [_i64, _p, _p, _p, C.c_int, C.c_int, C.c_double, C.c_int,
                                      _i64, _i64, C.c_int, _pp]),
    "hymor_op_create_receptivity": (C.c_int, [_i64, _p, _p, _p, C.c_int, C.c_int, C.c_double,
                                              C.c_int, _i64, _i64, _i64, _i64, _p, C.c_int,
                                              _pp]),
    "hymor_op_create_bsr": (C.c_int, [_i64, _i64, C.c_int, _i64, _p, _p, _p, _i64, _p, _p, _p,
                                      C.c_int, C.c_double, C.c_int, C.c_int, _pp]),
    "hymor_op_destroy": (C.c_int, [_p]),
    "hymor_op_info": (C.c_int, [_p, _pi64, C.c_int]),
    "hymor_op_set_value_prec": (C.c_int, [_p, C.c_int]),
    "hymor_spmv": (C.c_int, [_p, C.c_int, _p, _p, C.c_int]),
    "hymor_spmv_dev": (C.c_int, [_p, C.c_int, _p, _p, C.c_int]),
    "hymor_expm_taylor": (C.c_int, [_p, C.c_int, C.c_int, _i64, C.c_double, C.c_int, _p, _p,
                                    C.c_int]),
    "hymor_expm_taylor_dev": (C.c_int, [_p, C.c_int, C.c_int, _i64, C.c_double, C.c_int, _p, _p,
                                        C.c_int]),
    "hymor_expm_taylor_dev_noalloc": (C.c_int, [_p, C.c_int, C.c_int, _i64, C.c_double, C.c_int,
                                                _p, _p, C.c_int]),
    "hymor_expm_taylor_batched": (C.c_int, [_p, C.c_int, C.c_int, _i64, C.c_double, C.c_int, _i64,
                                            _p, _i64, _p, _i64, C.c_int]),
    "hymor_expm_taylor_batched_dev": (C.c_int, [_p, C.c_int, C.c_int, _i64, C.c_double, C.c_int,
                                                _i64, _p, _i64, _p, _i64, C.c_int]),
    "hymor_expm_taylor_batched_coldiag_dev": (C.c_int, [_p, C.c_int, C.c_int, _i64, C.c_double,
                                                        C.c_int, _i64, _p, _i64, _i64, _p, _i64,
                                                        _p, _i64]),
    "hymor_expm_taylor_vjp_values": (C.c_int, [_p, C.c_int, _i64, C.c_double, _i64, _p, _p, _p,
                                               _p, C.c_int, _p]),
    "hymor_pattern_create": (C.c_int, [_p, _p, _p, C.c_int, _pp]),
    "hymor_pattern_destroy": (C.c_int, [_p]),
    "hymor_pattern_nnz": (C.c_int, [_p, _pi64]),
    "hymor_expm_taylor_vjp_values_dev": (C.c_int, [_p, _p, C.c_int, _i64, C.c_double, _i64, _p,
                                                   _p, _p]),
    "hymor_chu_sandwich": (C.c_int, [_p, _p, C.c_int, _i64, C.c_double, _i64, _p, _p, C.c_int]),
    "hymor_chu_sandwich_dev": (C.c_int, [_p, _p, C.c_int, _i64, C.c_double, _i64, _p, _p,
                                         C.c_int]),
    "hymor_chu_sandwich_batched_dev": (C.c_int, [_p, _p, _p, _p, C.c_int, _i64, C.c_double, _i64,
                                                 _i64, _p, _i64, _p, _i64, C.c_int]),
    "hymor_op_synchronize": (C.c_int, [_p]),
    "hymor_dev_alloc": (C.c_int, [_i64, _pp]),
    "hymor_dev_free": (C.c_int, [_p]),
    "hymor_memcpy_h2d": (C.c_int, [_p, _p, _i64]),
    "hymor_memcpy_d2h": (C.c_int, [_p, _p, _i64]),
    "hymor_memcpy_d2d": (C.c_int, [_p, _p, _i64]),
    "hymor_dev_fill_uniform": (C.c_int, [_p, _i64, C.c_uint64]),
    "hymor_bench_term_kernel": (C.c_int, [_p, C.c_int, C.c_int, C.c_int, C.c_int, _pd, _pd]),
    "hymor_op_autotune": (C.c_int, [_p, C.c_int, C.c_int, _pd]),
    "hymor_op_set_graphs": (C.c_int, [_p, C.c_int]),
    "hymor_op_set_batch_kernel": (C.c_int, [_p, C.c_int]),
    "hymor_op_autotune_batch": (C.c_int, [_p, C.c_int, C.c_int, C.POINTER(C.c_int), _pd]),
    "hymor_op_get_tuning": (C.c_int, [_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "hymor_op_set_tuning": (C.c_int, [_p, C.c_int, C.c_int, C.c_int]),
    "hymor_bench_copy": (C.c_int, [_i64, C.c_int, _pd]),
    "hymor_bench_read": (C.c_int, [_i64, C.c_int, _pd]),
    "hymor_bench_stream_mix": (C.c_int, [_i64, _i64, C.c_int, _pd, _pd]),
    "hymor_krylov_create": (C.c_int, [_i64, C.c_int, C.c_int, _pp]),
    "hymor_krylov_destroy": (C.c_int, [_p]),
    "hymor_krylov_column": (C.c_int, [_p, C.c_int, _pp]),
    "hymor_krylov_set_column": (C.c_int, [_p, C.c_int, _p]),
    "hymor_krylov_get_columns": (C.c_int, [_p, C.c_int, C.c_int, _p]),
    "hymor_krylov_orthonormalize": (C.c_int, [_p, C.c_int, _p, _pd]),
    "hymor_krylov_orthonormalize_block": (C.c_int, [_p, C.c_int, C.c_int, _p, _i64, _pd]),
    "hymor_krylov_combine": (C.c_int, [_p, C.c_int, C.c_int, _pd]),
    "hymor_krylov_ritz": (C.c_int, [_p, C.c_int, C.c_int, _pd, _p]),
    "hymor_krylov_copy_column": (C.c_int, [_p, C.c_int, C.c_int]),
    "hymor_bench_krylov_projection": (C.c_int, [_i64, C.c_int, C.c_int, C.c_int, _pd]),
    "hymor_weights_create_csc": (C.c_int, [_i64, C.c_int, _p, _p, _p, C.c_int, C.c_int, _i64,
                                           _i64, _pp]),
    "hymor_weights_destroy": (C.c_int, [_p]),
    "hymor_energy_track": (C.c_int, [_p, _p, C.c_int, _i64, C.c_double, C.c_int, _p, C.c_int, _p,
                                     _p, _i64, _i64, _p]),
}

_lib = None


class HymorError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise HymorError(f"{LIB_PATH} is missing: build it with `python -m hymor_amd.build`")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().hymor_last_error().decode(errors="replace")
        raise HymorError(f"hymor error {rc}: {msg}")


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def device_available() -> bool:
    return bool(load().hymor_device_available())
