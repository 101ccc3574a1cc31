"""CPU-side checks of the C ABI: the library loads, exports every symbol that
include/hymor.h declares, and fails loudly (no CPU fallback) without a GPU."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from hymor_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "hymor.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hymor_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_symbols():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_gpu_means_a_loud_error():
    lib = _lib.load()
    if lib.hymor_device_available():
        pytest.skip("a GPU is present")
    from hymor_amd.operators import stencil_operator
    from hymor_amd.core import DeviceOperator
    op = stencil_operator(4, 3, 4, seed=0)
    with pytest.raises(_lib.HymorError, match="device"):
        DeviceOperator.from_blocks(op)


def test_argument_errors_are_reported():
    lib = _lib.load()
    h = ctypes.c_void_p()
    colptr = np.array([1, 1], dtype=np.int64)
    rowval = np.zeros(0, dtype=np.int64)
    nz = np.zeros(0)
    rc = lib.hymor_op_create_csc(1, _lib.ptr(colptr), _lib.ptr(rowval), _lib.ptr(nz), 1, 0, 1.0,
                                 5, 1, 1, 0, ctypes.byref(h))
    assert rc == -1
    assert b"vars_per_cell" in lib.hymor_last_error()
    assert not h.value


def test_new_entry_points_check_their_arguments():
    """The precision switch, the device fill, the copies and the probes
    reject null or malformed arguments before touching a device."""
    lib = _lib.load()
    assert lib.hymor_op_set_value_prec(None, 0) == -1
    assert b"null" in lib.hymor_last_error()
    assert lib.hymor_dev_fill_uniform(None, 8, 0) == -1
    assert lib.hymor_memcpy_d2d(None, None, 8) == -1
    ms = ctypes.c_double()
    assert lib.hymor_bench_read(15, 1, ctypes.byref(ms)) == -1
    assert b"multiple of 16" in lib.hymor_last_error()
    assert lib.hymor_bench_stream_mix(1024, 0, 1, ctypes.byref(ms), ctypes.byref(ms)) == -1
    assert b"16 MiB" in lib.hymor_last_error()
    assert lib.hymor_bench_term_kernel(None, 0, 1, 2, 1, ctypes.byref(ms), ctypes.byref(ms)) == -1


def test_receptivity_constructor_checks_its_arguments():
    """hymor_op_create_receptivity rejects a system without freestream DOFs,
    a wavenumber count below one and non-finite wavenumbers before it
    touches a device."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    n = 4 * 2 * 2 + 8
    colptr = np.ones(n + 1, dtype=np.int64)
    rowval = np.zeros(0, dtype=np.int64)
    nz = np.zeros(0)
    w = np.array([1.0, np.nan])

    def create(n_l, nwave, wv):
        return lib.hymor_op_create_receptivity(n, _lib.ptr(colptr), _lib.ptr(rowval),
                                               _lib.ptr(nz), 1, 0, 1.0, 4, 2, 2, n_l, nwave,
                                               _lib.ptr(wv), 0, ctypes.byref(h))

    assert create(n, 2, w) == -1
    assert b"freestream DOFs" in lib.hymor_last_error()
    assert create(16, 0, w) == -1
    assert b"nwave" in lib.hymor_last_error()
    assert create(16, 2, w) == -1
    assert b"finite" in lib.hymor_last_error()
    assert not h.value
