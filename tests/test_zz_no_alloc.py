"""hymor_expm_taylor_dev_noalloc: the Taylor action on the library's static
scratch array, called through the C ABI, against hymor_expm_taylor_dev (the
entry point it mirrors, same kernels in the same order: equal bits).

Two calls in a row, the second on a smaller operator, so stale scratch
contents from the first call would show; the first no-alloc call on each
handle comes before any call that builds the handle's own workspace, and the
handle's device memory does not grow across it."""
import sys

import numpy as np
import pytest

from hymor_amd import _lib
from hymor_amd import operators as ops
from hymor_amd._lib import C128, F64, check
from hymor_amd.core import _FORMS, DeviceOperator, DeviceVector

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _info_bytes(dev):
    import ctypes as C
    out = (C.c_int64 * 10)()
    check(_lib.load().hymor_op_info(dev.handle, out, 10))
    return int(out[9])


CASES = [
    # (label, operator, complex vectors)
    ("60x40 b8, real", lambda: ops.stencil_operator(60, 40, 8, seed=71, row_sum=2.0), False),
    ("12x10 b4 + 12 shock DOFs + 3 wavenumbers, complex",
     lambda: ops.receptivity_extension(
         ops.stencil_operator(12, 10, 4, seed=72, n_shock=12, row_sum=2.0), [0.0, 0.7, 2.1]),
     True),
]


def test_noalloc_equals_dev_entry_point():
    lib = _lib.load()
    for label, make, cplx in CASES:
        op = make()
        dtype = np.complex128 if cplx else np.float64
        abi = C128 if cplx else F64
        x = np.random.default_rng(73).standard_normal(op.n).astype(dtype)
        if cplx:
            x += 1j * np.random.default_rng(74).standard_normal(op.n)
        with DeviceOperator.from_blocks(op) as dev:
            dx = DeviceVector.from_host(x)
            dy_na, dy = DeviceVector(op.n, dtype), DeviceVector(op.n, dtype)
            try:
                for form in ("horner", "powers", "normalised"):
                    for tr in (0, 1):
                        say(f"[no-alloc] {label}: n={op.n} form={form} transpose={tr}")
                        before = _info_bytes(dev)
                        check(lib.hymor_expm_taylor_dev_noalloc(dev.handle, tr, 5, 3, 0.1,
                                                                _FORMS[form], dx.ptr, dy_na.ptr,
                                                                abi))
                        dev.synchronize()
                        assert _info_bytes(dev) == before
                        got = dy_na.download()
                        check(lib.hymor_expm_taylor_dev(dev.handle, tr, 5, 3, 0.1, _FORMS[form],
                                                        dx.ptr, dy.ptr, abi))
                        dev.synchronize()
                        np.testing.assert_array_equal(got, dy.download())
            finally:
                dx.free()
                dy_na.free()
                dy.free()


def test_noalloc_rejects_inputs_beyond_its_scratch():
    say("[no-alloc] 600x300 b8 with complex vectors: beyond the scratch limit")
    op = ops.config_operator("C5")
    lib = _lib.load()
    with DeviceOperator.from_blocks(op, with_adjoint=False) as dev:
        dx, dy = DeviceVector(op.n, np.complex128), DeviceVector(op.n, np.complex128)
        try:
            rc = lib.hymor_expm_taylor_dev_noalloc(dev.handle, 0, 5, 1, 0.1, _FORMS["horner"],
                                                   dx.ptr, dy.ptr, C128)
            assert rc == _lib.ERR_ARG
            assert "scratch limit" in lib.hymor_last_error().decode()
        finally:
            dx.free()
            dy.free()
