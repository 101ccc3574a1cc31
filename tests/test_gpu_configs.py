"""Full-size configurations C3 and C4 (sizes the oracle cannot run a full
action on in seconds) through size-independent properties and single
oracle SpMVs.

* C3: 400 x 300 cells, b = 4, 400 shock DOFs (N = 480 400): one SpMV
  against the oracle, the adjoint dot-product identity of the Taylor action,
  linearity.  (The Chu sandwich, transient growth and the gain gradient at
  this size are in test_gpu_fullsize.py.)
* C4: 500 x 250 cells, b = 4, 32 freestream wavenumbers
  w_k = 0.1 * 100^((k-1)/31) (N_ext = 564 000 complex): the factored
  receptivity action equals the assembled extended operator's action on the
  same vector, forward and adjoint, and one assembled SpMV equals the
  oracle's.
"""
import sys

import numpy as np
import pytest

from hymor_amd import operators as ops
from hymor_amd.core import DeviceOperator
from hymor_amd.receptivity import FactoredReceptivity, split_extended

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_c3_properties(orc):
    say("[C3 400x300 b4 + 400 shock DOFs] building")
    op = ops.config_operator("C3")
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    r = np.random.default_rng(31)
    x, y = r.standard_normal(op.n), r.standard_normal(op.n)
    say(f"[C3] n={op.n}: SpMV, adjoint identity, linearity")
    with DeviceOperator.from_blocks(op) as dev:
        assert relerr(dev.spmv(x), orc.csc_mul(A, x)) <= 1e-14
        T = lambda v, tr=False: dev.expm_taylor(v, order=5, nsub=6, dt=1 / 64, transpose=tr)
        Tx, Ty_t = T(x), T(y, tr=True)
        assert abs(Tx @ y - x @ Ty_t) <= 1e-12 * abs(Tx @ y)
        a, b = 0.3, -1.7
        assert relerr(T(a * x + b * y), a * Tx + b * T(y)) <= 1e-13


def test_c4_factored_receptivity_full_size(orc):
    say("[C4 500x250 b4, 32 wavenumbers] building the extended operator")
    base = ops.config_operator("C4")
    w = 0.1 * 100.0 ** (np.arange(32) / 31.0)
    L_ = ops.receptivity_extension(base, w)
    A_ = L_.to_scipy_varmajor()
    n_L, p = base.n, 4 * base.nchi
    L, B = split_extended(A_, n_L, p, w)
    r = np.random.default_rng(32)
    x = r.standard_normal(A_.shape[0]) + 1j * r.standard_normal(A_.shape[0])
    dt = 1 / 64
    say(f"[C4] n_ext={A_.shape[0]}: assembled vs factored action, and one SpMV vs the oracle")
    with DeviceOperator.from_blocks(L_, scale=dt) as dev:
        ref = dev.expm_taylor(x, order=5, nsub=2, dt=1.0)
        ref_t = dev.expm_taylor(x, order=5, nsub=2, dt=1.0, transpose=True)
        Ao = orc.CSC.from_scipy(A_)
        assert relerr(dev.spmv(x) / dt, orc.csc_mul(Ao, x)) <= 1e-13
    fr = FactoredReceptivity(L, B, w, nchi=base.nchi, neta=base.neta, vars_per_cell=4,
                             scale=dt)
    try:
        assert relerr(fr.expm(x, order=5, nsub=2), ref) <= 1e-12
        assert relerr(fr.expm(x, order=5, nsub=2, transpose=True), ref_t) <= 1e-12
    finally:
        fr.close()
