"""Full-size configurations C2, C3 and C4 (sizes the oracle cannot run a full
action on in seconds) through size-independent properties and single
oracle SpMVs.

* C2: 200 x 300 cells, b = 4, 200 shock DOFs (N = 240 200): 16-vector
  sweeps (FP64 and FP32 values) against the plain action, one batched SpMV
  column against the oracle.
* C3: 400 x 300 cells, b = 4, 400 shock DOFs (N = 480 400): one SpMV
  against the oracle, the adjoint dot-product identity of the Taylor action,
  linearity.  (The Chu sandwich, transient growth and the gain gradient at
  this size are in test_gpu_fullsize.py.)
* C4: 500 x 250 cells, b = 4, 32 freestream wavenumbers
  w_k = 0.1 * 100^((k-1)/31) (N_ext = 564 000 complex): the factored
  receptivity actions (the extended operator from its parts, and the 32
  per-wavenumber systems) equal the assembled extended operator's action on
  the same vector, forward and adjoint, and one assembled SpMV equals the
  oracle's; the Horner terms with the replicated rows fused into one launch
  (one vector and 16-vector batches) equal the powers form's separate
  launches, and the fused adjoint satisfies the dot-product identity;
  FREESTREAM_RECEPTIVITY at this size (device Krylov, one vector and
  blocks of 4) finds the same gains with D-orthonormal modes.
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


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 2e-6)])
def test_c2_batched_columns(orc, prec, tol):
    """C2 (200 x 300 cells, b = 4, 200 shock DOFs, N = 240 200): 16 vectors
    per Horner sweep -- the shock rows and the shock-column slices in the
    batch kernel's leading workgroups -- equal the plain FP64 action column
    by column, forward and adjoint (FP32-stored values: within their
    rounding band); one SpMV column against the oracle."""
    say(f"[C2 200x300 b4 + 200 shock DOFs] 16-vector sweeps, values {prec}")
    op = ops.config_operator("C2")
    X = np.asfortranarray(np.random.default_rng(41).standard_normal((op.n, 16)))
    kw = dict(order=12, nsub=4, dt=1 / 64, form="horner")
    with DeviceOperator.from_blocks(op) as d64, \
            DeviceOperator.from_blocks(op, value_prec=prec) as dev:
        for tr in (False, True):
            Y = dev.expm_taylor_batched(X, transpose=tr, **kw)
            for j in (0, 7, 15):
                want = d64.expm_taylor(X[:, j], transpose=tr, **kw)
                assert relerr(Y[:, j], want) <= tol, (tr, j)
        if prec == "f64":
            A = orc.CSC.from_arrays(*op.to_csc_varmajor())
            Y1 = dev.expm_taylor_batched(X, order=1, nsub=1, dt=1.0) - X
            assert relerr(Y1[:, 3], orc.csc_mul(A, X[:, 3])) <= 1e-13


@pytest.fixture(scope="module")
def c4_extended():
    say("[C4 500x250 b4, 32 wavenumbers] building the extended operator")
    base = ops.config_operator("C4")
    w = 0.1 * 100.0 ** (np.arange(32) / 31.0)
    L_ = ops.receptivity_extension(base, w)
    return base, w, L_, L_.to_scipy_varmajor()


def test_c4_factored_receptivity_full_size(orc, c4_extended):
    base, w, L_, A_ = c4_extended
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
    for per_wavenumber in (False, True):
        fr = FactoredReceptivity(L, B, w, nchi=base.nchi, neta=base.neta, vars_per_cell=4,
                                 scale=dt, per_wavenumber=per_wavenumber)
        try:
            assert relerr(fr.expm(x, order=5, nsub=2), ref) <= 1e-12, per_wavenumber
            assert relerr(fr.expm(x, order=5, nsub=2, transpose=True), ref_t) <= 1e-12, \
                per_wavenumber
        finally:
            fr.close()


def test_c4_freestream_receptivity_solve(c4_extended):
    """FREESTREAM_RECEPTIVITY at full size (564 000 complex unknowns; the
    extended operator uploaded from its parts, B kept once): the device
    Krylov-Schur on the Cholesky-reduced Hermitian problem, one vector and
    blocks of 4 (batched complex sandwiches with the freestream zeroing),
    finds the same 4 gains (1e-9); the modes are D-orthonormal and satisfy
    the reduced problem to the solver's tolerance."""
    from hymor_amd import stability as st
    from hymor_amd.baseflow import synthetic_solution
    base, w, L_, A_ = c4_extended
    say("[C4] FREESTREAM_RECEPTIVITY k=4 T=1, blocks of 1 and 4")
    s = st.CFL_TIMESTEP(synthetic_solution(500, 250, seed=11, perturb_shock=False, cfl=100.0,
                                           max_dt=1.0 / 64))
    tol = 1e-10
    _, G1, _ = st.FREESTREAM_RECEPTIVITY(A_, s, 4, 1.0, w, tol=tol, krylov="device", ncv=24,
                                         maxiter=400, factored=True, block=1)
    V, G4, T_opt = st.FREESTREAM_RECEPTIVITY(A_, s, 4, 1.0, w, tol=tol, krylov="device",
                                             ncv=24, maxiter=400, factored=True, block=4)
    say(f"[C4] gains {G4}")
    assert np.all(G4 > 0) and np.all(np.diff(G4) <= 0)
    assert np.all(np.abs(G4 - G1) <= 1e-9 * np.abs(G1)), (G1, G4)
    R_ = st.CONSTRUCT_R_(s, w)
    D_ = (R_.T @ st.CONSTRUCT_M_INFTY_(s, [1, 1, 1, 1], 1.0, w, True) @ R_).tocsc()
    Gm = V.conj().T @ (D_ @ V)
    assert np.abs(np.diag(Gm) - 1.0).max() <= 1e-10
    assert np.abs(Gm - np.diag(np.diag(Gm))).max() <= 1e-6


def test_c4_fused_terms_full_size(c4_extended):
    base, w, L_, A_ = c4_extended
    r = np.random.default_rng(33)
    n = A_.shape[0]
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    y = r.standard_normal(n) + 1j * r.standard_normal(n)
    X = r.standard_normal((n, 16)) + 1j * r.standard_normal((n, 16))
    say(f"[C4] n_ext={n}: fused Horner terms vs the powers form; fused adjoint identity")
    with DeviceOperator.from_blocks(L_, scale=1 / 64) as dev:
        T = lambda v, form, tr=False: dev.expm_taylor(v, order=6, nsub=3, dt=1.0, form=form,
                                                      transpose=tr)
        assert relerr(T(x, "horner"), T(x, "powers")) <= 1e-13
        Y = dev.expm_taylor_batched(X, order=6, nsub=3, dt=1.0, form="horner")
        P = dev.expm_taylor_batched(X, order=6, nsub=3, dt=1.0, form="powers")
        for j in range(16):
            assert relerr(Y[:, j], P[:, j]) <= 1e-13, j
        # <T x, y> = <x, T^H y>
        lhs = np.vdot(y, T(x, "horner"))
        rhs = np.vdot(T(y, "horner", tr=True), x)
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
