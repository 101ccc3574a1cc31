"""Solver legs and the headline workload at their full sizes, through the C ABI.

* 600 x 300 cells, b = 8, N = 1 440 000 (103.3M nonzeros, the benchmark's
  workload):
  - one forward and one adjoint SpMV against the oracle (1e-14);
  - the short action T_12(A/64)^2 against the oracle's POWER_METHOD!
    restatement, forward and adjoint (1e-12);
  - the full action T_12(A/64)^64: adjoint dot-product identity, linearity,
    Horner and powers forms equal;
  - 16-vector batches: FP64 values equal the single-vector columns (1e-13),
    FP32 values inside the FP32 band (1e-12 < err <= 2e-6);
  - 128 Krylov-sized vectors as 8 batched sweeps of 16 against the
    single-vector action of every column.
* 100 x 80 cells, b = 4 (N = 32 000): EIGENVALUES_GPU_ARNOLDI with 10 modes
  and Krylov dimension 60, through the device Krylov-Schur and through
  ARPACK (the reference's eigensolver), on an operator of that shape whose
  spectrum is exact (one-way coupling): eigenvalues 1e-8, Ritz residuals.
* 200 x 300 cells, b = 4, 200 shock DOFs (N = 240 200): 20 modes, Krylov
  dimension 80, on the same kind of operator with the shock bands' shapes.
* 400 x 300 cells, b = 4, 400 shock DOFs (N = 480 400) with Chu weights
  from a base flow of that grid: the Chu sandwich is symmetric, positive
  semi-definite and equals its three stages run separately;
  TRANSIENT_GROWTH_DOWNSTREAM for T in {0.5, 1, 2, 4} returns D-orthonormal
  modes whose residuals meet the solve tolerance; the gain gradient agrees
  with central differences (h = 1e-6) on 16 random directions; the gradient
  with respect to the operator's 19.3M entries (checkpointed every 8
  substeps) satisfies the scaling identity <dJ/dA, A> = dJ(sA)/ds at s = 1
  and does not depend on the checkpoint interval.
"""
import math
import sys

import numpy as np
import pytest

from hymor_amd import operators as ops
from hymor_amd import stability as st
from hymor_amd.adjoint import GainFunctional
from hymor_amd.baseflow import synthetic_solution
from hymor_amd.core import (DeviceMatrix, DeviceOperator, chu_sandwich_batched_ptr,
                            expm_taylor_batched_ptr)

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---------------------------------------------------------------- 600 x 300, b = 8

@pytest.fixture(scope="module")
def c5(orc):
    say("[600x300 b8, N=1440000] building the operator and its reference CSC")
    op = ops.config_operator("C5")
    A = orc.CSC.from_arrays(*op.to_csc_varmajor_blocks())
    dev = DeviceOperator.from_blocks(op)
    yield op, A, dev
    dev.close()


def test_c5_spmv_vs_oracle(c5, orc):
    op, A, dev = c5
    x = np.random.default_rng(51).standard_normal(op.n)
    say("[600x300 b8] forward SpMV vs oracle")
    assert relerr(dev.spmv(x), orc.csc_mul(A, x)) <= 1e-14
    say("[600x300 b8] adjoint SpMV vs oracle")
    At = orc.csc_adjoint(A)
    assert relerr(dev.spmv(x, transpose=True), orc.csc_mul(At, x)) <= 1e-14


def test_c5_short_action_vs_oracle(c5, orc):
    op, A, dev = c5
    x = np.random.default_rng(52).standard_normal(op.n)
    say("[600x300 b8] T_12(A/64)^2 vs oracle (powers and Horner forms)")
    ref = orc.power_method(A.scaled(1.0 / 64), x, 2, order=12)
    for form in ("powers", "horner"):
        got = dev.expm_taylor(x, order=12, nsub=2, dt=1.0 / 64, form=form)
        assert relerr(got, ref) <= 1e-12, form
    say("[600x300 b8] adjoint T_12(A'/64)^2 vs oracle")
    ref_t = orc.power_method(orc.csc_adjoint(A).scaled(1.0 / 64), x, 2, order=12)
    got_t = dev.expm_taylor(x, order=12, nsub=2, dt=1.0 / 64, form="powers", transpose=True)
    assert relerr(got_t, ref_t) <= 1e-12


def test_c5_full_action_properties(c5):
    op, A, dev = c5
    r = np.random.default_rng(53)
    x, y = r.standard_normal(op.n), r.standard_normal(op.n)
    say("[600x300 b8] T_12(A/64)^64: adjoint identity, linearity, forms")

    def T(v, tr=False, form="horner"):
        return dev.expm_taylor(v, order=12, nsub=64, dt=1.0 / 64, form=form, transpose=tr)

    Tx, Ty = T(x), T(y)
    # <T x, y> = <x, T' y>, to rounding of the two n-term sums
    assert abs(Tx @ y - x @ T(y, tr=True)) <= 1e-12 * np.linalg.norm(Tx) * np.linalg.norm(y)
    a, b = 0.7, -2.3
    assert relerr(T(a * x + b * y), a * Tx + b * Ty) <= 1e-13
    assert relerr(T(x, form="powers"), Tx) <= 1e-12


def test_c5_batched_columns(c5):
    op, A, dev = c5
    nb = 16
    X = np.asfortranarray(np.random.default_rng(54).standard_normal((op.n, nb)))
    kw = dict(order=12, nsub=64, dt=1.0 / 64, form="horner")
    say("[600x300 b8] 16-vector sweep, FP64 values, vs the single-vector action")
    Y = dev.expm_taylor_batched(X, **kw)
    ref = [dev.expm_taylor(X[:, j], **kw) for j in range(nb)]
    for j in range(nb):
        assert relerr(Y[:, j], ref[j]) <= 1e-13, j
    say("[600x300 b8] 16-vector sweep, FP32 values")
    with DeviceOperator.from_blocks(op, value_prec="f32", with_adjoint=False) as d32:
        Y32 = d32.expm_taylor_batched(X, **kw)
    for j in range(nb):
        assert 1e-12 < relerr(Y32[:, j], ref[j]) <= 2e-6, j


def test_c5_128_vectors_in_batches_of_16(c5):
    """128 basis-sized vectors resident on the device, applied as 8 batched
    sweeps of 16 consecutive columns (FP32 and FP64 values), each column
    against the plain single-vector FP64 action."""
    op, A, dev = c5
    n, nv, nb = op.n, 128, 16
    X = np.asfortranarray(np.random.default_rng(55).standard_normal((n, nv)))
    kw = dict(order=12, nsub=64, dt=1.0 / 64, form="horner")
    dX = DeviceMatrix.from_host(X)
    dY = DeviceMatrix(n, nv)
    col = 8 * n
    try:
        with DeviceOperator.from_blocks(op, value_prec="f32", with_adjoint=False) as d32:
            outs = {}
            for name, d in (("f64", dev), ("f32", d32)):
                say(f"[600x300 b8] 128 vectors, 8 sweeps of 16, {name} values")
                for b0 in range(0, nv, nb):
                    expm_taylor_batched_ptr(d, dX.ptr.value + b0 * col, n, nb,
                                            dY.ptr.value + b0 * col, n, **kw)
                d.synchronize()
                outs[name] = dY.download()
        say("[600x300 b8] 128 single-vector actions")
        for j in range(nv):
            ref = dev.expm_taylor(X[:, j], **kw)
            assert relerr(outs["f64"][:, j], ref) <= 1e-13, j
            assert 1e-12 < relerr(outs["f32"][:, j], ref) <= 2e-6, j
    finally:
        dX.free()
        dY.free()


# ---------------------------------------------------------------- eigenvalues

def _exact_eigs(lam, dt, n_t, k):
    mu = sum((lam * dt) ** j / math.factorial(j) for j in range(6)) ** n_t
    return np.log(mu[np.argsort(-np.abs(mu), kind="stable")][: 3 * k].astype(complex)) / (n_t * dt)


def _check_eigs(op, lam, D, V, k, dt, tol_eig, tol_res):
    """Every returned eigenvalue is one of the exact ones; the leading k
    exact ones are all found; Ritz residuals ||T v - mu v|| <= tol_res |mu|
    with T = T_5(A dt)^n applied on the device."""
    n_t = int(round(0.5 / dt))
    exact = _exact_eigs(lam, dt, n_t, k)
    assert np.all(np.isfinite(D)), D
    for d in D:
        assert np.min(np.abs(exact - d)) <= tol_eig * abs(d), (d, exact[:k + 2])
    for e in exact[:k]:
        assert np.min(np.abs(D - e)) <= tol_eig * abs(e), (e, D)
    mu = np.exp(D * n_t * dt)
    with DeviceOperator.from_blocks(op, scale=dt, with_adjoint=False) as dA:
        T = lambda v: dA.expm_taylor(v, order=5, nsub=n_t, dt=1.0, form="powers")
        for j in range(k):
            v = V[:, j] / np.linalg.norm(V[:, j])
            Tv = T(v.real) + 1j * T(v.imag)
            assert np.linalg.norm(Tv - mu[j] * v) <= tol_res * abs(mu[j]), j


def _s_dict(nchi, neta, dt):
    return {"mesh": {"Nchi": nchi, "Neta": neta}, "time_integration": {"dt": dt},
            "stability_analysis": {"perturb_shock": False,
                                   "eigenvalue_solver": "GPU_TIMESTEPPER_ARNOLDI"}}


@pytest.mark.parametrize("krylov", ["device", "arpack"])
def test_c1_eigenvalues_k10_m60(krylov):
    nchi, neta, k, dt = 100, 80, 10, 1.0 / 64
    say(f"[100x80 b4, N=32000] EIGENVALUES_GPU_ARNOLDI k=10 ncv=60 ({krylov})")
    op, lam = ops.marching_operator(nchi, neta, 4, seed=11, n_lead=12)
    V, D = st.EIGENVALUES_GPU_ARNOLDI(op, _s_dict(nchi, neta, dt), k, krylov=krylov, ncv=60,
                                      tol=1e-11, v0=np.ones(op.n))
    _check_eigs(op, lam, D, V, k, dt, 1e-8, 1e-9)


def test_c2_eigenvalues_k20_m80():
    nchi, neta, k, dt = 200, 300, 20, 1.0 / 64
    say("[200x300 b4 + 200 shock DOFs, N=240200] EIGENVALUES_GPU_ARNOLDI k=20 ncv=80 (device)")
    op, lam = ops.marching_operator(nchi, neta, 4, seed=12, n_lead=24, bulk=(-6.0, -3.0),
                                    n_shock=nchi)
    s = _s_dict(nchi, neta, dt)
    V, D = st.EIGENVALUES_GPU_ARNOLDI(op, s, k, krylov="device", ncv=80, tol=1e-11)
    say(f"[200x300 b4] {s['timings']['eigenvalues_krylov']['applications']} actions, "
        f"{s['timings']['eigenvalues']:.2f} s")
    _check_eigs(op, lam, D, V, k, dt, 1e-8, 1e-9)


def test_c2_batched_eigenvalues_k20():
    """The same 20 modes by block Krylov-Schur on 4-vector sweeps with FP64
    values (batched_exp_eigs): every expansion one batched sweep over the
    shock operator, its shock rows in the batch kernel's launch; the same
    eigenvalue and Ritz-residual checks."""
    nchi, neta, k, dt = 200, 300, 20, 1.0 / 64
    say("[200x300 b4 + 200 shock DOFs] batched_exp_eigs k=20, blocks of 4, FP64 values")
    op, lam = ops.marching_operator(nchi, neta, 4, seed=12, n_lead=24, bulk=(-6.0, -3.0),
                                    n_shock=nchi)
    s = _s_dict(nchi, neta, dt)
    V, D, info = st.batched_exp_eigs(op, s, k, block=4, value_prec="f64", tol=1e-11,
                                     maxiter=300)
    say(f"[200x300 b4] {info['phase1']['applications']} actions, {info['seconds']:.2f} s")
    _check_eigs(op, lam, D, V, k, dt, 1e-8, 1e-9)


# ---------------------------------------------------------------- transient growth

@pytest.fixture(scope="module")
def c3():
    say("[400x300 b4 + 400 shock DOFs, N=480400] operator, base flow, Chu weights")
    nchi, neta = 400, 300
    s = synthetic_solution(nchi, neta, seed=7, perturb_shock=True, cfl=100.0, max_dt=1.0 / 64)
    s = st.CFL_TIMESTEP(s)
    op = ops.config_operator("C3")
    A = op.to_scipy_varmajor()
    R = st.CONSTRUCT_R(s)
    C = (R.T @ st.CONSTRUCT_M(s, [1, 1, 1, 1]) @ R).tocsc()
    D = (R.T @ st.CONSTRUCT_M_HAT(s, [1, 1, 1, 1]) @ R).tocsc()
    return s, op, A, C, D


def test_c3_sandwich_symmetric_psd(c3):
    s, op, A, C, D = c3
    dt = s["time_integration"]["dt"]
    n_t = 16
    nchi, neta = s["mesh"]["Nchi"], s["mesh"]["Neta"]
    r = np.random.default_rng(56)
    x, y = r.standard_normal(op.n), r.standard_normal(op.n)
    say(f"[400x300 b4] Chu sandwich, n_t={n_t}: symmetry, PSD, stages")
    with DeviceOperator.from_blocks(op, scale=dt) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC:
        S = lambda v: dA.chu_sandwich(dC, v, order=5, nsub=n_t, dt=1.0)
        Sx, Sy = S(x), S(y)
        assert abs(Sx @ y - x @ Sy) <= 1e-12 * np.linalg.norm(Sx) * np.linalg.norm(y)
        for v in (x, y, x - 2 * y):
            assert v @ S(v) >= 0.0
        # the three stages run separately (forward sweep, weight, adjoint sweep)
        fx = dA.expm_taylor(x, order=5, nsub=n_t, dt=1.0, form="normalised")
        st3 = dA.expm_taylor(C @ fx, order=5, nsub=n_t, dt=1.0, form="normalised",
                             transpose=True)
        assert relerr(Sx, st3) <= 1e-12


@pytest.mark.parametrize("T", [0.5, 1.0, 2.0, 4.0])
def test_c3_transient_growth(c3, T):
    s, op, A, C, D = c3
    k, tol = 4, 1e-8
    say(f"[400x300 b4] TRANSIENT_GROWTH_DOWNSTREAM T={T} k={k} (device Krylov)")
    V, G, T_opt = st.TRANSIENT_GROWTH_DOWNSTREAM(A, s, k, T, tol=tol, krylov="device",
                                                 ncv=24, maxiter=400)
    dt_cfl = s["time_integration"]["dt"]
    n_t = int(round(T / dt_cfl))
    dt = T / n_t
    say(f"[400x300 b4] T={T}: gains {G}")
    assert np.all(np.isfinite(G)) and np.all(G > 0), G
    assert np.all(np.diff(G) <= 0) and T_opt == pytest.approx(n_t * dt, rel=1e-15)
    # D-orthonormal modes: unit D-norm to rounding; mutual D-orthogonality
    # to the accuracy of Ritz vectors of nearly equal gains
    Gm = V.T @ (D @ V)
    assert np.abs(np.diag(Gm) - 1.0).max() <= 1e-10
    assert np.abs(Gm - np.diag(np.diag(Gm))).max() <= 1e-6
    nchi, neta = s["mesh"]["Nchi"], s["mesh"]["Neta"]
    Linv, _ = st.block_cholesky_inverse(D)
    with DeviceOperator.from_blocks(op, scale=dt) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC:
        for j in range(k):
            v = V[:, j]
            Sv = dA.chu_sandwich(dC, v, order=5, nsub=n_t, dt=1.0)
            # with D = L L', the solve works on L^-1 S L^-T y = G y (v =
            # L^-T y, |y| = 1) to residual tol * G: ||L^-1 (S v - G D v)||
            res = Linv @ (Sv - G[j] * (D @ v))
            assert np.linalg.norm(res) <= 10 * tol * G[j], j


def test_c3_batched_sandwich_columns(c3):
    """The sandwich on 4 and 8 columns at once between the Cholesky factors
    of R'M_hat R (hymor_chu_sandwich_batched_dev: the shock rows and the
    shock-column slices in the batch kernel's launch) equals, column by
    column, L^-1 times the single-vector sandwich of L^-T x, to 1e-12."""
    s, op, A, C, D = c3
    dt = s["time_integration"]["dt"]
    n_t = 8
    nchi, neta = s["mesh"]["Nchi"], s["mesh"]["Neta"]
    Linv, LinvH = st.block_cholesky_inverse(D)
    say(f"[400x300 b4] batched Chu sandwich with factors, n_t={n_t}")
    with DeviceOperator.from_blocks(op, scale=dt) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC, \
            DeviceOperator.from_scipy(Linv, nchi=nchi, neta=neta) as dL, \
            DeviceOperator.from_scipy(LinvH, nchi=nchi, neta=neta) as dLH:
        for nrhs in (4, 8):
            X = np.asfortranarray(np.random.default_rng(nrhs).standard_normal((op.n, nrhs)))
            dX, dY = DeviceMatrix.from_host(X), DeviceMatrix(op.n, nrhs)
            try:
                chu_sandwich_batched_ptr(dA, dC, dX.ptr, op.n, nrhs, dY.ptr, op.n, order=5,
                                         nsub=n_t, pre=dLH, post=dL)
                dA.synchronize()
                Y = dY.download()
            finally:
                dX.free()
                dY.free()
            for j in (0, nrhs - 1):
                want = Linv @ dA.chu_sandwich(dC, LinvH @ X[:, j], order=5, nsub=n_t, dt=1.0)
                assert relerr(Y[:, j], want) <= 1e-12, (nrhs, j)


def test_c3_transient_growth_block_sizes_agree(c3):
    """Block Krylov-Schur on blocks of 4 (batched sandwiches) and the
    single-vector solver find the same 4 gains at T = 1 (tolerance 1e-10)."""
    s, op, A, C, D = c3
    say("[400x300 b4] TRANSIENT_GROWTH_DOWNSTREAM T=1, blocks of 1 and 4")
    _, G1, _ = st.TRANSIENT_GROWTH_DOWNSTREAM(A, s, 4, 1.0, tol=1e-10, krylov="device", ncv=24,
                                              maxiter=400, block=1)
    _, G4, _ = st.TRANSIENT_GROWTH_DOWNSTREAM(A, s, 4, 1.0, tol=1e-10, krylov="device", ncv=24,
                                              maxiter=400, block=4)
    assert np.all(np.abs(G4 - G1) <= 1e-9 * np.abs(G1)), (G1, G4)


def test_c3_gain_gradient_fd(c3):
    s, op, A, C, D = c3
    dt = s["time_integration"]["dt"]
    n_t = 32
    nchi, neta = s["mesh"]["Nchi"], s["mesh"]["Neta"]
    r = np.random.default_rng(57)
    v = r.standard_normal(op.n)
    say(f"[400x300 b4] gain gradient vs central differences, 16 directions, n_t={n_t}")
    with DeviceOperator.from_blocks(op, scale=dt) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC:
        F = GainFunctional(dA, dC, D, n_t)
        G, g = F.value_and_grad(v)
        assert G > 0
        # unit directions: 15 random ones (their derivatives are small,
        # about |g| / sqrt(n)) and the gradient's own (derivative |g|); the
        # difference quotient resolves them to O(h^2) + O(eps G / h), far
        # below 1e-6 |g|, while a wrong gradient misses by about |g . d|
        h = 1e-6
        gn = np.linalg.norm(g)
        dirs = [r.standard_normal(op.n) for _ in range(15)] + [g.copy()]
        for i, d in enumerate(dirs):
            d /= np.linalg.norm(d)
            fd = (F.value(v + h * d) - F.value(v - h * d)) / (2 * h)
            an = g @ d
            assert abs(fd - an) <= 1e-6 * gn, (i, fd, an)
        assert abs(dirs[-1] @ g - gn) <= 1e-12 * gn


def test_c3_operator_gradient_scaling_identity(c3):
    """J(A) = <wbar, T_5(dt A)^n x> is a function of s A for J(s) := J(sA),
    so <dJ/dA, A> = dJ/ds at s = 1, which central differences of the device
    action in the per-call dt resolve (s dt A = (s dt) A): the contraction
    of all 19.3M gradient entries with the operator's values against a
    difference quotient of two actions (1e-7 relative); checkpoints every 8
    and every 3 substeps give the same gradient (1e-13)."""
    s, op, A, C, D = c3
    dt = s["time_integration"]["dt"]
    n_t = 16
    r = np.random.default_rng(58)
    x, wbar = r.standard_normal(op.n), r.standard_normal(op.n)
    Acsc = A.tocsc()
    Acsc.sort_indices()
    colptr = Acsc.indptr.astype(np.int64)
    rowval = Acsc.indices.astype(np.int64)
    say(f"[400x300 b4] operator gradient on {Acsc.nnz} entries, n_t={n_t}: scaling identity")
    with DeviceOperator.from_scipy(Acsc, nchi=s["mesh"]["Nchi"], neta=s["mesh"]["Neta"]) as dA:
        g8 = dA.expm_taylor_vjp_values(x, wbar, colptr, rowval, order=5, nsub=n_t, dt=dt,
                                       ckpt_every=8, index_base=0)
        g3 = dA.expm_taylor_vjp_values(x, wbar, colptr, rowval, order=5, nsub=n_t, dt=dt,
                                       ckpt_every=3, index_base=0)
        assert relerr(g3, g8) <= 1e-13
        eps = 1e-4
        J = lambda sc: wbar @ dA.expm_taylor(x, order=5, nsub=n_t, dt=dt * sc, form="horner")
        dJ = (J(1 + eps) - J(1 - eps)) / (2 * eps)
        contraction = g8 @ Acsc.data
        assert abs(contraction - dJ) <= 1e-7 * abs(dJ), (contraction, dJ)
