"""Device Krylov basis: block Gram-Schmidt (V^T W on the matrix cores)
against a float64 numpy reference, and the batched eigensolver on an operator
with an exactly known spectrum."""
import math
import sys

import numpy as np
import pytest

from hymor_amd import operators as ops
from hymor_amd import stability as st
from hymor_amd.core import DeviceMatrix
from hymor_amd.krylov import DeviceKrylov

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,j,s", [(1440, 0, 8), (1440, 5, 8), (1003, 37, 16), (100003, 20, 3),
                                   (4096, 63, 1), (998, 16, 16)])
def test_block_orthonormalize(n, j, s):
    """Columns j..j+s-1 come out orthonormal to the basis and each other,
    W = V[:, :j+s] H holds, and the top j rows of H are V^T W (CGS2 sums).
    Odd n exercises the unaligned loads, n % 16 != 0 the tail rows."""
    say(f"[block gs] n={n} j={j} s={s}")
    r = np.random.default_rng(n + j + s)
    K = DeviceKrylov(n, j + s + 1)
    try:
        Q, _ = np.linalg.qr(r.standard_normal((n, max(j, 1))))
        for c in range(j):
            K.set_column(c, Q[:, c])
        W0 = r.standard_normal((n, s))
        W0[:, -1] += Q[:, : min(j, 3)].sum(axis=1) * 5.0 if j else 0.0   # a component inside span(V)
        dW = DeviceMatrix.from_host(W0)
        H = K.orthonormalize_block(j, dW.ptr, n, s)
        dW.free()
        V = K.get_columns(0, j + s)
        assert H.shape == (j + s, s)
        G = V.T @ V
        assert np.abs(G - np.eye(j + s)).max() <= 1e-13
        assert np.abs(V @ H - W0).max() <= 1e-12 * np.abs(W0).max()
        if j:
            assert np.allclose(H[:j], Q[:, :j].T @ W0, rtol=0, atol=1e-12 * np.abs(W0).max())
        assert np.all(np.diag(H[j:]) > 0) and np.allclose(np.tril(H[j:], -1), 0.0)
    finally:
        K.close()


@pytest.mark.parametrize("n", [1003, 100003])
def test_dependent_vectors_are_flagged(n):
    """A vector inside span(V) (exactly, or up to a rounding-level part) is
    reported dependent -- norm 0 and a zero column -- by both
    orthonormalisations; a small but genuine new direction (1e-8 relative)
    is normalised as usual."""
    say(f"[dependence] n={n}")
    r = np.random.default_rng(n)
    j = 3
    Q, _ = np.linalg.qr(r.standard_normal((n, j)))
    coef = np.array([1.0, 2.0, -1.5])
    cases = [("exact", Q @ coef, True),
             ("rounding", Q @ coef + 1e-17 * r.standard_normal(n), True),
             ("small", Q @ coef + 1e-8 * r.standard_normal(n) / np.sqrt(n), False)]
    for name, w, dep in cases:
        K = DeviceKrylov(n, j + 2)
        try:
            for c in range(j):
                K.set_column(c, Q[:, c])
            K.set_column(j + 1, w)
            K.set_column(j, np.full(n, 7.0))            # stale contents of column j
            h = K.orthonormalize(j, K.column(j + 1))
            col = K.get_columns(j, 1)[:, 0]
            assert np.allclose(h[:j], Q.T @ w, rtol=0, atol=1e-12), name
            if dep:
                assert h[j] == 0.0 and not np.any(col), name
            else:
                assert h[j] > 0.0 and abs(np.linalg.norm(col) - 1.0) <= 1e-12, name
                assert np.abs(Q.T @ col).max() <= 1e-12, name
        finally:
            K.close()
    # block form: one column inside span(V), one repeating an earlier column
    s = 4
    W0 = r.standard_normal((n, s))
    W0[:, 1] = Q @ coef
    W0[:, 3] = 2.0 * W0[:, 0]
    K = DeviceKrylov(n, j + s + 1)
    try:
        for c in range(j):
            K.set_column(c, Q[:, c])
        dW = DeviceMatrix.from_host(W0)
        H = K.orthonormalize_block(j, dW.ptr, n, s)
        dW.free()
        d = np.diag(H[j:])
        assert d[1] == 0.0 and d[3] == 0.0 and d[0] > 0 and d[2] > 0
        V = K.get_columns(j, s)
        assert not np.any(V[:, 1]) and not np.any(V[:, 3])
        assert abs(np.linalg.norm(V[:, 0]) - 1) <= 1e-12 and abs(np.linalg.norm(V[:, 2]) - 1) <= 1e-12
    finally:
        K.close()


def test_batched_eigs_exact_spectrum():
    """Batched block Krylov-Schur (8 vectors per sweep) on FP64 values, and
    on FP32 values with FP64 refinement, against the exact spectrum of a
    one-way (block triangular) operator; plain FP64 Krylov-Schur alongside."""
    nchi, neta, k = 40, 30, 10
    op, lam = ops.marching_operator(nchi, neta, 8, seed=3)
    dt = 1.0 / 64
    n_t = 32
    mu = sum((lam * dt) ** j / math.factorial(j) for j in range(6)) ** n_t
    exact = np.log(mu[np.argsort(-np.abs(mu), kind="stable")][: 3 * k]) / (n_t * dt)

    def err(D):
        return max(np.min(np.abs(exact - d)) / abs(d) for d in D)

    def sd():
        return {"mesh": {"Nchi": nchi, "Neta": neta}, "time_integration": {"dt": dt},
                "stability_analysis": {"perturb_shock": False,
                                       "eigenvalue_solver": "GPU_TIMESTEPPER_ARNOLDI"}}
    say("[batched eigs] marching operator 40x30 b=8")
    _, D, _ = st.batched_exp_eigs(op, sd(), k, block=8, value_prec="f64", tol=1e-11)
    assert err(D) <= 1e-10
    _, D, info = st.batched_exp_eigs(op, sd(), k, block=8, value_prec="f32", tol=1e-11)
    assert err(D) <= 1e-10 and "phase2" in info
    _, D, _ = st.batched_exp_eigs(op, sd(), k, block=8, value_prec="f32", tol=1e-11,
                                  refine=False)
    assert 1e-12 < err(D) <= 1e-6          # FP32 storage alone: about its rounding
    _, D = st.EIGENVALUES_GPU_ARNOLDI(op, sd(), k, krylov="device", tol=1e-11, ncv=25)
    assert err(D) <= 1e-10


def test_device_uniform_fill():
    """Start blocks made on the device: values in [-1, 1), a function of
    (seed, index) only, mean and variance of the uniform distribution."""
    import ctypes as C
    from hymor_amd import _lib
    from hymor_amd.core import DeviceMatrix
    n = 1 << 20
    d = DeviceMatrix(n, 3)
    try:
        lib = _lib.load()
        for j, seed in enumerate((7, 7, 8)):
            _lib.check(lib.hymor_dev_fill_uniform(C.c_void_p(d.ptr.value + j * n * 8), n, seed))
        X = d.download()
    finally:
        d.free()
    assert np.all(X >= -1.0) and np.all(X < 1.0)
    np.testing.assert_array_equal(X[:, 0], X[:, 1])
    assert np.mean(X[:, 0] == X[:, 2]) < 1e-5
    assert abs(X[:, 0].mean()) < 5e-3 and abs(X[:, 0].var() - 1 / 3) < 5e-3
    assert abs(np.corrcoef(X[:, 0], X[:, 2])[0, 1]) < 5e-3
