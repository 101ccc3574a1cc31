"""Gradient of the Taylor action with respect to the operator's entries
(hymor_expm_taylor_vjp_values: the forward state kept every ckpt_every
substeps, each group of states recomputed from its checkpoint in the
backward sweep), through the C ABI:

* against the oracle's restatement (every forward state kept), itself
  pinned by central differences and divided-difference closed forms
  (tests/test_oracle_pins.py): 1e-12, on b = 4 operators with shock bands and
  b = 8 operators, orders 1, 5 and 12, checkpoint intervals from 1 to past
  the sweep, on the operator's pattern plus entries outside it;
* the gain's sensitivity dG/dA (GainFunctional.operator_gradient) against
  central differences of G over 16 random directions in the pattern and
  against the oracle's (1e-12);
* argument checks.
The 400 x 300-cell case (N = 480 400) is in test_gpu_fullsize.py."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from hymor_amd import _lib
from hymor_amd import operators as ops
from hymor_amd import stability as st
from hymor_amd.adjoint import GainFunctional
from hymor_amd.baseflow import synthetic_solution
from hymor_amd.core import DeviceOperator

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _pattern(A, extra, seed):
    """1-based CSC pattern of A plus `extra` random slots outside it."""
    n = A.shape[0]
    rng = np.random.default_rng(seed)
    P = sp.csc_matrix((np.ones(A.nnz), A.nonzero()), shape=A.shape)
    E = sp.csc_matrix((np.ones(extra), (rng.integers(0, n, extra), rng.integers(0, n, extra))),
                      shape=A.shape)
    P = sp.csc_matrix(P + E)
    P.sort_indices()
    return P.indptr.astype(np.int64) + 1, P.indices.astype(np.int64) + 1


CASES = [dict(nchi=10, neta=7, b=4, seed=61, n_shock=10),
         dict(nchi=6, neta=5, b=8, seed=62, n_shock=0),
         dict(nchi=9, neta=4, b=4, seed=63, n_shock=0, periodic_chi=True)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['nchi']}x{c['neta']}b{c['b']}")
@pytest.mark.parametrize("order,nsub,ckpt", [(5, 7, 1), (5, 7, 3), (5, 7, 20), (12, 4, 2),
                                             (1, 3, 2), (3, 19, 1)])
def test_vjp_values_vs_oracle(orc, case, order, nsub, ckpt):
    c = dict(case)
    say(f"[operator gradient] {c} order={order} nsub={nsub} ckpt_every={ckpt}")
    op = ops.stencil_operator(c.pop("nchi"), c.pop("neta"), c.pop("b"), c.pop("seed"),
                              row_sum=2.0, **c)
    A = op.to_scipy_varmajor()
    colptr, rowval = _pattern(A, 12, 7)
    r = np.random.default_rng(order + nsub)
    x, wbar = r.standard_normal(op.n), r.standard_normal(op.n)
    dt = 0.05
    with DeviceOperator.from_blocks(op) as dev:
        g = dev.expm_taylor_vjp_values(x, wbar, colptr, rowval, order=order, nsub=nsub, dt=dt,
                                       ckpt_every=ckpt)
    Ao = orc.CSC.from_scipy(A).scaled(dt)
    ref = dt * orc.taylor_vjp_values(Ao, x, wbar, nsub, order, colptr, rowval)
    assert relerr(g, ref) <= 1e-12


def test_vjp_values_zero_based_pattern_and_edges(orc):
    """A 0-based pattern gives the 1-based pattern's gradient; nsub = 0
    gives zeros (w = x does not depend on A); a complex operator and a bad
    checkpoint interval are refused."""
    op = ops.stencil_operator(7, 5, 4, 64, n_shock=7, row_sum=2.0)
    A = op.to_scipy_varmajor()
    colptr, rowval = _pattern(A, 5, 8)
    r = np.random.default_rng(9)
    x, wbar = r.standard_normal(op.n), r.standard_normal(op.n)
    with DeviceOperator.from_blocks(op) as dev:
        g1 = dev.expm_taylor_vjp_values(x, wbar, colptr, rowval, order=5, nsub=3, dt=0.05)
        g0 = dev.expm_taylor_vjp_values(x, wbar, colptr - 1, rowval - 1, order=5, nsub=3, dt=0.05,
                                        index_base=0)
        assert np.array_equal(g0, g1)
        z = dev.expm_taylor_vjp_values(x, wbar, colptr, rowval, order=5, nsub=0, dt=0.05)
        assert not z.any()
        with pytest.raises(_lib.HymorError, match="ckpt_every"):
            dev.expm_taylor_vjp_values(x, wbar, colptr, rowval, order=5, nsub=3, ckpt_every=0)
    L_ = ops.receptivity_extension(ops.stencil_operator(5, 4, 4, 65), [0.5, 1.0])
    with DeviceOperator.from_blocks(L_) as dev:
        lib = _lib.load()
        n = dev.n
        cp = np.ones(n + 1, dtype=np.int64)
        rv = np.zeros(0, dtype=np.int64)
        v = np.zeros(n)
        out = np.zeros(1)
        rc = lib.hymor_expm_taylor_vjp_values(dev.handle, 5, 1, 1.0, 8, _lib.ptr(v), _lib.ptr(v),
                                              _lib.ptr(cp), _lib.ptr(rv), 1, _lib.ptr(out))
        assert rc == _lib.ERR_ARG and b"real operators" in lib.hymor_last_error()


def test_gain_operator_gradient_fd(orc):
    """dG/dA of the transient-growth gain G = <w, C w> / <v, D v>, w = T(A)^n v
    (A the stored A*dt): against central differences of G over 16 random
    unit directions in the pattern (relative 1e-6 of |dG/dA|) and against
    the oracle's pullback of the same cotangent (1e-12)."""
    nchi, neta = 9, 7
    s = st.CFL_TIMESTEP(synthetic_solution(nchi, neta, seed=5, perturb_shock=True))
    dt = s["time_integration"]["dt"]
    base = ops.stencil_operator(nchi, neta, 4, seed=21, n_shock=nchi, row_sum=0.5 / dt)
    R = st.CONSTRUCT_R(s)
    C = (R.T @ st.CONSTRUCT_M(s, [1, 1, 1, 1]) @ R).tocsc()
    D = (R.T @ st.CONSTRUCT_M_HAT(s, [1, 1, 1, 1]) @ R).tocsc()
    As = sp.csc_matrix(base.to_scipy_varmajor() * dt)
    As.sort_indices()
    n, n_t = As.shape[0], 6
    colptr, rowval = As.indptr.astype(np.int64) + 1, As.indices.astype(np.int64) + 1
    rows = rowval - 1
    cols = np.repeat(np.arange(n), np.diff(colptr))
    r = np.random.default_rng(66)
    v = r.standard_normal(n)
    say("[operator gradient] gain dG/dA vs central differences")

    def gain(M):
        with DeviceOperator.from_scipy(M, nchi=nchi, neta=neta) as dA, \
                DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC:
            return GainFunctional(dA, dC, D, n_t).value(v)

    with DeviceOperator.from_scipy(As, nchi=nchi, neta=neta) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC:
        G, g = GainFunctional(dA, dC, D, n_t).operator_gradient(v, colptr, rowval,
                                                                ckpt_every=4)
    # the oracle's pullback of the cotangent 2 C w / <v, D v>
    Ao = orc.CSC.from_scipy(As)
    w = orc.power_method(Ao, v, n_t, 5)
    den = v @ (D @ v)
    assert abs(G - w @ (C @ w) / den) <= 1e-12 * G
    ref = orc.taylor_vjp_values(Ao, v, 2 * (C @ w) / den, n_t, 5, colptr, rowval)
    assert relerr(g, ref) <= 1e-12
    h = 1e-6
    gn = np.linalg.norm(g)
    for i in range(16):
        d = r.standard_normal(rows.shape[0])
        d /= np.linalg.norm(d)
        Dm = sp.csc_matrix((d, (rows, cols)), shape=As.shape)
        fd = (gain(As + h * Dm) - gain(As - h * Dm)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-6 * gn, (i, fd, g @ d)


def test_vjp_values_on_a_device_pattern(orc):
    """hymor_pattern_create + hymor_expm_taylor_vjp_values_dev (pattern,
    vectors and gradient on the device) give the host entry point's bits, on
    repeated calls with the same pattern and different vectors (the gradient
    array overwritten each time); GainFunctional.operator_gradient_dev gives
    operator_gradient's G and dG/dA; a pattern of another shape is refused."""
    from hymor_amd.core import DevicePattern, DeviceVector
    op = ops.stencil_operator(10, 7, 4, 67, n_shock=10, row_sum=2.0)
    A = op.to_scipy_varmajor()
    colptr, rowval = _pattern(A, 9, 10)
    say("[operator gradient] device-resident pattern")
    with DeviceOperator.from_blocks(op) as dev, DevicePattern(dev, colptr, rowval) as pat:
        assert pat.nnz == rowval.shape[0]
        dg = DeviceVector(pat.nnz)
        try:
            for seed in (1, 2):
                r = np.random.default_rng(seed)
                x, wbar = r.standard_normal(op.n), r.standard_normal(op.n)
                dx, dw = DeviceVector.from_host(x), DeviceVector.from_host(wbar)
                try:
                    dev.expm_taylor_vjp_values_dev(pat, dx, dw, dg, order=5, nsub=7, dt=0.05,
                                                   ckpt_every=3)
                finally:
                    dx.free()
                    dw.free()
                want = dev.expm_taylor_vjp_values(x, wbar, colptr, rowval, order=5, nsub=7,
                                                  dt=0.05, ckpt_every=3)
                np.testing.assert_array_equal(dg.download(), want)
        finally:
            dg.free()
    other = ops.stencil_operator(9, 7, 4, 68, n_shock=9)
    with DeviceOperator.from_blocks(op) as dev, DeviceOperator.from_blocks(other) as dev2, \
            DevicePattern(dev, colptr, rowval) as pat:
        v = DeviceVector(dev2.n)
        try:
            with pytest.raises(_lib.HymorError, match="another shape"):
                dev2.expm_taylor_vjp_values_dev(pat, v, v, v, order=5, nsub=2)
        finally:
            v.free()
    # the gain's dG/dA through the device pattern
    nchi, neta = 9, 7
    s = st.CFL_TIMESTEP(synthetic_solution(nchi, neta, seed=5, perturb_shock=True))
    dt = s["time_integration"]["dt"]
    base = ops.stencil_operator(nchi, neta, 4, seed=21, n_shock=nchi, row_sum=0.5 / dt)
    R = st.CONSTRUCT_R(s)
    C = (R.T @ st.CONSTRUCT_M(s, [1, 1, 1, 1]) @ R).tocsc()
    D = (R.T @ st.CONSTRUCT_M_HAT(s, [1, 1, 1, 1]) @ R).tocsc()
    As = sp.csc_matrix(base.to_scipy_varmajor() * dt)
    As.sort_indices()
    cp, rv = As.indptr.astype(np.int64) + 1, As.indices.astype(np.int64) + 1
    v = np.random.default_rng(66).standard_normal(As.shape[0])
    with DeviceOperator.from_scipy(As, nchi=nchi, neta=neta) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC, \
            DevicePattern(dA, cp, rv) as pat:
        F = GainFunctional(dA, dC, D, 6)
        G, g = F.operator_gradient(v, cp, rv, ckpt_every=4)
        dg = DeviceVector(pat.nnz)
        try:
            G2 = F.operator_gradient_dev(v, pat, dg, ckpt_every=4)
            assert G2 == G
            np.testing.assert_array_equal(dg.download(), g)
            # every state kept: nothing recomputed; the entries' sums are
            # grouped over other segments of states, so equal to rounding
            F.operator_gradient_dev(v, pat, dg, ckpt_every=1)
            assert relerr(dg.download(), g) <= 1e-12
        finally:
            dg.free()

@pytest.mark.parametrize("b,n_shock", [(8, 0), (4, 9)])
def test_device_pattern_blocks_give_the_per_entry_bits(orc, b, n_shock):
    """A device pattern groups its flow entries by b x b block (b 4 and 8);
    entries of blocks holding fewer than b, those touching extra DOFs and a
    repeated (row, column) stay per entry.  The gradient equals the host
    entry point's (every entry per entry) bit for bit, and the oracle's."""
    from hymor_amd.core import DevicePattern, DeviceVector
    op = ops.stencil_operator(9, 6, b, 71, n_shock=n_shock, row_sum=2.0)
    A = op.to_scipy_varmajor()
    colptr, rowval = _pattern(A, 40, 11)
    # repeat the first entry of column 3
    j = 3
    e0 = colptr[j] - 1
    rowval = np.insert(rowval, e0, rowval[e0])
    colptr = colptr.copy()
    colptr[j + 1:] += 1
    say(f"[operator gradient] block map b={b} n_shock={n_shock}")
    r = np.random.default_rng(b + n_shock)
    x, wbar = r.standard_normal(op.n), r.standard_normal(op.n)
    with DeviceOperator.from_blocks(op) as dev, DevicePattern(dev, colptr, rowval) as pat:
        want = dev.expm_taylor_vjp_values(x, wbar, colptr, rowval, order=4, nsub=5, dt=0.05,
                                          ckpt_every=2)
        dx, dw = DeviceVector.from_host(x), DeviceVector.from_host(wbar)
        dg = DeviceVector(pat.nnz)
        try:
            dev.expm_taylor_vjp_values_dev(pat, dx, dw, dg, order=4, nsub=5, dt=0.05,
                                           ckpt_every=2)
            got = dg.download()
        finally:
            for v in (dx, dw, dg):
                v.free()
    np.testing.assert_array_equal(got, want)
    assert got[e0] == got[e0 + 1]
    Ao = orc.CSC.from_scipy(A).scaled(0.05)
    ref = 0.05 * orc.taylor_vjp_values(Ao, x, wbar, 5, 4, colptr, rowval)
    assert relerr(got, ref) <= 1e-12
