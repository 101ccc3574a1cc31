"""The shock bands in the single-vector term kernel's launch (k_term_split_p
with the extra-DOF rows in leading workgroups and the flow rows' shock
columns as dense slice panels), through the C ABI: SpMV and Taylor actions,
forward and adjoint, FP64 and FP32 values, against the oracle's
restatement and against the two-launch term (block kernel, then k_extra)
that a nine-wave tuning runs; captured sweeps replay the same bits."""
import sys

import numpy as np
import pytest

from hymor_amd import operators as ops
from hymor_amd.core import DeviceOperator

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


CASES = [dict(nchi=10, neta=7, b=4, seed=101), dict(nchi=6, neta=5, b=8, seed=102),
         dict(nchi=40, neta=9, b=4, seed=103), dict(nchi=7, neta=3, b=8, seed=104)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['nchi']}x{c['neta']}b{c['b']}")
def test_shock_bands_in_the_term_launch(orc, case, prec):
    c = dict(case)
    say(f"[shock bands fused] {c} values={prec}")
    op = ops.stencil_operator(c["nchi"], c["neta"], c["b"], c["seed"], n_shock=c["nchi"],
                              row_sum=2.0)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    At = orc.csc_adjoint(A)
    r = np.random.default_rng(c["seed"])
    x = r.standard_normal(op.n)
    dt, nsub = 0.05, 4
    tol_spmv, tol = (1e-14, 1e-12) if prec == "f64" else (2e-6, 2e-6)
    with DeviceOperator.from_blocks(op, value_prec=prec) as dev:
        assert dev.tuning == (3, 1, 1)
        fused = {}
        for tr, M in ((False, A), (True, At)):
            y = dev.spmv(x, transpose=tr)
            assert relerr(y, orc.csc_mul(M, x)) <= tol_spmv, tr
            fused[(tr, "spmv")] = y
            for form in ("powers", "horner", "normalised"):
                got = dev.expm_taylor(x, order=5, nsub=nsub, dt=dt, form=form, transpose=tr)
                ref = orc.power_method(M.scaled(dt), x, nsub)
                assert relerr(got, ref) <= tol, (tr, form)
                fused[(tr, form)] = got
        dev.set_graphs(True)
        for _ in range(3):
            assert np.array_equal(dev.expm_taylor(x, order=5, nsub=nsub, dt=dt, form="horner"),
                                  fused[(False, "horner")])
        dev.set_graphs(False)
        # the two-launch term: nine waves per slice keep k_extra
        dev.set_tuning(3, 9, 1)
        for tr in (False, True):
            assert relerr(dev.spmv(x, transpose=tr), fused[(tr, "spmv")]) <= 1e-13, tr
            for form in ("powers", "horner", "normalised"):
                two = dev.expm_taylor(x, order=5, nsub=nsub, dt=dt, form=form, transpose=tr)
                assert relerr(two, fused[(tr, form)]) <= 1e-13, (tr, form)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("nrhs", [2, 3, 4, 8, 16, 32])
@pytest.mark.parametrize("case", [(24, 30, 4), (70, 6, 4), (70, 3, 8)],
                         ids=lambda c: f"{c[0]}x{c[1]}b{c[2]}")
def test_shock_bands_in_batched_terms(orc, case, nrhs, prec):
    """Batched sweeps over shock operators with the extra-DOF rows in the
    batch kernel's launch (k_term_lds at 2 and 4 doubles per row,
    k_spmm_quad4c at 8+ at b = 4; 216-1190 entries per row) and the flow
    rows' shock entries in a k_extra launch after it.  Every column against
    the oracle, forward and adjoint (FP32-stored values: within their
    rounding band); the two-launch term (a nine-wave tuning at 2 and 4
    doubles, batch kernel 7 at 8+) to 1e-13, with FP32 values too (the whole
    rows' block entries carry the block kernel's FP32 rounding); graph
    replays give the eager bits."""
    nchi, neta, b = case
    say(f"[shock bands batched] {nchi}x{neta} b={b} nrhs={nrhs} values={prec}")
    op = ops.stencil_operator(nchi, neta, b, 105 + nchi + b, n_shock=nchi, row_sum=2.0)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    At = orc.csc_adjoint(A)
    X = np.asfortranarray(np.random.default_rng(nrhs).standard_normal((op.n, nrhs)))
    kw = dict(order=5, nsub=3, dt=0.05, form="horner")
    tol = 1e-12 if prec == "f64" else 2e-6
    with DeviceOperator.from_blocks(op, value_prec=prec) as dev:
        got = {}
        for tr, M in ((False, A), (True, At)):
            Y = dev.expm_taylor_batched(X, transpose=tr, **kw)
            Y0 = dev.expm_taylor_batched(X, order=5, nsub=0, dt=0.05, transpose=tr)
            np.testing.assert_array_equal(Y0, X)
            for j in range(nrhs):
                ref = orc.power_method(M.scaled(0.05), X[:, j], 3)
                assert relerr(Y[:, j], ref) <= tol, (tr, j)
            got[tr] = Y
        dev.set_graphs(True)
        for _ in range(3):
            for tr in (False, True):
                np.testing.assert_array_equal(dev.expm_taylor_batched(X, transpose=tr, **kw),
                                              got[tr])
        dev.set_graphs(False)
        if nrhs <= 4:
            dev.set_tuning(3, 9, 1)
        else:
            dev.set_batch_kernel(7)
        for tr in (False, True):
            two = dev.expm_taylor_batched(X, transpose=tr, **kw)
            assert relerr(two, got[tr]) <= 1e-13, tr


@pytest.mark.parametrize("nrhs", [2, 3, 4, 8])
@pytest.mark.parametrize("b", [4, 8])
def test_shock_rows_at_batched_widths(orc, b, nrhs):
    """Batched sweeps over a shock operator whose shock rows are long (30
    eta stations: 216 entries at b = 4): the extra rows at 2..8 doubles per
    row take one lane per entry (8 doubles at b = 8: the tile row layout);
    every column against the oracle, forward and adjoint, SpMV and Horner
    action."""
    nchi, neta = 24, 30
    say(f"[shock rows batched] b={b} nrhs={nrhs}")
    op = ops.stencil_operator(nchi, neta, b, 105 + b, n_shock=nchi, row_sum=2.0)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    At = orc.csc_adjoint(A)
    X = np.asfortranarray(np.random.default_rng(nrhs).standard_normal((op.n, nrhs)))
    with DeviceOperator.from_blocks(op) as dev:
        for tr, M in ((False, A), (True, At)):
            Y = dev.expm_taylor_batched(X, order=5, nsub=3, dt=0.05, form="horner", transpose=tr)
            Y0 = dev.expm_taylor_batched(X, order=5, nsub=0, dt=0.05, transpose=tr)
            np.testing.assert_array_equal(Y0, X)
            for j in range(nrhs):
                ref = orc.power_method(M.scaled(0.05), X[:, j], 3)
                assert relerr(Y[:, j], ref) <= 1e-12, (tr, j)
