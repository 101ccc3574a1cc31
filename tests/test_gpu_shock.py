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
