"""GPU parity: libhymor_hip.so (through the C ABI) against the CPU oracle.

Tolerances (relative 2-norm unless stated):
  * FP64 storage vs the FP64 oracle: 1e-12 for a few substeps, 1e-11 for the
    768-SpMV action on the 100 x 80-cell, b = 4 operator (summation order and FMA differ from the
    reference's column scatter, so agreement is to rounding, not bits);
  * FP32 storage (FP64 accumulation): 2e-6 -- the values themselves carry a
    2^-24 relative rounding;
  * true exponential of a block-circulant operator vs its FFT closed form:
    1e-10 (the accuracy target for exp(A*T)v).
"""
import sys

import numpy as np
import pytest

from hymor_amd import operators as ops
from hymor_amd.core import (DeviceMatrix, DeviceOperator, DeviceVector, chu_sandwich_batched_ptr,
                            expm_taylor_dev)

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = {
    "grid20x16_b4": dict(nchi=20, neta=16, b=4, seed=1),
    "shock12x10_b4": dict(nchi=12, neta=10, b=4, seed=2, n_shock=12),
    "grid10x8_b8": dict(nchi=10, neta=8, b=8, seed=3),
    "periodic9x6_b4": dict(nchi=9, neta=6, b=4, seed=4, periodic_chi=True),
    "shock7x5_b8": dict(nchi=7, neta=5, b=8, seed=5, n_shock=7),
}


def make_case(name, row_sum=2.0):
    c = dict(CASES[name])
    return ops.stencil_operator(c.pop("nchi"), c.pop("neta"), c.pop("b"), c.pop("seed"),
                                row_sum=row_sum, **c)


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request, orc):
    say(f"[case {request.param}] building operator")
    op = make_case(request.param)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    dev = DeviceOperator.from_csc(op.n, A.colptr, A.rowval, A.nzval, index_base=1,
                                  vars_per_cell=op.b, nchi=op.nchi, neta=op.neta)
    yield request.param, op, A, dev
    dev.close()


def test_spmv_forward_and_adjoint(case, orc):
    name, op, A, dev = case
    x = np.random.default_rng(0).standard_normal(op.n)
    say(f"[{name}] spmv")
    assert relerr(dev.spmv(x), orc.csc_mul(A, x)) <= 1e-14
    At = orc.csc_adjoint(A)
    assert relerr(dev.spmv(x, transpose=True), orc.csc_mul(At, x)) <= 1e-14


def test_power_method_parity(case, orc):
    name, op, A, dev = case
    x = np.random.default_rng(1).standard_normal(op.n)
    dt = 0.1
    say(f"[{name}] POWER_METHOD! order 5 x 7 substeps")
    ref = orc.power_method(A.scaled(dt), x, 7, order=5)
    got = dev.expm_taylor(x, order=5, nsub=7, dt=dt, form="powers")
    assert relerr(got, ref) <= 1e-12


def test_adjoint_taylor_normalised_form(case, orc):
    name, op, A, dev = case
    x = np.random.default_rng(2).standard_normal(op.n)
    dt = 0.1
    say(f"[{name}] adjoint Taylor, normalised recursion")
    ref = orc.power_method(orc.csc_adjoint(A).scaled(dt), x, 4, order=6)
    got = dev.expm_taylor(x, order=6, nsub=4, dt=dt, form="normalised", transpose=True)
    assert relerr(got, ref) <= 1e-12


def test_bsr_route_equals_csc_route(case):
    name, op, A, dev = case
    say(f"[{name}] create_bsr vs create_csc")
    x = np.random.default_rng(3).standard_normal(op.n)
    with DeviceOperator.from_blocks(op) as d2:
        assert np.array_equal(d2.expm_taylor(x, order=5, nsub=3, dt=0.1),
                              dev.expm_taylor(x, order=5, nsub=3, dt=0.1))
        assert np.array_equal(d2.spmv(x, transpose=True), dev.spmv(x, transpose=True))


def test_device_resident_path_matches_host_path(case):
    name, op, A, dev = case
    say(f"[{name}] device-resident action")
    x = np.random.default_rng(4).standard_normal(op.n)
    dx = DeviceVector.from_host(x)
    dy = DeviceVector(op.n)
    expm_taylor_dev(dev, dx, dy, order=5, nsub=5, dt=0.1)
    dev.synchronize()
    assert np.array_equal(dy.download(), dev.expm_taylor(x, order=5, nsub=5, dt=0.1))


def _chu_weight(op, seed):
    """A 4x4 (b x b) block-diagonal SPD weight shaped like R'MR, plus a
    diagonal on the extra DOFs."""
    rng = np.random.default_rng(seed)
    b = op.b
    G = rng.standard_normal((op.ncell, b, b))
    blocks = np.einsum("cij,ckj->cik", G, G) + b * np.eye(b)
    brow_ptr = np.arange(op.ncell + 1, dtype=np.int64)
    bcol = np.arange(op.ncell, dtype=np.int32)
    xr = op.b * op.ncell + np.arange(op.n_extra)
    return ops.BlockOperator(op.nchi, op.neta, b, op.n_extra, brow_ptr, bcol, blocks,
                             xr.astype(np.int64), xr.astype(np.int64), rng.uniform(0.5, 1.5, op.n_extra))


def test_chu_sandwich_parity(case, orc):
    name, op, A, dev = case
    say(f"[{name}] EXPONENTIAL! sandwich")
    W = _chu_weight(op, 5)
    Cw = orc.CSC.from_arrays(*W.to_csc_varmajor())
    dt, nt = 0.1, 3
    x = np.random.default_rng(6).standard_normal(op.n)
    At = orc.csc_adjoint(A)
    ref = orc.construct_c(x, A.scaled(dt), At.scaled(dt), Cw, nt, False, 0)
    with DeviceOperator.from_csc(op.n, Cw.colptr, Cw.rowval, Cw.nzval, vars_per_cell=op.b,
                                 nchi=op.nchi, neta=op.neta) as dW:
        got = dev.chu_sandwich(dW, x, order=5, nsub=nt, dt=dt)
        assert relerr(got, ref) <= 1e-12
        n0 = op.b * op.ncell // 2
        ref0 = orc.construct_c(x, A.scaled(dt), At.scaled(dt), Cw, nt, True, n0)
        got0 = dev.chu_sandwich(dW, x, order=5, nsub=nt, dt=dt, n_set0=n0)
        assert not np.any(got0[:n0])
        assert relerr(got0, ref0) <= 1e-12


def _batched_sandwich_ref(orc, A, dt, nt, Cw, X, n0, Pin=None, Pout=None):
    At = orc.csc_adjoint(A)
    out = []
    for j in range(X.shape[1]):
        x = X[:, j] if Pin is None else orc.csc_mul(Pin, X[:, j])
        y = orc.construct_c(x, A.scaled(dt), At.scaled(dt), Cw, nt, n0 > 0, n0)
        out.append(y if Pout is None else orc.csc_mul(Pout, y))
    return np.stack(out, axis=1)


@pytest.mark.parametrize("nrhs", [3, 8])
def test_chu_sandwich_batched_parity(case, orc, nrhs):
    """hymor_chu_sandwich_batched_dev against the oracle's CONSTRUCT_C per
    column: with and without the two factor operators, the zeroing on a
    var-major prefix that splits cells (so it lands inside cell-major rows),
    3 columns (padded width 4) and 8 (the tile row layout at b = 8)."""
    name, op, A, dev = case
    say(f"[{name}] batched sandwich nrhs={nrhs}")
    W = _chu_weight(op, 5)
    Cw = orc.CSC.from_arrays(*W.to_csc_varmajor())
    P1 = orc.CSC.from_arrays(*_chu_weight(op, 13).to_csc_varmajor())
    P2 = orc.CSC.from_arrays(*_chu_weight(op, 14).to_csc_varmajor())
    dt, nt = 0.1, 3
    X = np.asfortranarray(np.random.default_rng(15).standard_normal((op.n, nrhs)))
    n0 = op.b * op.ncell // 2 + 1
    mk = lambda M: DeviceOperator.from_csc(op.n, M.colptr, M.rowval, M.nzval, vars_per_cell=op.b,
                                           nchi=op.nchi, neta=op.neta)
    with mk(Cw) as dW, mk(P1) as dP1, mk(P2) as dP2:
        dX = DeviceMatrix.from_host(X)
        dY = DeviceMatrix(op.n, nrhs)
        try:
            for pre, post, z in ((None, None, 0), (None, None, n0), (dP1, dP2, n0),
                                 (dP1, None, 0)):
                chu_sandwich_batched_ptr(dev, dW, dX.ptr, op.n, nrhs, dY.ptr, op.n, order=5,
                                         nsub=nt, dt=dt, n_set0=z, pre=pre, post=post)
                dev.synchronize()
                got = dY.download()
                ref = _batched_sandwich_ref(orc, A, dt, nt, Cw, X, z,
                                            P1 if pre is not None else None,
                                            P2 if post is not None else None)
                for j in range(nrhs):
                    assert relerr(got[:, j], ref[:, j]) <= 1e-12, (pre, post, z, j)
                if z and post is None:
                    assert not np.any(got[:z])
        finally:
            dX.free()
            dY.free()


def test_f32_storage_error_bound(case, orc):
    name, op, A, dev = case
    say(f"[{name}] FP32-stored values")
    x = np.random.default_rng(7).standard_normal(op.n)
    with DeviceOperator.from_blocks(op, value_prec="f32") as d32:
        ref = orc.power_method(A.scaled(0.1), x, 5)
        err = relerr(d32.expm_taylor(x, order=5, nsub=5, dt=0.1), ref)
        assert 1e-12 < err <= 2e-6       # really FP32 values, bounded error


def test_mixed_widths_on_one_handle(case):
    """Plain and batched calls interleaved on one handle give the bits a
    fresh handle gives: the extra rows' partial sums of one width or layout
    never leak into another (shock rows exercise them)."""
    name, op, A, dev = case
    say(f"[{name}] mixed widths")
    r = np.random.default_rng(12)
    x = r.standard_normal(op.n)
    X = np.asfortranarray(r.standard_normal((op.n, 8)))
    X2 = np.asfortranarray(r.standard_normal((op.n, 2)))
    kw = dict(order=5, nsub=2, dt=0.1)
    with DeviceOperator.from_blocks(op) as fresh:
        refb = fresh.expm_taylor_batched(X, **kw)
    with DeviceOperator.from_blocks(op) as fresh:
        ref2 = fresh.expm_taylor_batched(X2, **kw)
    with DeviceOperator.from_blocks(op) as fresh:
        ref1 = fresh.expm_taylor(x, **kw)
    with DeviceOperator.from_blocks(op) as d:
        for _ in range(2):
            np.testing.assert_array_equal(d.expm_taylor_batched(X, **kw), refb)
            np.testing.assert_array_equal(d.expm_taylor(x, **kw), ref1)
            np.testing.assert_array_equal(d.expm_taylor_batched(X2, **kw), ref2)


def test_value_prec_switch_matches_fresh_upload(case):
    """An FP64 upload switched to FP32 reads the same rounded values, in the
    same interleave, as an FP32 upload: identical bits for the plain and the
    batched (tile-order) actions; switching back restores the FP64 bits."""
    name, op, A, dev = case
    say(f"[{name}] value precision switch")
    r = np.random.default_rng(11)
    x = r.standard_normal(op.n)
    X = np.asfortranarray(r.standard_normal((op.n, 8)))
    kw = dict(order=5, nsub=3, dt=0.1)
    with DeviceOperator.from_blocks(op, value_prec="f64") as d64, \
            DeviceOperator.from_blocks(op, value_prec="f32") as d32:
        ref64 = d64.expm_taylor(x, **kw)
        ref64t = d64.expm_taylor(x, transpose=True, **kw)
        ref64b = d64.expm_taylor_batched(X, **kw)
        ref32 = d32.expm_taylor(x, **kw)
        ref32t = d32.expm_taylor(x, transpose=True, **kw)
        ref32b = d32.expm_taylor_batched(X, **kw)
        bytes0 = d64.device_bytes
        d64.set_value_prec("f32")
        assert d64.value_prec == "f32" and d64.device_bytes > bytes0
        np.testing.assert_array_equal(d64.expm_taylor(x, **kw), ref32)
        np.testing.assert_array_equal(d64.expm_taylor(x, transpose=True, **kw), ref32t)
        np.testing.assert_array_equal(d64.expm_taylor_batched(X, **kw), ref32b)
        d64.set_value_prec("f64")
        assert d64.value_prec == "f64"
        np.testing.assert_array_equal(d64.expm_taylor(x, **kw), ref64)
        np.testing.assert_array_equal(d64.expm_taylor(x, transpose=True, **kw), ref64t)
        np.testing.assert_array_equal(d64.expm_taylor_batched(X, **kw), ref64b)
        d64.set_value_prec("f32")      # the cast copy is reused
        np.testing.assert_array_equal(d64.expm_taylor(x, **kw), ref32)
        with pytest.raises(RuntimeError, match="FP32"):
            d32.set_value_prec("f64")


@pytest.mark.parametrize("nrhs", [1, 2, 4, 8, 16])
def test_horner_form(case, orc, nrhs):
    """The Horner form evaluates POWER_METHOD!'s polynomial: the oracle's
    powers-form action on every column at 1e-12, forward and adjoint; graph
    replays give the eager bits."""
    name, op, A, dev = case
    say(f"[{name}] Horner form, {nrhs} vector(s)")
    X = np.asfortranarray(np.random.default_rng(31).standard_normal((op.n, nrhs)))
    kw = dict(order=5, nsub=3, dt=0.1, form="horner")
    for tr, M in ((False, A), (True, orc.csc_adjoint(A))):
        if nrhs == 1:
            got = dev.expm_taylor(X[:, 0], transpose=tr, **kw)[:, None]
        else:
            got = dev.expm_taylor_batched(X, transpose=tr, **kw)
        for j in range(nrhs):
            assert relerr(got[:, j], orc.power_method(M.scaled(0.1), X[:, j], 3)) <= 1e-12
    with DeviceOperator.from_blocks(op) as g:
        g.set_graphs(True)
        want = g.expm_taylor_batched(X, **kw) if nrhs > 1 else g.expm_taylor(X[:, 0], **kw)
        for _ in range(3):
            again = g.expm_taylor_batched(X, **kw) if nrhs > 1 else g.expm_taylor(X[:, 0], **kw)
            np.testing.assert_array_equal(again, want)


# ------------------------------------------------------------------ complex (receptivity)

@pytest.fixture(scope="module")
def recept(orc):
    say("[receptivity] building extended complex operator")
    base = ops.stencil_operator(10, 8, 4, seed=8, n_shock=10, row_sum=2.0)
    L_ = ops.receptivity_extension(base, [0.0, 0.7, 2.1, 5.0])
    A = orc.CSC.from_arrays(*L_.to_csc_varmajor())
    dev = DeviceOperator.from_csc(L_.n, A.colptr, A.rowval, A.nzval, vars_per_cell=4,
                                  nchi=10, neta=8)
    yield base, L_, A, dev
    dev.close()


def _cvec(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def test_complex_spmv_and_taylor(recept, orc):
    base, L_, A, dev = recept
    x = _cvec(L_.n, 9)
    say("[receptivity] complex spmv / Taylor")
    assert relerr(dev.spmv(x), orc.csc_mul(A, x)) <= 1e-14
    assert relerr(dev.expm_taylor(x, order=5, nsub=3, dt=0.1, form="horner"),
                  orc.power_method(A.scaled(0.1), x, 3)) <= 1e-12
    assert relerr(dev.spmv(x, transpose=True), orc.csc_mul(orc.csc_adjoint(A), x)) <= 1e-14
    ref = orc.power_method(A.scaled(0.05), x, 6)
    assert relerr(dev.expm_taylor(x, order=5, nsub=6, dt=0.05), ref) <= 1e-12


def test_complex_sandwich_with_freestream_zeroing(recept, orc):
    base, L_, A, dev = recept
    say("[receptivity] CONSTRUCT_C with N_set_0 zeroing")
    W = _chu_weight(ops.BlockOperator(L_.nchi, L_.neta, 4, L_.n_extra, base.brow_ptr, base.bcol,
                                      base.bval), 10)
    Cw = orc.CSC.from_arrays(*W.to_csc_varmajor())
    n0 = 4 * base.ncell
    x = _cvec(L_.n, 11)
    At = orc.csc_adjoint(A)
    ref = orc.construct_c(x, A.scaled(0.05), At.scaled(0.05), Cw, 4, True, n0)
    with DeviceOperator.from_csc(L_.n, Cw.colptr, Cw.rowval, Cw.nzval, vars_per_cell=4,
                                 nchi=L_.nchi, neta=L_.neta) as dW:
        got = dev.chu_sandwich(dW, x, order=5, nsub=4, dt=0.05, n_set0=n0)
    assert not np.any(got[:n0])
    assert relerr(got, ref) <= 1e-12


def test_complex_batched_sandwich_with_factors(recept, orc):
    """The batched sandwich on 4 complex columns of the receptivity operator
    (width 8) with real factor operators and the freestream zeroing, against
    the oracle's CONSTRUCT_C per column."""
    base, L_, A, dev = recept
    say("[receptivity] batched CONSTRUCT_C with factors")
    shape = ops.BlockOperator(L_.nchi, L_.neta, 4, L_.n_extra, base.brow_ptr, base.bcol,
                              base.bval)
    Cw = orc.CSC.from_arrays(*_chu_weight(shape, 10).to_csc_varmajor())
    P1 = orc.CSC.from_arrays(*_chu_weight(shape, 16).to_csc_varmajor())
    P2 = orc.CSC.from_arrays(*_chu_weight(shape, 17).to_csc_varmajor())
    n0 = 4 * base.ncell
    X = np.asfortranarray(np.stack([_cvec(L_.n, 20 + j) for j in range(4)], axis=1))
    mk = lambda M: DeviceOperator.from_csc(L_.n, M.colptr, M.rowval, M.nzval, vars_per_cell=4,
                                           nchi=L_.nchi, neta=L_.neta)
    with mk(Cw) as dW, mk(P1) as dP1, mk(P2) as dP2:
        dX = DeviceMatrix.from_host(X)
        dY = DeviceMatrix(L_.n, 4, np.complex128)
        try:
            chu_sandwich_batched_ptr(dev, dW, dX.ptr, L_.n, 4, dY.ptr, L_.n, order=5, nsub=4,
                                     dt=0.05, n_set0=n0, pre=dP1, post=dP2)
            dev.synchronize()
            got = dY.download()
        finally:
            dX.free()
            dY.free()
    ref = _batched_sandwich_ref(orc, A, 0.05, 4, Cw, X, n0, P1, P2)
    for j in range(4):
        assert relerr(got[:, j], ref[:, j]) <= 1e-12, j


# ------------------------------------------------------------------ closed forms

def test_true_exponential_block_circulant(orc):
    say("[circulant] T_12(A T/64)^64 vs exact exp(A T)")
    for b in (4, 8):
        op, blocks = ops.circulant_operator(12, 9, b, seed=3, row_sum=8.0)
        perm = ops.varmajor_to_cellmajor_index(op.ncell, b, op.n)
        x = np.random.default_rng(12).standard_normal(op.n)
        x_cm = np.empty_like(x)
        x_cm[perm] = x
        ref = orc.circulant_expm(blocks, op.nchi, op.neta, x_cm, 1.0)[perm]
        with DeviceOperator.from_blocks(op) as dev:
            got = dev.expm_taylor(x, order=12, nsub=64, dt=1.0 / 64, form="normalised")
            got_h = dev.expm_taylor(x, order=12, nsub=64, dt=1.0 / 64, form="horner")
        assert relerr(got, ref) <= 1e-10 and relerr(got_h, ref) <= 1e-10


# ------------------------------------------------------------------ full config sizes

@pytest.fixture(scope="module")
def c1(orc):
    say("[C1 100x80 b4] building")
    op = ops.config_operator("C1")
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    dev = DeviceOperator.from_csc(op.n, A.colptr, A.rowval, A.nzval, vars_per_cell=4, nchi=100,
                                  neta=80)
    yield op, A, dev
    dev.close()


def test_c1_action_vs_oracle(c1, orc):
    op, A, dev = c1
    say("[C1] T_12(A/64)^64 v vs oracle")
    x = np.random.default_rng(1).standard_normal(op.n)
    ref = orc.power_method(A.scaled(1.0 / 64), x, 64, order=12)
    got = dev.expm_taylor(x, order=12, nsub=64, dt=1.0 / 64, form="powers")
    assert relerr(got, ref) <= 1e-11
    got = dev.expm_taylor(x, order=12, nsub=64, dt=1.0 / 64, form="horner")
    assert relerr(got, ref) <= 1e-11


def test_c1_linearity_composition_adjoint(c1):
    op, A, dev = c1
    say("[C1] linearity / composition / adjoint dot product")
    r = np.random.default_rng(2)
    x, y = r.standard_normal(op.n), r.standard_normal(op.n)
    T = lambda v, n=8, tr=False: dev.expm_taylor(v, order=12, nsub=n, dt=1 / 64, transpose=tr)
    assert relerr(T(2.0 * x - 3.0 * y), 2.0 * T(x) - 3.0 * T(y)) <= 1e-13
    assert relerr(T(T(x, 3), 5), T(x, 8)) <= 1e-13
    lhs, rhs = T(x) @ y, x @ T(y, tr=True)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_c2_shock_config_properties(orc):
    say("[C2 200x300 b4 + 200 shock DOFs] properties at full size")
    op = ops.config_operator("C2")
    with DeviceOperator.from_blocks(op) as dev:
        r = np.random.default_rng(3)
        x, y = r.standard_normal(op.n), r.standard_normal(op.n)
        T = lambda v, tr=False: dev.expm_taylor(v, order=5, nsub=4, dt=1 / 64, transpose=tr)
        lhs, rhs = T(x) @ y, x @ T(y, tr=True)
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
        # one SpMV against the oracle at full size
        A = orc.CSC.from_arrays(*op.to_csc_varmajor())
        assert relerr(dev.spmv(x), orc.csc_mul(A, x)) <= 1e-14


def test_argument_errors_on_gpu():
    op = make_case("grid20x16_b4")
    say("[errors] adjoint missing / dtype mismatch")
    from hymor_amd._lib import HymorError
    with DeviceOperator.from_blocks(op, with_adjoint=False) as dev:
        with pytest.raises(HymorError, match="adjoint"):
            dev.expm_taylor(np.zeros(op.n), transpose=True)
        with pytest.raises(HymorError, match="order"):
            dev.expm_taylor(np.zeros(op.n), order=0)


def test_forward_after_adjoint_with_extras(orc):
    """Regression: the forward and adjoint operators list different extra
    rows (A12 vs A21'); interleaving them must not leak partial sums."""
    op = make_case("shock12x10_b4")
    say("[shock12x10_b4] adjoint spmv, then forward Taylor, then adjoint Taylor")
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    At = orc.csc_adjoint(A)
    x = np.random.default_rng(13).standard_normal(op.n)
    with DeviceOperator.from_blocks(op) as dev:
        for _ in range(2):
            dev.spmv(x, transpose=True)
            got = dev.expm_taylor(x, order=5, nsub=3, dt=0.1)
            assert relerr(got, orc.power_method(A.scaled(0.1), x, 3)) <= 1e-13
            got = dev.expm_taylor(x, order=5, nsub=3, dt=0.1, transpose=True)
            assert relerr(got, orc.power_method(At.scaled(0.1), x, 3)) <= 1e-13


VARIANTS = [(0, 1, 0), (0, 2, 1), (0, 4, 0), (0, 4, 1), (1, 2, 0), (1, 2, 1), (1, 4, 0),
            (1, 4, 1), (2, 3, 0), (2, 3, 1), (2, 9, 0), (2, 9, 1), (3, 3, 0), (3, 3, 1),
            (3, 9, 0), (3, 9, 1), (3, 1, 1), (3, 2, 0)]


@pytest.mark.parametrize("waves,nt", [(3, 0), (3, 1), (9, 1), (1, 1), (1, 0), (2, 1)])
@pytest.mark.parametrize("b,nrhs", [(4, 2), (4, 4), (8, 2), (8, 4), (4, 8), (4, 16)])
@pytest.mark.parametrize("kind", ["shock", "wide"])
def test_lds_staged_small_batches(orc, kind, b, nrhs, waves, nt):
    """k_term_lds (x blocks staged through LDS) at 2 and 4 doubles per row:
    the oracle's action on every column, forward and adjoint, and with 3 or 9
    waves per slice the same bits as k_term_split_p with the same waves (same
    summation order)."""
    say(f"[lds] {kind} b={b} nrhs={nrhs} waves={waves} nt={nt}")
    if kind == "shock":
        op = ops.stencil_operator(12, 10, b, seed=30 + b, n_shock=12, row_sum=2.0)
    else:
        op = random_pattern_operator(9, 7, b, 19, seed=40 + b)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    At = orc.csc_adjoint(A)
    X = np.asfortranarray(np.random.default_rng(22).standard_normal((op.n, nrhs)))
    kw = dict(order=5, nsub=3, dt=0.1)
    with DeviceOperator.from_blocks(op) as dev, DeviceOperator.from_blocks(op) as ref:
        dev.set_tuning(3, waves, nt)
        ref.set_tuning(2, waves if waves in (3, 9) else 3, nt)
        if nrhs >= 8:      # b = 4: batch kernel 0 sends 8 doubles per row to the term kernels
            dev.set_batch_kernel(0)
            ref.set_batch_kernel(0)
        for tr, M in ((False, A), (True, At)):
            got = dev.expm_taylor_batched(X, transpose=tr, **kw)
            if waves in (3, 9) and nrhs <= 8:
                np.testing.assert_array_equal(got,
                                              ref.expm_taylor_batched(X, transpose=tr, **kw))
            for j in range(nrhs):
                assert relerr(got[:, j], orc.power_method(M.scaled(0.1), X[:, j], 3)) <= 1e-12


@pytest.mark.parametrize("split,waves,nt", VARIANTS)
def test_launch_variants_agree(orc, split, waves, nt):
    """Every Taylor-term launch variant the autotuner may pick matches the
    oracle (real with shock extras, and complex)."""
    say(f"[variants] split={split} waves={waves} nt={nt}")
    op = make_case("shock12x10_b4")
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    x = np.random.default_rng(14).standard_normal(op.n)
    with DeviceOperator.from_blocks(op) as dev:
        dev.set_tuning(split, waves, nt)
        assert dev.tuning == (split, waves, nt)
        got = dev.expm_taylor(x, order=5, nsub=4, dt=0.1)
        assert relerr(got, orc.power_method(A.scaled(0.1), x, 4)) <= 1e-12
        got = dev.expm_taylor(x, order=5, nsub=4, dt=0.1, transpose=True)
        assert relerr(got, orc.power_method(orc.csc_adjoint(A).scaled(0.1), x, 4)) <= 1e-12
    op8 = make_case("grid10x8_b8")
    A8 = orc.CSC.from_arrays(*op8.to_csc_varmajor())
    xc = np.random.default_rng(15).standard_normal(op8.n) * (1 + 1j)
    with DeviceOperator.from_blocks(op8) as dev:
        dev.set_tuning(split, waves, nt)
        got = dev.expm_taylor(xc, order=5, nsub=3, dt=0.1)
        assert relerr(got, orc.power_method(A8.scaled(0.1), xc, 3)) <= 1e-12


def test_autotune_picks_a_valid_variant(orc):
    op = ops.config_operator("C1")
    say("[autotune] C1")
    with DeviceOperator.from_blocks(op) as dev:
        ms, tun = dev.autotune(reps=3)
        assert ms > 0 and tun[0] in (0, 1, 2, 3)
        x = np.random.default_rng(16).standard_normal(op.n)
        A = orc.CSC.from_arrays(*op.to_csc_varmajor())
        got = dev.expm_taylor(x, order=12, nsub=4, dt=1 / 64)
        assert relerr(got, orc.power_method(A.scaled(1 / 64), x, 4, order=12)) <= 1e-12


def test_graph_replays_match_eager(orc):
    """The captured Taylor sweep, replayed many times with eager work and
    synchronisations in between, gives bit-identical results to the eager
    launches; a parameter change recaptures."""
    op = make_case("shock12x10_b4")
    say("[graphs] replays vs eager")
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    r = np.random.default_rng(17)
    xs = [r.standard_normal(op.n) for _ in range(4)]
    with DeviceOperator.from_blocks(op) as eager, DeviceOperator.from_blocks(op) as g:
        g.set_graphs(True)
        for rep in range(3):
            for x in xs:
                want = eager.expm_taylor(x, order=5, nsub=6, dt=0.1)
                got = g.expm_taylor(x, order=5, nsub=6, dt=0.1)
                assert np.array_equal(got, want), rep
                eager.synchronize()
            # change of parameters -> recapture, then back
            got = g.expm_taylor(xs[0], order=4, nsub=3, dt=0.1, transpose=True)
            assert np.array_equal(got, eager.expm_taylor(xs[0], order=4, nsub=3, dt=0.1,
                                                         transpose=True))
        X = np.stack(xs, axis=1)
        for rep in range(3):
            assert np.array_equal(g.expm_taylor_batched(X, order=5, nsub=2, dt=0.1),
                                  eager.expm_taylor_batched(X, order=5, nsub=2, dt=0.1))
        assert relerr(g.expm_taylor(xs[1], order=5, nsub=6, dt=0.1),
                      orc.power_method(A.scaled(0.1), xs[1], 6)) <= 1e-12


def random_pattern_operator(nchi, neta, b, per_row, seed):
    """Block operator with `per_row` random block columns per cell (plus the
    diagonal), wider than any 3x3 stencil row: exercises slot loops that
    need more than one group of loads per wave."""
    r = np.random.default_rng(seed)
    ncell = nchi * neta
    cols = [np.unique(np.concatenate([[i], r.choice(ncell, per_row, replace=False)]))
            for i in range(ncell)]
    brow_ptr = np.zeros(ncell + 1, np.int64)
    brow_ptr[1:] = np.cumsum([len(c) for c in cols])
    bcol = np.concatenate(cols).astype(np.int32)
    bval = r.standard_normal((bcol.shape[0], b, b)) / (b * per_row)
    return ops.BlockOperator(nchi, neta, b, 0, brow_ptr, bcol, bval)


@pytest.mark.parametrize("split,waves,nt", VARIANTS)
@pytest.mark.parametrize("b", [4, 8])
def test_launch_variants_wide_rows(orc, split, waves, nt, b):
    """Rows of up to 20 blocks (several load groups per wave, ragged widths
    inside a slice) for every launch variant, forward and adjoint."""
    say(f"[variants wide] b={b} split={split} waves={waves} nt={nt}")
    op = random_pattern_operator(9, 7, b, 19, seed=b)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    x = np.random.default_rng(21).standard_normal(op.n)
    with DeviceOperator.from_blocks(op) as dev:
        dev.set_tuning(split, waves, nt)
        got = dev.expm_taylor(x, order=5, nsub=3, dt=0.1)
        assert relerr(got, orc.power_method(A.scaled(0.1), x, 3)) <= 1e-12
        got = dev.expm_taylor(x, order=5, nsub=3, dt=0.1, transpose=True)
        assert relerr(got, orc.power_method(orc.csc_adjoint(A).scaled(0.1), x, 3)) <= 1e-12
