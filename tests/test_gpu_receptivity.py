"""Factored receptivity action (N per-wavenumber systems in one batched
sweep) against the oracle's POWER_METHOD! restatement on the reference's
extended operator L_ = [L repeat(B,1,N); 0 kron(diag(-i w), I)], forward and
adjoint, at FP64 tolerance (1e-12)."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from hymor_amd import operators as ops
from hymor_amd.receptivity import FactoredReceptivity, split_extended

def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
@pytest.mark.parametrize("N,shock", [(3, True), (8, False), (32, True)])
def test_factored_matches_extended_operator(orc, N, shock):
    nchi, neta = 10, 7
    say(f"[factored receptivity] N={N} shock={shock}")
    base = ops.stencil_operator(nchi, neta, 4, seed=40, n_shock=nchi if shock else 0,
                                row_sum=2.0)
    w = np.linspace(0.0, 6.0, N)
    L_ = ops.receptivity_extension(base, w)
    A_ = L_.to_scipy_varmajor()
    n_L, p = base.n, 4 * nchi
    L, B = split_extended(A_, n_L, p, w)
    r = np.random.default_rng(N)
    x = r.standard_normal(A_.shape[0]) + 1j * r.standard_normal(A_.shape[0])
    Ao = orc.CSC.from_scipy(A_)
    dt, nsub = 0.05, 5
    fr = FactoredReceptivity(L, B, w, nchi=nchi, neta=neta, vars_per_cell=4, scale=dt)
    try:
        got = fr.expm(x, order=5, nsub=nsub)
        ref = orc.power_method(Ao.scaled(dt), x, nsub)
        assert relerr(got, ref) <= 1e-12
        got_t = fr.expm(x, order=5, nsub=nsub, transpose=True)
        ref_t = orc.power_method(orc.csc_adjoint(Ao).scaled(dt), x, nsub)
        assert relerr(got_t, ref_t) <= 1e-12
    finally:
        fr.close()


def test_split_extended_rejects_non_reference_structure():
    base = ops.stencil_operator(5, 4, 4, seed=41)
    w = [0.0, 1.0]
    A_ = ops.receptivity_extension(base, w).to_scipy_varmajor().tolil()
    A_[base.n + 1, base.n + 1] = 7.0         # break the -i w diagonal
    with pytest.raises(ValueError, match="freestream block"):
        split_extended(sp.csc_matrix(A_), base.n, 20, w)


class _HostSystems(FactoredReceptivity):
    """The packing and unpacking of FactoredReceptivity.expm with the batched
    device sweep replaced by a scipy restatement of the per-column systems
    (test-only, to check the factorisation on the CPU)."""

    @staticmethod
    def _make_device(S, nchi, neta, vars_per_cell, scale, value_prec):
        class _Host:
            def close(self):
                pass
        h = _Host()
        h.S = S * scale
        return h

    def _run(self, X, order, nsub, dt, form, transpose):
        S = self.dev.S.conj().T.tocsr() if transpose else self.dev.S.tocsr()
        d = self.diag.reshape(-1, 2)
        dk = d[:, 0] + 1j * d[:, 1]
        if transpose:
            dk = dk.conj()
        rows = np.arange(self.n_sys) >= self.n_L
        Y = np.array(X, dtype=np.complex128)
        for _ in range(nsub):
            v = Y.copy()
            t = Y.copy()
            f = 1.0
            for j in range(1, order + 1):
                t = S @ t + rows[:, None] * t * dk[None, :]
                t = t * dt
                f *= j
                v = v + t / f
            Y = v
        return Y


@pytest.mark.parametrize("shock", [False, True])
def test_factorisation_on_host(orc, shock):
    """expm()'s column packing, the per-column diagonal and the flow-part sum
    (forward) / broadcast (adjoint) reproduce POWER_METHOD! on the reference's
    extended operator; the device sweep is checked against the same oracle
    by test_factored_matches_extended_operator."""
    nchi, neta, N = 6, 5, 4
    base = ops.stencil_operator(nchi, neta, 4, seed=42, n_shock=nchi if shock else 0,
                                row_sum=2.0)
    w = [0.0, 0.5, 2.0, 3.0]
    A_ = ops.receptivity_extension(base, w).to_scipy_varmajor()
    L, B = split_extended(A_, base.n, 4 * nchi, w)
    r = np.random.default_rng(3)
    x = r.standard_normal(A_.shape[0]) + 1j * r.standard_normal(A_.shape[0])
    Ao = orc.CSC.from_scipy(A_)
    fr = _HostSystems(L, B, w, nchi=nchi, neta=neta, vars_per_cell=4, scale=0.05)
    got = fr.expm(x, order=5, nsub=3)
    assert relerr(got, orc.power_method(Ao.scaled(0.05), x, 3)) <= 1e-13
    got_t = fr.expm(x, order=5, nsub=3, transpose=True)
    assert relerr(got_t, orc.power_method(orc.csc_adjoint(Ao).scaled(0.05), x, 3)) <= 1e-13


class _RealValues(FactoredReceptivity):
    """[L B; 0 0] uploaded with real values (HYMOR_F64): the per-column -i w
    diagonal must still act on the complex columns."""

    @staticmethod
    def _make_device(S, nchi, neta, vars_per_cell, scale, value_prec):
        from hymor_amd.core import DeviceOperator
        assert S.imag.nnz == 0 or abs(S.imag).max() == 0
        return DeviceOperator.from_scipy(sp.csc_matrix(S.real), nchi=nchi, neta=neta,
                                         vars_per_cell=vars_per_cell, scale=scale,
                                         value_prec=value_prec)


@pytest.mark.gpu
@pytest.mark.parametrize("N,shock", [(3, True), (32, False)])
def test_factored_real_upload_equals_complex_upload(orc, N, shock):
    """The same per-wavenumber system uploaded as real and as complex values
    gives the same factored action (forward and adjoint), and both match
    POWER_METHOD! on the extended operator."""
    nchi, neta = 8, 6
    say(f"[factored receptivity, real upload] N={N} shock={shock}")
    base = ops.stencil_operator(nchi, neta, 4, seed=43, n_shock=nchi if shock else 0,
                                row_sum=2.0)
    w = np.linspace(0.2, 5.0, N)
    A_ = ops.receptivity_extension(base, w).to_scipy_varmajor()
    n_L, p = base.n, 4 * nchi
    L, B = split_extended(A_, n_L, p, w)
    r = np.random.default_rng(50 + N)
    x = r.standard_normal(A_.shape[0]) + 1j * r.standard_normal(A_.shape[0])
    Ao = orc.CSC.from_scipy(A_)
    dt, nsub = 0.05, 4
    fc = FactoredReceptivity(L, B, w, nchi=nchi, neta=neta, vars_per_cell=4, scale=dt)
    fr = _RealValues(L, B, w, nchi=nchi, neta=neta, vars_per_cell=4, scale=dt)
    try:
        for tr in (False, True):
            got_r = fr.expm(x, order=5, nsub=nsub, transpose=tr)
            got_c = fc.expm(x, order=5, nsub=nsub, transpose=tr)
            M = orc.csc_adjoint(Ao) if tr else Ao
            ref = orc.power_method(M.scaled(dt), x, nsub)
            assert relerr(got_r, got_c) <= 1e-14
            assert relerr(got_r, ref) <= 1e-12
    finally:
        fr.close()
        fc.close()


@pytest.mark.gpu
@pytest.mark.parametrize("N,shock,nrhs", [(3, True, 8), (8, False, 16), (32, True, 8)])
def test_assembled_wide_batches_replicated_blocks(orc, N, shock, nrhs):
    """The assembled extended operator at 16 and 32 doubles per row, where
    the library applies the N copies of B once (column sums of the
    freestream blocks forward, shared rows in the adjoint): every column
    equals the oracle's POWER_METHOD! on L_ (1e-12) and the single-vector
    action, which keeps the copies (1e-13)."""
    from hymor_amd.core import DeviceOperator
    nchi, neta = 9, 6
    say(f"[assembled receptivity, wide batches] N={N} shock={shock} nrhs={nrhs}")
    base = ops.stencil_operator(nchi, neta, 4, seed=44, n_shock=nchi if shock else 0,
                                row_sum=2.0)
    L_ = ops.receptivity_extension(base, np.linspace(0.0, 5.0, N))
    A_ = L_.to_scipy_varmajor()
    Ao = orc.CSC.from_scipy(A_)
    Aot = orc.csc_adjoint(Ao)
    r = np.random.default_rng(60 + N)
    X = r.standard_normal((L_.n, nrhs)) + 1j * r.standard_normal((L_.n, nrhs))
    dt, nsub = 0.05, 3
    with DeviceOperator.from_blocks(L_, scale=dt) as dev:
        for tr, M in ((False, Ao), (True, Aot)):
            Y = dev.expm_taylor_batched(X, order=5, nsub=nsub, dt=1.0, transpose=tr)
            for j in range(nrhs):
                ref = orc.power_method(M.scaled(dt), X[:, j], nsub)
                assert relerr(Y[:, j], ref) <= 1e-12, (tr, j)
                single = dev.expm_taylor(X[:, j], order=5, nsub=nsub, dt=1.0, transpose=tr)
                assert relerr(Y[:, j], single) <= 1e-13, (tr, j)


@pytest.mark.gpu
@pytest.mark.parametrize("N,nsub,prec", [(2, 3, "f64"), (9, 4, "f64"), (32, 2, "f64"),
                                         (9, 3, "f32")])
def test_assembled_single_vector_fused_horner(orc, N, nsub, prec):
    """One complex vector in Horner form on the assembled extended operator
    without shock DOFs: the library fuses the replicated freestream rows into
    the term kernel (one launch per term; the column sums of each new w
    carried to the next term and substep).  Equal to the oracle's
    POWER_METHOD! (1e-12 with FP64 values, 2e-6 with FP32), identical on
    three replays of the captured sweep, and equal to the two-launch term
    that a two-wave tuning runs (1e-13); the fused term reports the bytes of
    the compressed entries it reads."""
    from hymor_amd.core import DeviceOperator
    nchi, neta = 11, 7
    say(f"[assembled receptivity, fused Horner] N={N} nsub={nsub} values={prec}")
    base = ops.stencil_operator(nchi, neta, 4, seed=45, row_sum=2.0)
    L_ = ops.receptivity_extension(base, np.linspace(0.1, 6.0, N))
    Ao = orc.CSC.from_scipy(L_.to_scipy_varmajor())
    r = np.random.default_rng(70 + N)
    x = r.standard_normal(L_.n) + 1j * r.standard_normal(L_.n)
    dt, order = 0.05, 6
    ref = orc.power_method(Ao.scaled(dt), x, nsub, order)
    tol = 1e-12 if prec == "f64" else 2e-6
    with DeviceOperator.from_blocks(L_, scale=dt, value_prec=prec) as dev:
        assert dev.tuning == (3, 1, 1)
        fused = dev.expm_taylor(x, order=order, nsub=nsub, dt=1.0, form="horner")
        assert relerr(fused, ref) <= tol
        _, by_fused = dev.bench_term_kernel(reps=2, nv=2, form="horner")
        dev.set_graphs(True)
        for _ in range(3):
            y = dev.expm_taylor(x, order=order, nsub=nsub, dt=1.0, form="horner")
            assert np.array_equal(y, fused)
        dev.set_graphs(False)
        dev.set_tuning(3, 2, 1)
        two = dev.expm_taylor(x, order=order, nsub=nsub, dt=1.0, form="horner")
        _, by_two = dev.bench_term_kernel(reps=2, nv=2, form="horner")
        assert relerr(two, ref) <= tol
        assert relerr(fused, two) <= 1e-13
        # the copies of B are read once per term instead of N times
        assert by_fused < by_two


@pytest.mark.gpu
def test_fused_horner_real_values(orc):
    """Real extra values (the freestream diagonal -w made real) with two real
    vectors per sweep: the fused term's real-value form against the
    oracle, column by column."""
    from hymor_amd.core import DeviceOperator
    nchi, neta = 9, 6
    say("[fused Horner, real values, two vectors]")
    base = ops.stencil_operator(nchi, neta, 4, seed=46, row_sum=2.0)
    A = ops.receptivity_extension(base, [0.4, 1.1, 2.0, 3.3, 4.1]).to_scipy_varmajor()
    A = (A.real + A.imag).tocsc()
    Ao = orc.CSC.from_scipy(A).scaled(0.05)
    X = np.random.default_rng(71).standard_normal((A.shape[0], 2))
    with DeviceOperator.from_scipy(A, nchi=nchi, neta=neta, vars_per_cell=4, scale=0.05) as dev:
        Y = dev.expm_taylor_batched(X, order=5, nsub=3, dt=1.0, form="horner")
    for v in range(2):
        assert relerr(Y[:, v], orc.power_method(Ao, X[:, v], 3, 5)) <= 1e-12, v


@pytest.mark.gpu
@pytest.mark.parametrize("nrhs,prec", [(4, "f64"), (8, "f64"), (16, "f64"), (32, "f64"),
                                       (8, "f32")])
def test_assembled_batches_fused_horner(orc, nrhs, prec):
    """nrhs complex vectors (8-64 doubles per row) in Horner form on the
    assembled extended operator: the b = 4 batch kernel runs with the
    replicated rows fused in (one launch per term).  Every column equals the
    oracle's POWER_METHOD! (1e-12 with FP64 values, 2e-6 with FP32) and the
    powers form, which keeps separate launches (1e-13); three graph replays
    give the same bits; at 8 doubles per row the fused term reports the
    compressed entries' bytes (the separate launches read B's copies)."""
    from hymor_amd.core import DeviceOperator
    nchi, neta = 10, 7
    say(f"[assembled receptivity, fused Horner batches] nrhs={nrhs} values={prec}")
    base = ops.stencil_operator(nchi, neta, 4, seed=47, row_sum=2.0)
    L_ = ops.receptivity_extension(base, np.linspace(0.1, 6.0, 12))
    Ao = orc.CSC.from_scipy(L_.to_scipy_varmajor())
    r = np.random.default_rng(80 + nrhs)
    X = r.standard_normal((L_.n, nrhs)) + 1j * r.standard_normal((L_.n, nrhs))
    dt, nsub, order = 0.05, 3, 5
    tol = 1e-12 if prec == "f64" else 2e-6
    with DeviceOperator.from_blocks(L_, scale=dt, value_prec=prec) as dev:
        Y = dev.expm_taylor_batched(X, order=order, nsub=nsub, dt=1.0, form="horner")
        P = dev.expm_taylor_batched(X, order=order, nsub=nsub, dt=1.0, form="powers")
        for j in range(nrhs):
            ref = orc.power_method(Ao.scaled(dt), X[:, j], nsub, order)
            assert relerr(Y[:, j], ref) <= tol, j
            assert relerr(Y[:, j], P[:, j]) <= 1e-13, j
        dev.set_graphs(True)
        for _ in range(3):
            assert np.array_equal(
                dev.expm_taylor_batched(X, order=order, nsub=nsub, dt=1.0, form="horner"), Y)
        dev.set_graphs(False)
        if nrhs == 4:
            _, by_h = dev.bench_term_kernel(reps=2, nv=8, form="horner")
            _, by_p = dev.bench_term_kernel(reps=2, nv=8, form="powers")
            # Horner moves 3 vectors, powers 4; beyond that, B's copies
            n8 = L_.n * 8 * 8.0
            assert by_h - 3 * n8 < by_p - 4 * n8


@pytest.mark.gpu
@pytest.mark.parametrize("nrhs", [1, 4, 16])
def test_adjoint_fused_term(orc, nrhs):
    """The adjoint of the assembled extended operator without shock DOFs:
    its only extra rows are the shared rows B^H x (+ the conjugate
    diagonal) of the freestream copies, evaluated in the term kernel's own
    launch for every Taylor form.  Horner and powers forms against the
    oracle's POWER_METHOD! on L_^H (1e-12); the single vector also against
    the two-launch term another tuning runs (1e-13), and the adjoint SpMV
    against scipy (1e-13)."""
    from hymor_amd.core import DeviceOperator
    nchi, neta = 10, 6
    say(f"[assembled receptivity, fused adjoint] nrhs={nrhs}")
    base = ops.stencil_operator(nchi, neta, 4, seed=48, row_sum=2.0)
    L_ = ops.receptivity_extension(base, np.linspace(0.1, 6.0, 10))
    A_ = L_.to_scipy_varmajor()
    Aot = orc.csc_adjoint(orc.CSC.from_scipy(A_))
    r = np.random.default_rng(90 + nrhs)
    X = r.standard_normal((L_.n, nrhs)) + 1j * r.standard_normal((L_.n, nrhs))
    dt, nsub, order = 0.05, 3, 5
    with DeviceOperator.from_blocks(L_, scale=dt) as dev:
        for form in ("horner", "powers"):
            if nrhs == 1:
                Y = dev.expm_taylor(X[:, 0], order=order, nsub=nsub, dt=1.0, form=form,
                                    transpose=True)[:, None]
            else:
                Y = dev.expm_taylor_batched(X, order=order, nsub=nsub, dt=1.0, form=form,
                                            transpose=True)
            for j in range(nrhs):
                ref = orc.power_method(Aot.scaled(dt), X[:, j], nsub, order)
                assert relerr(Y[:, j], ref) <= 1e-12, (form, j)
        if nrhs == 1:
            y = dev.spmv(X[:, 0], transpose=True)
            ref = dt * (A_.conj().T @ X[:, 0])
            assert relerr(y, ref) <= 1e-13
            fused = dev.expm_taylor(X[:, 0], order=order, nsub=nsub, dt=1.0, form="horner",
                                    transpose=True)
            dev.set_tuning(3, 2, 1)
            two = dev.expm_taylor(X[:, 0], order=order, nsub=nsub, dt=1.0, form="horner",
                                  transpose=True)
            assert relerr(fused, two) <= 1e-13
