"""Receptivity actions against the oracle's POWER_METHOD! restatement on the
reference's extended operator L_ = [L repeat(B,1,N); 0 kron(diag(-i w), I)],
forward and adjoint, at FP64 tolerance (1e-12): the extended operator built
from its parts (B kept once), the N per-wavenumber systems in one batched
sweep, and the assembled upload with its copies of B found and kept once."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from hymor_amd import operators as ops
from hymor_amd.receptivity import (FactoredReceptivity, per_omega_system, receptivity_operator,
                                   split_extended)

def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relerr(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
@pytest.mark.parametrize("per_wavenumber", [False, True])
@pytest.mark.parametrize("N,shock", [(3, True), (8, False), (32, True)])
def test_factored_matches_extended_operator(orc, N, shock, per_wavenumber):
    nchi, neta = 10, 7
    say(f"[factored receptivity] N={N} shock={shock} per_wavenumber={per_wavenumber}")
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
    fr = FactoredReceptivity(L, B, w, nchi=nchi, neta=neta, vars_per_cell=4, scale=dt,
                             per_wavenumber=per_wavenumber)
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
    fr = _HostSystems(L, B, w, nchi=nchi, neta=neta, vars_per_cell=4, scale=0.05,
                      per_wavenumber=True)
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
    fc = FactoredReceptivity(L, B, w, nchi=nchi, neta=neta, vars_per_cell=4, scale=dt,
                             per_wavenumber=True)
    fr = _RealValues(L, B, w, nchi=nchi, neta=neta, vars_per_cell=4, scale=dt,
                     per_wavenumber=True)
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


@pytest.mark.gpu
@pytest.mark.parametrize("N,shock,prec", [(3, True, "f64"), (8, False, "f64"), (32, False, "f64"),
                                          (32, True, "f64"), (12, False, "f32")])
def test_receptivity_operator_from_parts(orc, N, shock, prec):
    """The extended operator built from [L B; 0 0] and the wavenumbers
    (hymor_op_create_receptivity, B kept once at every width) against the
    oracle's POWER_METHOD! on the assembled L_ and its adjoint: SpMV (1e-14),
    one complex vector in the powers, normalised and Horner forms, and 2-32
    complex vectors per sweep in the Horner and powers forms (1e-12 with FP64
    values, 2e-6 with FP32), each equal to the assembled upload (1e-13 FP64);
    three graph replays give the same bits; a real-valued upload of the
    system gives the same bits as a complex one."""
    from hymor_amd.core import DeviceOperator
    nchi, neta = 9, 6
    say(f"[receptivity operator from parts] N={N} shock={shock} values={prec}")
    base = ops.stencil_operator(nchi, neta, 4, seed=49, n_shock=nchi if shock else 0,
                                row_sum=2.0)
    w = np.linspace(0.0, 6.0, N)
    L_ = ops.receptivity_extension(base, w)
    A_ = L_.to_scipy_varmajor()
    Ao = orc.CSC.from_scipy(A_)
    Aot = orc.csc_adjoint(Ao)
    L, B = split_extended(A_, base.n, 4 * nchi, w)
    r = np.random.default_rng(100 + N)
    n = A_.shape[0]
    X = r.standard_normal((n, 16)) + 1j * r.standard_normal((n, 16))
    dt, nsub, order = 0.05, 3, 5
    tol = 1e-12 if prec == "f64" else 2e-6
    same = 1e-13 if prec == "f64" else 1e-6
    kw = dict(nchi=nchi, neta=neta, vars_per_cell=4, scale=dt, value_prec=prec)
    with receptivity_operator(L, B, w, **kw) as dev, \
            DeviceOperator.from_blocks(L_, scale=dt, value_prec=prec) as asm:
        assert dev.n == n and dev.is_complex
        # B once: the device holds one copy of its entries, the assembly N
        assert dev.nnz_extra < asm.nnz_extra or N == 1
        for tr, M in ((False, Ao), (True, Aot)):
            y = dev.spmv(X[:, 0], transpose=tr)
            if prec == "f64":
                assert relerr(y / dt, orc.csc_mul(M, X[:, 0])) <= 1e-14, tr
            for form in ("powers", "normalised", "horner"):
                got = dev.expm_taylor(X[:, 0], order=order, nsub=nsub, dt=1.0, form=form,
                                      transpose=tr)
                ref = orc.power_method(M.scaled(dt), X[:, 0], nsub, order)
                assert relerr(got, ref) <= tol, (tr, form)
                a = asm.expm_taylor(X[:, 0], order=order, nsub=nsub, dt=1.0, form=form,
                                    transpose=tr)
                assert relerr(got, a) <= same, (tr, form)
            for m in (2, 4, 8, 16):
                for form in ("horner", "powers"):
                    Y = dev.expm_taylor_batched(X[:, :m], order=order, nsub=nsub, dt=1.0,
                                                form=form, transpose=tr)
                    Ya = asm.expm_taylor_batched(X[:, :m], order=order, nsub=nsub, dt=1.0,
                                                 form=form, transpose=tr)
                    for j in range(m):
                        ref = orc.power_method(M.scaled(dt), X[:, j], nsub, order)
                        assert relerr(Y[:, j], ref) <= tol, (tr, m, form, j)
                        assert relerr(Y[:, j], Ya[:, j]) <= same, (tr, m, form, j)
        base_y = dev.expm_taylor(X[:, 1], order=order, nsub=nsub, dt=1.0, form="horner")
        dev.set_graphs(True)
        for _ in range(3):
            assert np.array_equal(
                dev.expm_taylor(X[:, 1], order=order, nsub=nsub, dt=1.0, form="horner"), base_y)
        dev.set_graphs(False)
    if prec == "f64":
        with receptivity_operator(sp.csc_matrix(L.real), sp.csc_matrix(B.real), w,
                                  **kw) as real:
            assert not np.iscomplexobj(sp.csc_matrix(L.real).data)
            for tr in (False, True):
                with receptivity_operator(L, B, w, **kw) as cplx:
                    want = cplx.expm_taylor_batched(X[:, :4], order=order, nsub=nsub, dt=1.0,
                                                    form="horner", transpose=tr)
                got = real.expm_taylor_batched(X[:, :4], order=order, nsub=nsub, dt=1.0,
                                               form="horner", transpose=tr)
                assert np.array_equal(got, want), tr


@pytest.mark.gpu
def test_receptivity_operator_rejects_a_per_column_diagonal():
    """The operator from parts carries its wavenumbers: the factored entry
    point's per-column diagonal is refused with a clear error."""
    from hymor_amd import _lib
    from hymor_amd.core import DeviceMatrix
    nchi, neta = 5, 4
    base = ops.stencil_operator(nchi, neta, 4, seed=50)
    w = [0.5, 1.5]
    A_ = ops.receptivity_extension(base, w).to_scipy_varmajor()
    L, B = split_extended(A_, base.n, 4 * nchi, w)
    with receptivity_operator(L, B, w, nchi=nchi, neta=neta) as dev:
        X = DeviceMatrix.from_host(np.zeros((dev.n, 2), dtype=np.complex128))
        Y = DeviceMatrix(dev.n, 2, np.complex128)
        d = np.zeros(4)
        try:
            rc = _lib.load().hymor_expm_taylor_batched_coldiag_dev(
                dev.handle, 0, 5, 1, 1.0, 1, 2, d.ctypes.data, base.n, dev.n, X.ptr, dev.n,
                Y.ptr, dev.n)
            assert rc == _lib.ERR_ARG
            assert b"carries its wavenumbers" in _lib.load().hymor_last_error()
        finally:
            X.free()
            Y.free()
