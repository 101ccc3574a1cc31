"""Pins of the CPU oracle (oracle/hymor_oracle.c) against independent facts.

The reference ships no golden vectors for this path (it has no test suite),
so every oracle function is pinned here by something that does not share its
code: dense numpy algebra, scipy's expm (Pade, not Taylor), closed forms of
block-circulant / diagonal / nilpotent / rotation operators, and identities
(composition, adjoint dot products).  Each pin is chosen so that a plausible
slip (wrong factorial, transposed operand, off-by-one index base, dropped
zeroing, missing conjugate) fails it.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

from hymor_amd import operators as ops


def rand_sparse(n, density, seed, cplx=False):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng, format="csc")
    A.data = rng.standard_normal(A.data.shape[0])
    if cplx:
        A = A.astype(np.complex128)
        A.data = A.data + 1j * rng.standard_normal(A.data.shape[0])
    return A


# ---------------------------------------------------------------- SpMV

@pytest.mark.parametrize("cplx", [False, True])
def test_csc_mul_matches_dense(orc, cplx):
    A = rand_sparse(57, 0.1, 1, cplx)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(57) + (1j * rng.standard_normal(57) if cplx else 0)
    y = orc.csc_mul(orc.CSC.from_scipy(A), x)
    assert np.allclose(y, A.toarray() @ x, rtol=1e-13, atol=1e-13)


def test_csc_mul_column_scatter_order_is_exact_for_integers(orc):
    # integer-valued data: any summation order is exact, so the result must
    # equal dense algebra bit for bit (pins indexing, not rounding)
    rng = np.random.default_rng(3)
    A = sp.random(40, 40, density=0.2, random_state=rng, format="csc")
    A.data = rng.integers(-5, 6, A.data.shape[0]).astype(float)
    x = rng.integers(-7, 8, 40).astype(float)
    assert np.array_equal(orc.csc_mul(orc.CSC.from_scipy(A), x), A.toarray() @ x)


@pytest.mark.parametrize("cplx", [False, True])
def test_csc_adjoint_is_conjugate_transpose(orc, cplx):
    A = rand_sparse(33, 0.15, 4, cplx)
    At = orc.csc_adjoint(orc.CSC.from_scipy(A)).to_scipy().toarray()
    assert np.array_equal(At, A.toarray().conj().T)


# ---------------------------------------------------------------- POWER_METHOD!

def test_power_method_diagonal_closed_form(orc):
    lam = np.linspace(-2.0, 0.5, 25)
    dt, n, p = 0.1, 7, 5
    A = sp.diags(lam * dt, format="csc")
    v = np.linspace(1, 2, 25)
    got = orc.power_method(orc.CSC.from_scipy(A), v, n, order=p)
    t5 = sum((lam * dt) ** j / math.factorial(j) for j in range(p + 1))
    assert np.allclose(got, t5 ** n * v, rtol=1e-14, atol=0)


def test_power_method_zero_is_identity(orc):
    A = sp.csc_matrix((10, 10))
    v = np.arange(10.0)
    assert np.array_equal(orc.power_method(orc.CSC.from_scipy(A), v, 3), v)


def test_power_method_nilpotent_terminates_to_expm(orc):
    # strictly upper triangular 6x6: N^6 = 0, so T_5(N) == exp(N) exactly
    rng = np.random.default_rng(5)
    N = np.triu(rng.standard_normal((6, 6)), 1) * 0.5
    v = rng.standard_normal(6)
    got = orc.power_method(orc.CSC.from_scipy(sp.csc_matrix(N)), v, 1, order=5)
    assert np.allclose(got, sl.expm(N) @ v, rtol=1e-13, atol=1e-14)


def test_power_method_rotation_blocks(orc):
    # 2x2 rotation generator: T_5(A dt) acts as the complex scalar T_5(i w dt)
    w, dt, n = 1.3, 0.05, 11
    A = sp.csc_matrix(np.array([[0.0, -w], [w, 0.0]]) * dt)
    z = sum((1j * w * dt) ** j / math.factorial(j) for j in range(6)) ** n
    got = orc.power_method(orc.CSC.from_scipy(A), np.array([1.0, 0.0]), n)
    assert np.allclose(got, [z.real, z.imag], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("cplx", [False, True])
def test_power_method_matches_dense_polynomial(orc, cplx):
    A = rand_sparse(48, 0.08, 6, cplx) * 0.3
    rng = np.random.default_rng(7)
    v = rng.standard_normal(48) + (1j * rng.standard_normal(48) if cplx else 0)
    P = orc.taylor_poly_dense(A.toarray(), 5)
    ref = np.linalg.matrix_power(P, 4) @ v
    got = orc.power_method(orc.CSC.from_scipy(A), v, 4)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


def test_power_method_composition(orc):
    A = orc.CSC.from_scipy(rand_sparse(64, 0.06, 8) * 0.2)
    v = np.random.default_rng(9).standard_normal(64)
    a = orc.power_method(A, v, 5)
    b = orc.power_method(A, orc.power_method(A, v, 2), 3)
    assert np.allclose(a, b, rtol=1e-13, atol=1e-14)


def test_power_method_order_parameter(orc):
    # order p must produce the degree-p polynomial (pins the factorial/loop)
    A = rand_sparse(30, 0.1, 10) * 0.4
    v = np.random.default_rng(11).standard_normal(30)
    for p in (1, 2, 5, 9):
        ref = orc.taylor_poly_dense(A.toarray(), p) @ v
        got = orc.power_method(orc.CSC.from_scipy(A), v, 1, order=p)
        assert np.allclose(got, ref, rtol=1e-13, atol=1e-14), p


# ---------------------------------------------------------------- layout

def test_block_operator_varmajor_layout():
    op = ops.stencil_operator(5, 4, 4, seed=3, n_shock=5)
    Acm = op.to_scipy_cellmajor().toarray()
    Avm = op.to_scipy_varmajor().toarray()
    perm = ops.varmajor_to_cellmajor_index(op.ncell, op.b, op.n)
    # A_vm[i, j] == A_cm[perm[i], perm[j]]
    assert np.array_equal(Avm, Acm[np.ix_(perm, perm)])
    # var-major index of (var k, cell c) is k*ncell + c (LINEARIZE_...:276-285)
    c, k = 7, 2
    assert perm[k * op.ncell + c] == c * op.b + k


def test_stencil_block_count_matches_reference_stencil():
    op = ops.stencil_operator(100, 80, 4, seed=1)
    assert op.nnzb == (3 * 100 - 2) * (3 * 80 - 2)     # 70 924 blocks for C1
    assert op.nnzb * 16 == 1_134_784


def test_shock_band_shapes():
    nchi, neta, b = 6, 5, 4
    op = ops.stencil_operator(nchi, neta, b, seed=2, n_shock=nchi)
    nf = b * nchi * neta
    r, c = op.xrow, op.xcol
    a12 = (r < nf) & (c >= nf)
    a21 = (r >= nf) & (c < nf)
    a22 = (r >= nf) & (c >= nf)
    assert a12.sum() == 4 * nchi * nchi          # dense 4*Nchi x Nchi
    assert a21.sum() == nchi * 8 * nchi          # Nchi x 8*Nchi
    assert a22.sum() == nchi * nchi


def test_receptivity_extension_structure():
    base = ops.stencil_operator(4, 3, 4, seed=1)
    w = [0.0, 1.5, 3.0]
    L_ = ops.receptivity_extension(base, w)
    p = 4 * 4
    assert L_.n == base.n + len(w) * p
    A = L_.to_scipy_varmajor().toarray()
    n0 = base.n
    for k, wk in enumerate(w):
        blk = A[n0 + k * p:n0 + (k + 1) * p, n0 + k * p:n0 + (k + 1) * p]
        assert np.array_equal(blk, -1j * wk * np.eye(p))            # kron(diag(-i w), I)
        assert np.array_equal(A[:n0, n0 + k * p:n0 + (k + 1) * p], A[:n0, n0:n0 + p])  # repeat(B)
    assert not np.any(A[n0:, :n0])                                  # zero lower-left block


# ---------------------------------------------------------------- block-circulant closed form

@pytest.mark.parametrize("b", [4, 8])
def test_power_method_block_circulant_closed_form(orc, b):
    nchi, neta = 6, 5
    op, blocks = ops.circulant_operator(nchi, neta, b, seed=1, row_sum=4.0)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    rng = np.random.default_rng(12)
    v_vm = rng.standard_normal(op.n)
    perm = ops.varmajor_to_cellmajor_index(op.ncell, b, op.n)
    v_cm = np.empty_like(v_vm)
    v_cm[perm] = v_vm
    dt, nsub, p = 0.05, 6, 5
    got_vm = orc.power_method(A.scaled(dt), v_vm, nsub, order=p)
    ref_cm = orc.circulant_taylor(blocks, nchi, neta, v_cm, dt, p, nsub)
    assert np.linalg.norm(got_vm - ref_cm[perm]) <= 1e-13 * np.linalg.norm(ref_cm)


def test_taylor12_approaches_true_exponential(orc):
    # the benchmark's discretisation T_12(A T/64)^64 against exact exp(A T)
    nchi, neta, b = 6, 6, 4
    op, blocks = ops.circulant_operator(nchi, neta, b, seed=2, row_sum=8.0)
    A = orc.CSC.from_arrays(*op.to_csc_varmajor())
    v_vm = np.random.default_rng(13).standard_normal(op.n)
    perm = ops.varmajor_to_cellmajor_index(op.ncell, b, op.n)
    v_cm = np.empty_like(v_vm)
    v_cm[perm] = v_vm
    T = 1.0
    got = orc.power_method(A.scaled(T / 64), v_vm, 64, order=12)
    ref = orc.circulant_expm(blocks, nchi, neta, v_cm, T)[perm]
    assert np.linalg.norm(got - ref) <= 1e-10 * np.linalg.norm(ref)


# ---------------------------------------------------------------- EXPONENTIAL! / CONSTRUCT_C

def _spd_blockdiag(n, b, seed):
    rng = np.random.default_rng(seed)
    blocks = []
    for _ in range(n // b):
        G = rng.standard_normal((b, b))
        blocks.append(G @ G.T + b * np.eye(b))
    return sp.block_diag(blocks, format="csc")


@pytest.mark.parametrize("cplx", [False, True])
def test_exponential_matches_dense_sandwich(orc, cplx):
    n = 40
    A = rand_sparse(n, 0.08, 14, cplx) * 0.25
    Cw = _spd_blockdiag(n, 4, 15)
    Ao = orc.CSC.from_scipy(A)
    At = orc.csc_adjoint(Ao)
    rng = np.random.default_rng(16)
    x = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    got = orc.exponential(Ao, At, orc.CSC.from_scipy(Cw), x, 3)
    Q = np.linalg.matrix_power(orc.taylor_poly_dense(A.toarray(), 5), 3)
    ref = Q.conj().T @ (Cw.toarray() @ (Q @ x))
    assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


def test_exponential_adjoint_dot_product_and_symmetry(orc):
    n = 64
    A = orc.CSC.from_scipy(rand_sparse(n, 0.06, 17) * 0.3)
    At = orc.csc_adjoint(A)
    rng = np.random.default_rng(18)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    # <T(A) x, y> == <x, T(A') y>
    lhs = orc.power_method(A, x, 4) @ y
    rhs = x @ orc.power_method(At, y, 4)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    Cw = orc.CSC.from_scipy(_spd_blockdiag(n, 4, 19))
    Sx = orc.exponential(A, At, Cw, x, 4)
    Sy = orc.exponential(A, At, Cw, y, 4)
    assert abs(Sx @ y - x @ Sy) <= 1e-12 * abs(Sx @ y)     # symmetric
    assert x @ Sx > 0                                       # positive definite


def test_construct_c_zeroing(orc):
    n, n0 = 48, 32
    A = orc.CSC.from_scipy(rand_sparse(n, 0.1, 20) * 0.3)
    At = orc.csc_adjoint(A)
    Cw = orc.CSC.from_scipy(_spd_blockdiag(n, 4, 21))
    x = np.random.default_rng(22).standard_normal(n)
    y = orc.construct_c(x, A, At, Cw, 2, True, n0)
    assert not np.any(y[:n0])
    x0 = x.copy()
    x0[:n0] = 0
    ref = orc.exponential(A, At, Cw, x0, 2)
    assert np.array_equal(y[n0:], ref[n0:])
    assert np.array_equal(orc.construct_c(x, A, At, Cw, 2, False, n0), orc.exponential(A, At, Cw, x, 2))


# ---------------------------------------------------------------- energies

@pytest.mark.parametrize("cplx", [False, True])
def test_linear_integration_energy_matches_dense(orc, cplx):
    n, nt = 32, 5
    A = rand_sparse(n, 0.1, 23) * 0.2
    if cplx:
        A = A.astype(np.complex128)
        A.setdiag(A.diagonal() - 0.3j)
    Ws = [_spd_blockdiag(n, 4, 24 + k) for k in range(4)]
    rng = np.random.default_rng(30)
    x = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    E, vfin = orc.linear_integration_energy(orc.CSC.from_scipy(A), [orc.CSC.from_scipy(W) for W in Ws],
                                            x, nt)
    P = orc.taylor_poly_dense(A.toarray(), 5)
    v = x.copy()
    for i in range(nt):
        v = P @ v
        vr = v.real
        for k, W in enumerate(Ws):
            assert abs(E[k, i + 1] - vr @ (W @ vr)) <= 1e-12 * abs(vr @ (W @ vr))
    assert np.all(E[:, 0] == 0)
    assert np.allclose(vfin, v, rtol=1e-12, atol=1e-13)


def test_taylor_csr_omp_matches_power_method(orc):
    op = ops.stencil_operator(12, 10, 4, seed=5, row_sum=1.0)
    A = op.to_scipy_varmajor().tocsr() * 0.5
    v = np.random.default_rng(31).standard_normal(op.n)
    got, used = orc.taylor_csr_omp(A, v, 3, 5, nthreads=2)
    ref = orc.power_method(orc.CSC.from_scipy(A.tocsc()), v, 3)
    assert used >= 1
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------- operator gradient

def _pattern(A):
    """1-based CSC arrays of A's pattern plus a few entries outside it."""
    B = sp.csc_matrix(A, dtype=np.float64).tolil()
    n = A.shape[0]
    rng = np.random.default_rng(11)
    for i, j in zip(rng.integers(0, n, 6), rng.integers(0, n, 6)):
        if B[i, j] == 0:
            B[i, j] = 1e-300        # marks the slot; the value is not used
    B = sp.csc_matrix(B)
    B.sort_indices()
    return B.indptr.astype(np.int64) + 1, B.indices.astype(np.int64) + 1


def _j(orc, A, x, wbar, n_t, p):
    return float(wbar @ orc.power_method(orc.CSC.from_scipy(A), x, n_t, p))


@pytest.mark.parametrize("n_t,p", [(1, 1), (3, 5), (4, 12)])
def test_taylor_vjp_central_differences(orc, n_t, p):
    """taylor_vjp_values against central differences of power_method's
    J = <wbar, P(A)^n x> in 16 random directions over A's pattern and six
    entries outside it (a transposed operand, a dropped term or a wrong
    factorial moves the gradient far outside the 1e-7 band)."""
    A = rand_sparse(40, 0.12, 21) * 0.2
    A = sp.csc_matrix(A)
    rng = np.random.default_rng(22)
    x, wbar = rng.standard_normal(40), rng.standard_normal(40)
    colptr, rowval = _pattern(A)
    g = orc.taylor_vjp_values(orc.CSC.from_scipy(A), x, wbar, n_t, p, colptr, rowval)
    rows = rowval - 1
    cols = np.repeat(np.arange(40), np.diff(colptr))
    eps = 1e-5
    for d in range(16):
        delta = rng.standard_normal(rows.shape[0])
        D = sp.csc_matrix((delta, (rows, cols)), shape=A.shape)
        fd = (_j(orc, A + eps * D, x, wbar, n_t, p) - _j(orc, A - eps * D, x, wbar, n_t, p)) / (2 * eps)
        assert abs(g @ delta - fd) <= 1e-7 * max(1.0, abs(fd)), (d, g @ delta, fd)


def test_taylor_vjp_closed_forms(orc):
    """Diagonal A = diag(a): the derivative of the matrix polynomial
    phi(A) = P(A)^n at a diagonal matrix has entries given by divided
    differences (Daleckii-Krein), so dJ/dA_ij = wbar_i x_j (phi(a_i) -
    phi(a_j)) / (a_i - a_j) off the diagonal and wbar_i x_i phi'(a_i) =
    wbar_i x_i n P(a_i)^(n-1) P'(a_i) on it, P'(a) = sum_{j<p} a^j / j!;
    one step of order 1 (P(A) = I + A): dJ/dA_ij = wbar_i x_j on any
    pattern."""
    rng = np.random.default_rng(23)
    n, n_t, p = 30, 6, 5
    a = rng.uniform(-0.5, 0.3, n)
    A = sp.diags(a, format="csc")
    A = A + sp.csc_matrix(([0.0], ([0], [1])), shape=(n, n))   # an empty slot in (0, 1)
    x, wbar = rng.standard_normal(n), rng.standard_normal(n)
    colptr, rowval = _pattern(A)
    g = orc.taylor_vjp_values(orc.CSC.from_scipy(sp.diags(a, format="csc")), x, wbar, n_t, p,
                              colptr, rowval)
    rows = rowval - 1
    cols = np.repeat(np.arange(n), np.diff(colptr))
    P = sum(a ** j / math.factorial(j) for j in range(p + 1))
    dP = sum(a ** j / math.factorial(j) for j in range(p))
    phi = P ** n_t
    diag = rows == cols
    assert diag.sum() == n and (~diag).sum() >= 3
    with np.errstate(invalid="ignore", divide="ignore"):
        dd = np.where(diag, n_t * P[rows] ** (n_t - 1) * dP[rows],
                      (phi[rows] - phi[cols]) / (a[rows] - a[cols]))
    want = wbar[rows] * x[cols] * dd
    assert np.allclose(g, want, rtol=1e-11, atol=1e-13)
    B = rand_sparse(25, 0.2, 24)
    colptr, rowval = _pattern(B)
    x, wbar = rng.standard_normal(25), rng.standard_normal(25)
    g1 = orc.taylor_vjp_values(orc.CSC.from_scipy(B), x, wbar, 1, 1, colptr, rowval)
    rows = rowval - 1
    cols = np.repeat(np.arange(25), np.diff(colptr))
    assert np.allclose(g1, wbar[rows] * x[cols], rtol=1e-15, atol=0)
