"""CPU oracle for HYMOR's exp(A*dt) Taylor path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  It is the checker, never the product path.

Two layers:
  * ctypes wrappers over hymor_oracle.c, a line-by-line restatement of the
    reference's CPU code path on its own operator format (1-based,
    variable-major Julia CSC):
      - csc_mul        SparseArrays mul!(y, A, x, 1.0, 0.0)  (column scatter)
      - csc_adjoint    sparse(A_dt')   TRANSIENT_GROWTH_DOWNSTREAM.jl:75
      - power_method   POWER_METHOD!   Eigenvalues/ARNOLDI_POWER_GPU.jl:64-82
      - exponential    EXPONENTIAL!    Transient_growth_downstream/CONSTRUCT_C.jl:78-111
      - construct_c    CONSTRUCT_C     CONSTRUCT_C.jl:33-58 (zeroing of x[1:N_set_0])
      - linear_integration_energy
                       LINEAR_INTEGRATION_AND_ENERGY_GPU_NO_SAMPLES
                       Freestream_receptivity/LINEAR_INTEGRATION_AND_GAINS.jl:436-477
  * taylor_vjp_values: the gradient of <wbar, P(A)^n x> (POWER_METHOD!'s
    polynomial) with respect to A's entries, every forward state kept (no
    reference counterpart: the reference has no derivatives);
  * closed forms that pin those restatements independently of them
    (block-circulant operators diagonalised by the 2-D DFT, dense matrix
    polynomials, diagonal and nilpotent operators).

Parity status per function (what pins it) is recorded in the tests that use
it: tests/test_oracle_pins.py.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        p = C.c_void_p
        i64 = C.c_int64
        lib.oc_csc_mul_f64.argtypes = [i64, p, p, p, p, p]
        lib.oc_csc_mul_c128.argtypes = [i64, p, p, p, p, p]
        lib.oc_csc_adjoint.argtypes = [i64, p, p, p, C.c_int, p, p, p]
        lib.oc_power_method_f64.argtypes = [i64, p, p, p, C.c_int, i64, p, p, p]
        lib.oc_power_method_c128.argtypes = [i64, p, p, p, C.c_int, i64, p, p, p]
        lib.oc_exponential_f64.argtypes = [i64, p, p, p, p, p, p, p, p, p, C.c_int, i64, p, p, p]
        lib.oc_exponential_c128.argtypes = [i64, p, p, p, p, p, p, p, p, p, C.c_int, C.c_int, i64,
                                            p, p, p]
        lib.oc_linear_integration_energy.argtypes = [i64, p, p, p, C.c_int, C.c_int, i64, p, p, p,
                                                     C.c_int, p, p]
        lib.oc_taylor_csr_omp.argtypes = [i64, p, p, p, C.c_int, i64, p, p, p, p, C.c_int]
        lib.oc_taylor_csr_omp.restype = C.c_int
        _lib = lib
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data)


@dataclass
class CSC:
    """Julia SparseMatrixCSC fields, 1-based int64 indices."""

    n: int
    colptr: np.ndarray
    rowval: np.ndarray
    nzval: np.ndarray

    @property
    def is_complex(self):
        return np.iscomplexobj(self.nzval)

    @classmethod
    def from_scipy(cls, A):
        A = A.tocsc()
        A.sort_indices()
        cplx = np.iscomplexobj(A.data)
        return cls(A.shape[0], (A.indptr.astype(np.int64) + 1), (A.indices.astype(np.int64) + 1),
                   np.ascontiguousarray(A.data, dtype=np.complex128 if cplx else np.float64))

    @classmethod
    def from_arrays(cls, colptr, rowval, nzval):
        colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        cplx = np.iscomplexobj(nzval)
        return cls(colptr.shape[0] - 1, colptr, np.ascontiguousarray(rowval, dtype=np.int64),
                   np.ascontiguousarray(nzval, dtype=np.complex128 if cplx else np.float64))

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval, self.rowval - 1, self.colptr - 1), shape=(self.n, self.n))

    def scaled(self, s):
        return CSC(self.n, self.colptr, self.rowval, self.nzval * s)

    def as_complex(self):
        return CSC(self.n, self.colptr, self.rowval, self.nzval.astype(np.complex128))


def csc_mul(A: CSC, x):
    lib = load()
    if A.is_complex or np.iscomplexobj(x):
        Ac = A if A.is_complex else A.as_complex()
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        lib.oc_csc_mul_c128(A.n, _p(Ac.colptr), _p(Ac.rowval), _p(Ac.nzval), _p(x), _p(y))
        return y
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib.oc_csc_mul_f64(A.n, _p(A.colptr), _p(A.rowval), _p(A.nzval), _p(x), _p(y))
    return y


def csc_adjoint(A: CSC) -> CSC:
    lib = load()
    nnz = int(A.colptr[-1] - 1)
    tcp = np.empty(A.n + 1, dtype=np.int64)
    trv = np.empty(nnz, dtype=np.int64)
    tnz = np.empty(nnz, dtype=A.nzval.dtype)
    lib.oc_csc_adjoint(A.n, _p(A.colptr), _p(A.rowval), _p(A.nzval), int(A.is_complex), _p(tcp),
                       _p(trv), _p(tnz))
    return CSC(A.n, tcp, trv, tnz)


def power_method(A: CSC, v, timesteps: int, order: int = 5):
    """POWER_METHOD!(v, A, timesteps, vout, term) -- returns the new v."""
    lib = load()
    if A.is_complex or np.iscomplexobj(v):
        Ac = A if A.is_complex else A.as_complex()
        v = np.array(v, dtype=np.complex128)
        vout = np.empty_like(v)
        term = np.empty_like(v)
        lib.oc_power_method_c128(A.n, _p(Ac.colptr), _p(Ac.rowval), _p(Ac.nzval), order, timesteps,
                                 _p(v), _p(vout), _p(term))
        return v
    v = np.array(v, dtype=np.float64)
    vout = np.empty_like(v)
    term = np.empty_like(v)
    lib.oc_power_method_f64(A.n, _p(A.colptr), _p(A.rowval), _p(A.nzval), order, timesteps, _p(v),
                            _p(vout), _p(term))
    return v


def taylor_vjp_values(A: CSC, x, wbar, timesteps: int, order: int, colptr, rowval):
    """Gradient of J = <wbar, P(A)^timesteps x>, P(A) = sum_{j <= order} A^j / j!
    (POWER_METHOD!'s polynomial, ARNOLDI_POWER_GPU.jl:70-79, with A the
    scaled operator A*dt), with respect to the entries (rowval[e], column of
    e) of a 1-based CSC pattern (colptr, rowval), real operators.  The
    reverse sweep of the polynomial: with lambda_n = wbar and x_{k+1} =
    P(A) x_k,
        d<l, A^j x> = sum_{m < j} <(A')^m l, dA A^(j-1-m) x>,
        grad += sum_{m < order} u_m[row] z_m[col],  u_m = (A')^m lambda_{k+1},
        z_m = sum_{r < order-m} A^r x_k / (m+r+1)!,   lambda_k = P(A') lambda_{k+1},
    every forward state kept, the SpMVs those of csc_mul on A and its
    adjoint (pinned by central differences of power_method and by closed
    forms in tests/test_oracle_pins.py)."""
    At = csc_adjoint(A)
    p = int(order)
    colptr = np.asarray(colptr, dtype=np.int64)
    rows = np.asarray(rowval, dtype=np.int64) - 1
    cols = np.repeat(np.arange(colptr.shape[0] - 1), np.diff(colptr))
    fact = [1.0]
    for j in range(1, p + 1):
        fact.append(fact[-1] / j)
    xs = [np.array(x, dtype=np.float64)]
    for _ in range(int(timesteps)):
        xs.append(power_method(A, xs[-1], 1, p))
    lam = np.array(wbar, dtype=np.float64)
    grad = np.zeros(rows.shape[0])
    for k in reversed(range(int(timesteps))):
        y = [xs[k]]
        for _ in range(1, p):
            y.append(csc_mul(A, y[-1]))
        u = [lam]
        for _ in range(p):
            u.append(csc_mul(At, u[-1]))
        for m in range(p):
            z = np.zeros_like(lam)
            for r in range(p - m):
                z = z + fact[m + r + 1] * y[r]
            grad += u[m][rows] * z[cols]
        lam = np.zeros_like(lam)
        for m in range(p + 1):
            lam = lam + fact[m] * u[m]
    return grad


def exponential(A: CSC, At: CSC, Cw: CSC, v, timesteps: int, order: int = 5):
    """EXPONENTIAL!(v, A, At, C_no_exp, timesteps, vout, term)."""
    lib = load()
    if A.is_complex or np.iscomplexobj(v):
        Ac = A if A.is_complex else A.as_complex()
        Atc = At if At.is_complex else At.as_complex()
        v = np.array(v, dtype=np.complex128)
        vout = np.empty_like(v)
        term = np.empty_like(v)
        lib.oc_exponential_c128(A.n, _p(Ac.colptr), _p(Ac.rowval), _p(Ac.nzval), _p(Atc.colptr),
                                _p(Atc.rowval), _p(Atc.nzval), _p(Cw.colptr), _p(Cw.rowval),
                                _p(Cw.nzval), int(Cw.is_complex), order, timesteps, _p(v),
                                _p(vout), _p(term))
        return v
    v = np.array(v, dtype=np.float64)
    vout = np.empty_like(v)
    term = np.empty_like(v)
    lib.oc_exponential_f64(A.n, _p(A.colptr), _p(A.rowval), _p(A.nzval), _p(At.colptr),
                           _p(At.rowval), _p(At.nzval), _p(Cw.colptr), _p(Cw.rowval), _p(Cw.nzval),
                           order, timesteps, _p(v), _p(vout), _p(term))
    return v


def construct_c(x, A: CSC, At: CSC, Cw: CSC, n_t: int, freestream_disturbances: bool,
                n_set_0: int, order: int = 5):
    """CONSTRUCT_C (CONSTRUCT_C.jl:33-58): zero x[1:N_set_0] before/after."""
    x = np.array(x)
    if freestream_disturbances:
        x[:n_set_0] = 0
    y = exponential(A, At, Cw, x, n_t, order)
    if freestream_disturbances:
        y[:n_set_0] = 0
    return y


def linear_integration_energy(A: CSC, weights, x, n_t: int, order: int = 5):
    """Energies E_w[0..n_t] of LINEAR_INTEGRATION_AND_ENERGY_GPU_NO_SAMPLES.

    Returns (E, v_final) with E of shape (len(weights), n_t+1); E[:,0] = 0
    as in the reference (its caller fills entry 0, :168-173).
    """
    lib = load()
    cplx = A.is_complex or np.iscomplexobj(x)
    Ac = A.as_complex() if (cplx and not A.is_complex) else A
    v = np.array(x, dtype=np.complex128 if cplx else np.float64)
    nw = len(weights)
    cps = (C.c_void_p * nw)(*[w.colptr.ctypes.data for w in weights])
    rvs = (C.c_void_p * nw)(*[w.rowval.ctypes.data for w in weights])
    nzs = (C.c_void_p * nw)(*[w.nzval.ctypes.data for w in weights])
    E = np.zeros((nw, n_t + 1))
    lib.oc_linear_integration_energy(A.n, _p(Ac.colptr), _p(Ac.rowval), _p(Ac.nzval), int(cplx),
                                     order, n_t, cps, rvs, nzs, nw, _p(v), _p(E))
    return E, v


class CsrArrays:
    """A scipy CSR matrix's arrays in the C form the OpenMP baseline reads
    (int64 indices), converted once so that timed calls do not convert."""

    def __init__(self, A_csr):
        self.shape = A_csr.shape
        self.indptr = np.ascontiguousarray(A_csr.indptr, dtype=np.int64)
        self.indices = np.ascontiguousarray(A_csr.indices, dtype=np.int64)
        self.data = np.ascontiguousarray(A_csr.data, dtype=np.float64)


def taylor_csr_omp(A_csr, v, timesteps: int, order: int, nthreads: int = 0):
    """Row-parallel CPU form of the same action (timing baseline only).

    Normalised recursion; A_csr is a scipy CSR matrix (already scaled by dt)
    or its CsrArrays.  Returns (result, threads_used).
    """
    lib = load()
    if not isinstance(A_csr, CsrArrays):
        A_csr = CsrArrays(A_csr)
    rp, ci, va = A_csr.indptr, A_csr.indices, A_csr.data
    v = np.array(v, dtype=np.float64)
    w0 = np.empty_like(v)
    w1 = np.empty_like(v)
    vout = np.empty_like(v)
    used = lib.oc_taylor_csr_omp(A_csr.shape[0], _p(rp), _p(ci), _p(va), order, timesteps, _p(v),
                                 _p(w0), _p(w1), _p(vout), nthreads)
    return v, used


# --------------------------------------------------------------------------
# Closed forms
# --------------------------------------------------------------------------

def taylor_poly_dense(M, order: int):
    """sum_{j=0}^{order} M^j / j! for a dense (stack of) square matrices."""
    M = np.asarray(M)
    n = M.shape[-1]
    eye = np.broadcast_to(np.eye(n, dtype=M.dtype), M.shape).copy()
    P = eye.copy()
    term = eye
    for j in range(1, order + 1):
        term = term @ M / j
        P = P + term
    return P


def circulant_symbol(blocks: dict, nchi: int, neta: int):
    """S_hat[q, p] = sum B_{di,dj} exp(2 pi i (di p/nchi + dj q/neta))."""
    b = next(iter(blocks.values())).shape[0]
    p = np.arange(nchi)
    q = np.arange(neta)
    S = np.zeros((neta, nchi, b, b), dtype=np.complex128)
    for (di, dj), B in blocks.items():
        ph = np.exp(2j * np.pi * (di * p[None, :] / nchi + dj * q[:, None] / neta))
        S += ph[:, :, None, None] * B[None, None, :, :]
    return S


def _apply_symbol(Mq, v_cellmajor, nchi, neta, b):
    u = np.asarray(v_cellmajor).reshape(neta, nchi, b)
    uh = np.fft.fft2(u, axes=(0, 1))
    wh = np.einsum("qpij,qpj->qpi", Mq, uh)
    w = np.fft.ifft2(wh, axes=(0, 1))
    if not np.iscomplexobj(v_cellmajor):
        w = w.real
    return w.reshape(-1)


def circulant_expm(blocks, nchi, neta, v_cellmajor, T):
    """Exact exp(A*T) v for the block-circulant operator (cell-major v)."""
    import scipy.linalg as sl
    S = circulant_symbol(blocks, nchi, neta)
    b = S.shape[-1]
    E = sl.expm((S * T).reshape(-1, b, b)).reshape(S.shape)
    return _apply_symbol(E, v_cellmajor, nchi, neta, b)


def circulant_taylor(blocks, nchi, neta, v_cellmajor, dt, order, nsub):
    """T_order(A*dt)^nsub v for the block-circulant operator, per wavenumber."""
    S = circulant_symbol(blocks, nchi, neta)
    b = S.shape[-1]
    P = taylor_poly_dense(S * dt, order)
    Pn = np.linalg.matrix_power(P.reshape(-1, b, b), nsub).reshape(S.shape)
    return _apply_symbol(Pn, v_cellmajor, nchi, neta, b)


def circulant_eigenvalues(blocks, nchi, neta):
    S = circulant_symbol(blocks, nchi, neta)
    b = S.shape[-1]
    return np.linalg.eigvals(S.reshape(-1, b, b)).reshape(-1)


# --------------------------------------------------------------------------
# Second statement of the Chu energy (pins the package's R'MR constructors)
# --------------------------------------------------------------------------

def chu_components(rho0, rho_u0, rho_v0, p0, a0, gam0, prho, prho_u, prho_v, prho_E):
    """GET_CHU_COMPONENTS (utils/Energy_budgets/GET_CHU_COMPONENTS.jl:1-62):
    conservative perturbations -> (p', u', v', s'/R), elementwise."""
    u0 = rho_u0 / rho0
    v0 = rho_v0 / rho0
    pu = (prho_u - u0 * prho) / rho0
    pv = (prho_v - v0 * prho) / rho0
    t1 = -u0 * prho_u
    t2 = -v0 * prho_v
    t3 = (u0 ** 2 + v0 ** 2) * prho / 2
    pp = (gam0 - 1) * (prho_E + t1 + t2 + t3)
    ps = pp / p0 / (gam0 - 1) - prho / rho0 * gam0 / (gam0 - 1)
    return pp, pu, pv, ps


def energy_norm(rho0, rho_u0, rho_v0, p0, a0, gam0, prho, prho_u, prho_v, prho_E):
    """GET_ENERGY_NORM (utils/Energy_budgets/GET_ENERGY_NORM.jl:1-54):
    acoustic, kinetic and entropic Chu energy densities (real parts)."""
    pp, pu, pv, ps = chu_components(rho0, rho_u0, rho_v0, p0, a0, gam0, prho, prho_u, prho_v,
                                    prho_E)
    return {
        "acoustic": rho0 * a0 ** 2 / (2 * (gam0 * p0) ** 2) * np.real(pp) ** 2,
        "kinetic": rho0 / 2 * np.real(pu) ** 2 + rho0 / 2 * np.real(pv) ** 2,
        "entropic": (gam0 - 1) * p0 / (2 * gam0) * np.real(ps) ** 2,
    }
