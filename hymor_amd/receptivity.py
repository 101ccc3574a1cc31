"""Freestream-receptivity operator from its parts.

The reference builds the extended operator (LINEARIZE_L_.jl:58-67)

    L_ = [ L   B  B ... B ;
           0   kron(diag(-i w_1 .. -i w_N), I_p) ]

with N copies of the upstream coupling B, and applies T_5(L_ dt)^n to one
vector of length n_L + N p (Freestream_receptivity/FREESTREAM_RECEPTIVITY.jl,
CONSTRUCT_C.jl:78-111).  Two device forms keep B once:

* ``receptivity_operator`` (the default of FactoredReceptivity): L_ itself,
  built by hymor_op_create_receptivity from one per-wavenumber system
  [L B; 0 0] and the wavenumbers.  Forward, the flow rows read B times the
  column sums S[q] = sum_k f_k[q] of the freestream blocks (the sum over the
  N copies done once per term); adjoint, B' is evaluated once for the N
  blocks.  A term costs one flow SpMV, as for the assembled upload.
* per-wavenumber systems (``per_wavenumber=True``): since L_ is linear and
  the freestream blocks do not interact, the action splits exactly into N
  systems

      S_k = [ L  B ; 0  -i w_k I_p ]       (size n_L + p)

  that share L and B: with x = [q; f_1 .. f_N], column k carries
  [q d_k1; f_k] (the flow part rides with the first wavenumber), and
  afterwards q_out = sum_k q_k, f_k,out = f_k.  For the adjoint L_' the flow
  part is shared instead: every column carries [y_q; y_fk], the flow result
  is that of any column and f_k,out comes from column k.  All N systems run
  as one batched sweep (one complex column each, up to 32) with the
  per-column diagonal -i w_k on the freestream rows.  Each column pays its
  own flow SpMV (N per term); what it adds is each wavenumber's flow
  response q_k.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

import ctypes as C

from . import _lib
from ._lib import C128, F64, check, ptr
from .core import _FORMS, _PRECS, DeviceMatrix, DeviceOperator


def per_omega_system(L, B):
    """[L B; 0 0] (scipy, var-major flow part first) of size n_L + p."""
    L = sp.csc_matrix(L)
    B = sp.csc_matrix(B)
    nL, p = B.shape
    if L.shape != (nL, nL):
        raise ValueError("B must have as many rows as L")
    return sp.bmat([[L, B], [sp.csc_matrix((p, nL)), sp.csc_matrix((p, p))]], format="csc")


def split_extended(A_, n_L: int, p: int, w_infty):
    """Recover (L, B) from an extended operator in the reference layout and
    check that it has the reference structure (repeated B, -i w diagonal)."""
    A_ = sp.csc_matrix(A_)
    N = len(w_infty)
    if A_.shape[0] != n_L + N * p:
        raise ValueError("size does not match n_L + N p")
    L = A_[:n_L, :n_L]
    B = A_[:n_L, n_L:n_L + p]
    for k in range(1, N):
        Bk = A_[:n_L, n_L + k * p:n_L + (k + 1) * p]
        if abs(Bk - B).max() != 0:
            raise ValueError("the freestream columns are not repeated copies of B")
    # compared sparse: the block is N*p square (64 000 at 32 wavenumbers)
    D = sp.csr_matrix(A_[n_L:, n_L:])
    w = np.asarray(w_infty, dtype=float)
    want = sp.kron(sp.diags(-1j * w), sp.identity(p), format="csr")
    diff = (D - want).tocoo()
    tol = 1e-14 * max(1.0, float(np.abs(w).max()))
    if diff.nnz and np.abs(diff.data).max() > tol:
        raise ValueError("the freestream block is not kron(diag(-i w), I)")
    if A_[n_L:, :n_L].nnz and abs(A_[n_L:, :n_L]).max() != 0:
        raise ValueError("the lower-left block must be zero")
    return L, B


def receptivity_operator(L, B, w_infty, *, nchi: int, neta: int, vars_per_cell: int = 4,
                         scale: float = 1.0, value_prec: str = "f64") -> DeviceOperator:
    """The extended operator L_ on the device from its parts
    (hymor_op_create_receptivity): the per-wavenumber system [L B; 0 0] is
    uploaded once and the N wavenumbers give the -i w_k diagonal; B is kept
    once.  The handle is a complex DeviceOperator of size n_L + N p that
    serves every action (single vector, batched, adjoint, Chu sandwich)."""
    S = per_omega_system(L, B)
    S.sort_indices()
    w = np.ascontiguousarray(w_infty, dtype=np.float64)
    if w.ndim != 1 or w.shape[0] < 1:
        raise ValueError("one or more wavenumbers")
    cplx = np.iscomplexobj(S.data)
    nz = np.ascontiguousarray(S.data, dtype=np.complex128 if cplx else np.float64)
    colptr = np.ascontiguousarray(S.indptr, dtype=np.int64)
    rowval = np.ascontiguousarray(S.indices, dtype=np.int64)
    h = C.c_void_p()
    check(_lib.load().hymor_op_create_receptivity(
        S.shape[0], ptr(colptr), ptr(rowval), ptr(nz), 0, C128 if cplx else F64, float(scale),
        vars_per_cell, nchi, neta, L.shape[0], w.shape[0], ptr(w), _PRECS[value_prec],
        C.byref(h)))
    return DeviceOperator(h)


class FactoredReceptivity:
    """Device form of T_p(L_ dt)^n and its adjoint for the extended
    receptivity operator, without materialising the N copies of B."""

    def __init__(self, L, B, w_infty, *, nchi: int, neta: int, vars_per_cell: int = 4,
                 scale: float = 1.0, value_prec: str = "f64", per_wavenumber: bool = False):
        """L: (n_L x n_L) flow operator in the reference's var-major order,
        flow DOFs (vars_per_cell per cell) first, then any shock DOFs;
        B: (n_L x p) freestream coupling of one wavenumber.
        per_wavenumber: run the N per-wavenumber systems as N columns of one
        batched sweep instead of the extended operator."""
        self.w = np.asarray(w_infty, dtype=float)
        self.N = self.w.shape[0]
        self.per_wavenumber = bool(per_wavenumber)
        if self.N < 1 or (self.per_wavenumber and self.N > 32):
            raise ValueError("1 to 32 wavenumbers (32 per sweep in per-wavenumber form)")
        self.n_L = L.shape[0]
        self.p = B.shape[1]
        self.scale = float(scale)
        if self.n_L < vars_per_cell * nchi * neta:
            raise ValueError("L is smaller than the flow part of the grid")
        self.n_sys = self.n_L + self.p
        self.n_ext = self.n_L + self.N * self.p
        if self.per_wavenumber:
            S = per_omega_system(L, B).astype(np.complex128)
            self.dev = self._make_device(S, nchi, neta, vars_per_cell, scale, value_prec)
            # the -i w_k diagonal in the units of the stored (scaled) values
            self.diag = np.ascontiguousarray(
                np.stack([np.zeros(self.N), -self.w * self.scale], axis=1).reshape(-1))
        else:
            self.dev = receptivity_operator(L, B, self.w, nchi=nchi, neta=neta,
                                            vars_per_cell=vars_per_cell, scale=scale,
                                            value_prec=value_prec)

    @staticmethod
    def _make_device(S, nchi, neta, vars_per_cell, scale, value_prec):
        return DeviceOperator.from_scipy(S, nchi=nchi, neta=neta, vars_per_cell=vars_per_cell,
                                         scale=scale, value_prec=value_prec)

    def close(self):
        self.dev.close()

    def _run(self, X, order, nsub, dt, form, transpose):
        dX = DeviceMatrix.from_host(X)
        dY = DeviceMatrix(self.n_sys, self.N, np.complex128)
        try:
            check(_lib.load().hymor_expm_taylor_batched_coldiag_dev(
                self.dev.handle, int(bool(transpose)), int(order), int(nsub), float(dt),
                _FORMS[form], self.N, self.diag.ctypes.data, self.n_L, self.n_sys, dX.ptr,
                self.n_sys, dY.ptr, self.n_sys))
            return dY.download()
        finally:
            dX.free()
            dY.free()

    def expm(self, x, *, order=5, nsub=1, dt=1.0, form="powers", transpose=False):
        """T_order(L_ dt)^nsub x (transpose: the adjoint), x in the reference's
        extended layout [q; f_1 .. f_N]."""
        x = np.asarray(x, dtype=np.complex128)
        nL, p, N = self.n_L, self.p, self.N
        if x.shape != (self.n_ext,):
            raise ValueError("x must have length n_L + N p")
        if not self.per_wavenumber:
            return self.dev.expm_taylor(x, order=order, nsub=nsub, dt=dt, form=form,
                                        transpose=transpose)
        f = x[nL:].reshape(N, p)
        X = np.zeros((self.n_sys, N), dtype=np.complex128, order="F")
        X[nL:, :] = f.T
        if transpose:
            X[:nL, :] = x[:nL, None]
        else:
            X[:nL, 0] = x[:nL]
        Y = self._run(X, order, nsub, dt, form, transpose)
        out = np.empty(self.n_ext, dtype=np.complex128)
        out[:nL] = Y[:nL, 0] if transpose else Y[:nL, :].sum(axis=1)
        out[nL:] = Y[nL:, :].T.reshape(-1)
        return out

    def expm_batched(self, X, *, order=5, nsub=1, dt=1.0, form="powers", transpose=False):
        """The action on the columns of X (n_ext x m, m <= 32 complex
        vectors), one batched sweep of the extended operator."""
        if self.per_wavenumber:
            raise ValueError("batches of extended vectors run on the extended operator")
        return self.dev.expm_taylor_batched(X, order=order, nsub=nsub, dt=dt, form=form,
                                            transpose=transpose)
