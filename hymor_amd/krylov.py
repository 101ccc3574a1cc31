"""Krylov-Schur eigensolver over a device-resident basis.

The reference calls ARPACK (Arpack.jl) on the host; every matvec copies the
vector to the GPU and back (utils/Stability_analysis/Eigenvalues/
ARNOLDI_POWER_GPU.jl:42-44) and all orthogonalisation runs in host BLAS.
Here the basis V lives in HBM (``hymor_krylov``): the operator is applied
device-to-device, classical Gram-Schmidt with reorthogonalisation (CGS2) runs
as tall-skinny GPU kernels, restarts rotate the basis on the GPU, and the
host only handles the (m+1) x m Hessenberg matrix -- real Schur form with
reordering (Stewart's Krylov-Schur restart), Ritz values and residual
estimates.

``which="LM"`` (largest magnitude) is the selection all three reference
drivers use.  Convergence: ||A y - theta y|| <= tol * |theta| for each of the
k wanted Ritz pairs, the estimate ARPACK uses.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.linalg as sl

from . import _lib
from ._lib import C128, F64, check, ptr
from .core import DeviceMatrix, DeviceVector


class DeviceKrylov:
    """A device basis of m_max+1 vectors of length n (real or complex)."""

    def __init__(self, n: int, m_max: int, complex_: bool = False):
        self.n, self.m_max = n, m_max
        self.cplx = complex_
        self.dtype = np.complex128 if complex_ else np.float64
        self._h = C.c_void_p()
        check(_lib.load().hymor_krylov_create(n, m_max, C128 if complex_ else F64,
                                              C.byref(self._h)))

    def close(self):
        if self._h is not None and self._h.value:
            check(_lib.load().hymor_krylov_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def column(self, j: int) -> C.c_void_p:
        p = C.c_void_p()
        check(_lib.load().hymor_krylov_column(self._h, j, C.byref(p)))
        return p

    def set_column(self, j: int, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        check(_lib.load().hymor_krylov_set_column(self._h, j, ptr(x)))

    def get_columns(self, j0: int, ncols: int):
        out = np.empty((ncols, self.n), dtype=self.dtype)
        check(_lib.load().hymor_krylov_get_columns(self._h, j0, ncols, ptr(out)))
        return out.T

    def orthonormalize(self, j: int, d_w) -> np.ndarray:
        """CGS2 of device vector d_w against columns 0..j-1, stored as column
        j; returns the j coefficients followed by the norm."""
        nc = 2 if self.cplx else 1
        h = np.zeros(nc * (j + 1))
        check(_lib.load().hymor_krylov_orthonormalize(self._h, j, d_w, h.ctypes.data_as(
            C.POINTER(C.c_double))))
        if self.cplx:
            hc = h.view(np.complex128)
            return hc
        return h

    def orthonormalize_block(self, j: int, d_W, ldw: int, s: int) -> np.ndarray:
        """Block CGS2 of the s device columns d_W (leading dimension ldw)
        against columns 0..j-1, then among themselves, stored as columns
        j..j+s-1; returns the (j+s) x s coefficients (W = V[:, :j+s] H).
        Real bases only."""
        if self.cplx:
            raise ValueError("block orthonormalisation covers real bases")
        h = np.zeros((s, j + s))
        check(_lib.load().hymor_krylov_orthonormalize_block(
            self._h, j, s, d_W, ldw, h.ctypes.data_as(C.POINTER(C.c_double))))
        return h.T

    def combine(self, Q: np.ndarray):
        """V[:, :k] <- V[:, :m] Q for Q of shape (m, k)."""
        m, k = Q.shape
        Qf = np.asfortranarray(Q, dtype=self.dtype)
        check(_lib.load().hymor_krylov_combine(self._h, m, k,
                                               Qf.ctypes.data_as(C.POINTER(C.c_double))))

    def ritz(self, Q: np.ndarray) -> np.ndarray:
        """V[:, :m] Q (n x k) to the host, leaving V unchanged."""
        m, k = Q.shape
        Qf = np.asfortranarray(Q, dtype=self.dtype)
        out = np.empty((k, self.n), dtype=self.dtype)
        check(_lib.load().hymor_krylov_ritz(self._h, m, k,
                                            Qf.ctypes.data_as(C.POINTER(C.c_double)), ptr(out)))
        return out.T

    def copy_column(self, src: int, dst: int):
        check(_lib.load().hymor_krylov_copy_column(self._h, src, dst))

    def set_column_dev(self, j: int, d_src: C.c_void_p):
        """Column j <- n values at device address d_src."""
        check(_lib.load().hymor_memcpy_d2d(self.column(j), d_src,
                                           self.n * np.dtype(self.dtype).itemsize))


    def set_column_random(self, j: int, seed: int):
        """Column j <- uniform [-1, 1) entries (real and imaginary parts),
        generated on the device from (seed, index)."""
        ncomp = 2 if self.cplx else 1
        check(_lib.load().hymor_dev_fill_uniform(self.column(j), self.n * ncomp, int(seed)))


def _ritz_vectors(K: "DeviceKrylov", Q: np.ndarray, complex_: bool) -> np.ndarray:
    """Unit Ritz vectors V Q (n x k, complex128).  V is orthonormal, so the
    columns are normalised in the small space; a real basis takes the real
    and imaginary parts of Q separately (the imaginary part only if any)."""
    Q = Q / np.linalg.norm(Q, axis=0)
    if complex_:
        return K.ritz(Q)
    Yt = K.ritz(np.ascontiguousarray(Q.real)).T.astype(np.complex128)
    if np.any(Q.imag):
        Yt.imag = K.ritz(np.ascontiguousarray(Q.imag)).T
    return Yt.T


def _fresh_direction(K: "DeviceKrylov", j: int, seed: int, tries: int = 3):
    """Column j <- a random unit vector orthogonal to columns 0..j-1."""
    for t in range(tries):
        K.set_column_random(j, seed + 104729 * t)
        if K.orthonormalize(j, K.column(j))[j] != 0:
            return
    raise RuntimeError("no direction orthogonal to the Krylov basis was found")


def _select_lm(evals: np.ndarray, k: int) -> float:
    """Magnitude threshold so that the k largest-|lambda| values are kept."""
    mags = np.sort(np.abs(evals))[::-1]
    return mags[min(k, len(mags)) - 1]


def krylov_schur(apply_dev, n: int, k: int, m: int, *, complex_: bool = False, tol: float = 1e-10,
                 maxiter: int = 300, v0=None, return_vectors: bool = True, seed: int = 0):
    """Leading-magnitude eigenpairs of a linear operator given as
    ``apply_dev(d_in, d_out)`` (device pointers, synchronous).

    Returns (theta, Y, info): theta the k Ritz values sorted by decreasing
    magnitude, Y the Ritz vectors (n x k, or None), info a dict with
    iterations, operator applications and residual estimates.  Raises
    RuntimeError if not converged within maxiter restarts (the callers turn
    that into NaN results, as the reference does).
    """
    if not 1 <= k < m:
        raise ValueError("need 1 <= k < m")
    if m >= n:
        raise ValueError("Krylov dimension must be smaller than n")
    dtype = np.complex128 if complex_ else np.float64
    K = DeviceKrylov(n, m, complex_)
    scratch = DeviceVector(n, dtype)
    try:
        if v0 is None:                                     # random start on the device
            K.set_column_random(0, seed)
        else:
            K.set_column(0, np.asarray(v0, dtype=dtype))
        if K.orthonormalize(0, K.column(0))[0] == 0:        # normalise in place
            raise ValueError("the start vector is zero")
        H = np.zeros((m + 1, m), dtype=dtype)
        p = 0
        n_apply = 0
        n_fresh = 0
        for it in range(1, maxiter + 1):
            for j in range(p, m):
                apply_dev(K.column(j), scratch.ptr)
                n_apply += 1
                h = K.orthonormalize(j + 1, scratch.ptr)
                H[: j + 2, j] = h[: j + 2]
                if h[j + 1] == 0:
                    # V[:, :j+1] spans an invariant subspace (H[j+1, j] = 0
                    # keeps the Krylov relation exact): continue from a
                    # random direction orthogonal to it, as ARPACK does
                    n_fresh += 1
                    _fresh_direction(K, j + 1, 7919 * (seed + 1) + n_fresh)
            beta = H[m, m - 1]
            Hm = H[:m, :m]
            evals = np.linalg.eigvals(Hm)
            thr = _select_lm(evals, k)
            if complex_:
                T, Z, sdim = sl.schur(Hm, output="complex", sort=lambda z: abs(z) >= thr * (1 - 1e-12))
            else:
                T, Z, sdim = sl.schur(Hm, output="real",
                                      sort=lambda re, im: abs(complex(re, im)) >= thr * (1 - 1e-12))
            p_keep = max(sdim, k)
            # Ritz pairs of the leading block and their residual estimates
            w, Yt = sl.eig(T[:p_keep, :p_keep])
            order = np.argsort(-np.abs(w), kind="stable")[:k]
            w = w[order]
            Yt = Yt[:, order]
            zlast = Z[m - 1, :p_keep] @ Yt                     # last row of Z*y
            res = np.abs(beta) * np.abs(zlast)
            conv = res <= tol * np.maximum(np.abs(w), np.finfo(float).tiny)
            if np.all(conv):
                info = {"iterations": it, "applications": n_apply, "residual_estimates": res}
                Y = None
                if return_vectors:
                    Y = _ritz_vectors(K, Z[:, :p_keep] @ Yt, complex_)
                return w, Y, info
            # Krylov-Schur restart: keep the p_keep leading Schur vectors
            p = min(p_keep + (k - int(conv.sum())) // 2, m - 2)
            p = max(p, p_keep)
            if not complex_ and p < m and p > p_keep:
                # do not split a 2x2 block of the real Schur form
                if abs(T[p, p - 1]) > 0:
                    p += 1
            p = min(p, m - 1)
            K.combine(Z[:, :p])
            K.copy_column(m, p)
            Hn = np.zeros((m + 1, m), dtype=dtype)
            Hn[:p, :p] = T[:p, :p]
            Hn[p, :p] = beta * Z[m - 1, :p]
            H = Hn
        raise RuntimeError(f"Krylov-Schur did not converge in {maxiter} restarts")
    finally:
        scratch.free()
        K.close()


def block_krylov_schur(apply_block, n: int, k: int, m: int, s: int, *, complex_: bool = False,
                       tol: float = 1e-10, maxiter: int = 200, V0=None, basis: int = 0,
                       return_vectors: bool = True, seed: int = 0):
    """Block Krylov-Schur: ``apply_block(d_in, ld, nvec, d_out)`` applies the
    operator to nvec consecutive device columns (leading dimension ld) at
    once -- the batched SpMM action -- and every expansion step feeds it the
    s newest basis vectors.  The s results are appended by sequential CGS2
    (the same device orthonormaliser as krylov_schur), so the projected
    matrix is band Hessenberg with s sub-diagonals.

    V0 (n x s host array or a DeviceMatrix, optional) seeds the first
    block, e.g. a basis from an earlier, cheaper solve; a seed of fewer than
    s columns is padded with random directions.  Returns (theta, Y, info) as
    krylov_schur; with basis > 0, info["basis"] also holds a DeviceMatrix
    (n x about `basis` columns, owned by the caller) with an orthonormal
    basis of the Schur subspace of the `basis` largest-magnitude Ritz values
    -- the seed for a refinement on a more accurate operator;
    return_vectors=False skips the Ritz vectors (Y is None); `seed` selects
    the random start directions.
    """
    if not (1 <= k and k + s <= m):
        raise ValueError("need k + s <= m")
    dtype = np.complex128 if complex_ else np.float64
    K = DeviceKrylov(n, m + s, complex_)
    W = DeviceMatrix(n, s, dtype)
    try:
        colbytes = n * np.dtype(dtype).itemsize
        on_device = V0 is not None and hasattr(V0, "nrhs")   # a DeviceMatrix
        n_seed = 0 if V0 is None else (V0.nrhs if on_device else np.shape(V0)[1])
        if not on_device and V0 is not None:
            V0 = np.asarray(V0, dtype=dtype)
        H = np.zeros((m + s, m + s), dtype=dtype)
        # orthonormal first block (dependent seed columns are replaced)
        for i in range(s):
            if i >= n_seed:                   # random directions (pad a short seed)
                K.set_column_random(i, 2 * s * seed + i)
            elif on_device:
                K.set_column_dev(i, C.c_void_p(V0.ptr.value + i * colbytes))
            else:
                K.set_column(i, V0[:, i])
            h = K.orthonormalize(i, K.column(i))
            if abs(h[i]) == 0:
                K.set_column_random(i, 2 * s * seed + s + i)
                K.orthonormalize(i, K.column(i))
        j = 0              # first column whose image is not yet in the basis
        n_apply = 0
        # the wanted set is widened by one block where the basis leaves room:
        # with clustered magnitudes the k-th and (k+1)-th Ritz values
        # otherwise trade places between restarts, and each restart drops the
        # directions that were converging
        k_int = max(k, min(k + s, m - 2 * s))
        for it in range(1, maxiter + 1):
            while j + s <= m:
                apply_block(K.column(j), n, s, W.ptr)
                n_apply += s
                if not complex_:
                    # one block CGS2: a pass over the basis per projection
                    Hb = K.orthonormalize_block(j + s, W.ptr, n, s)
                    for i in range(s):
                        if Hb[j + s + i, i] == 0.0:
                            raise RuntimeError("block Krylov-Schur: the block lost rank")
                    H[: j + 2 * s, j:j + s] = Hb
                else:
                    for i in range(s):
                        wcol = C.c_void_p(W.ptr.value + i * colbytes)
                        h = K.orthonormalize(j + s + i, wcol)
                        H[: j + s + i + 1, j + i] = h[: j + s + i + 1]
                j += s
            mm = j
            Hm = H[:mm, :mm]
            Bres = H[mm:mm + s, :mm]          # residual rows
            evals = np.linalg.eigvals(Hm)
            thr = _select_lm(evals, k_int)
            if complex_:
                T, Z, sdim = sl.schur(Hm, output="complex", sort=lambda z: abs(z) >= thr * (1 - 1e-12))
            else:
                T, Z, sdim = sl.schur(Hm, output="real",
                                      sort=lambda re, im: abs(complex(re, im)) >= thr * (1 - 1e-12))
            p_keep = max(sdim, k_int)
            w, Yt = sl.eig(T[:p_keep, :p_keep])
            order = np.argsort(-np.abs(w), kind="stable")[:k_int]
            w, Yt = w[order], Yt[:, order]
            res = np.linalg.norm(Bres @ (Z[:, :p_keep] @ Yt), axis=0)
            conv = res <= tol * np.maximum(np.abs(w), np.finfo(float).tiny)
            if np.all(conv[:k]):
                w, Yt, res = w[:k], Yt[:, :k], res[:k]
                Y = _ritz_vectors(K, Z[:, :p_keep] @ Yt, complex_) if return_vectors else None
                info = {"iterations": it, "applications": n_apply, "residual_estimates": res}
                if basis:
                    nb = min(basis, mm)
                    thr_b = _select_lm(evals, nb) * (1 - 1e-12)
                    if complex_:
                        _, Zb, sd = sl.schur(Hm, output="complex", sort=lambda z: abs(z) >= thr_b)
                    else:
                        _, Zb, sd = sl.schur(Hm, output="real",
                                             sort=lambda re, im: abs(complex(re, im)) >= thr_b)
                    nb = min(max(sd, nb), mm)
                    K.combine(Zb[:, :nb])          # V[:, :nb] <- the Schur basis
                    Bd = DeviceMatrix(n, nb, dtype)
                    check(_lib.load().hymor_memcpy_d2d(Bd.ptr, K.column(0), nb * colbytes))
                    info["basis"] = Bd
                return w, Y, info
            # restart: keep p Schur vectors, then the residual block
            p = max(p_keep, min(p_keep + (k_int - int(conv.sum())), mm - s))
            if not complex_ and p < mm and p > p_keep and abs(T[p, p - 1]) > 0:
                p += 1
            p = min(p, mm - 1)
            Bz = Bres @ Z[:, :p]
            K.combine(Z[:, :p])
            for i in range(s):
                K.copy_column(mm + i, p + i)
            H = np.zeros((m + s, m + s), dtype=dtype)
            H[:p, :p] = T[:p, :p]
            H[p:p + s, :p] = Bz
            j = p
        raise RuntimeError(f"block Krylov-Schur did not converge in {maxiter} restarts")
    finally:
        W.free()
        K.close()
