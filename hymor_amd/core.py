"""Device operator handles over the C ABI (include/hymor.h).

``DeviceOperator`` replaces the reference's ``CuSparseMatrixCSC(A*dt)``
uploads (utils/Stability_analysis/Eigenvalues/EIGENVALUES.jl:139,
Transient_growth_downstream/TRANSIENT_GROWTH_DOWNSTREAM.jl:78-80) and the
Taylor loops that run on them.  Vectors are in the reference's variable-major
order; complex vectors are numpy complex128 (interleaved re/im, as Julia's
ComplexF64).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (C128, F64, FORM_HORNER, FORM_NORMALISED, FORM_POWERS, PREC_F32, PREC_F64,
                   check, ptr)

_FORMS = {"normalised": FORM_NORMALISED, "powers": FORM_POWERS, "horner": FORM_HORNER}
_PRECS = {"f64": PREC_F64, "f32": PREC_F32}


def _vec(x, complex_op: bool):
    x = np.asarray(x)
    if complex_op or np.iscomplexobj(x):
        return np.ascontiguousarray(x, dtype=np.complex128), C128
    return np.ascontiguousarray(x, dtype=np.float64), F64


class DeviceOperator:
    """A linearised operator resident on the GPU (forward and adjoint)."""

    def __init__(self, handle: C.c_void_p, keep=None):
        self._h = handle
        info = (C.c_int64 * 10)()
        check(_lib.load().hymor_op_info(self._h, info, 10))
        (self.n, self.b, self.ncell, self.nnzb, self.nnzb_padded, self.nnz_extra,
         self.extra_rows, cplx, prec, self.device_bytes) = [int(v) for v in info]
        self.is_complex = bool(cplx)
        self.value_prec = "f32" if prec == PREC_F32 else "f64"

    # ------------------------------------------------------------ creation
    @classmethod
    def from_csc(cls, n, colptr, rowval, nzval, *, index_base=1, scale=1.0, vars_per_cell=4,
                 nchi, neta, value_prec="f64"):
        lib = _lib.load()
        colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        rowval = np.ascontiguousarray(rowval, dtype=np.int64)
        cplx = np.iscomplexobj(nzval)
        nzval = np.ascontiguousarray(nzval, dtype=np.complex128 if cplx else np.float64)
        h = C.c_void_p()
        check(lib.hymor_op_create_csc(n, ptr(colptr), ptr(rowval), ptr(nzval), index_base,
                                      C128 if cplx else F64, float(scale), vars_per_cell, nchi,
                                      neta, _PRECS[value_prec], C.byref(h)))
        return cls(h)

    @classmethod
    def from_scipy(cls, A, *, nchi, neta, vars_per_cell=4, scale=1.0, value_prec="f64"):
        """From a scipy sparse matrix in the reference's var-major order."""
        A = A.tocsc()
        A.sort_indices()
        return cls.from_csc(A.shape[0], A.indptr, A.indices, A.data, index_base=0, scale=scale,
                            vars_per_cell=vars_per_cell, nchi=nchi, neta=neta,
                            value_prec=value_prec)

    @classmethod
    def from_blocks(cls, op, *, scale=1.0, value_prec="f64", with_adjoint=True):
        """From a hymor_amd.operators.BlockOperator (cell-major blocks)."""
        lib = _lib.load()
        brow_ptr = np.ascontiguousarray(op.brow_ptr, dtype=np.int64)
        bcol = np.ascontiguousarray(op.bcol, dtype=np.int32)
        if np.iscomplexobj(op.bval) and np.any(op.bval.imag != 0):
            raise ValueError("block values must be real; put imaginary parts in the extras")
        bval = np.ascontiguousarray(np.real(op.bval), dtype=np.float64)
        cplx = op.is_complex
        xrow = np.ascontiguousarray(op.xrow, dtype=np.int64)
        xcol = np.ascontiguousarray(op.xcol, dtype=np.int64)
        xval = np.ascontiguousarray(op.xval, dtype=np.complex128 if cplx else np.float64)
        h = C.c_void_p()
        check(lib.hymor_op_create_bsr(op.nchi, op.neta, op.b, op.n_extra, ptr(brow_ptr), ptr(bcol),
                                      ptr(bval), xrow.shape[0], ptr(xrow), ptr(xcol), ptr(xval),
                                      C128 if cplx else F64, float(scale), _PRECS[value_prec],
                                      1 if with_adjoint else 0, C.byref(h)))
        return cls(h)

    def set_value_prec(self, value_prec: str):
        """Switch the block values the kernels read between "f64" and "f32"
        in place.  An FP64 upload keeps its values and casts an FP32 copy on
        the device the first time; an FP32 upload cannot go back to FP64."""
        check(_lib.load().hymor_op_set_value_prec(self._h, _PRECS[value_prec]))
        self._refresh_info()
        return self

    def _refresh_info(self):
        info = (C.c_int64 * 10)()
        check(_lib.load().hymor_op_info(self._h, info, 10))
        self.device_bytes = int(info[9])
        self.value_prec = "f32" if int(info[8]) == PREC_F32 else "f64"

    # ------------------------------------------------------------ lifetime
    def close(self):
        if self._h is not None and self._h.value:
            check(_lib.load().hymor_op_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------ actions
    def spmv(self, x, transpose=False):
        x, dt = _vec(x, self.is_complex)
        y = np.empty_like(x)
        check(_lib.load().hymor_spmv(self._h, int(bool(transpose)), ptr(x), ptr(y), dt))
        return y

    def expm_taylor(self, x, *, order=5, nsub=1, dt=1.0, form="powers", transpose=False):
        """T_order(A*dt)^nsub x (transpose: the adjoint operator)."""
        x, dty = _vec(x, self.is_complex)
        y = np.empty_like(x)
        check(_lib.load().hymor_expm_taylor(self._h, int(bool(transpose)), int(order), int(nsub),
                                            float(dt), _FORMS[form], ptr(x), ptr(y), dty))
        return y

    def expm_taylor_batched(self, X, *, order=5, nsub=1, dt=1.0, form="powers", transpose=False):
        """T_order(A*dt)^nsub applied to every column of X (n x nrhs) in one
        SpMM sweep per Taylor term."""
        X = np.asarray(X)
        cplx = self.is_complex or np.iscomplexobj(X)
        X = np.asfortranarray(X, dtype=np.complex128 if cplx else np.float64)
        if X.ndim != 2 or X.shape[0] != self.n:
            raise ValueError("X must be n x nrhs")
        Y = np.empty_like(X, order="F")
        check(_lib.load().hymor_expm_taylor_batched(
            self._h, int(bool(transpose)), int(order), int(nsub), float(dt), _FORMS[form],
            X.shape[1], ptr(X), self.n, ptr(Y), self.n, C128 if cplx else F64))
        return Y

    def expm_taylor_vjp_values(self, x, wbar, colptr, rowval, *, order=5, nsub=1, dt=1.0,
                               ckpt_every=8, index_base=1):
        """Gradient of <wbar, T_order(dt A)^nsub x> with respect to the
        stored operator's entries on the CSC pattern (colptr, rowval) (real
        operators; hymor_expm_taylor_vjp_values, checkpointed every
        ckpt_every substeps)."""
        if self.is_complex:
            raise ValueError("the operator gradient is for real operators")
        x = np.ascontiguousarray(x, dtype=np.float64)
        wbar = np.ascontiguousarray(wbar, dtype=np.float64)
        colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        rowval = np.ascontiguousarray(rowval, dtype=np.int64)
        if x.shape != (self.n,) or wbar.shape != (self.n,) or colptr.shape != (self.n + 1,):
            raise ValueError("x, wbar need n entries and colptr n + 1")
        grad = np.empty(rowval.shape[0], dtype=np.float64)
        check(_lib.load().hymor_expm_taylor_vjp_values(
            self._h, int(order), int(nsub), float(dt), int(ckpt_every), ptr(x), ptr(wbar),
            ptr(colptr), ptr(rowval), int(index_base), ptr(grad)))
        return grad

    def expm_taylor_vjp_values_dev(self, pattern: "DevicePattern", d_x, d_wbar, d_grad, *,
                                   order=5, nsub=1, dt=1.0, ckpt_every=8):
        """expm_taylor_vjp_values on a DevicePattern with device vectors
        (raw pointers or DeviceVector) and a device gradient of
        pattern.nnz doubles (hymor_expm_taylor_vjp_values_dev)."""
        ptr_of = lambda d: d.ptr if hasattr(d, "ptr") else d
        check(_lib.load().hymor_expm_taylor_vjp_values_dev(
            self._h, pattern.handle, int(order), int(nsub), float(dt), int(ckpt_every),
            ptr_of(d_x), ptr_of(d_wbar), ptr_of(d_grad)))

    def chu_sandwich(self, weight: "DeviceOperator", x, *, order=5, nsub=1, dt=1.0, n_set0=0):
        """T(A'dt)^n * C * T(A dt)^n * x with CONSTRUCT_C's zeroing."""
        x, dty = _vec(x, self.is_complex)
        y = np.empty_like(x)
        check(_lib.load().hymor_chu_sandwich(self._h, weight._h, int(order), int(nsub), float(dt),
                                             int(n_set0), ptr(x), ptr(y), dty))
        return y

    def synchronize(self):
        check(_lib.load().hymor_op_synchronize(self._h))

    def autotune(self, reps=10, nv=None):
        """Pick the fastest Taylor-term launch variant for this operator;
        returns (best ms per launch, (split, waves, nt))."""
        nv = nv or (2 if self.is_complex else 1)
        ms = C.c_double()
        check(_lib.load().hymor_op_autotune(self._h, nv, reps, C.byref(ms)))
        return ms.value, self.tuning

    @property
    def tuning(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(_lib.load().hymor_op_get_tuning(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_graphs(self, enable: bool):
        """Replay the Taylor sweep from a captured HIP graph (see hymor.h)."""
        check(_lib.load().hymor_op_set_graphs(self._h, int(bool(enable))))

    def set_batch_kernel(self, kernel):
        """Batch kernel: 0 vector FMA, 1 matrix cores, 2/3 matrix cores with
        two/one cells per wave, 4/5 matrix cores with grouped operand loads,
        6 matrix cores on the tile layout (default: at b = 8 two cells per
        product at 16+ doubles per row, one cell per wave at 8; at b = 4 the
        four-cells-per-wave 4x4x4 matrix-core kernel), 7 the 4x4x4
        matrix-core form (one cell per wave at b = 4), 8 one cell per wave on
        the tile layout (6 at b = 4).  True/False map to 6/0."""
        if kernel is True:
            kernel = 6
        elif kernel is False:
            kernel = 0
        check(_lib.load().hymor_op_set_batch_kernel(self._h, int(kernel)))

    def autotune_batch(self, nv=16, reps=5):
        """Pick the fastest batch kernel at width nv; returns (id, ms)."""
        k = C.c_int()
        ms = C.c_double()
        check(_lib.load().hymor_op_autotune_batch(self._h, nv, reps, C.byref(k), C.byref(ms)))
        return k.value, ms.value

    def set_tuning(self, split, waves, nt):
        check(_lib.load().hymor_op_set_tuning(self._h, int(split), int(waves), int(nt)))

    def bench_term_kernel(self, reps=20, transpose=False, nv=None, form="powers"):
        """(ms per launch, algorithmic bytes per launch) of a middle Taylor
        term of `form` at nv doubles per row."""
        nv = nv or (2 if self.is_complex else 1)
        ms = C.c_double()
        by = C.c_double()
        check(_lib.load().hymor_bench_term_kernel(self._h, int(bool(transpose)), nv,
                                                  _FORMS[form], reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value


class DeviceVector:
    """A var-major vector in device memory, for the *_dev entry points."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = n
        self.dtype = np.dtype(dtype)
        self.nbytes = n * self.dtype.itemsize
        self._p = C.c_void_p()
        check(_lib.load().hymor_dev_alloc(self.nbytes, C.byref(self._p)))

    @classmethod
    def from_host(cls, x):
        x = np.ascontiguousarray(x)
        d = cls(x.shape[0], x.dtype)
        d.upload(x)
        return d

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        check(_lib.load().hymor_memcpy_h2d(self._p, ptr(x), self.nbytes))

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(_lib.load().hymor_memcpy_d2h(ptr(out), self._p, self.nbytes))
        return out

    @property
    def ptr(self):
        return self._p

    @property
    def abi_dtype(self):
        return C128 if self.dtype == np.complex128 else F64

    def free(self):
        if self._p is not None and self._p.value:
            check(_lib.load().hymor_dev_free(self._p))
        self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceMatrix:
    """A column-major n x nrhs block of var-major vectors in device memory."""

    def __init__(self, n: int, nrhs: int, dtype=np.float64):
        self.n, self.nrhs = n, nrhs
        self.dtype = np.dtype(dtype)
        self.nbytes = n * nrhs * self.dtype.itemsize
        self._p = C.c_void_p()
        check(_lib.load().hymor_dev_alloc(self.nbytes, C.byref(self._p)))

    @classmethod
    def from_host(cls, X):
        X = np.asfortranarray(X)
        d = cls(X.shape[0], X.shape[1], X.dtype)
        check(_lib.load().hymor_memcpy_h2d(d._p, ptr(X), d.nbytes))
        return d

    def download(self):
        out = np.empty((self.n, self.nrhs), dtype=self.dtype, order="F")
        check(_lib.load().hymor_memcpy_d2h(ptr(out), self._p, self.nbytes))
        return out

    @property
    def ptr(self):
        return self._p

    @property
    def abi_dtype(self):
        return C128 if self.dtype == np.complex128 else F64

    def free(self):
        if self._p is not None and self._p.value:
            check(_lib.load().hymor_dev_free(self._p))
        self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def expm_taylor_batched_dev(op: DeviceOperator, X: DeviceMatrix, Y: DeviceMatrix, *, order=5,
                            nsub=1, dt=1.0, form="powers", transpose=False):
    """Asynchronous device-resident batched action; op.synchronize() waits."""
    check(_lib.load().hymor_expm_taylor_batched_dev(
        op.handle, int(bool(transpose)), int(order), int(nsub), float(dt), _FORMS[form], X.nrhs,
        X.ptr, X.n, Y.ptr, Y.n, X.abi_dtype))


def expm_taylor_dev(op: DeviceOperator, x: DeviceVector, y: DeviceVector, *, order=5, nsub=1,
                    dt=1.0, form="powers", transpose=False):
    """Asynchronous device-resident action; call op.synchronize() to wait."""
    check(_lib.load().hymor_expm_taylor_dev(op.handle, int(bool(transpose)), int(order), int(nsub),
                                            float(dt), _FORMS[form], x.ptr, y.ptr, x.abi_dtype))


class DeviceWeights:
    """Block-diagonal Chu weights C_w = R'M_wR resident on the GPU (1..8 of
    them), for energy tracking.  Built from CSC matrices in the reference's
    layout (scipy sparse, var-major)."""

    def __init__(self, mats, *, nchi, neta, vars_per_cell=4):
        lib = _lib.load()
        self._keep = []
        cps, rvs, nzs = [], [], []
        n = mats[0].shape[0]
        for M in mats:
            M = M.tocsc()
            M.sort_indices()
            if M.shape != (n, n):
                raise ValueError("all weights must have the same shape")
            a = (np.ascontiguousarray(M.indptr, dtype=np.int64),
                 np.ascontiguousarray(M.indices, dtype=np.int64),
                 np.ascontiguousarray(M.data, dtype=np.float64))
            self._keep.append(a)
            cps.append(a[0].ctypes.data)
            rvs.append(a[1].ctypes.data)
            nzs.append(a[2].ctypes.data)
        k = len(mats)
        self.n, self.nmat = n, k
        self._h = C.c_void_p()
        check(lib.hymor_weights_create_csc(n, k, (C.c_void_p * k)(*cps), (C.c_void_p * k)(*rvs),
                                           (C.c_void_p * k)(*nzs), 0, vars_per_cell, nchi, neta,
                                           C.byref(self._h)))

    def close(self):
        if self._h is not None and self._h.value:
            check(_lib.load().hymor_weights_destroy(self._h))
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h


def energy_track(op: DeviceOperator, weights: DeviceWeights, x, *, n_t, order=5, dt=1.0,
                 form="powers", n_skip=0, n_samples=0):
    """n_t Taylor steps from x with the weights' energies after every step.

    Returns (E, x_final, samples): E has shape (nmat, n_t+1), E[:, 0] the
    energies of x; samples (n, count) holds the states after every n_skip-th
    step (at most n_samples of them), or None.
    """
    x, dty = _vec(x, op.is_complex)
    E = np.zeros((weights.nmat, n_t + 1))
    xf = np.empty_like(x)
    samp = None
    sp_ptr = None
    if n_samples > 0:
        samp = np.zeros((n_samples, x.shape[0]), dtype=x.dtype)
        sp_ptr = ptr(samp)
    check(_lib.load().hymor_energy_track(op.handle, weights.handle, int(order), int(n_t),
                                         float(dt), _FORMS[form], ptr(x), dty, ptr(E), ptr(xf),
                                         int(max(n_skip, 1)), int(n_samples), sp_ptr))
    return E, xf, (samp.T if samp is not None else None)


class DevicePattern:
    """A CSC pattern of an operator of `op`'s shape kept on the device for
    repeated operator gradients (hymor_pattern_create)."""

    def __init__(self, op: DeviceOperator, colptr, rowval, *, index_base=1):
        colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        rowval = np.ascontiguousarray(rowval, dtype=np.int64)
        if colptr.shape != (op.n + 1,):
            raise ValueError("colptr needs n + 1 entries")
        self._h = C.c_void_p()
        check(_lib.load().hymor_pattern_create(op.handle, ptr(colptr), ptr(rowval),
                                               int(index_base), C.byref(self._h)))
        nnz = C.c_int64()
        check(_lib.load().hymor_pattern_nnz(self._h, C.byref(nnz)))
        self.nnz = nnz.value

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            check(_lib.load().hymor_pattern_destroy(self._h))
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def spmv_dev(op: DeviceOperator, d_x, d_y, *, transpose=False, complex_=None):
    """Asynchronous y = A x (or A' x) between device pointers (var-major)."""
    cplx = op.is_complex if complex_ is None else complex_
    check(_lib.load().hymor_spmv_dev(op.handle, int(bool(transpose)), d_x, d_y,
                                     C128 if cplx else F64))


def expm_taylor_ptr(op: DeviceOperator, d_x, d_y, *, order=5, nsub=1, dt=1.0, form="powers",
                    transpose=False, complex_=None):
    """expm_taylor_dev on raw device pointers (e.g. Krylov basis columns)."""
    cplx = op.is_complex if complex_ is None else complex_
    check(_lib.load().hymor_expm_taylor_dev(op.handle, int(bool(transpose)), int(order), int(nsub),
                                            float(dt), _FORMS[form], d_x, d_y,
                                            C128 if cplx else F64))


def chu_sandwich_ptr(op: DeviceOperator, weight: DeviceOperator, d_x, d_y, *, order=5, nsub=1,
                     dt=1.0, n_set0=0, complex_=None):
    """hymor_chu_sandwich_dev on raw device pointers."""
    cplx = op.is_complex if complex_ is None else complex_
    check(_lib.load().hymor_chu_sandwich_dev(op.handle, weight.handle, int(order), int(nsub),
                                             float(dt), int(n_set0), d_x, d_y,
                                             C128 if cplx else F64))


def chu_sandwich_batched_ptr(op: DeviceOperator, weight: DeviceOperator, d_X, ldx: int,
                             nrhs: int, d_Y, ldy: int, *, order=5, nsub=1, dt=1.0, n_set0=0,
                             pre: DeviceOperator = None, post: DeviceOperator = None,
                             complex_=None):
    """hymor_chu_sandwich_batched_dev: Y = post Z T(A')^n C T(A)^n Z pre X on
    nrhs device columns (pre / post None: identity)."""
    cplx = op.is_complex if complex_ is None else complex_
    check(_lib.load().hymor_chu_sandwich_batched_dev(
        op.handle, weight.handle, pre.handle if pre is not None else None,
        post.handle if post is not None else None, int(order), int(nsub), float(dt),
        int(n_set0), int(nrhs), d_X, int(ldx), d_Y, int(ldy), C128 if cplx else F64))


def expm_taylor_batched_ptr(op: DeviceOperator, d_X, ldx: int, nrhs: int, d_Y, ldy: int, *,
                            order=5, nsub=1, dt=1.0, form="powers", transpose=False,
                            complex_=None):
    """hymor_expm_taylor_batched_dev on raw device pointers (e.g. a block of
    Krylov basis columns)."""
    cplx = op.is_complex if complex_ is None else complex_
    check(_lib.load().hymor_expm_taylor_batched_dev(
        op.handle, int(bool(transpose)), int(order), int(nsub), float(dt), _FORMS[form],
        int(nrhs), d_X, int(ldx), d_Y, int(ldy), C128 if cplx else F64))
