"""Exact gradients of the transient-growth / receptivity gain.

The reference maximises the gain G(v) = (v' S v) / (v' D v) with
S = T_5(A'dt)^n C T_5(A dt)^n by ARPACK
(Transient_growth_downstream/TRANSIENT_GROWTH_DOWNSTREAM.jl:103,
Freestream_receptivity/FREESTREAM_RECEPTIVITY.jl:107); S is applied by
EXPONENTIAL! (CONSTRUCT_C.jl:78-111) and, for receptivity, CONSTRUCT_C zeroes
the first N_set_0 entries before and after (P S P).  Because the discrete
adjoint of a polynomial in A is the same polynomial in A'
(p(A)' = p(A')), the gradient needs no tape and no stored substeps:

    grad G(v) = 2 (P S P v - G(v) D v) / (v' D v)

-- one Chu sandwich (forward Taylor sweep, Chu weight, adjoint Taylor sweep
on the GPU) plus one block-diagonal apply of D.  For complex v the same
expression is the gradient with respect to (Re v, Im v) packed as a complex
vector.

``GainFunctional`` exposes value / value_and_grad; ``gain_autograd`` wraps it
as a torch.autograd.Function (PyTorch's counterpart of a ChainRules rrule)
when torch is importable.

The gradient with respect to the operator's entries (the gain's
sensitivity to the Jacobian, ``GainFunctional.operator_gradient``) does need
the forward states: with w = T(A)^n P v,

    dG/dA = d<2 C w / (v' D v), T(A)^n P v>/dA    (C, D fixed),

the pullback of the Taylor sweep for the cotangent 2 C w / (v' D v), which
hymor_expm_taylor_vjp_values evaluates with checkpointed substeps (a state
kept every ckpt_every substeps, each group recomputed from its checkpoint
in the backward sweep; ckpt_every=1 keeps every state and recomputes none,
the fastest form where nsub vectors fit in device memory).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .core import DeviceOperator


class GainFunctional:
    """G(v) = (Pv)' S (Pv) / (v' D v) for a device operator of A*dt, a device
    Chu weight C (R'MR) and a host matrix D (R'M_hat R)."""

    def __init__(self, devA: DeviceOperator, devC: DeviceOperator, D, n_t: int, *, n_set0: int = 0,
                 order: int = 5):
        self.devA, self.devC = devA, devC
        self.D = sp.csr_matrix(D)
        self.n_t, self.n_set0, self.order = int(n_t), int(n_set0), int(order)

    def _S(self, v):
        return self.devA.chu_sandwich(self.devC, v, order=self.order, nsub=self.n_t, dt=1.0,
                                      n_set0=self.n_set0)

    def value(self, v) -> float:
        v = np.asarray(v)
        num = np.vdot(v, self._S(v)).real
        den = np.vdot(v, self.D @ v).real
        return float(num / den)

    def value_and_grad(self, v):
        v = np.asarray(v)
        Sv = self._S(v)
        Dv = self.D @ v
        den = np.vdot(v, Dv).real
        G = np.vdot(v, Sv).real / den
        return float(G), 2.0 * (Sv - G * Dv) / den


    def operator_gradient(self, v, colptr, rowval, *, index_base=1, ckpt_every=8):
        """(G(v), dG/dA) with dG/dA on the CSC pattern (colptr, rowval) of
        the stored operator (A*dt as uploaded; real operators)."""
        v = np.array(v, dtype=np.float64)
        v[:self.n_set0] = 0.0
        w = self.devA.expm_taylor(v, order=self.order, nsub=self.n_t, dt=1.0, form="horner")
        Cw = self.devC.spmv(w)
        vo = np.asarray(v)
        den = float(vo @ (self.D @ vo))
        G = float(w @ Cw) / den
        grad = self.devA.expm_taylor_vjp_values(v, 2.0 * Cw / den, colptr, rowval,
                                                order=self.order, nsub=self.n_t, dt=1.0,
                                                ckpt_every=ckpt_every, index_base=index_base)
        return G, grad

    def operator_gradient_dev(self, v, pattern, d_grad, *, ckpt_every=8):
        """operator_gradient on a DevicePattern kept between calls: dG/dA is
        written to the device array d_grad (pattern.nnz doubles, in the
        pattern's CSC order) and never crosses to the host; returns G(v)."""
        from .core import DeviceVector
        v = np.array(v, dtype=np.float64)
        v[:self.n_set0] = 0.0
        w = self.devA.expm_taylor(v, order=self.order, nsub=self.n_t, dt=1.0, form="horner")
        Cw = self.devC.spmv(w)
        den = float(v @ (self.D @ v))
        G = float(w @ Cw) / den
        dx, dw = DeviceVector.from_host(v), DeviceVector.from_host(2.0 * Cw / den)
        try:
            self.devA.expm_taylor_vjp_values_dev(pattern, dx, dw, d_grad, order=self.order,
                                                 nsub=self.n_t, dt=1.0, ckpt_every=ckpt_every)
        finally:
            dx.free()
            dw.free()
        return G


def gain_autograd(functional: GainFunctional):
    """torch.autograd.Function computing G(v) for a torch float64 / complex128
    CPU tensor v, with the exact gradient above in backward."""
    import torch

    class _Gain(torch.autograd.Function):
        @staticmethod
        def forward(ctx, v):
            G, g = functional.value_and_grad(v.detach().cpu().numpy())
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(g)))
            return torch.tensor(G, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            # for complex v torch expects dL/dRe + i dL/dIm, which is g
            (g,) = ctx.saved_tensors
            return grad_out * g

    return _Gain.apply
