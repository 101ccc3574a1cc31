"""Gradient of the gain G(v) = (Pv)' S (Pv) / (v' D v) through the device
adjoint (Chu sandwich), checked
  * against central finite differences (h = 1e-6) on 16 random directions:
    relative agreement 1e-6 (the O(h^2) truncation and O(eps/h) rounding of
    the difference quotient, not of the gradient);
  * against the same formula evaluated with the oracle's restatement of the
    reference's own adjoint sweep (EXPONENTIAL!): 1e-12;
  * through torch autograd."""
import sys

import numpy as np
import pytest

from hymor_amd import operators as ops
from hymor_amd import stability as st
from hymor_amd.adjoint import GainFunctional, gain_autograd
from hymor_amd.baseflow import synthetic_solution
from hymor_amd.core import DeviceOperator

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(cplx):
    nchi, neta = 9, 7
    s = st.CFL_TIMESTEP(synthetic_solution(nchi, neta, seed=5, perturb_shock=True))
    dt = s["time_integration"]["dt"]
    base = ops.stencil_operator(nchi, neta, 4, seed=21, n_shock=nchi, row_sum=0.5 / dt)
    if cplx:
        w = [0.0, 1.5]
        op = ops.receptivity_extension(base, w)
        R = st.CONSTRUCT_R_(s, w)
        C = R.T @ st.CONSTRUCT_M_(s, [1, 1, 1, 1], w) @ R
        D = R.T @ st.CONSTRUCT_M_INFTY_(s, [1, 1, 1, 1], 1.0, w, True) @ R
        n0 = 4 * nchi * neta
    else:
        op = base
        R = st.CONSTRUCT_R(s)
        C = R.T @ st.CONSTRUCT_M(s, [1, 1, 1, 1]) @ R
        D = R.T @ st.CONSTRUCT_M_HAT(s, [1, 1, 1, 1]) @ R
        n0 = 0
    A = op.to_scipy_varmajor()
    n_t = 4
    return s, A, C.tocsc(), D.tocsc(), dt, n_t, n0


@pytest.mark.parametrize("cplx", [False, True])
def test_gain_gradient_fd_and_reference_adjoint(orc, cplx):
    say(f"[adjoint] gain gradient, complex={cplx}")
    s, A, C, D, dt, n_t, n0 = _setup(cplx)
    nchi, neta = s["mesh"]["Nchi"], s["mesh"]["Neta"]
    n = A.shape[0]
    r = np.random.default_rng(3)
    v = r.standard_normal(n) + (1j * r.standard_normal(n) if cplx else 0)
    with DeviceOperator.from_scipy(A, nchi=nchi, neta=neta, scale=dt) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC:
        F = GainFunctional(dA, dC, D, n_t, n_set0=n0)
        G, g = F.value_and_grad(v)
        # the reference's adjoint sweep (oracle restatement of EXPONENTIAL!)
        Ao = orc.CSC.from_scipy(A * dt)
        Sv = orc.construct_c(v, Ao, orc.csc_adjoint(Ao), orc.CSC.from_scipy(C), n_t, cplx, n0)
        den = np.vdot(v, D @ v).real
        G_ref = np.vdot(v, Sv).real / den
        g_ref = 2 * (Sv - G_ref * (D @ v)) / den
        assert abs(G - G_ref) <= 1e-12 * abs(G_ref)
        assert np.linalg.norm(g - g_ref) <= 1e-12 * np.linalg.norm(g_ref)
        # central differences on 16 random directions
        h = 1e-6
        for i in range(16):
            d = r.standard_normal(n) + (1j * r.standard_normal(n) if cplx else 0)
            d /= np.linalg.norm(d)
            fd = (F.value(v + h * d) - F.value(v - h * d)) / (2 * h)
            an = np.real(np.vdot(g, d)) if cplx else g @ d
            assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3 * np.linalg.norm(g)), (i, fd, an)


def test_gain_torch_autograd():
    torch = pytest.importorskip("torch")
    say("[adjoint] torch autograd wrapper")
    s, A, C, D, dt, n_t, n0 = _setup(False)
    nchi, neta = s["mesh"]["Nchi"], s["mesh"]["Neta"]
    with DeviceOperator.from_scipy(A, nchi=nchi, neta=neta, scale=dt) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC:
        F = GainFunctional(dA, dC, D, n_t)
        gain = gain_autograd(F)
        v = torch.tensor(np.random.default_rng(4).standard_normal(A.shape[0]), requires_grad=True)
        out = 3.0 * gain(v)
        out.backward()
        _, g = F.value_and_grad(v.detach().numpy())
        assert np.allclose(v.grad.numpy(), 3.0 * g, rtol=1e-13, atol=0)
