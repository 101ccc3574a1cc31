"""The physical channel-flow operator (hymor_amd.channel) on the GPU.

* The 99 x 80-cell channel operator (uniform grid -- Nchi divisible by 3
  for the periodic colouring -- M 0.5, Re 2000, Pr 0.72): the device
  Taylor action T_5(A dt)^n at the CFL step vs the oracle's POWER_METHOD!
  (1e-11: the operator is far from normal, ||A dt||_inf ~ 13).
* A small physical channel through the device eigenvalue driver, against
  the exact exponential transform of its dense spectrum.  (The tutorial
  case itself is pinned on the CPU by shift-invert, as the tutorial does:
  its near-neutral acoustic modes cluster in |exp(lambda T)|, which the
  exponential-transform Arnoldi resolves only slowly.)
"""
import sys

import numpy as np
import pytest

import hymor_amd.channel as ch
from hymor_amd import stability as st
from hymor_amd.core import DeviceOperator

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_config1_action_vs_oracle(orc):
    c = ch.ChannelCase(nchi=99, neta=80, eta_refinement_power=1.0, Mach=0.5, Re=2000.0, Pr=0.72)
    op, A, s = ch.channel_operator(c)
    s = st.CFL_TIMESTEP(s)
    dt = s["time_integration"]["dt"]
    say(f"[channel C1] n={op.n} nnz={A.nnz} dt={dt:.3e}")
    x = np.random.default_rng(2).standard_normal(op.n)
    Ao = orc.CSC.from_scipy(A)
    ref = orc.power_method(Ao.scaled(dt), x, 40)
    with DeviceOperator.from_blocks(op, scale=dt) as dev:
        got = dev.expm_taylor(x, order=5, nsub=40, dt=1.0, form="powers")
        got_t = dev.expm_taylor(x, order=5, nsub=40, dt=1.0, form="powers", transpose=True)
    assert np.linalg.norm(got - ref) <= 1e-11 * np.linalg.norm(ref)
    ref_t = orc.power_method(orc.csc_adjoint(Ao).scaled(dt), x, 40)
    assert np.linalg.norm(got_t - ref_t) <= 1e-11 * np.linalg.norm(ref_t)


def test_small_channel_eigenvalues_through_device_arnoldi():
    """The device exponential-transform Krylov-Schur on a physical channel
    operator small enough for a dense eigensolve (24 x 12 cells, M 0.5,
    Re 2000): every returned eigenvalue equals log(T_5(lambda dt)^n)/(n dt)
    of an eigenvalue lambda of the dense matrix, to 1e-8."""
    import math
    c = ch.ChannelCase(nchi=24, neta=12, eta_refinement_power=1.0, Mach=0.5, Re=2000.0,
                       Pr=0.72)
    op, A, s = ch.channel_operator(c)
    s = st.CFL_TIMESTEP(s)
    dt = s["time_integration"]["dt"]
    n_t = int(round(0.5 / dt))
    say(f"[channel small] n={op.n} dt={dt:.3e} n_t={n_t}: dense eigenvalues")
    lam = np.linalg.eigvals(A.toarray())
    mu = sum((lam * dt) ** j / math.factorial(j) for j in range(6)) ** n_t
    order = np.argsort(-np.abs(mu), kind="stable")
    exact = np.log(mu[order][:40].astype(complex)) / (n_t * dt)
    say("[channel small] device Krylov-Schur")
    # |exp(lambda T)| of the leading acoustic pairs differ by ~1e-3: a slow
    # but steady Krylov-Schur case (about 9 000 actions at ncv = 40)
    V, D = st.EIGENVALUES_GPU_ARNOLDI(op, s, 4, krylov="device", tol=1e-11, ncv=40,
                                      maxiter=1500)
    say(f"[channel small] D = {D}")
    assert np.all(np.isfinite(D)), D
    for d in D:
        assert np.min(np.abs(exact - d)) <= 1e-8 * max(abs(d), 1.0), (d, exact[:8])

