"""Channel-flow operator generator (hymor_amd.channel) on the CPU.

Pins, each able to catch a different mistake:
* the reference tutorial's known answer (tutorials/channel_flow/README.md:
  most unstable eigenvalue 0.000247 + 0.248408i at 180 x 60, Re 7500,
  M 0.05): reproduced to 5e-6 -- the forward-difference noise of the
  reference's own step (1e-8) moves it by ~2e-5 between steps 1e-8 and
  1e-9, so this pins the discretisation and the Jacobian assembly;
* CHECK_LINEARIZATION_L-style consistency: A dq against f(q + dq) - f(q);
* the forward-difference Jacobian against the exact complex-step Jacobian
  of the same residual, column by column for one colour;
* discrete conservation: the volume-weighted mass residual sums to zero
  (periodic chi, impermeable walls).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import hymor_amd.channel as ch


@pytest.fixture(scope="module")
def tutorial():
    return ch.channel_operator(ch.ChannelCase())


def test_tutorial_eigenvalue(tutorial):
    op, A, s = tutorial
    assert A.shape == (43200, 43200)
    vals, _ = sla.eigs(A.astype(complex), k=1, sigma=0.0 + 0.25j)
    lam = vals[0]
    assert abs(lam - (0.000247 + 0.248408j)) <= 5e-6, lam
    # Nektar++ / Canuto et al.: 0.000224 + 0.24989i, 0.6 % away at this grid
    assert abs(lam - (0.000224 + 0.24989j)) / abs(0.24989) <= 0.007


def _flat(f):
    return np.concatenate([a.T.reshape(-1) for a in f])     # var-major, chi fastest


def test_linearization_consistency():
    c = ch.ChannelCase(nchi=24, neta=10)
    op, A, s = ch.channel_operator(c)
    mesh, fs = ch.make_mesh(c), ch.freestream(c)
    state = ch.base_state(c, mesh, fs)
    f0 = ch.rhs(state, mesh, fs)
    r = np.random.default_rng(0)
    dq = 1e-6 * r.standard_normal(A.shape[0]) * np.repeat([1, 1, 1, 700], c.nchi * c.neta)
    pert = [np.array(a) for a in state]
    for k in range(4):
        blk = dq[k * c.nchi * c.neta:(k + 1) * c.nchi * c.neta].reshape(c.neta, c.nchi).T
        pert[k][1:-1, 1:-1] += blk
    df = _flat(ch.rhs(pert, mesh, fs)) - _flat(f0)
    assert np.linalg.norm(A @ dq - df) <= 1e-3 * np.linalg.norm(df)


def test_fd_jacobian_vs_complex_step():
    c = ch.ChannelCase(nchi=12, neta=8)
    op, A, s = ch.channel_operator(c)
    mesh, fs = ch.make_mesh(c), ch.freestream(c)
    state = ch.base_state(c, mesh, fs)
    n = c.nchi * c.neta
    h = 1e-30
    for var in range(4):
        mask = ch._colour_mask(5, c.nchi, c.neta)
        pert = [np.array(a, dtype=complex) for a in state]
        pert[var][1:-1, 1:-1] = pert[var][1:-1, 1:-1] + mask * (1j * h)
        exact = np.imag(_flat(ch.rhs(pert, mesh, fs))) / h
        e = np.zeros(4 * n)
        e[var * n:(var + 1) * n] = mask.T.reshape(-1)
        fd = A @ e
        scale = np.abs(exact).max()
        assert np.abs(fd - exact).max() <= 1e-4 * scale, var


def test_mass_conservation():
    c = ch.ChannelCase(nchi=30, neta=12)
    mesh, fs = ch.make_mesh(c), ch.freestream(c)
    state = [np.array(a) for a in ch.base_state(c, mesh, fs)]
    r = np.random.default_rng(1)
    for a in state[:3]:
        a[1:-1, 1:-1] += 1e-3 * r.standard_normal(a[1:-1, 1:-1].shape)
    f = ch.rhs(state, mesh, fs)
    total = (f[0] * mesh["volume"]).sum()
    assert abs(total) <= 1e-13 * np.abs(f[0] * mesh["volume"]).sum()
    assert abs(mesh["volume"].sum() - c.Lx * c.Ly) <= 1e-12 * c.Lx * c.Ly


def test_rejects_non_periodic_colouring():
    with pytest.raises(ValueError, match="divisible by 3"):
        ch.linearize(ch.ChannelCase(nchi=100, neta=8))
