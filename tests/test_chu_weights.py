"""The package's restatements of CONSTRUCT_R / CONSTRUCT_M / CONSTRUCT_M_HAT /
CONSTRUCT_M_INFTY (hymor_amd.stability) against the reference's independent
second statement of the same Chu energy, GET_ENERGY_NORM +
GET_CHU_COMPONENTS (restated in the oracle): for any real perturbation v,
v' R' M R v must equal the volume-weighted sum of the energy densities.
A wrong row of R, a wrong weight or a wrong index order fails it."""
import numpy as np
import pytest

from hymor_amd import stability as st
from hymor_amd.baseflow import synthetic_solution

NORMS = {"acoustic": [1, 0, 0, 0], "kinetic": [0, 1, 1, 0], "entropic": [0, 0, 0, 1]}


def _fields(s):
    v = s["var"]
    f = lambda a: st._vec(st._interior(a))
    return (f(v["rho"]), f(v["rho_u"]), f(v["rho_v"]), f(v["p"]), f(v["a"]), f(v["gamma_star"]))


@pytest.mark.parametrize("shock", [False, True])
def test_chu_quadratic_form_equals_energy_norm(orc, shock):
    nchi, neta = 7, 5
    s = synthetic_solution(nchi, neta, seed=2, perturb_shock=shock)
    R = st.CONSTRUCT_R(s)
    ncell = nchi * neta
    x = np.random.default_rng(1).standard_normal(R.shape[0])
    pr = [x[k * ncell:(k + 1) * ncell] for k in range(4)]
    E = orc.energy_norm(*_fields(s), *pr)
    vol = st._vec(s["mesh"]["volume"])
    fc = st._vec(s["shock"]["flow_cells"])
    for name, norms in NORMS.items():
        q = x @ (R.T @ (st.CONSTRUCT_M(s, norms) @ (R @ x)))
        assert q == pytest.approx(np.sum(E[name] * vol * fc), rel=1e-12)
        qh = x @ (R.T @ (st.CONSTRUCT_M_HAT(s, norms) @ (R @ x)))
        tail = 0.0
        if shock:
            ci = s["shock"]["cell_indices"][:, 0] - 1
            # CONSTRUCT_M_HAT adds the shock weights bt_area/2 whatever the norms
            tail = np.sum(s["mesh"]["bt_area"][np.arange(nchi), ci] / 2 * x[4 * ncell:] ** 2)
        assert qh - tail == pytest.approx(np.sum(E[name] * vol), rel=1e-12)


def test_freestream_weights_equal_energy_norm(orc):
    nchi, neta = 6, 4
    s = synthetic_solution(nchi, neta, seed=3, perturb_shock=True)
    fs = s["freestream"]
    Rinf = st.CONSTRUCT_R_INFTY(s)
    x = np.random.default_rng(2).standard_normal(4 * nchi)
    pr = [x[k * nchi:(k + 1) * nchi] for k in range(4)]
    one = np.ones(nchi)
    E = orc.energy_norm(fs["rho_0"] * one, fs["rho_u_0"] * one, fs["rho_v_0"] * one,
                        fs["p_0"] * one, fs["a_0"] * one, fs["gamma_star"] * one, *pr)
    ci = s["shock"]["cell_indices"][:, 0] - 1
    i = np.arange(nchi)
    vmag = np.hypot(fs["rho_u_0"], fs["rho_v_0"])
    proj = (s["mesh"]["bt_x_normal"][i, ci] * fs["rho_u_0"] + s["mesh"]["bt_y_normal"][i, ci] *
            fs["rho_v_0"]) / vmag
    arc = s["mesh"]["bt_area"][i, ci] * np.abs(proj)
    T = 2.5
    for name, norms in NORMS.items():
        M = st.CONSTRUCT_M_INFTY(s, norms, T, False)
        q = x @ (Rinf.T @ (M @ (Rinf @ x)))
        assert q == pytest.approx(np.sum(E[name] * arc * T), rel=1e-12)


def test_receptivity_weights_block_structure():
    nchi, neta = 5, 4
    w = [0.0, 1.0, 2.0]
    s = synthetic_solution(nchi, neta, seed=4, perturb_shock=True)
    R_ = st.CONSTRUCT_R_(s, w)
    n0 = st.CONSTRUCT_R(s).shape[0]
    assert R_.shape[0] == n0 + 3 * 4 * nchi
    Mi = st.CONSTRUCT_M_INFTY_(s, [1, 1, 1, 1], 1.0, w, False).toarray()
    blk = st.CONSTRUCT_M_INFTY(s, [1, 1, 1, 1], 1.0, False).toarray()
    p = 4 * nchi
    assert np.allclose(Mi[n0:n0 + p, n0:n0 + p], blk)              # w = 0: full weight
    assert np.allclose(Mi[n0 + p:n0 + 2 * p, n0 + p:n0 + 2 * p], blk / 2)   # w != 0: halved
    M_ = st.CONSTRUCT_M_(s, [1, 1, 1, 1], w).toarray()
    assert not np.any(M_[n0:, n0:])                                  # zero on freestream DOFs


def test_cfl_timestep_matches_direct_formula():
    s = synthetic_solution(6, 5, seed=5)
    dt = st.CFL_TIMESTEP(s)["time_integration"]["dt"]
    assert 0 < dt <= s["time_integration"]["max_dt"]
    s2 = synthetic_solution(6, 5, seed=5)
    s2["time_integration"]["CFL"] *= 0.5
    assert st.CFL_TIMESTEP(s2)["time_integration"]["dt"] == pytest.approx(dt * 0.5, rel=1e-14) or \
        st.CFL_TIMESTEP(s2)["time_integration"]["dt"] == s2["time_integration"]["max_dt"]
