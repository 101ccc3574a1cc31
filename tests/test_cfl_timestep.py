"""CFL_TIMESTEP (utils/Time_marching/CFL_TIMESTEP.jl:41-138) pinned by closed
forms on a uniform Cartesian grid, where every term of the reference's
formula can be written down by hand:

    lam_c = (|u . n_lr| + c) * dy + (|u . n_bt| + c) * dx        (:80-102)
    lam_v = 2 max(4/3, gamma/Pr) / (rho Re) * (dy^2 + dx^2) / V  (:104-112)
    dt1   = CFL / max((lam_c + lam_v) * flow_cells / V)          (:115-116)
    dt2   = CFL / max(|f_rho_v| + |f_rho_u|) / rho, |f_rho| / rho,
                      |f_rho_E| / rho_E)                         (:119-130)
    dt    = min(dt1, dt2, max_dt)                                (:133)

Each case is built so that a plausible slip changes dt: left-right and
bottom-top faces swapped (dx != dy, u != v), the two faces of a cell summed
instead of maxed (uniform flow: twice the value) or minned (a velocity
ramp), the viscous prefactor branch (4/3 vs gamma/Pr), the sound speed
taken over interior cells only (the maximum sits in a ghost cell), a sum
instead of a max over the flux limits, and the flow-cell mask (which always
applies to dt1 and applies to dt2 only with a fitted shock).
"""
import numpy as np
import pytest

from hymor_amd import stability as st

NCHI, NETA = 5, 4
DX, DY = 0.5, 0.25            # bt faces have area dx, lr faces dy
VOL = DX * DY


def uniform_state(*, rho=2.0, u=2.0, v=1.0, a=3.0, gamma=1.4, pr=0.75, re=100.0, rho_e=10.0,
                  cfl=0.8, max_dt=1.0, flux=(0.01, 0.02, -0.03, 0.5), shock=True):
    g = (NCHI + 2, NETA + 2)
    var = {"rho": np.full(g, rho), "rho_u": np.full(g, rho * u), "rho_v": np.full(g, rho * v),
           "rho_E": np.full(g, rho_e), "a": np.full(g, a), "gamma_star": np.full(g, gamma),
           "Re_flow": np.full(g, re), "Pr_flow": np.full(g, pr), "p": np.ones(g)}
    mesh = {"Nchi": NCHI, "Neta": NETA,
            "lr_x_normal": np.ones((NCHI + 1, NETA)), "lr_y_normal": np.zeros((NCHI + 1, NETA)),
            "lr_area": np.full((NCHI + 1, NETA), DY),
            "bt_x_normal": np.zeros((NCHI, NETA + 1)), "bt_y_normal": np.ones((NCHI, NETA + 1)),
            "bt_area": np.full((NCHI, NETA + 1), DX),
            "volume": np.full((NCHI, NETA), VOL)}
    fr, fu, fv, fe = flux
    fl = {"rho": np.full((NCHI, NETA), fr), "rho_u": np.full((NCHI, NETA), fu),
          "rho_v": np.full((NCHI, NETA), fv), "rho_E": np.full((NCHI, NETA), fe)}
    return {"var": var, "mesh": mesh, "flux": fl,
            "shock": {"enabled": shock, "flow_cells": np.ones((NCHI, NETA))},
            "time_integration": {"CFL": cfl, "max_dt": max_dt}}


def dt_of(s):
    return st.CFL_TIMESTEP(s)["time_integration"]["dt"]


def test_uniform_flow_convective_and_viscous():
    # lam_c = (2 + 3) * 0.25 + (1 + 3) * 0.5 = 3.25
    # nu = max(4/3, 1.4/0.75) / (2 * 100) = (28/15) / 200
    # lam_v = 2 nu (0.25^2 + 0.5^2) / 0.125 = 2 nu * 2.5
    nu = (28.0 / 15.0) / 200.0
    lam = 3.25 + 5.0 * nu
    want = 0.8 / (lam / 0.125)
    assert dt_of(uniform_state()) == pytest.approx(want, rel=1e-13)
    # flux limit (0.05 -> dt2 = 16) and max_dt (1.0) do not bind here
    assert want < 1.0


def test_viscous_prefactor_four_thirds_branch():
    # gamma/Pr = 1.4 / (7/6) = 1.2 < 4/3; a low Reynolds number makes the
    # viscous term dominate: nu = (4/3) / (2 * 0.05)
    nu = (4.0 / 3.0) / (2.0 * 0.05)
    lam = 3.25 + 5.0 * nu
    s = uniform_state(pr=7.0 / 6.0, re=0.05)
    assert dt_of(s) == pytest.approx(0.8 / (lam / 0.125), rel=1e-13)


def test_faces_are_maxed_and_sound_speed_includes_ghosts():
    # rho_u ramps along chi: u(i) = 1 + 0.3 i over ghost-inclusive index i,
    # so the face between i and i+1 carries u = 1 + 0.3 (i + 0.5); the worst
    # cell (i = NCHI) takes its right face, 1 + 0.3 (NCHI + 0.5) = 2.65.
    # The sound speed is 3 except 3.5 in one ghost cell: c = 3.5.
    s = uniform_state()
    i = np.arange(NCHI + 2, dtype=float)[:, None]
    s["var"]["rho_u"] = s["var"]["rho"] * (1.0 + 0.3 * i) * np.ones((1, NETA + 2))
    s["var"]["a"][0, 0] = 3.5
    nu = (28.0 / 15.0) / 200.0
    lam = (2.65 + 3.5) * DY + (1.0 + 3.5) * DX + 5.0 * nu
    assert dt_of(s) == pytest.approx(0.8 / (lam / VOL), rel=1e-13)


def test_flux_limit_is_a_max_over_three_ratios():
    # (|f_v| + |f_u|) / rho = 0.025, |f_rho| / rho = 0.005,
    # |f_E| / rho_E = 1000 / 10 = 100  ->  dt2 = 0.8 / 100
    s = uniform_state(flux=(0.01, 0.02, -0.03, 1000.0))
    assert dt_of(s) == pytest.approx(0.008, rel=1e-13)
    # the momentum ratio binding instead: (|-300| + |100|) / 2 = 200
    s = uniform_state(flux=(0.01, 100.0, -300.0, 0.5))
    assert dt_of(s) == pytest.approx(0.8 / 200.0, rel=1e-13)


def test_max_dt_binds():
    assert dt_of(uniform_state(max_dt=1e-3)) == 1e-3


@pytest.mark.parametrize("shock", [True, False])
def test_flow_cell_mask(shock):
    # one cell outside the shock layer: a tiny volume (its lam/V would set
    # dt1) and a huge energy flux (its ratio would set dt2)
    s = uniform_state(shock=shock)
    s["shock"]["flow_cells"][0, NETA - 1] = 0.0
    s["mesh"]["volume"][0, NETA - 1] = 1e-6
    s["flux"]["rho_E"][0, NETA - 1] = 1e6
    nu = (28.0 / 15.0) / 200.0
    dt1 = 0.8 / ((3.25 + 5.0 * nu) / 0.125)
    if shock:
        # the mask removes the cell from both limits
        assert dt_of(s) == pytest.approx(dt1, rel=1e-13)
    else:
        # dt1 is always masked; the flux limit is not without a fitted shock
        assert dt_of(s) == pytest.approx(0.8 / (1e6 / 10.0), rel=1e-13)
