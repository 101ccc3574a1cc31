"""Synthetic base-flow solution dictionaries shaped like HYMOR's ``s``.

The stability entry points take the solver's nested ``Dict{String,Any}``
(utils/Initialization/INITIALIZATION.jl) and read base-flow fields from it:
``s["var"]`` holds (Nchi+2) x (Neta+2) arrays with one ghost layer,
``s["mesh"]`` face normals/areas and cell volumes, ``s["shock"]`` the
flow-cell mask and shock cell indices, ``s["freestream"]`` the upstream
state.  This module builds a dict with those keys and shapes, filled with a
smooth, physically plausible synthetic flow (positive density, pressure,
sound speed), so the host-side constructors (R, M, M_hat, CFL dt) can run
without the nonlinear solver.  Index conventions follow Julia: arrays are
stored Fortran-flattened (chi fastest), and ``shock.cell_indices`` is 1-based.
"""
from __future__ import annotations

import numpy as np


def synthetic_solution(nchi: int, neta: int, seed: int = 0, *, perturb_shock: bool = False,
                       mach: float = 6.0, gamma: float = 1.4, cfl: float = 2.0,
                       max_dt: float = 1.0, solver: str = "GPU_TIMESTEPPER_ARNOLDI") -> dict:
    rng = np.random.Generator(np.random.PCG64(30_000 + seed))
    X, Y = np.meshgrid(np.linspace(0, 1, nchi + 2), np.linspace(0, 1, neta + 2), indexing="ij")
    wob = lambda s=0.05: 1.0 + s * rng.standard_normal((nchi + 2, neta + 2))
    rho = (1.0 + 0.5 * X + 0.3 * Y * Y) * wob()
    u = (0.8 - 0.3 * Y) * wob()
    v = (0.1 * np.sin(3 * X)) * wob()
    p = (1.0 / (gamma * mach ** 2) * (1.0 + 2.0 * X)) * wob()
    gstar = gamma * np.ones_like(rho) - 0.02 * X
    a = np.sqrt(gstar * p / rho)
    rho_E = p / (gstar - 1.0) + 0.5 * rho * (u * u + v * v)
    var = {
        "rho": rho, "rho_u": rho * u, "rho_v": rho * v, "rho_E": rho_E, "p": p, "a": a,
        "gamma_star": gstar, "Re_flow": 1e5 * np.ones_like(rho), "Pr_flow": 0.72 * np.ones_like(rho),
    }
    dx, dy = 1.0 / nchi, 1.0 / neta
    th_lr = 0.1 * rng.standard_normal((nchi + 1, neta))
    th_bt = 0.1 * rng.standard_normal((nchi, neta + 1))
    mesh = {
        "Nchi": nchi, "Neta": neta,
        "lr_x_normal": np.cos(th_lr), "lr_y_normal": np.sin(th_lr),
        "lr_area": dy * (1.0 + 0.1 * rng.random((nchi + 1, neta))),
        "bt_x_normal": -np.sin(th_bt), "bt_y_normal": np.cos(th_bt),
        "bt_area": dx * (1.0 + 0.1 * rng.random((nchi, neta + 1))),
        "volume": dx * dy * (1.0 + 0.2 * rng.random((nchi, neta))),
    }
    flow_cells = np.ones((nchi, neta))
    # the last eta row lies outside the shock for a few chi stations
    flow_cells[: max(1, nchi // 5), neta - 1] = 0.0
    shock = {
        "enabled": True,
        "flow_cells": flow_cells,
        "cell_indices": np.full((nchi, 1), neta, dtype=np.int64),   # 1-based eta index
        "spline_param": 0.5,
    }
    flux = {k: 1e-2 * rng.standard_normal((nchi, neta)) for k in ("rho", "rho_u", "rho_v", "rho_E")}
    rho0, u0, p0 = 1.0, 1.0, 1.0 / (gamma * mach ** 2)
    freestream = {"rho_0": rho0, "rho_u_0": rho0 * u0, "rho_v_0": 0.0, "p_0": p0,
                  "gamma_star": gamma, "a_0": np.sqrt(gamma * p0 / rho0)}
    return {
        "var": var, "mesh": mesh, "shock": shock, "flux": flux, "freestream": freestream,
        "time_integration": {"CFL": cfl, "max_dt": max_dt},
        "stability_analysis": {"perturb_shock": perturb_shock, "eigenvalue_solver": solver},
    }
