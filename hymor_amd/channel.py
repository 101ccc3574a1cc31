"""Channel-flow operator generator.

The linearised operator A of the reference's 2-D calorically perfect
Navier-Stokes finite-volume discretisation, for the channel geometry of
tutorials/channel_flow (periodic in chi, no-slip adiabatic walls in eta):

- mesh: utils/Mesh/GENERATE_MESH.jl (corners, face normals and areas,
  shoelace volumes), GO_TO_PHYSICAL_SPACE.jl:121-132 (channel mapping with
  wall refinement), GET_EXTENDED_CELL_CENTERS.jl (ghost centres);
- freestream and base state: utils/Initialization/SET_FREESTREAM_PROPERTIES.jl
  (non-dimensionalisation), INITIAL_SOLUTION.jl:43-57 (Poiseuille profile),
  VERIFY_INPUT_DATA.jl:110-113 (rho = T = 1 defaults);
- thermodynamics: utils/Operators/UPDATE_THERMODYNAMIC_PROPERTIES.jl,
  UPDATE_SOUND_SPEED.jl, UPDATE_CHEMISTRY_EQUILIBRIUM.jl (gamma*, cv*, mu*,
  k* constant without chemistry);
- boundary conditions: utils/Operators/APPLY_BOUNDARY_CONDITIONS.jl
  (periodic chi :183-195, :372-384; no-slip adiabatic eta :449-459,
  :640-650), chi sides first, then eta;
- fluxes: FLUX_MASS.jl, FLUX_MOMENTUM.jl, FLUX_TOTAL_ENERGY.jl (central face
  averages, mass-weighted face velocity), VISCOUS_FLUX.jl (face-normal /
  tangential derivatives with corner averages);
- Jacobian: utils/Stability_analysis/Modal_stability_analysis/
  LINEARIZE_EQUATIONS_NO_DISCONTINUITY.jl: forward differences
  (f(q + eps e_P) - f(q)) / eps with the absolute perturbation eps of
  stability_analysis.perturbation_magnitude, over the 9 colours of
  GET_PERTURBATION_MATRIX (:959, every third cell in chi and eta) for each
  of the 4 conserved variables, each response scattered to the 3 x 3
  neighbourhood of its perturbed cell (GET_OUTPUT_INDEXING_A11 :306, with
  the chi-periodic wrap), explicit zeros kept as sparse(I, J, V) keeps them.

Every arithmetic expression keeps the reference's operand order, so that
the rounding of the base-state residual, which the forward differences
divide by eps, follows the reference's.  The operator comes back both as a
hymor_amd.operators.BlockOperator (cell-major 4 x 4 blocks, for the device)
and as the reference's var-major CSC layout.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from .operators import BlockOperator


@dataclass
class ChannelCase:
    """tutorials/channel_flow/input_file.jl defaults."""

    nchi: int = 180
    neta: int = 60
    Lx: float = 2 * np.pi
    Ly: float = 2.0
    eta_refinement_power: float = 1.6
    channel_angle: float = 0.0
    u: float = 1.0
    v: float = 0.0
    gamma: float = 1.4
    Mach: float = 0.05
    Pr: float = 0.73
    Re: float = 7500.0
    rho: float = 1.0           # VERIFY_INPUT_DATA.jl:110 default
    T: float = 1.0             # VERIFY_INPUT_DATA.jl:113 default
    CFL: float = 2.0
    max_dt: float = 0.3
    perturbation_magnitude: float = 1e-8


# ------------------------------------------------------------------ mesh

def _corners(c: ChannelCase):
    """Corner coordinates (nchi+1, neta+1): GENERATE_MESH.jl with the
    channel mapping of GO_TO_PHYSICAL_SPACE.jl:121-132."""
    chi = np.arange(c.nchi + 1, dtype=np.float64) / c.nchi      # range(0, 1, length=N+1)
    eta = np.arange(c.neta + 1, dtype=np.float64) / c.neta
    chi2, eta2 = np.meshgrid(chi, eta, indexing="ij")
    p = c.eta_refinement_power
    eta_r = np.empty_like(eta2)
    low = eta2 <= 0.5
    eta_r[low] = 2 ** (p - 1) * eta2[low] ** p
    eta_r[~low] = 1 - 2 ** (p - 1) * (1 - eta2[~low]) ** p
    ca, sa = np.cos(c.channel_angle), np.sin(c.channel_angle)
    x = chi2 * ca * c.Lx - eta_r * sa * c.Ly
    y = eta_r * ca * c.Ly + chi2 * sa * c.Lx
    return x, y


def make_mesh(c: ChannelCase) -> dict:
    xc, yc = _corners(c)
    N, M = c.nchi, c.neta
    # left-right faces (constant chi), normals towards +chi
    dx = xc[:, 1:] - xc[:, :-1]
    dy = yc[:, 1:] - yc[:, :-1]
    lr_a = np.sqrt(dx ** 2 + dy ** 2)
    lr_xn = -dy / lr_a
    lr_yn = dx / lr_a
    ex = np.empty_like(lr_a)
    ey = np.empty_like(lr_a)
    ex[:-1] = xc[1:, :-1] - xc[:-1, :-1]
    ey[:-1] = yc[1:, :-1] - yc[:-1, :-1]
    ex[-1] = xc[-1, :-1] - xc[-2, :-1]
    ey[-1] = yc[-1, :-1] - yc[-2, :-1]
    sgn = np.sign(lr_xn * ex + lr_yn * ey)
    lr_xn, lr_yn = lr_xn * sgn, lr_yn * sgn
    # bottom-top faces (constant eta), normals towards +eta
    dx = xc[1:, :] - xc[:-1, :]
    dy = yc[1:, :] - yc[:-1, :]
    bt_a = np.sqrt(dx ** 2 + dy ** 2)
    bt_xn = dy / bt_a
    bt_yn = -dx / bt_a
    ex = np.empty_like(bt_a)
    ey = np.empty_like(bt_a)
    ex[:, :-1] = xc[:-1, 1:] - xc[:-1, :-1]
    ey[:, :-1] = yc[:-1, 1:] - yc[:-1, :-1]
    ex[:, -1] = xc[:-1, -1] - xc[:-1, -2]
    ey[:, -1] = yc[:-1, -1] - yc[:-1, -2]
    sgn = np.sign(bt_xn * ex + bt_yn * ey)
    bt_xn, bt_yn = bt_xn * sgn, bt_yn * sgn
    # shoelace volumes over corners (i,j), (i+1,j), (i+1,j+1), (i,j+1)
    x1, y1 = xc[:-1, :-1], yc[:-1, :-1]
    x2, y2 = xc[1:, :-1], yc[1:, :-1]
    x3, y3 = xc[1:, 1:], yc[1:, 1:]
    x4, y4 = xc[:-1, 1:], yc[:-1, 1:]
    vol = -x1 * y2 + x2 * y1
    vol = vol - x2 * y3 + x3 * y2
    vol = vol - x3 * y4 + x4 * y3
    vol = vol - x4 * y1 + x1 * y4
    volume = np.abs(vol / 2)
    # cell centres and ghost centres (GET_EXTENDED_CELL_CENTERS.jl)
    xm = (xc[:-1, :-1] + xc[1:, :-1] + xc[:-1, 1:] + xc[1:, 1:]) / 4
    ym = (yc[:-1, :-1] + yc[1:, :-1] + yc[:-1, 1:] + yc[1:, 1:]) / 4
    xE = np.zeros((N + 2, M + 2))
    yE = np.zeros((N + 2, M + 2))
    xE[1:-1, 1:-1], yE[1:-1, 1:-1] = xm, ym
    for arr, cr in ((xE, xc), (yE, yc)):
        arr[1:-1, 0] = (cr[:-1, 0] + cr[1:, 0]) / 2
        arr[1:-1, -1] = (cr[:-1, -1] + cr[1:, -1]) / 2
        arr[0, 1:-1] = (cr[0, :-1] + cr[0, 1:]) / 2
        arr[-1, 1:-1] = (cr[-1, :-1] + cr[-1, 1:]) / 2
        arr[0, 0], arr[0, -1] = cr[0, 0], cr[0, -1]
        arr[-1, 0], arr[-1, -1] = cr[-1, 0], cr[-1, -1]
        arr[:, 0] = 2 * arr[:, 0] - arr[:, 1]
        arr[:, -1] = 2 * arr[:, -1] - arr[:, -2]
        arr[0, :] = 2 * arr[0, :] - arr[1, :]
        arr[-1, :] = 2 * arr[-1, :] - arr[-2, :]
    mesh = {"Nchi": N, "Neta": M, "x_corner": xc, "y_corner": yc, "x": xm, "y": ym,
            "x_Ext": xE, "y_Ext": yE, "lr_area": lr_a, "lr_x_normal": lr_xn,
            "lr_y_normal": lr_yn, "bt_area": bt_a, "bt_x_normal": bt_xn, "bt_y_normal": bt_yn,
            "volume": volume}
    return mesh


# ------------------------------------------------------------------ state

def freestream(c: ChannelCase) -> dict:
    """SET_FREESTREAM_PROPERTIES.jl without chemistry."""
    U = np.sqrt(c.u ** 2 + c.v ** 2)
    fs = {"u": c.u, "v": c.v, "gamma": c.gamma, "Mach": c.Mach, "Pr": c.Pr, "Re": c.Re,
          "rho": c.rho, "T": c.T, "U": U}
    fs["kin"] = 0.5 * U ** 2
    fs["velocity_factor"] = U
    fs["energy_factor"] = U ** 2
    fs["rho_factor"] = c.rho
    fs["e"] = fs["energy_factor"] * 1 / (c.gamma - 1) / c.gamma / c.Mach ** 2
    fs["a"] = U / c.Mach
    fs["E"] = fs["e"] + fs["kin"]
    fs["gamma_star"] = c.gamma
    fs["cv"] = fs["e"] / c.T
    fs["L"] = 1.0
    fs["mu"] = c.rho * U * fs["L"] / c.Re
    fs["k"] = c.gamma * fs["cv"] * fs["mu"] / c.Pr
    fs["p"] = c.rho * fs["e"] * (fs["gamma_star"] - 1)
    fs["rho_u_0"] = c.rho * c.u / (U * fs["rho_factor"])
    fs["rho_v_0"] = c.rho * c.v / (U * fs["rho_factor"])
    fs["rho_0"] = c.rho / fs["rho_factor"]
    fs["rho_E_0"] = c.rho * fs["E"] / (fs["energy_factor"] * fs["rho_factor"])
    fs["p_0"] = fs["p"] / (fs["rho_factor"] * fs["velocity_factor"] ** 2)
    fs["a_0"] = fs["a"] / fs["velocity_factor"]
    return fs


def base_state(c: ChannelCase, mesh: dict, fs: dict):
    """INITIAL_SOLUTION.jl:43-57: Poiseuille profile u = y(2 - y), rho and
    the internal energy uniform; ghosts included."""
    shape = mesh["x_Ext"].shape
    y = 2 * mesh["y_Ext"] / c.Ly
    profile = y * (2 - y)
    rho = np.full(shape, fs["rho_0"])
    rho_v = np.zeros(shape)
    rho_u = fs["rho_u_0"] * profile
    rho_e = fs["rho_E_0"] - 0.5 * (fs["rho_u_0"] ** 2 + fs["rho_v_0"] ** 2) / fs["rho_0"]
    rho_E = rho_e + 0.5 * rho_u ** 2 / rho
    return rho, rho_u, rho_v, rho_E


# ------------------------------------------------------------------ right-hand side

def _thermo_and_bcs(rho, rho_u, rho_v, rho_E, gam, cvs):
    """Pressure and temperature everywhere, then the ghost layers:
    periodic chi (both sides), then no-slip adiabatic eta walls."""
    e = (rho_E - (rho_u ** 2 + rho_v ** 2) / rho / 2) / rho
    p = rho * e * (gam - 1)
    T = e / cvs
    for a in (rho_u, rho_v, rho, rho_E, p, T):        # chi0: ghost = last interior column
        a[0, :] = a[-2, :]
    for a in (rho_u, rho_v, rho, rho_E, p, T):        # chi1: ghost = first interior column
        a[-1, :] = a[1, :]
    rho[:, 0] = rho[:, 1]                              # eta0 wall
    rho_u[:, 0] = -rho_u[:, 1]
    rho_v[:, 0] = -rho_v[:, 1]
    rho_E[:, 0] = 2 * rho_E[:, 1] - rho_E[:, 2]
    p[:, 0] = 2 * p[:, 1] - p[:, 2]
    T[:, 0] = T[:, 1]
    rho[:, -1] = rho[:, -2]                            # eta1 wall
    rho_u[:, -1] = -rho_u[:, -2]
    rho_v[:, -1] = -rho_v[:, -2]
    rho_E[:, -1] = 2 * rho_E[:, -2] - rho_E[:, -3]
    p[:, -1] = 2 * p[:, -2] - p[:, -3]
    T[:, -1] = T[:, -2]
    return p, T


def rhs(state, mesh: dict, fs: dict):
    """d/dt (rho, rho_u, rho_v, rho_E) on the interior cells: PDE.jl for the
    2-D calorically perfect gas.  `state` holds the four (nchi+2, neta+2)
    arrays; they are copied (their ghosts are rewritten).  Complex input
    is carried through (complex-step derivatives)."""
    rho, rho_u, rho_v, rho_E = (np.array(a, dtype=np.result_type(a, np.float64))
                                for a in state)
    gam = np.full(rho.shape, fs["gamma_star"])
    cvs = np.ones(rho.shape)
    mus = np.ones(rho.shape)
    ks = np.ones(rho.shape)
    p, T = _thermo_and_bcs(rho, rho_u, rho_v, rho_E, gam, cvs)
    lr_xn, lr_yn, lr_a = mesh["lr_x_normal"], mesh["lr_y_normal"], mesh["lr_area"]
    bt_xn, bt_yn, bt_a = mesh["bt_x_normal"], mesh["bt_y_normal"], mesh["bt_area"]
    vol = mesh["volume"]

    def lr(a):          # left / right neighbours of the chi faces
        return a[:-1, 1:-1], a[1:, 1:-1]

    def bt(a):          # below / above neighbours of the eta faces
        return a[1:-1, :-1], a[1:-1, 1:]

    def div(f1, f2):    # out-minus-in over both face families, per volume
        out = -f1[1:, :]
        out = out + f1[:-1, :]
        out = out - f2[:, 1:]
        out = out + f2[:, :-1]
        return out / vol

    # FLUX_MASS.jl
    uL, uR = lr(rho_u)
    vL, vR = lr(rho_v)
    m1 = ((uL + uR) * lr_xn / 2 + (vL + vR) * lr_yn / 2) * lr_a
    uB, uT = bt(rho_u)
    vB, vT = bt(rho_v)
    m2 = ((uB + uT) * bt_xn / 2 + (vB + vT) * bt_yn / 2) * bt_a
    f_rho = div(m1, m2)
    # FLUX_MOMENTUM.jl, FLUX_TOTAL_ENERGY.jl
    rL, rR = lr(rho)
    rB, rT = bt(rho)
    pL, pR = lr(p)
    pB, pT = bt(p)
    vel1 = (uL + uR) / (rR + rL) * lr_xn + (vL + vR) / (rR + rL) * lr_yn
    vel2 = (uB + uT) / (rT + rB) * bt_xn + (vB + vT) / (rT + rB) * bt_yn
    f_ru = div(((uL + uR) / 2 * vel1 + (pL + pR) * lr_xn / 2) * lr_a,
               ((uB + uT) / 2 * vel2 + (pB + pT) * bt_xn / 2) * bt_a)
    f_rv = div(((vL + vR) / 2 * vel1 + (pL + pR) * lr_yn / 2) * lr_a,
               ((vB + vT) / 2 * vel2 + (pB + pT) * bt_yn / 2) * bt_a)
    EL, ER = lr(rho_E)
    EB, ET = bt(rho_E)
    f_rE = div((((EL + ER) / 2 + (pL + pR) / 2) * vel1) * lr_a,
               (((EB + ET) / 2 + (pB + pT) / 2) * vel2) * bt_a)
    # VISCOUS_FLUX.jl
    Re, Pr = fs["Re"], fs["Pr"]
    xE, yE = mesh["x_Ext"], mesh["y_Ext"]
    xc, yc = mesh["x_corner"], mesh["y_corner"]
    u = rho_u / rho
    v = rho_v / rho

    def corner_avg(a):
        return (a[:-1, :-1] + a[1:, :-1] + a[:-1, 1:] + a[1:, 1:]) / 4

    u_c, v_c, T_c = corner_avg(u), corner_avg(v), corner_avg(T)
    d_bt = np.sqrt((yE[1:-1, 1:] - yE[1:-1, :-1]) ** 2 + (xE[1:-1, 1:] - xE[1:-1, :-1]) ** 2)
    d_lr = np.sqrt((yE[1:, 1:-1] - yE[:-1, 1:-1]) ** 2 + (xE[1:, 1:-1] - xE[:-1, 1:-1]) ** 2)
    len_bt = np.sqrt((xc[1:, :] - xc[:-1, :]) ** 2 + (yc[1:, :] - yc[:-1, :]) ** 2)
    len_lr = np.sqrt((xc[:, 1:] - xc[:, :-1]) ** 2 + (yc[:, 1:] - yc[:, :-1]) ** 2)
    c1x = -(xE[1:, 1:-1] - xE[:-1, 1:-1]) / d_lr
    c1y = -(yE[1:, 1:-1] - yE[:-1, 1:-1]) / d_lr
    c2x = (xE[1:-1, 1:] - xE[1:-1, :-1]) / d_bt
    c2y = (yE[1:-1, 1:] - yE[1:-1, :-1]) / d_bt
    cn2_fn2 = c2x * bt_xn + c2y * bt_yn
    cn2_ft2 = c2x * bt_yn - c2y * bt_xn
    cn1_fn1 = -c1x * lr_xn - c1y * lr_yn
    cn1_ft1 = c1x * lr_yn - c1y * lr_xn

    def grads(a, a_c):
        """(d/dc2, d/dc1, d/dt2, d/dt1) of a cell field with corner values."""
        d_c2 = (a[1:-1, 1:] - a[1:-1, :-1]) / d_bt
        d_c1 = -(a[1:, 1:-1] - a[:-1, 1:-1]) / d_lr
        d_t2 = -(a_c[1:, :] - a_c[:-1, :]) / len_bt
        d_t1 = (a_c[:, 1:] - a_c[:, :-1]) / len_lr
        return d_c2, d_c1, d_t2, d_t1

    du_c2, du_c1, du_t2, du_t1 = grads(u, u_c)
    dv_c2, dv_c1, dv_t2, dv_t1 = grads(v, v_c)
    den2 = -cn2_fn2
    du_n2 = (cn2_ft2 * du_t2 - du_c2) / den2
    dv_n2 = (cn2_ft2 * dv_t2 - dv_c2) / den2
    dut_dt2 = du_t2 * bt_yn - dv_t2 * bt_xn
    dun_dt2 = du_t2 * bt_xn + dv_t2 * bt_yn
    dut_dn2 = du_n2 * bt_yn - dv_n2 * bt_xn
    dun_dn2 = du_n2 * bt_xn + dv_n2 * bt_yn
    du_n1 = (du_c1 - du_t1 * cn1_ft1) / cn1_fn1
    dv_n1 = (dv_c1 - dv_t1 * cn1_ft1) / cn1_fn1
    dun_dn1 = -du_n1 * lr_xn - dv_n1 * lr_yn
    dut_dn1 = du_n1 * lr_yn - dv_n1 * lr_xn
    dun_dt1 = -du_t1 * lr_xn - dv_t1 * lr_yn
    dut_dt1 = du_t1 * lr_yn - dv_t1 * lr_xn
    mu1 = (mus[1:, 1:-1] + mus[:-1, 1:-1]) / 2
    mu2 = (mus[1:-1, 1:] + mus[1:-1, :-1]) / 2
    div1 = dut_dt1 + dun_dn1
    tau11 = mu1 / Re * (2 * dun_dn1 - (2 / 3) * div1)
    tau12 = mu1 / Re * (dun_dt1 + dut_dn1)
    div2 = dut_dt2 + dun_dn2
    tau22 = mu2 / Re * (2 * dun_dn2 - (2 / 3) * div2)
    tau21 = mu2 / Re * (dut_dn2 + dun_dt2)
    t1x = -(tau11 * lr_xn - tau12 * lr_yn)
    t1y = -(tau11 * lr_yn + tau12 * lr_xn)
    t2x = tau22 * bt_xn + tau21 * bt_yn
    t2y = tau22 * bt_yn - tau21 * bt_xn

    def vdiv(g1, g2):   # in-minus-out on chi faces plus out-minus-in on eta faces
        return ((g1[:-1, :] * lr_a[:-1, :] - g1[1:, :] * lr_a[1:, :]) +
                (g2[:, 1:] * bt_a[:, 1:] - g2[:, :-1] * bt_a[:, :-1]))

    f_ru = f_ru + vdiv(t1x, t2x) / vol
    f_rv = f_rv + vdiv(t1y, t2y) / vol
    ir1 = 1 / (rho[:-1, 1:-1] + rho[1:, 1:-1])
    ir2 = 1 / (rho[1:-1, :-1] + rho[1:-1, 1:])
    u1 = (rho_u[:-1, 1:-1] + rho_u[1:, 1:-1]) * ir1
    u2 = (rho_u[1:-1, :-1] + rho_u[1:-1, 1:]) * ir2
    v1 = (rho_v[:-1, 1:-1] + rho_v[1:, 1:-1]) * ir1
    v2 = (rho_v[1:-1, :-1] + rho_v[1:-1, 1:]) * ir2
    w1 = u1 * t1x + v1 * t1y
    w2 = u2 * t2x + v2 * t2y
    f_rE = f_rE + vdiv(w1, w2) / vol
    g = fs["gamma_star"]
    k2 = (ks[1:-1, 1:] + ks[1:-1, :-1]) / 2 * g / Re / Pr
    k1 = (ks[1:, 1:-1] + ks[:-1, 1:-1]) / 2 * g / Re / Pr
    dT_c2, dT_c1, dT_t2, dT_t1 = grads(T, T_c)
    dT_n2 = (cn2_ft2 * dT_t2 - dT_c2) / den2
    dT_n1 = (dT_c1 - dT_t1 * cn1_ft1) / cn1_fn1
    q2 = k2 * dT_n2
    q1 = k1 * dT_n1
    heat = ((q2[:, 1:] * bt_a[:, 1:] - q2[:, :-1] * bt_a[:, :-1]) -
            (q1[1:, :] * lr_a[1:, :] - q1[:-1, :] * lr_a[:-1, :]))
    f_rE = f_rE + heat / vol
    return f_rho, f_ru, f_rv, f_rE


# ------------------------------------------------------------------ Jacobian

def _colour_mask(index: int, nchi: int, neta: int) -> np.ndarray:
    """GET_PERTURBATION_MATRIX.jl:959 (1-based index 1..9): every third cell,
    chi start ((index - 1) mod 3) + 1, eta start ceil(index / 3)."""
    si = index % 3 if index % 3 else 3
    sj = -(-index // 3)
    m = np.zeros((nchi, neta), dtype=bool)
    m[si - 1::3, sj - 1::3] = True
    return m


def linearize(c: ChannelCase):
    """Forward-difference Jacobian of rhs() about the base state.

    Returns (op, s): a BlockOperator with the 4 x 4 cell blocks (row cell
    <- column cell, variables rho, rho_u, rho_v, rho_E) including explicit
    zero blocks of the 3 x 3 stencil, and the solution dict (mesh, var,
    flux, freestream, time_integration, boundary_conditions, shock) in the
    reference's layout for CFL_TIMESTEP and the Chu weights."""
    if c.nchi % 3:
        raise ValueError("a chi-periodic grid needs Nchi divisible by 3 (VERIFY_INPUT_DATA.jl:228)")
    mesh = make_mesh(c)
    fs = freestream(c)
    state = base_state(c, mesh, fs)
    f0 = rhs(state, mesh, fs)
    N, M = c.nchi, c.neta
    eps = c.perturbation_magnitude
    ncell = N * M
    # blocks[cell_out, cell_in] for the 9 offsets, as dense (ncell, 9, 4, 4)
    offs = [(di, dj) for dj in (-1, 0, 1) for di in (-1, 0, 1)]
    blk = np.zeros((N, M, 9, 4, 4))      # [i_out, j_out, offset of in-cell, var_out, var_in]
    present = np.zeros((N, M, 9), dtype=bool)
    ii, jj = np.meshgrid(np.arange(N), np.arange(M), indexing="ij")
    for var_in in range(4):
        for colour in range(1, 10):
            mask = _colour_mask(colour, N, M)
            pert = [np.array(a) for a in state]
            inner = pert[var_in][1:-1, 1:-1]
            pert[var_in][1:-1, 1:-1] = inner + mask * eps
            f1 = rhs(pert, mesh, fs)
            resp = [(f1[k] - f0[k]) / eps for k in range(4)]
            # each output cell (i, j) sees exactly one perturbed cell of this
            # colour in its 3 x 3 neighbourhood, at offset (di, dj)
            for o, (di, dj) in enumerate(offs):
                si = (ii + di) % N                    # chi-periodic
                sj = jj + dj
                ok = (sj >= 0) & (sj < M)
                hit = np.zeros_like(ok)
                hit[ok] = mask[si[ok], sj[ok]]
                for var_out in range(4):
                    blk[hit, o, var_out, var_in] = resp[var_out][hit]
                present[:, :, o] |= ok
    # cell-major BSR over cells (cell = i + j*N), block columns ascending
    rows, cols, vals = [], [], []
    for o, (di, dj) in enumerate(offs):
        sj = jj + dj
        ok = (sj >= 0) & (sj < M) & present[:, :, o]
        src = ((ii + di) % N) + sj * N
        dst = ii + jj * N
        rows.append(dst[ok])
        cols.append(src[ok])
        vals.append(blk[:, :, o][ok])
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    vals = np.concatenate(vals)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    brow_ptr = np.zeros(ncell + 1, dtype=np.int64)
    np.cumsum(np.bincount(rows, minlength=ncell), out=brow_ptr[1:])
    op = BlockOperator(N, M, 4, 0, brow_ptr, cols.astype(np.int32), vals)
    s = _solution_dict(c, mesh, fs, state, f0)
    return op, s


def _solution_dict(c, mesh, fs, state, f0):
    rho, rho_u, rho_v, rho_E = (np.array(a) for a in state)
    gam = np.full(rho.shape, fs["gamma_star"])
    p, T = _thermo_and_bcs(rho, rho_u, rho_v, rho_E, gam, np.ones(rho.shape))
    var = {"rho": rho, "rho_u": rho_u, "rho_v": rho_v, "rho_E": rho_E, "p": p, "T": T,
           "gamma_star": gam, "cv_star": np.ones(rho.shape), "mu_star": np.ones(rho.shape),
           "k_star": np.ones(rho.shape), "a": np.sqrt(gam * p / rho),
           "Re_flow": np.full(rho.shape, fs["Re"]), "Pr_flow": np.full(rho.shape, fs["Pr"])}
    flux = dict(zip(("rho", "rho_u", "rho_v", "rho_E"), f0))
    return {"mesh": mesh, "var": var, "flux": flux, "freestream": fs,
            "time_integration": {"CFL": c.CFL, "max_dt": c.max_dt},
            "boundary_conditions": {"boundary_chi0": {"name": "periodic"},
                                    "boundary_chi1": {"name": "periodic"},
                                    "boundary_eta0": {"name": "no_slip_adiabatic"},
                                    "boundary_eta1": {"name": "no_slip_adiabatic"}},
            "shock": {"enabled": False, "flow_cells": np.ones((c.nchi, c.neta))},
            "stability_analysis": {"perturbation_magnitude": c.perturbation_magnitude,
                                   "perturb_shock": False,
                                   "eigenvalue_solver": "GPU_TIMESTEPPER_ARNOLDI"}}


def channel_operator(c: ChannelCase | None = None):
    """(BlockOperator, scipy CSC in the reference's var-major order, s)."""
    c = c or ChannelCase()
    op, s = linearize(c)
    colptr, rowval, nzval = op.to_csc_varmajor(one_based=False)
    A = sp.csc_matrix((nzval, rowval, colptr), shape=(op.n, op.n))
    return op, A, s
