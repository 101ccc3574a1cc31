"""Host-side mirror of HYMOR's stability entry points over libhymor_hip.

Same names, argument meaning and error behaviour as the Julia functions in
utils/Stability_analysis/ (Julia's ``!`` suffix becomes ``_``).  The Taylor
actions run on the GPU through the C ABI; the one-time host constructions
(Chu weights, CFL time step) are numpy restatements of the reference's own
code; the Krylov driver is ARPACK, as in the reference (Arpack.jl there,
scipy's bundled ARPACK-ng here), with the same which/tol/maxiter settings.

Matrices passed in are scipy sparse matrices in the reference's layout
(variable-major state, chi-fastest cells, 4 variables per cell, optional
Nchi shock DOFs, optional freestream blocks).  Base-flow dictionaries follow
the reference's ``s`` (see hymor_amd.baseflow for the keys used).  Array
fields use Julia's column-major flattening (chi fastest).
"""
from __future__ import annotations

import time
import warnings

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import _lib
from .core import (DeviceOperator, DeviceVector, DeviceWeights, chu_sandwich_batched_ptr,
                   chu_sandwich_ptr, energy_track, expm_taylor_batched_ptr, expm_taylor_ptr,
                   spmv_dev)
from .krylov import block_krylov_schur, krylov_schur

ORDER = 5            # Taylor order hard-coded by the reference
B = 4                # conserved variables per cell in the reference


def _vec(a):
    """Julia vec(): column-major flattening (chi fastest)."""
    return np.asarray(a).reshape(-1, order="F")


def _interior(a):
    return np.asarray(a)[1:-1, 1:-1]


def _use_gpu():
    # CUDA.functional() in the reference; here the engine has no CPU path,
    # so a missing GPU is an error at the first device call.
    return _lib.device_available()


# ====================================================================== dt

def CFL_TIMESTEP(s):
    """Time step from the CFL condition (utils/Time_marching/CFL_TIMESTEP.jl:41-138).

    Sets ``s["time_integration"]["dt"]`` and returns ``s``.  If ``s`` holds
    no base-flow fields, an existing ``dt`` is kept (callers that already
    know dt pass it that way).
    """
    ti = s["time_integration"]
    if "var" not in s:
        if "dt" not in ti:
            raise KeyError("s has neither base-flow fields nor time_integration.dt")
        return s
    var, mesh, shock, flux = s["var"], s["mesh"], s["shock"], s["flux"]
    rho, rho_u, rho_v, rho_E = var["rho"], var["rho_u"], var["rho_v"], var["rho_E"]
    a, gam, Re, Pr = var["a"], var["gamma_star"], var["Re_flow"], var["Pr_flow"]
    lr_xn, lr_yn, lr_a = mesh["lr_x_normal"], mesh["lr_y_normal"], mesh["lr_area"]
    bt_xn, bt_yn, bt_a = mesh["bt_x_normal"], mesh["bt_y_normal"], mesh["bt_area"]
    vol = mesh["volume"]
    fc = shock["flow_cells"]
    c = np.max(np.abs(a))                                                  # :80
    rs = rho[:-1, 1:-1] + rho[1:, 1:-1]                                    # :84-88
    ul = (rho_u[:-1, 1:-1] + rho_u[1:, 1:-1]) / rs
    vl = (rho_v[:-1, 1:-1] + rho_v[1:, 1:-1]) / rs
    lam_lr = (np.abs(ul * lr_xn + vl * lr_yn) + c) * lr_a
    rs = rho[1:-1, :-1] + rho[1:-1, 1:]                                    # :91-95
    ub = (rho_u[1:-1, :-1] + rho_u[1:-1, 1:]) / rs
    vb = (rho_v[1:-1, :-1] + rho_v[1:-1, 1:]) / rs
    lam_bt = (np.abs(ub * bt_xn + vb * bt_yn) + c) * bt_a
    lam_c = np.maximum(lam_lr[:-1, :], lam_lr[1:, :]) + np.maximum(lam_bt[:, :-1], lam_bt[:, 1:])
    ri, gi, Rei, Pri = _interior(rho), _interior(gam), _interior(Re), _interior(Pr)
    A_xi = 0.5 * (lr_a[:-1, :] + lr_a[1:, :])                              # :108-112
    A_eta = 0.5 * (bt_a[:, :-1] + bt_a[:, 1:])
    nu = np.maximum(4.0 / 3.0, gi / Pri) / (ri * Rei)
    lam_v = 2 * nu * (A_xi ** 2 + A_eta ** 2) / vol
    dt1 = ti["CFL"] / np.max((lam_c + lam_v) * fc / vol)                   # :115-116
    rEi = _interior(rho_E)
    if shock["enabled"]:                                                   # :121-129
        mu = np.max((np.abs(flux["rho_v"]) + np.abs(flux["rho_u"])) / ri * fc)
        mr = np.max(np.abs(flux["rho"] / ri * fc))
        mE = np.max(np.abs(flux["rho_E"] / rEi * fc))
    else:
        mu = np.max((np.abs(flux["rho_v"]) + np.abs(flux["rho_u"])) / ri)
        mr = np.max(np.abs(flux["rho"] / ri))
        mE = np.max(np.abs(flux["rho_E"] / rEi))
    dt2 = ti["CFL"] / max(mu, mr, mE)
    ti["dt"] = min(dt1, dt2, ti["max_dt"])                                 # :133
    return s


# ====================================================================== Chu weights

# structural pattern of R: the 12 entries CONSTRUCT_R.jl:51-101 assigns
_R_PATTERN = np.array([[1, 1, 1, 1], [1, 1, 0, 0], [1, 0, 1, 0], [1, 1, 1, 1]], dtype=bool)


def _r_blocks(gam, p, rho, u, v):
    """Per-point 4x4 conservative -> (p, u, v, S) Jacobian, rows/cols
    (p,u,v,S) x (rho, rho_u, rho_v, rho_E) (CONSTRUCT_R.jl:51-101)."""
    m = gam.shape[0]
    R = np.zeros((m, 4, 4))
    R[:, 0, 0] = (gam - 1) * (u ** 2 + v ** 2) / 2
    R[:, 0, 1] = -u * (gam - 1)
    R[:, 0, 2] = -v * (gam - 1)
    R[:, 0, 3] = gam - 1
    R[:, 1, 0] = -u / rho
    R[:, 1, 1] = 1.0 / rho
    R[:, 2, 0] = -v / rho
    R[:, 2, 2] = 1.0 / rho
    R[:, 3, 0] = (u ** 2 + v ** 2) / (2 * p) - gam / (gam - 1) / rho
    R[:, 3, 1] = -u / p
    R[:, 3, 2] = -v / p
    R[:, 3, 3] = 1.0 / p
    return R


def _blocks_varmajor(R, npts, n_total, diag_tail=None):
    """Assemble per-point 4x4 blocks into a var-major sparse matrix over
    4*npts unknowns (+ an optional diagonal on the tail)."""
    pts = np.arange(npts)
    rows = (np.arange(4)[None, :, None] * npts + pts[:, None, None]) + np.zeros((1, 1, 4), int)
    cols = (np.arange(4)[None, None, :] * npts + pts[:, None, None]) + np.zeros((1, 4, 1), int)
    r, c, v = rows.reshape(-1), cols.reshape(-1), R.reshape(-1)
    keep = np.broadcast_to(_R_PATTERN, R.shape).reshape(-1)   # entries the reference sets
    r, c, v = r[keep], c[keep], v[keep]
    if diag_tail is not None:
        t = np.arange(diag_tail.shape[0]) + 4 * npts
        r, c, v = np.concatenate([r, t]), np.concatenate([c, t]), np.concatenate([v, diag_tail])
    return sp.csc_matrix((v, (r, c)), shape=(n_total, n_total))


def _flow(s):
    var = s["var"]
    gam = _vec(_interior(var["gamma_star"]))
    p = _vec(_interior(var["p"]))
    rho = _vec(_interior(var["rho"]))
    u = _vec(_interior(var["rho_u"])) / rho
    v = _vec(_interior(var["rho_v"])) / rho
    a = _vec(_interior(var["a"]))
    return gam, p, rho, u, v, a


def _n_state(s):
    nx, ny = s["mesh"]["Nchi"], s["mesh"]["Neta"]
    return 4 * nx * ny + (nx if s["stability_analysis"]["perturb_shock"] else 0)


def CONSTRUCT_R(s):
    """Conservative -> Chu-variable transform (CONSTRUCT_R.jl:24-133)."""
    nx, ny = s["mesh"]["Nchi"], s["mesh"]["Neta"]
    gam, p, rho, u, v, _ = _flow(s)
    R = _r_blocks(gam, p, rho, u, v)
    tail = np.ones(nx) if s["stability_analysis"]["perturb_shock"] else None   # :114-121
    return _blocks_varmajor(R, nx * ny, _n_state(s), tail)


def _m_diag(rho, a, gam, p, vol, norms):
    """Diagonal Chu weights per point, (npts, 4) (CONSTRUCT_M.jl:50-75)."""
    w = np.zeros((rho.shape[0], 4))
    if norms[0]:
        w[:, 0] = rho * a ** 2 / (2 * (gam * p) ** 2) * vol
    if norms[1]:
        w[:, 1] = rho / 2 * vol
    if norms[2]:
        w[:, 2] = rho / 2 * vol
    if norms[3]:
        w[:, 3] = (gam - 1) * p / (2 * gam) * vol
    return w


def _diag_varmajor(w, n_total, tail=None, present=None):
    npts = w.shape[0]
    idx = (np.arange(4)[None, :] * npts + np.arange(npts)[:, None])
    cols = [k for k in range(4) if present is None or present[k]]
    r = idx[:, cols].reshape(-1)
    v = w[:, cols].reshape(-1)
    if tail is not None:
        t = np.arange(tail.shape[0]) + 4 * npts
        r, v = np.concatenate([r, t]), np.concatenate([v, tail])
    return sp.csc_matrix((v, (r, r)), shape=(n_total, n_total))


def CONSTRUCT_M(s, norms):
    """Chu energy weights, masked by flow_cells (CONSTRUCT_M.jl:27-94)."""
    gam, p, rho, u, v, a = _flow(s)
    vol = _vec(s["mesh"]["volume"]) * _vec(s["shock"]["flow_cells"])
    present = [bool(x) for x in norms]
    return _diag_varmajor(_m_diag(rho, a, gam, p, vol, present), _n_state(s), present=present)


def CONSTRUCT_M_HAT(s, norms):
    """Unmasked Chu weights plus shock weights bt_area/2 (CONSTRUCT_M_HAT.jl:25-105)."""
    gam, p, rho, u, v, a = _flow(s)
    vol = _vec(s["mesh"]["volume"])
    present = [bool(x) for x in norms]
    tail = None
    if s["stability_analysis"]["perturb_shock"]:
        nx = s["mesh"]["Nchi"]
        ci = np.asarray(s["shock"]["cell_indices"]).reshape(nx, -1)[:, 0].astype(int) - 1
        tail = s["mesh"]["bt_area"][np.arange(nx), ci] / 2                     # :79-82
    return _diag_varmajor(_m_diag(rho, a, gam, p, vol, present), _n_state(s), tail, present)


def _freestream_pts(s):
    fs = s["freestream"]
    nx = s["mesh"]["Nchi"]
    one = np.ones(nx)
    return (fs["gamma_star"] * one, fs["p_0"] * one, fs["rho_0"] * one,
            fs["rho_u_0"] / fs["rho_0"] * one, fs["rho_v_0"] / fs["rho_0"] * one,
            fs["a_0"] * one)


def CONSTRUCT_R_INFTY(s):
    """Freestream transform, 4*Nchi square (CONSTRUCT_R_.jl:55-145)."""
    gam, p, rho, u, v, _ = _freestream_pts(s)
    nx = s["mesh"]["Nchi"]
    return _blocks_varmajor(_r_blocks(gam, p, rho, u, v), nx, 4 * nx)


def CONSTRUCT_R_(s, w_infty):
    """blockdiag(R, R_infty x N) (CONSTRUCT_R_.jl:27-37)."""
    Rinf = CONSTRUCT_R_INFTY(s)
    return sp.block_diag([CONSTRUCT_R(s)] + [Rinf] * len(w_infty), format="csc")


def CONSTRUCT_M_(s, norms, w_infty):
    """blockdiag(M, 0) (CONSTRUCT_M_.jl:29-38)."""
    nx = s["mesh"]["Nchi"]
    M = CONSTRUCT_M(s, norms)
    z = 4 * nx * len(w_infty)
    return sp.block_diag([M, sp.csc_matrix((z, z))], format="csc")


def GET_T_REF(s):
    """Reference advection time: post-shock mass / inflow mass flux (GET_T_REF.jl:27-48)."""
    mass = np.sum(_interior(s["var"]["rho"]) * s["shock"]["flow_cells"] * s["mesh"]["volume"])
    fs = s["freestream"]
    nx = s["mesh"]["Nchi"]
    ci = np.asarray(s["shock"]["cell_indices"]).reshape(nx, -1)[:, 0].astype(int) - 1
    i = np.arange(nx)
    flux = 0.0
    for k in range(nx):           # sequential sum, as the reference's loop
        nd = fs["rho_u_0"] * s["mesh"]["bt_x_normal"][i[k], ci[k]] + \
            fs["rho_v_0"] * s["mesh"]["bt_y_normal"][i[k], ci[k]]
        flux += s["mesh"]["bt_area"][i[k], ci[k]] * abs(nd)
    return mass / flux


def CONSTRUCT_M_INFTY(s, norms, T, scaling_non_temporal):
    """Freestream inflow weights (CONSTRUCT_M_INFTY_.jl:72-164)."""
    gam, p, rho, u, v, a = _freestream_pts(s)
    nx = s["mesh"]["Nchi"]
    fs = s["freestream"]
    vmag = np.sqrt(fs["rho_u_0"] ** 2 + fs["rho_v_0"] ** 2)
    uh, vh = fs["rho_u_0"] / vmag, fs["rho_v_0"] / vmag
    ci = np.asarray(s["shock"]["cell_indices"]).reshape(nx, -1)[:, 0].astype(int) - 1
    i = np.arange(nx)
    proj = s["mesh"]["bt_x_normal"][i, ci] * uh + s["mesh"]["bt_y_normal"][i, ci] * vh
    arc = s["mesh"]["bt_area"][i, ci] * np.abs(proj)
    scaling = GET_T_REF(s) if scaling_non_temporal else T
    present = [bool(x) for x in norms]
    return _diag_varmajor(_m_diag(rho, a, gam, p, arc * scaling, present), 4 * nx,
                          present=present)


def CONSTRUCT_M_INFTY_(s, norms, T, w_infty, scaling_non_temporal):
    """blockdiag(M_hat, M_infty (halved for w != 0) x N) (CONSTRUCT_M_INFTY_.jl:32-50)."""
    Mh = CONSTRUCT_M_HAT(s, norms)
    Mi = CONSTRUCT_M_INFTY(s, norms, T, scaling_non_temporal)
    blocks = [Mh] + [Mi if w == 0 else Mi / 2 for w in w_infty]
    return sp.block_diag(blocks, format="csc")


# ====================================================================== Taylor actions

def _grid(s):
    return s["mesh"]["Nchi"], s["mesh"]["Neta"]


def upload(A, s, scale=1.0, value_prec="f64"):
    """CuSparseMatrixCSC(A*scale) -> a device operator (forward + adjoint)."""
    nx, ny = _grid(s)
    return DeviceOperator.from_scipy(A, nchi=nx, neta=ny, vars_per_cell=B, scale=scale,
                                     value_prec=value_prec)


def POWER_METHOD_(v, A: DeviceOperator, timesteps, vout=None, term=None):
    """POWER_METHOD! (Eigenvalues/ARNOLDI_POWER_GPU.jl:64-82): v <- T_5(A)^timesteps v
    in place; A is the device operator of A*dt.  vout/term are accepted for
    signature parity (the library owns its buffers)."""
    v[...] = A.expm_taylor(v, order=ORDER, nsub=int(timesteps), dt=1.0, form="powers")


def ARNOLDI_POWER_GPU(x, Adt: DeviceOperator, timesteps, use_gpu=True, x_buf=None, vout_buf=None,
                      term_buf=None, x_cpu_buf=None):
    """ARPACK callback adapter (ARNOLDI_POWER_GPU.jl:36-52): returns a new vector."""
    xc = np.ascontiguousarray(x)
    return Adt.expm_taylor(xc, order=ORDER, nsub=int(timesteps), dt=1.0, form="powers")


def LINEAR_INTEGRATION_GPU(x, n_t, order, Adt: DeviceOperator, use_gpu=True, v_buf=None,
                           vout_buf=None, term_buf=None):
    """PLOT_MODES.jl:330-351: T_order(A dt)^n_t x."""
    return Adt.expm_taylor(np.ascontiguousarray(x), order=int(order), nsub=int(n_t), dt=1.0,
                           form="powers")


def EXPONENTIAL_(v, A: DeviceOperator, At, C_no_exp: DeviceOperator, timesteps, vout=None,
                 term=None):
    """EXPONENTIAL! (CONSTRUCT_C.jl:78-111): v <- T5(A'dt)^n C T5(A dt)^n v in
    place.  ``At`` is accepted for signature parity: the handle of A*dt
    holds its own adjoint."""
    v[...] = A.chu_sandwich(C_no_exp, v, order=ORDER, nsub=int(timesteps), dt=1.0)


def CONSTRUCT_C(x, A_dt: DeviceOperator, At_dt, C_no_exp: DeviceOperator, n_t,
                freestream_disturbances, N_set_0, use_gpu=True, x_buf=None, vout_buf=None,
                term_buf=None, x_cpu_buf=None):
    """CONSTRUCT_C.jl:33-58, with the zeroing of x[1:N_set_0] before and after."""
    n0 = int(N_set_0) if freestream_disturbances else 0
    return A_dt.chu_sandwich(C_no_exp, np.ascontiguousarray(x), order=ORDER, nsub=int(n_t), dt=1.0,
                             n_set0=n0)


# ====================================================================== eigen drivers

def _sort_lm(vals, vecs):
    order = np.argsort(-np.abs(vals), kind="stable")
    return vals[order], vecs[:, order]


def EIGENVALUES(A, s, n_modes):
    """Dispatcher (EIGENVALUES.jl:37-49)."""
    solver = s["stability_analysis"]["eigenvalue_solver"]
    if solver == "CPU_LU":
        raise NotImplementedError("CPU_LU (shift-invert LU) is not part of the GPU engine; "
                                  "use GPU_TIMESTEPPER_ARNOLDI")
    if solver == "GPU_TIMESTEPPER_ARNOLDI":
        return EIGENVALUES_GPU_ARNOLDI(A, s, n_modes)
    raise ValueError(f"Unknown eigenvalue solver: {solver}")


def _device_op(A, s, scale, value_prec):
    """Device operator of A*scale from a scipy matrix in the reference's
    var-major order, or from a hymor_amd.operators.BlockOperator."""
    if hasattr(A, "brow_ptr"):
        return DeviceOperator.from_blocks(A, scale=scale, value_prec=value_prec,
                                          with_adjoint=False)
    return upload(A, s, scale=scale, value_prec=value_prec)


def EIGENVALUES_GPU_ARNOLDI(A, s, n_modes, *, T_ref=0.5, tol=1e-4, maxiter=None, ncv=None,
                            v0=None, krylov="arpack", seed=0, form="powers"):
    """Arnoldi on the exponential transform (EIGENVALUES.jl:112-205).

    dt from CFL_TIMESTEP, timesteps = round(T_ref/dt), A*dt uploaded once,
    ARPACK which=LM, tol=1e-4, maxiter=40k; discrete eigenvalues mapped by
    lambda = log(mu)/(dt*timesteps).  On any failure returns NaN V, D with a
    warning, as the reference does.

    krylov="device" replaces ARPACK by the Krylov-Schur solver over a
    device-resident basis (hymor_amd.krylov): same selection (largest
    magnitude) and convergence test, no host round trip per matvec; with no
    v0 its start vector is random on the device (`seed`).  `form` "horner"
    evaluates POWER_METHOD!'s polynomial by Horner's rule (three vector
    passes per term instead of four; equal to rounding).
    """
    s = CFL_TIMESTEP(s)
    dt = s["time_integration"]["dt"]
    timesteps = int(round(T_ref / dt))            # Julia round(Int, .) ties to even, as Python
    n = A.n if hasattr(A, "brow_ptr") else A.shape[0]
    k = n_modes
    t0 = time.time()
    dev = None
    try:
        dev = _device_op(A, s, dt, "f64")
        if krylov == "device":
            dev.set_graphs(True)      # every matvec repeats the same sweep
            def apply(din, dout):
                expm_taylor_ptr(dev, din, dout, order=ORDER, nsub=timesteps, form=form,
                                complex_=False)
                dev.synchronize()
            vals, vecs, kinfo = krylov_schur(apply, n, k, ncv or max(20, 2 * k + 1), tol=tol,
                                             maxiter=maxiter or 40 * k, v0=v0, seed=seed)
            s.setdefault("timings", {})["eigenvalues_krylov"] = kinfo
        else:
            op = spla.LinearOperator((n, n), dtype=np.float64,
                                     matvec=lambda x: ARNOLDI_POWER_GPU(x, dev, timesteps))
            vals, vecs = spla.eigs(op, k=k, which="LM", tol=tol, maxiter=maxiter or 40 * k,
                                   ncv=ncv, v0=v0)
        vals, vecs = _sort_lm(vals, vecs)
        D = np.log(vals.astype(np.complex128)) / dt / timesteps          # :164-168
        return vecs, D
    except Exception as e:   # the reference returns NaN on any failure (:184-203)
        warnings.warn(f"eigs did not converge: {e} -- returning NaN.")
        return np.full((n, k), np.nan), np.full(k, np.nan + 1j * np.nan)
    finally:
        if dev is not None:
            dev.close()
        s.setdefault("timings", {})["eigenvalues"] = time.time() - t0


def TRANSIENT_GROWTH_DOWNSTREAM(A, s, n_modes, T, *, tol=1e-4, maxiter=None, ncv=None, v0=None,
                                krylov="arpack", block=2):
    """Optimal transient growth (TRANSIENT_GROWTH_DOWNSTREAM.jl:34-153).

    Returns (V, D, T_opt): D are the gains, eigenvalues of
    exp(A'T) R'MR exp(AT) x = lambda R'M_hat R x (ARPACK, symmetric
    generalized mode, which=LM, tol=1e-4, maxiter=10k).

    krylov="device": Krylov-Schur on the Cholesky-reduced operator with a
    device basis; block > 1 expands it `block` vectors at a time through the
    batched sandwich (block=1: one sandwich per matvec).
    """
    k = n_modes
    s = CFL_TIMESTEP(s)
    n_t = int(round(T / s["time_integration"]["dt"]))
    dt = T / n_t
    T_opt = n_t * dt
    R = CONSTRUCT_R(s)
    norms = [1, 1, 1, 1]
    C_no_exp = (R.T @ CONSTRUCT_M(s, norms) @ R).tocsc()
    D_mat = (R.T @ CONSTRUCT_M_HAT(s, norms) @ R).tocsc()
    n = A.shape[0]
    devA = devC = None
    try:
        t0 = time.perf_counter()
        devA = upload(A, s, scale=dt)
        devC = upload(C_no_exp, s)
        s.setdefault("timings", {})["transient_growth_upload"] = time.perf_counter() - t0
        if krylov == "device":
            vals, vecs = _device_generalized(devA, devC, D_mat, s, n, k, n_t, 0, False, tol,
                                             maxiter or 10 * k, ncv, v0, block,
                                             s.setdefault("timings", {}).setdefault(
                                                 "transient_growth_krylov", {}))
            vals = np.real(vals)
        else:
            op = spla.LinearOperator((n, n), dtype=np.float64,
                                     matvec=lambda x: CONSTRUCT_C(x, devA, None, devC, n_t, False,
                                                                  0))
            vals, vecs = spla.eigsh(op, k=k, M=D_mat, which="LM", tol=tol,
                                    maxiter=maxiter or 10 * k, ncv=ncv, v0=v0)
        order = np.argsort(-np.abs(vals), kind="stable")
        return vecs[:, order], vals[order], T_opt
    except Exception as e:
        warnings.warn(f"eigs did not converge: {e} -- returning NaN.")
        return np.full((n, k), np.nan), np.full(k, np.nan), T_opt
    finally:
        for d in (devA, devC):
            if d is not None:
                d.close()


def _upload_extended_from_parts(A_, s, w_infty, dt):
    """The extended operator A_ * dt on the device from its parts: A_ is
    checked to have the reference structure (split_extended) and rebuilt by
    hymor_op_create_receptivity from [L B; 0 0] and the wavenumbers, with B
    kept once."""
    from .receptivity import receptivity_operator, split_extended
    nx, ny = _grid(s)
    N = len(w_infty)
    p = B * nx
    n_L = A_.shape[0] - N * p
    L, Bm = split_extended(A_, n_L, p, w_infty)
    return receptivity_operator(L, Bm, w_infty, nchi=nx, neta=ny, vars_per_cell=B, scale=dt)


def FREESTREAM_RECEPTIVITY(A_, s, n_modes, T, w_infty, *, tol=1e-4, maxiter=None, ncv=None,
                           v0=None, krylov="arpack", factored=False, block=4):
    """Freestream receptivity (FREESTREAM_RECEPTIVITY.jl:37-162).

    A_ is the complex extended operator [L B..B; 0 kron(diag(-i w), I)].
    Hermitian generalized ARPACK on CONSTRUCT_C with the first N_set_0 =
    4*Nchi*Neta entries zeroed; gains are real(lambda)/2.

    factored=True uploads A_ from its parts (one per-wavenumber system
    [L B; 0 0] and the wavenumbers, hymor_op_create_receptivity): B is never
    repeated on the device, and each action costs one flow SpMV per term.
    krylov="device" and block as in TRANSIENT_GROWTH_DOWNSTREAM (blocks of 4
    by default: on the 500 x 250-cell operator extended by 32 wavenumbers
    the Krylov stage takes 0.50 s against 0.69 s at 2 and 0.87 s at 1).
    """
    k = n_modes
    s = CFL_TIMESTEP(s)
    n_t = int(round(T / s["time_integration"]["dt"]))
    dt = T / n_t
    T_opt = n_t * dt
    R_ = CONSTRUCT_R_(s, w_infty)
    norms = [1, 1, 1, 1]
    C_ = (R_.T @ CONSTRUCT_M_(s, norms, w_infty) @ R_).tocsc()
    D_ = (R_.T @ CONSTRUCT_M_INFTY_(s, norms, T, w_infty, True) @ R_).tocsc()
    nx, ny = _grid(s)
    n_set_0 = 4 * nx * ny
    n = A_.shape[0]
    devA = devC = None
    try:
        if factored:
            devA = _upload_extended_from_parts(A_, s, w_infty, dt)
        else:
            devA = upload(sp.csc_matrix(A_, dtype=np.complex128), s, scale=dt)
        devC = upload(C_, s)
        if krylov == "device":
            vals, vecs = _device_generalized(devA, devC, D_, s, n, k, n_t, n_set_0, True, tol,
                                             maxiter or 10 * k, ncv, v0, block,
                                             s.setdefault("timings", {}).setdefault(
                                                 "receptivity_krylov", {}))
        else:
            op = spla.LinearOperator((n, n), dtype=np.complex128,
                                     matvec=lambda x: CONSTRUCT_C(x, devA, None, devC, n_t, True,
                                                                  n_set_0))
            vals, vecs = spla.eigsh(op, k=k, M=D_.astype(np.complex128), which="LM", tol=tol,
                                    maxiter=maxiter or 10 * k, ncv=ncv, v0=v0)
        D = np.real(vals) / 2                                            # :115-119
        order = np.argsort(-np.abs(D), kind="stable")
        return vecs[:, order], D[order], T_opt
    except Exception as e:
        warnings.warn(f"eigs did not converge: {e} -- returning NaN.")
        return np.full((n, k), np.nan), np.full(k, np.nan), T_opt
    finally:
        for d in (devA, devC):
            if d is not None:
                d.close()


# ====================================================================== energy tracking

def LINEAR_INTEGRATION_AND_ENERGY_GPU_NO_SAMPLES(x, n_t, dt, t_0, orderTaylor, Adt: DeviceOperator,
                                                 C_no_exp_all, C_no_exp_p, C_no_exp_k, C_no_exp_S,
                                                 use_gpu=True, v_buf=None, vout_buf=None,
                                                 term_buf=None, *, weights=None):
    """Forward integration with the four Chu energies per step
    (Freestream_receptivity/LINEAR_INTEGRATION_AND_GAINS.jl:436-477).

    Returns (t, E_all, E_p, E_k, E_S); as in the reference, entry 0 of each
    energy series is 0 (its caller fills it).  Adt is the device operator of
    A*dt; the C matrices are the scipy weights (or pass ``weights``, a
    DeviceWeights built from them, to reuse the upload).
    """
    t, E, _, _ = _energy_run(x, n_t, dt, t_0, orderTaylor, Adt,
                             (C_no_exp_all, C_no_exp_p, C_no_exp_k, C_no_exp_S), weights, 0, 0)
    return (t, *E)


def LINEAR_INTEGRATION_AND_ENERGY_GPU(x, n_t, N_samples, N_skip, dt, t_0, orderTaylor,
                                      Adt: DeviceOperator, C_no_exp_all, C_no_exp_p, C_no_exp_k,
                                      C_no_exp_S, use_gpu=True, v_buf=None, vout_buf=None,
                                      term_buf=None, *, weights=None):
    """As _NO_SAMPLES plus snapshots Vf[:, count] of the state after every
    N_skip-th step, N_samples columns (zeros where the run had fewer)
    (LINEAR_INTEGRATION_AND_GAINS.jl:375-427)."""
    t, E, Vf, _ = _energy_run(x, n_t, dt, t_0, orderTaylor, Adt,
                              (C_no_exp_all, C_no_exp_p, C_no_exp_k, C_no_exp_S), weights,
                              N_skip, N_samples)
    return (t, *E, Vf)


def LINEAR_INTEGRATION_AND_GAINS(freestream_disturbances, v, lam, T_opt, A_, s, chemistry, T_f,
                                 get_amplification_only, w_infty):
    """Energy gains of a perturbation integrated forward to T_f
    (Freestream_receptivity/LINEAR_INTEGRATION_AND_GAINS.jl:45-366): the Chu
    weights of the total, acoustic, kinetic and entropic energies (:66-114),
    one energy-tracked integration of v under T_5(A_ dt) with dt = T_f /
    round(T_f / dt_CFL) (:53-58, :141-145), the initial energies (:169-174),
    the scalings (:180-187: E(0), or for freestream disturbances the inflow
    energy rate v' D v times GET_T_REF and times t) and the returned
    max_gain dict (:341-363): "temporal" -- max over t >= t_10 of
    E(t) / scaling(t) -- and "non_temporal" -- max of E / scaling_ref --
    for "all", "acoustic", "kinetic", "entropic".

    Only the amplification path (get_amplification_only = True) is
    provided: the energy budgets of the other path (COMPUTE_BUDGETS_KINETIC,
    _ENTROPY, _SHOCK, ENERGY_INFLOW, :204-339) are flux and chemistry
    post-processing of the nonlinear solver, outside this engine; the
    figures are not drawn.  lam, T_opt and chemistry are accepted for the
    reference's signature (they feed only those parts)."""
    if not get_amplification_only:
        raise NotImplementedError(
            "LINEAR_INTEGRATION_AND_GAINS: the energy budgets (get_amplification_only=false) need "
            "the nonlinear solver's flux and chemistry post-processing (COMPUTE_BUDGETS_*), "
            "which this engine does not include; call it with get_amplification_only=true")
    orderTaylor = ORDER
    s = CFL_TIMESTEP(s)
    n_t = int(round(T_f / s["time_integration"]["dt"]))
    dt = T_f / n_t
    one, p_only, k_only, s_only = [1, 1, 1, 1], [1, 0, 0, 0], [0, 1, 1, 0], [0, 0, 0, 1]
    if freestream_disturbances:
        R_ = CONSTRUCT_R_(s, w_infty)
        quad = lambda M: (R_.T @ M @ R_).tocsc()
        C_all = quad(CONSTRUCT_M_(s, one, w_infty))
        D_mat = quad(CONSTRUCT_M_INFTY_(s, one, 1, w_infty, False))
        C_p = quad(CONSTRUCT_M_(s, p_only, w_infty))
        C_k = quad(CONSTRUCT_M_(s, k_only, w_infty))
        C_S = quad(CONSTRUCT_M_(s, s_only, w_infty))
    else:
        R = CONSTRUCT_R(s)
        quad = lambda M: (R.T @ M @ R).tocsc()
        C_all = quad(CONSTRUCT_M(s, one))
        D_mat = quad(CONSTRUCT_M_HAT(s, one))
        C_p = quad(CONSTRUCT_M(s, p_only))
        C_k = quad(CONSTRUCT_M(s, k_only))
        C_S = quad(CONSTRUCT_M(s, s_only))
    cplx = freestream_disturbances or np.iscomplexobj(v) or np.iscomplexobj(getattr(A_, "data", 0))
    devA = upload(sp.csc_matrix(A_, dtype=np.complex128 if cplx else np.float64), s, scale=dt)
    try:
        t, E_all, E_p, E_k, E_S = LINEAR_INTEGRATION_AND_ENERGY_GPU_NO_SAMPLES(
            np.asarray(v, dtype=np.complex128 if cplx else np.float64), n_t, dt, 0.0, orderTaylor,
            devA, C_all, C_p, C_k, C_S)
    finally:
        devA.close()
    v_real = np.real(np.asarray(v))
    for E, C in ((E_all, C_all), (E_p, C_p), (E_k, C_k), (E_S, C_S)):   # :169-174
        E[0] = float(v_real @ (C @ v_real))
    if freestream_disturbances:                                           # :180-183
        vv = np.asarray(v)
        dE_dt_inflow_time = float(np.real(np.vdot(vv, D_mat @ vv)))
        scaling_E_ref = dE_dt_inflow_time * GET_T_REF(s)
        with np.errstate(divide="ignore", invalid="ignore"):
            scaling_E = dE_dt_inflow_time * t
    else:                                                                 # :184-186
        scaling_E_ref = E_all[0]
        scaling_E = E_all[0]
    series = {"all": E_all, "acoustic": E_p, "kinetic": E_k, "entropic": E_S}
    start = 9                                          # start_index = 10 (1-based), :342
    sc_t = scaling_E[start:] if np.ndim(scaling_E) else scaling_E
    max_gain = {"temporal": {k: float(np.max(np.real(E[start:]) / sc_t)) for k, E in series.items()},
                "non_temporal": {k: float(np.max(np.real(E) / scaling_E_ref))
                                 for k, E in series.items()}}
    s.setdefault("timings", {})["linear_integration_steps"] = n_t
    return max_gain


def _energy_run(x, n_t, dt, t_0, order, Adt, mats, weights, n_skip, n_samples):
    own = weights is None
    if own:
        nc = Adt.ncell
        # the grid only matters through ncell for block-diagonal weights
        weights = DeviceWeights(list(mats), nchi=nc, neta=1, vars_per_cell=Adt.b)
    try:
        E, xf, samp = energy_track(Adt, weights, x, n_t=int(n_t), order=int(order), dt=1.0,
                                   form="powers", n_skip=int(n_skip), n_samples=int(n_samples))
    finally:
        if own:
            weights.close()
    E[:, 0] = 0.0                     # the reference leaves entry 1 to its caller
    t = t_0 + dt * np.arange(int(n_t) + 1)
    return t, [E[k] for k in range(E.shape[0])], samp, xf


# ====================================================================== device generalized solver

def block_cholesky_inverse(D):
    """For a symmetric positive definite D that is block diagonal up to a
    permutation (the Chu matrices R'M_hat R and R_'M_infty_R_: 4x4 per cell
    or freestream station, 1x1 per shock DOF), return (Linv, LinvH) with
    D = L L^H, as sparse matrices with the same block structure."""
    from scipy.sparse.csgraph import connected_components
    D = sp.csr_matrix(D)
    n = D.shape[0]
    ncomp, labels = connected_components(D, directed=False)
    order = np.argsort(labels, kind="stable")
    counts = np.bincount(labels, minlength=ncomp)
    starts = np.zeros(ncomp + 1, dtype=np.int64)
    np.cumsum(counts, out=starts[1:])
    pos = np.empty(n, dtype=np.int64)
    pos[order] = np.arange(n) - starts[labels[order]]
    coo = D.tocoo()
    rows, cols, vals = [], [], []
    for size in np.unique(counts):
        comps = np.nonzero(counts == size)[0]
        slot = np.full(ncomp, -1, dtype=np.int64)
        slot[comps] = np.arange(comps.shape[0])
        blk = np.zeros((comps.shape[0], size, size), dtype=D.dtype)
        sel = slot[labels[coo.row]] >= 0
        blk[slot[labels[coo.row[sel]]], pos[coo.row[sel]], pos[coo.col[sel]]] = coo.data[sel]
        Li = np.linalg.inv(np.linalg.cholesky(blk))
        members = order[starts[comps][:, None] + np.arange(size)[None, :]]   # (count, size)
        rows.append(np.repeat(members, size, axis=1).reshape(-1))
        cols.append(np.tile(members, (1, size)).reshape(-1))
        vals.append(Li.reshape(-1))
    Linv = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))
    return Linv, Linv.conj().T.tocsc()


def _device_generalized(devA, devC, D, s, n, k, n_t, n_set0, cplx, tol, maxiter, ncv, v0,
                        block=1, info=None):
    """Largest eigenpairs of T(A')^n C T(A)^n x = lambda D x (with the
    freestream zeroing) through the exact block-Cholesky reduction
    L^-1 (.) L^-H and Krylov-Schur on a device basis.  Returns eigenvectors
    x = L^-H y, D-orthonormal as ARPACK's generalized mode returns them.

    block > 1: block Krylov-Schur whose every expansion applies the reduced
    operator to `block` basis vectors in one batched sandwich
    (hymor_chu_sandwich_batched_dev, the factors applied inside); v0, if
    given, seeds the first column of the start block."""
    t0 = time.perf_counter()
    Linv, LinvH = block_cholesky_inverse(D)
    dLi = upload(Linv, s)
    dLiH = upload(LinvH, s)
    t_reduce = time.perf_counter() - t0
    dtype = np.complex128 if cplx else np.float64
    t1, t2 = DeviceVector(n, dtype), DeviceVector(n, dtype)
    try:
        if v0 is not None:
            v0 = np.asarray(v0, dtype=dtype)
        if block > 1:
            def apply_block(d_in, ld, nvec, d_out):
                chu_sandwich_batched_ptr(devA, devC, d_in, ld, nvec, d_out, n, order=ORDER,
                                         nsub=n_t, n_set0=n_set0, pre=dLiH, post=dLi,
                                         complex_=cplx)
                devA.synchronize()
            m = ncv or max(2 * k + 4 * block, 8 * block)
            vals, Y, kinfo = block_krylov_schur(apply_block, n, k, m, block, complex_=cplx,
                                                tol=tol, maxiter=maxiter,
                                                V0=None if v0 is None else v0[:, None])
        else:
            def apply(din, dout):
                spmv_dev(dLiH, din, t1.ptr, complex_=cplx)
                dLiH.synchronize()
                chu_sandwich_ptr(devA, devC, t1.ptr, t2.ptr, order=ORDER, nsub=n_t,
                                 n_set0=n_set0, complex_=cplx)
                devA.synchronize()
                spmv_dev(dLi, t2.ptr, dout, complex_=cplx)
                dLi.synchronize()
            vals, Y, kinfo = krylov_schur(apply, n, k, ncv or max(20, 2 * k + 1), complex_=cplx,
                                          tol=tol, maxiter=maxiter, v0=v0)
        if info is not None:
            info.update(kinfo, block=block, reduce_seconds=t_reduce,
                        solve_seconds=time.perf_counter() - t0 - t_reduce)
        X = LinvH @ Y
        if not cplx:
            X = np.real(X)
        return vals, X
    finally:
        t1.free()
        t2.free()
        dLi.close()
        dLiH.close()


# ====================================================================== batched eigen solve

def batched_exp_eigs(A, s, n_modes, *, T_ref=0.5, block=4, ncv=None, tol=1e-10,
                     tol_f32=1e-7, maxiter=200, order=ORDER, value_prec="f32", refine=True,
                     refine_block=16, refine_ncv=32, autotune=False, seed=0, form="horner",
                     timings=None):
    """Leading eigenvalues of the exponential transform with batched actions.

    Phase 1: block Krylov-Schur whose every expansion applies T_order(A dt)^n
    to `block` basis vectors in one SpMM sweep (block Gram-Schmidt on the
    device), block values stored in `value_prec` (FP32 by default) and
    accumulated in FP64, converged to tol_f32 on the stored operator.
    Phase 2 (refine, FP32 storage only): block Krylov-Schur on FP64-stored
    values, `refine_block` wide, seeded with phase 1's orthonormal Schur
    basis of its `refine_block` leading Ritz values, to tol.  FP32 storage
    moves the eigenvalues by about eps_32 * |A| (relative 3e-8 at |lambda| =
    0.05 on a unit-norm operator); from that seed a few FP64 sweeps remove
    it.  Both phases share one FP64 upload; phase 1 reads an FP32 copy cast
    on the device (DeviceOperator.set_value_prec).  Eigenvalues are mapped
    as EIGENVALUES_GPU_ARNOLDI does: lambda = log(mu) / (n dt).
    A is a scipy matrix in the reference's var-major order or a
    BlockOperator.  autotune: time the batch kernels at each block width
    first (else the default tile kernel); `seed` selects the random start
    block; `form` is the Taylor recursion ("horner": the same polynomial as
    POWER_METHOD!'s powers form, with three vector passes per term instead
    of four).  Returns (V, D, info); info holds
    per-phase iterations, operator applications and seconds.
    """
    s = CFL_TIMESTEP(s)
    dt = s["time_integration"]["dt"]
    timesteps = int(round(T_ref / dt))
    n = A.n if hasattr(A, "brow_ptr") else A.shape[0]
    k = n_modes
    m = ncv or max(2 * k + 4 * block, 8 * block)
    info = {}

    def make_apply(dev):
        def apply(d_in, ld, nvec, d_out):
            expm_taylor_batched_ptr(dev, d_in, ld, nvec, d_out, n, order=order, nsub=timesteps,
                                    form=form, complex_=False)
            dev.synchronize()
        return apply

    def prepare(dev, width):
        if autotune and width in (8, 16, 32, 64):
            dev.autotune_batch(nv=width, reps=3)
        dev.set_graphs(True)

    t0 = time.perf_counter()
    two_phase = refine and value_prec == "f32"
    # one FP64 upload serves both phases: FP32 values are cast on the device
    dev = _device_op(A, s, dt, "f64" if two_phase else value_prec)
    try:
        info["upload_seconds"] = time.perf_counter() - t0
        if two_phase:
            dev.set_value_prec("f32")
        prepare(dev, block)
        vals, Y, i1 = block_krylov_schur(make_apply(dev), n, k, m, block,
                                         tol=tol_f32 if two_phase else tol, maxiter=maxiter,
                                         basis=refine_block if two_phase else 0,
                                         return_vectors=not two_phase, seed=seed)
        seed = i1.pop("basis", None)
        info["phase1"] = dict(i1, seconds=time.perf_counter() - t0, value_prec=value_prec,
                              block=block)
        if two_phase:
            t1 = time.perf_counter()
            dev.set_value_prec("f64")
            prepare(dev, refine_block)
            try:
                # room for at least two blocks past the wanted set between
                # restarts (with one, every restart follows one expansion)
                m2 = (refine_ncv if k + refine_block <= refine_ncv
                      else k + 2 * refine_block)
                vals, Y, i2 = block_krylov_schur(make_apply(dev), n, k, m2, refine_block,
                                                 tol=tol, maxiter=maxiter, V0=seed)
            finally:
                seed.free()
            info["phase2"] = dict(i2, seconds=time.perf_counter() - t1, value_prec="f64",
                                  block=refine_block)
    finally:
        dev.close()
    info["seconds"] = time.perf_counter() - t0
    order_ = np.argsort(-np.abs(vals), kind="stable")
    vals, Y = vals[order_], Y[:, order_]
    D = np.log(vals.astype(np.complex128)) / dt / timesteps
    if timings is not None:
        timings.update(info)
    return Y, D, info
