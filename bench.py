"""Benchmark: exp(A*T)v actions/s and Taylor-term-kernel HBM bandwidth.

Workload (default): the N ~ 1.44M-unknown case the metric's targets are
quoted at -- a 600 x 300 cell grid with 8 conserved variables per cell
(8x8 blocks on the 3x3 stencil, 1 614 604 blocks, 103.3M nonzeros, FP64),
synthetic values (hymor_amd.operators, seed 1005), start vector N(0,1).
One step = one action w = T_12(A*T/64)^64 v with T = 1, i.e. 64 substeps of
a degree-12 Taylor polynomial = 768 fused SpMV+update kernels, with v and w
resident in HBM.  `value` is that plain FP64 form.  The same run also times
the batched form (16 vectors per sweep, block values stored in FP32,
accumulation in FP64) and reports it under "batched", with its error against
the plain FP64 action on the same vector.

Prints one JSON line (see the fields at the end).  --gpus N > 1 is not
supported: this engine runs on one GPU.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np
import scipy.sparse as sp

METRIC = "exp(AT)v actions/s + SpMV HBM GB/s (% peak), batched vs plain FP64 at 1e-8"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(op, order, nsub, serial_samples=3, omp_samples=5):
    """Time the oracle's CPU forms on a bounded sample of the same workload
    (single Taylor substeps, each timed on its own; the median is scaled to
    actions/s, so a noisy host moves the figure less)."""
    from oracle import oracle

    x = np.random.default_rng(5).standard_normal(op.n)
    log("cpu baseline: building the reference-layout CSC (var-major, 1-based)")
    A = oracle.CSC.from_arrays(*op.to_csc_varmajor_blocks())
    A = A.scaled(1.0 / nsub)
    log(f"cpu baseline: serial POWER_METHOD! restatement, {serial_samples} substeps")
    t_serial = []
    for _ in range(serial_samples):
        t0 = time.perf_counter()
        oracle.power_method(A, x, 1, order=order)
        t_serial.append(time.perf_counter() - t0)
    del A
    log(f"cpu baseline: row-parallel OpenMP CSR form, {omp_samples} substeps")
    csr = oracle.CsrArrays(op.to_csr_cellmajor() * (1.0 / nsub))
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    oracle.taylor_csr_omp(csr, x, 1, order, nthreads)      # spin up the thread pool
    t_omp = []
    for _ in range(omp_samples):
        t0 = time.perf_counter()
        _, used = oracle.taylor_csr_omp(csr, x, 1, order, nthreads)
        t_omp.append(time.perf_counter() - t0)
    med_omp, med_serial = float(np.median(t_omp)), float(np.median(t_serial))
    return {
        "value": 1.0 / (med_omp * nsub),
        "unit": "exp(AT)v actions/s",
        "cores": int(used),
        "kind": "port",
        "sample": (f"median of {omp_samples} single substeps of the same action ({order} SpMV "
                   f"+ updates each, after one warm-up substep), scaled x{nsub}; OpenMP "
                   f"row-parallel CSR over {used} threads"),
        "substep_seconds_each": t_omp,
        "serial_faithful_value": 1.0 / (med_serial * nsub),
        "serial_faithful_cores": 1,
        "serial_faithful_sample": (f"median of {serial_samples} single substeps, 1-based "
                                   "var-major CSC column-scatter SpMV, restating "
                                   "ARNOLDI_POWER_GPU.jl:64-82 with order 12"),
        "cpu_model": _cpu_model(),
    }


def eigen_leg(k=10, tol=1e-9, block=4, seeds=(0, 1, 2)):
    """Leading eigenvalues of the exponential transform (T_ref = 0.5,
    dt = 1/64, order 5, as EIGENVALUES_GPU_ARNOLDI) at the benchmark's size:
    a 600 x 300-cell, 8-variable operator with the 3x3 stencil's 9-block
    pattern (N = 1.44M, the same storage and kernel work as the main
    workload) whose coupling is one-way (block upper triangular), so its
    spectrum is known exactly (hymor_amd.operators.marching_operator).
    Plain FP64 device Krylov-Schur (single vector per application) versus
    the batched forms: block Krylov-Schur, `block` vectors per SpMM sweep,
    on FP64-stored values, and on FP32-stored values followed by a seeded
    FP64 refinement; every solve converges to the same residual tolerance
    `tol`.  Errors are relative to the exact eigenvalues."""
    import math

    from hymor_amd import operators as ops
    from hymor_amd import stability as st

    nchi, neta, b = 600, 300, 8
    op, lam = ops.marching_operator(nchi, neta, b, seed=5)
    dt = 1.0 / 64
    n_t = int(round(0.5 / dt))
    mu = sum((lam * dt) ** j / math.factorial(j) for j in range(6)) ** n_t
    exact = np.log(mu[np.argsort(-np.abs(mu), kind="stable")][: 3 * k]) / (n_t * dt)

    def err(D):
        return float(max(np.min(np.abs(exact - d)) / abs(d) for d in D))

    def s_dict():
        return {"mesh": {"Nchi": nchi, "Neta": neta}, "time_integration": {"dt": dt},
                "stability_analysis": {"perturb_shock": False,
                                       "eigenvalue_solver": "GPU_TIMESTEPPER_ARNOLDI"}}

    out = {"workload": ("10 leading eigenvalues of T_5(A/64)^32 (T_ref 0.5) of a one-way "
                        "(block upper triangular) 600x300-cell, b=8 operator with the 3x3 "
                        "stencil pattern, N=1440000; exact spectrum from its diagonal blocks"),
           "start_seeds": list(seeds), "tol": tol,
           "note": "seconds: median over the random starts, each solve from upload to "
                   "Ritz vectors; errors: worst over the starts"}

    def summarize(runs):
        secs = sorted(r["seconds"] for r in runs)
        return {"seconds": secs[len(secs) // 2], "seconds_each": [r["seconds"] for r in runs],
                "max_rel_eig_err": max(r["err"] for r in runs),
                "applications_each": [r["apps"] for r in runs]}

    plain = []
    for sd_ in seeds:
        s = s_dict()
        t0 = time.perf_counter()
        _, D = st.EIGENVALUES_GPU_ARNOLDI(op, s, k, krylov="device", tol=tol, ncv=25,
                                          maxiter=200, seed=sd_, form=FORM)
        plain.append({"seconds": time.perf_counter() - t0, "err": err(D),
                      "apps": s["timings"]["eigenvalues_krylov"]["applications"]})
    out["plain_fp64"] = dict(summarize(plain), krylov_dim=25)
    tp = out["plain_fp64"]["seconds"]
    log(f"eigen leg: plain {out['plain_fp64']}")
    for prec, refine in (("f64", True), ("f32", True), ("f32", False)):
        runs = []
        for sd_ in seeds:
            _, Db, info = st.batched_exp_eigs(op, s_dict(), k, block=block, value_prec=prec,
                                              tol=tol, refine=refine, seed=sd_, form=FORM)
            r = {"seconds": info["seconds"], "err": err(Db),
                 "apps": info["phase1"]["applications"]}
            if "phase2" in info:
                r["apps"] += info["phase2"]["applications"]
                r["fp32_phase_seconds"] = info["phase1"]["seconds"]
                r["refinement_seconds"] = info["phase2"]["seconds"]
                r["refinement_fp64_applications"] = info["phase2"]["applications"]
            runs.append(r)
        row = dict(summarize(runs), block=block)
        row["speedup_vs_plain_fp64"] = tp / row["seconds"]
        if "refinement_seconds" in runs[0]:
            row["fp32_phase_seconds_each"] = [r["fp32_phase_seconds"] for r in runs]
            row["refinement_seconds_each"] = [r["refinement_seconds"] for r in runs]
            row["refinement_fp64_applications_each"] = [r["refinement_fp64_applications"]
                                                        for r in runs]
        key = f"batched_{prec}_values" + ("" if refine else "_unrefined")
        out[key] = row
        log(f"eigen leg: {key} {row}")
    return out


FP64_MFMA_PEAK_TFS = 78.6


def transient_growth_leg(T=1.0, k=4, tol=1e-8, steps=5):
    """Optimal transient growth at the 400 x 300-cell, b = 4 size with 400
    shock DOFs (N = 480 400; Chu weights from a synthetic base flow of that
    grid, dt = 1/64): TRANSIENT_GROWTH_DOWNSTREAM for horizon T with the
    device Krylov solver (k gains; block Krylov-Schur on blocks of 2, and
    the single-vector solver for comparison, each split into upload,
    Cholesky reduction and Krylov seconds), and the Chu sandwich it applies
    (forward Taylor sweep, weight, adjoint sweep; T_5, n_t = T/dt
    substeps each way) timed alone, on 1 and on 2, 4 and 8 columns, with
    its weight stage timed as the
    sandwich at n_t = 0 (permutations and the block-diagonal weight); and
    the gain's gradient with respect to the operator's entries
    (GainFunctional.operator_gradient: one forward action, the weight, and
    the checkpointed pullback, every 8 substeps) on the operator's own
    pattern, host vectors in and out."""
    from hymor_amd import operators as ops
    from hymor_amd import stability as st
    from hymor_amd.baseflow import synthetic_solution
    from hymor_amd.core import (DeviceMatrix, DeviceOperator, DeviceVector,
                                chu_sandwich_batched_ptr, chu_sandwich_ptr)

    nchi, neta = 400, 300
    s = st.CFL_TIMESTEP(synthetic_solution(nchi, neta, seed=7, perturb_shock=True, cfl=100.0,
                                           max_dt=1.0 / 64))
    op = ops.config_operator("C3")
    A = op.to_scipy_varmajor()
    R = st.CONSTRUCT_R(s)
    C = (R.T @ st.CONSTRUCT_M(s, [1, 1, 1, 1]) @ R).tocsc()
    dt = s["time_integration"]["dt"]
    n_t = int(round(T / dt))
    out = {"workload": (f"400x300 cells, b=4, 400 shock DOFs, N={op.n}; Chu weights of a "
                        f"synthetic base flow; T={T}, dt={dt}, n_t={n_t}, order 5"),
           "k": k, "tol": tol}
    def solve(block):
        t0 = time.perf_counter()
        _, G, _ = st.TRANSIENT_GROWTH_DOWNSTREAM(A, s, k, T, tol=tol, krylov="device", ncv=24,
                                                 maxiter=400, block=block)
        secs = time.perf_counter() - t0
        tk = s["timings"]["transient_growth_krylov"]
        return G, {"seconds": secs, "upload_seconds": s["timings"]["transient_growth_upload"],
                   "reduce_seconds": tk["reduce_seconds"], "krylov_seconds": tk["solve_seconds"],
                   "applications": tk.get("applications"), "block": block}
    G, out["solve"] = solve(2)
    G1, out["solve_single_vector"] = solve(1)
    out["solve_seconds"] = out["solve"]["seconds"]
    out["gains"] = [float(g) for g in G]
    out["solve"]["max_rel_diff_vs_single_vector"] = float(np.max(np.abs(G - G1) / np.abs(G1)))
    x = np.random.default_rng(8).standard_normal(op.n)
    with DeviceOperator.from_blocks(op, scale=dt) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC:
        dx, dy = DeviceVector.from_host(x), DeviceVector(op.n)
        try:
            def timed(nsub):
                chu_sandwich_ptr(dA, dC, dx.ptr, dy.ptr, order=5, nsub=nsub)
                dA.synchronize()
                t1 = time.perf_counter()
                for _ in range(steps):
                    chu_sandwich_ptr(dA, dC, dx.ptr, dy.ptr, order=5, nsub=nsub)
                dA.synchronize()
                return (time.perf_counter() - t1) / steps
            ts, tw = timed(n_t), timed(0)
        finally:
            dx.free()
            dy.free()
        # the sandwich on blocks of columns (hymor_chu_sandwich_batched_dev)
        batched = {}
        for nrhs in (2, 4, 8):
            X = DeviceMatrix.from_host(
                np.asfortranarray(np.random.default_rng(nrhs).standard_normal((op.n, nrhs))))
            Y = DeviceMatrix(op.n, nrhs)
            try:
                chu_sandwich_batched_ptr(dA, dC, X.ptr, op.n, nrhs, Y.ptr, op.n, order=5,
                                         nsub=n_t)
                dA.synchronize()
                t1 = time.perf_counter()
                for _ in range(steps):
                    chu_sandwich_batched_ptr(dA, dC, X.ptr, op.n, nrhs, Y.ptr, op.n, order=5,
                                             nsub=n_t)
                dA.synchronize()
                tb = (time.perf_counter() - t1) / steps
            finally:
                X.free()
                Y.free()
            batched[str(nrhs)] = {"ms": tb * 1e3, "ms_per_vector": tb * 1e3 / nrhs}
    out["sandwich_ms"] = ts * 1e3
    out["sandwich_weight_stage_ms"] = tw * 1e3
    out["sandwich_terms"] = 2 * 5 * n_t
    out["sandwich_note"] = "device-resident vectors (hymor_chu_sandwich_dev)"
    out["sandwich_batched"] = batched
    log(f"transient growth: solve {out['solve_seconds']:.2f} s (blocks of 2; single vector "
        f"{out['solve_single_vector']['seconds']:.2f} s), gains {out['gains']}, "
        f"sandwich {ts * 1e3:.2f} ms, weight stage {tw * 1e3:.3f} ms, batched per vector "
        + ", ".join(f"{k_}: {v['ms_per_vector']:.2f} ms" for k_, v in batched.items()))
    # the gain's sensitivity to the operator's entries
    from hymor_amd.adjoint import GainFunctional
    D = (R.T @ st.CONSTRUCT_M_HAT(s, [1, 1, 1, 1]) @ R).tocsc()
    As = sp.csc_matrix(A * dt)
    As.sort_indices()
    colptr, rowval = As.indptr.astype(np.int64), As.indices.astype(np.int64)
    ckpt = 8
    from hymor_amd.core import DevicePattern
    with DeviceOperator.from_scipy(As, nchi=nchi, neta=neta) as dA, \
            DeviceOperator.from_scipy(C, nchi=nchi, neta=neta) as dC:
        F = GainFunctional(dA, dC, D, n_t)
        F.operator_gradient(x, colptr, rowval, index_base=0, ckpt_every=ckpt)
        t1 = time.perf_counter()
        G, g = F.operator_gradient(x, colptr, rowval, index_base=0, ckpt_every=ckpt)
        tg = time.perf_counter() - t1
        # the same on a pattern kept on the device, gradient left there
        with DevicePattern(dA, colptr, rowval, index_base=0) as pat:
            dg = DeviceVector(pat.nnz)
            try:
                F.operator_gradient_dev(x, pat, dg, ckpt_every=ckpt)
                t1 = time.perf_counter()
                F.operator_gradient_dev(x, pat, dg, ckpt_every=ckpt)
                tgd = time.perf_counter() - t1
                same = bool(np.array_equal(dg.download(), g))
                # every state kept (ckpt_every 1): the backward sweep
                # recomputes none
                F.operator_gradient_dev(x, pat, dg, ckpt_every=1)
                t1 = time.perf_counter()
                F.operator_gradient_dev(x, pat, dg, ckpt_every=1)
                tga = time.perf_counter() - t1
                ga = dg.download()
                rel_all = float(np.linalg.norm(ga - g) / np.linalg.norm(g))
            finally:
                dg.free()
    out["operator_gradient"] = {
        "ms": tg * 1e3, "entries": int(As.nnz), "ckpt_every": ckpt,
        "device_vectors": int(-(-n_t // ckpt) + (2 * 5 + 1) * ckpt + 5),
        "note": ("dG/dA on the operator's CSC pattern: forward action, Chu weight, checkpointed "
                 "pullback (hymor_expm_taylor_vjp_values); host vectors and gradient"),
        "device_pattern_ms": tgd * 1e3, "device_pattern_same_bits": same,
        "device_pattern_note": ("the pattern mapped and kept on the device once "
                                "(hymor_pattern_create), gradient left on the device "
                                "(hymor_expm_taylor_vjp_values_dev)"),
        "all_states_ms": tga * 1e3, "all_states_relerr": rel_all,
        "all_states_vectors": int(n_t + (2 * 5 + 1) * 16 + 5),
        "all_states_note": ("device pattern with every state kept (ckpt_every 1): nothing "
                            "recomputed, the forward powers batched over 16 states")}
    log(f"transient growth: operator gradient on {As.nnz} entries {tg * 1e3:.1f} ms, "
        f"{tgd * 1e3:.1f} ms on a device pattern (checkpoint every {ckpt} substeps), "
        f"{tga * 1e3:.1f} ms keeping every state")
    return out


def shock_config_leg(steps=5, order=12, nsub=64, T=1.0, k=20, ncv=80, cpu=True):
    """The 200 x 300-cell, b = 4 operator with 200 shock DOFs (N = 240 200,
    the shock bands of the reference's shock fitting):
      * the plain FP64 action T_12(A T/64)^64 on the GPU as the drivers run
        it (one HIP graph replay per action) and on the host (the oracle's
        OpenMP CSR form: one substep, scaled x64);
      * EIGENVALUES_GPU_ARNOLDI for 20 modes with Krylov dimension 80
        (T_ref = 0.5, dt = 1/64, order 5) through the device Krylov-Schur
        (residual tolerance 1e-9) and through ARPACK at the reference's
        tolerance 1e-4 (EIGENVALUES.jl:159), both applying the GPU action,
        on an operator of that shape and storage whose spectrum is exact
        (one-way coupling; its diagonal blocks' eigenvalues), with the
        worst relative eigenvalue error of each."""
    from hymor_amd import operators as ops
    from hymor_amd.core import (DeviceMatrix, DeviceOperator, DeviceVector,
                                expm_taylor_batched_ptr, expm_taylor_dev)

    nchi, neta = 200, 300
    op = ops.config_operator("C2")
    x = np.random.default_rng(9).standard_normal(op.n)
    out = {"workload": (f"200x300 cells, b=4, 200 shock DOFs, N={op.n}, "
                        f"nnz={op.nnzb * 16 + op.xrow.shape[0]}; action T_{order}(A T/{nsub})^"
                        f"{nsub} v, T={T}, {FORM} form, FP64"), "unit": "exp(AT)v actions/s"}
    with DeviceOperator.from_blocks(op, with_adjoint=False) as dev:
        dev.set_graphs(True)
        dx, dy = DeviceVector.from_host(x), DeviceVector(op.n)
        try:
            run = lambda: expm_taylor_dev(dev, dx, dy, order=order, nsub=nsub, dt=T / nsub,
                                          form=FORM)
            run()
            dev.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                run()
            dev.synchronize()
            ta = (time.perf_counter() - t0) / steps
            k_ms, k_bytes = dev.bench_term_kernel(reps=20, form=FORM)
            # 16 vectors per sweep (FP64 values): the shock rows and the
            # shock-column slices in the batch kernel's leading workgroups
            nb = 16
            X = DeviceMatrix.from_host(
                np.asfortranarray(np.random.default_rng(10).standard_normal((op.n, nb))))
            Y = DeviceMatrix(op.n, nb)
            try:
                runb = lambda: expm_taylor_batched_ptr(dev, X.ptr, op.n, nb, Y.ptr, op.n,
                                                       order=order, nsub=nsub, dt=T / nsub,
                                                       form=FORM)
                runb()
                dev.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    runb()
                dev.synchronize()
                tb = (time.perf_counter() - t0) / steps
            finally:
                X.free()
                Y.free()
        finally:
            dx.free()
            dy.free()
    out["gpu"] = {"value": 1.0 / ta, "ms_per_action": ta * 1e3,
                  "term_ms": k_ms, "term_GBps": k_bytes / (k_ms * 1e-3) / 1e9,
                  "note": "term: block kernel + the shock bands' extra rows (HIP events)"}
    out["gpu_batched16"] = {"value": nb / tb, "ms_per_sweep": tb * 1e3,
                            "vs_single": (nb / tb) * ta,
                            "note": "16 vectors per sweep, FP64 values"}
    if cpu:
        out["cpu"] = _cpu_action_sample(op, x, order, nsub, T)
        out["gpu_vs_cpu"] = out["gpu"]["value"] / out["cpu"]["value"]
    out["eigen"] = _eigen_timings(nchi, neta, k, ncv, seed=12, n_lead=24, bulk=(-6.0, -3.0),
                                  n_shock=nchi, tag="shock config")
    log(f"shock config: GPU {out['gpu']['value']:.1f} actions/s, "
        f"{out['gpu_batched16']['value']:.1f} in 16-vector sweeps"
        + (f", CPU {out['cpu']['value']:.3f} ({out['gpu_vs_cpu']:.0f}x)" if cpu else ""))
    return out


def _cpu_action_sample(op, x, order, nsub, T, samples=3):
    """The oracle's OpenMP CSR form of the action on the host: the median of
    `samples` single substeps, scaled to actions/s."""
    from oracle import oracle
    csr = oracle.CsrArrays(op.to_csr_cellmajor() * (T / nsub))
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    oracle.taylor_csr_omp(csr, x, 1, order, nthreads)
    t_omp = []
    for _ in range(samples):
        t0 = time.perf_counter()
        _, used = oracle.taylor_csr_omp(csr, x, 1, order, nthreads)
        t_omp.append(time.perf_counter() - t0)
    return {"value": 1.0 / (float(np.median(t_omp)) * nsub), "cores": int(used), "kind": "port",
            "sample": f"median of {samples} single substeps of the OpenMP CSR form, x{nsub}"}


def _eigen_timings(nchi, neta, k, ncv, *, seed, n_lead, bulk, n_shock, tag):
    """EIGENVALUES_GPU_ARNOLDI (T_ref = 0.5, dt = 1/64, order 5) for k modes
    with Krylov dimension ncv through the device Krylov-Schur (tolerance
    1e-9) and through ARPACK at the reference's tolerance 1e-4
    (EIGENVALUES.jl:159), both applying the GPU action, and batched_exp_eigs
    (block Krylov-Schur on 4-vector sweeps, FP64 values, tolerance 1e-9), on
    a b = 4 operator of the given shape with an exactly known spectrum
    (one-way coupling: its diagonal blocks' eigenvalues); seconds and the
    worst relative eigenvalue error of each."""
    import math
    from hymor_amd import operators as ops
    from hymor_amd import stability as st

    dt, T_ref = 1.0 / 64, 0.5
    n_t = int(round(T_ref / dt))
    eop, lam = ops.marching_operator(nchi, neta, 4, seed=seed, n_lead=n_lead, bulk=bulk,
                                     n_shock=n_shock)
    mu = sum((lam * dt) ** j / math.factorial(j) for j in range(6)) ** n_t
    exact = np.log(mu[np.argsort(-np.abs(mu), kind="stable")][: 3 * k].astype(complex)) / (
        n_t * dt)
    out = {"k": k, "ncv": ncv, "T_ref": T_ref, "dt": dt}
    for name, kw in (("device_krylov_schur", dict(krylov="device", tol=1e-9)),
                     ("arpack", dict(krylov="arpack", tol=1e-4, v0=np.ones(eop.n)))):
        sd = {"mesh": {"Nchi": nchi, "Neta": neta}, "time_integration": {"dt": dt},
              "stability_analysis": {"perturb_shock": False,
                                     "eigenvalue_solver": "GPU_TIMESTEPPER_ARNOLDI"}}
        V, D = st.EIGENVALUES_GPU_ARNOLDI(eop, sd, k, ncv=ncv, **kw)
        err = max(float(np.min(np.abs(exact - d)) / abs(d)) for d in D)
        res = {"seconds": sd["timings"]["eigenvalues"], "tol": kw["tol"],
               "max_rel_eig_err": err}
        if "eigenvalues_krylov" in sd["timings"]:
            res["applications"] = sd["timings"]["eigenvalues_krylov"].get("applications")
        out[name] = res
        log(f"{tag}: EIGENVALUES ({name}) {res}")
    # the same modes by block Krylov-Schur on batched sweeps (blocks of 4,
    # FP64 values: every expansion one 4-column sweep)
    sd = {"mesh": {"Nchi": nchi, "Neta": neta}, "time_integration": {"dt": dt},
          "stability_analysis": {"perturb_shock": False,
                                 "eigenvalue_solver": "GPU_TIMESTEPPER_ARNOLDI"}}
    _, D, info = st.batched_exp_eigs(eop, sd, k, block=4, value_prec="f64", tol=1e-9,
                                     maxiter=300)
    err = max(float(np.min(np.abs(exact - d)) / abs(d)) for d in D)
    res = {"seconds": info["seconds"], "tol": 1e-9, "max_rel_eig_err": err,
           "applications": info["phase1"].get("applications"), "block": 4}
    out["batched_block4_fp64"] = res
    log(f"{tag}: batched_exp_eigs (blocks of 4, FP64) {res}")
    return out


def small_config_leg(steps=20, order=12, nsub=64, T=1.0, cpu=True):
    """The launch-bound end: the same action T_12(A T/64)^64 v (768 Taylor
    terms) on the 100 x 80-cell, b = 4 operator (N = 32 000, 1.13M
    nonzeros, 9 MB of values: cache-resident), FP64, vectors resident on
    the device, in actions/s: kernel by kernel, replayed from a captured
    HIP graph, and through the no-alloc entry point (kernel by kernel on
    the library's static scratch); plus the term kernel's own event-timed
    duration, so that 768 x that duration against the action's time shows
    the launch cost; the host's OpenMP CSR form of the same action (the
    size the reference runs on the CPU); and EIGENVALUES_GPU_ARNOLDI for 10
    modes with Krylov dimension 60 on an operator of that shape with an
    exact spectrum."""
    import ctypes

    from hymor_amd import _lib
    from hymor_amd import operators as ops
    from hymor_amd.core import _FORMS, DeviceOperator, DeviceVector, expm_taylor_dev

    op = ops.config_operator("C1")
    x = np.random.default_rng(0).standard_normal(op.n)
    out = {"workload": ("100x80 cells, b=4 (4x4 blocks, 3x3 stencil), N=32000, nnz=1134784; "
                        f"action T_{order}(A T/{nsub})^{nsub} v, T={T}, {FORM} form, FP64"),
           "unit": "exp(AT)v actions/s", "steps": steps, "kernels_per_action": order * nsub}
    lib = _lib.load()
    with DeviceOperator.from_blocks(op, with_adjoint=False) as dev:
        # at this size the launch variants differ more than at N = 1.44M:
        # every form below runs the autotuner's pick
        tune_ms, pick = dev.autotune(reps=50)
        out["launch"] = {"split": pick[0], "waves": pick[1], "nontemporal": pick[2],
                         "autotune_ms_powers_term": tune_ms}
        dx, dy = DeviceVector.from_host(x), DeviceVector(op.n)
        try:
            def run_eager():
                expm_taylor_dev(dev, dx, dy, order=order, nsub=nsub, dt=T / nsub, form=FORM)

            def run_noalloc():
                _lib.check(lib.hymor_expm_taylor_dev_noalloc(
                    dev.handle, 0, order, nsub, T / nsub, _FORMS[FORM], dx.ptr, dy.ptr, _lib.F64))

            def timed(run):
                run()
                dev.synchronize()
                t0 = time.perf_counter()
                for _ in range(steps):
                    run()
                dev.synchronize()
                return (time.perf_counter() - t0) / steps

            dev.set_graphs(False)
            t_eager = timed(run_eager)
            y_eager = dy.download()
            dev.set_graphs(True)
            t_graph = timed(run_eager)
            y_graph = dy.download()
            dev.set_graphs(False)
            t_na = timed(run_noalloc)
            y_na = dy.download()
            k_ms, k_bytes = dev.bench_term_kernel(reps=200, form=FORM)
        finally:
            dx.free()
            dy.free()
    kern_s = order * nsub * k_ms * 1e-3
    for key, t, y in (("eager", t_eager, y_eager), ("graph", t_graph, y_graph),
                      ("noalloc", t_na, y_na)):
        out[key] = {"value": 1.0 / t, "ms_per_action": t * 1e3,
                    "kernel_time_fraction": kern_s / t,
                    "bitwise_equal_to_eager": bool(np.array_equal(y, y_eager))}
    out["term_kernel_us"] = k_ms * 1e3
    out["term_kernel_GBps"] = k_bytes / (k_ms * 1e-3) / 1e9
    out["term_kernel_algorithmic_bytes"] = k_bytes
    log(f"small config: eager {1 / t_eager:.1f}, graph {1 / t_graph:.1f}, noalloc "
        f"{1 / t_na:.1f} actions/s; term kernel {k_ms * 1e3:.2f} us "
        f"({out['term_kernel_GBps']:.0f} GB/s)")
    if cpu:
        out["cpu"] = _cpu_action_sample(op, x, order, nsub, T, samples=5)
        out["graph_vs_cpu"] = out["graph"]["value"] / out["cpu"]["value"]
        log(f"small config: CPU {out['cpu']['value']:.2f} actions/s "
            f"(graph replay {out['graph_vs_cpu']:.0f}x)")
    out["eigen"] = _eigen_timings(100, 80, 10, 60, seed=11, n_lead=12, bulk=(-3.0, -1.0),
                                  n_shock=0, tag="small config")
    return out


def krylov_projection_leg(n=1440000, j=32, widths=(4, 16), reps=20):
    """The dense Krylov projections of the batched eigen leg at its size:
    one block Gram-Schmidt pass of s new vectors against j basis vectors of
    length n (hymor_bench_krylov_projection): V^T W on the FP64 matrix
    cores, the fixed-order reduction of its partials, and W -= V h, also on
    the matrix cores.  Each kernel's time comes from HIP events over `reps`
    launches; GB/s and TF/s are algorithmic bytes and flops over that time,
    matrix-core utilisation the flops the matrix cores execute (products
    always 16 x 16 wide, so 512 n ceil(j / 16) for either kernel) over time
    and the FP64 matrix-core peak."""
    import ctypes

    from hymor_amd import _lib

    from hymor_amd.calibrate import mixed_bandwidth

    lib = _lib.load()
    cases = []
    for s in widths:
        out = (ctypes.c_double * 8)()
        _lib.check(lib.hymor_bench_krylov_projection(n, j, s, reps, out))
        ms_vtw, ms_red, ms_sub, by_vtw, fl_vtw, fl_mfma, by_sub, fl_sub = list(out)
        # streaming ceilings at each kernel's read/write mix (V^T W only
        # reads; W -= V h writes its s columns)
        mix_vtw = mixed_bandwidth(by_vtw, 0, 20)["GBps"]
        mix_sub = mixed_bandwidth(by_sub - s * n * 8.0, s * n * 8.0, 20)["GBps"]
        cases.append({
            "s": s,
            "vtw": {"kernel": "k_vtw_mfma<true, 2>", "ms": ms_vtw,
                    "GBps": by_vtw / (ms_vtw * 1e-3) / 1e9,
                    "frac_of_hbm_peak": by_vtw / (ms_vtw * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "useful_TFps": fl_vtw / (ms_vtw * 1e-3) / 1e12,
                    "mfma_TFps_issued": fl_mfma / (ms_vtw * 1e-3) / 1e12,
                    "mfma_utilisation": fl_mfma / (ms_vtw * 1e-3) / 1e12 / FP64_MFMA_PEAK_TFS,
                    "algorithmic_bytes": by_vtw, "flops": fl_vtw, "mfma_flops_issued": fl_mfma,
                    "same_mix_stream_GBps": mix_vtw,
                    "frac_of_same_mix_stream": by_vtw / (ms_vtw * 1e-3) / 1e9 / mix_vtw},
            "reduce": {"kernel": "k_reduce_vtw", "ms": ms_red},
            "sub": {"kernel": "k_sub_vh_mfma<8>", "ms": ms_sub,
                    "GBps": by_sub / (ms_sub * 1e-3) / 1e9,
                    "frac_of_hbm_peak": by_sub / (ms_sub * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "TFps": fl_sub / (ms_sub * 1e-3) / 1e12,
                    "mfma_utilisation": fl_mfma / (ms_sub * 1e-3) / 1e12 / FP64_MFMA_PEAK_TFS,
                    "algorithmic_bytes": by_sub, "flops": fl_sub, "mfma_flops_issued": fl_mfma,
                    "write_share": s * n * 8.0 / by_sub, "same_mix_stream_GBps": mix_sub,
                    "frac_of_same_mix_stream": by_sub / (ms_sub * 1e-3) / 1e9 / mix_sub},
        })
        log(f"krylov projection s={s}: V^T W {ms_vtw * 1e3:.1f} us "
            f"({cases[-1]['vtw']['GBps']:.0f} GB/s, MFMA {cases[-1]['vtw']['mfma_utilisation']:.1%}), "
            f"reduce {ms_red * 1e3:.1f} us, W -= V h {ms_sub * 1e3:.1f} us "
            f"({cases[-1]['sub']['GBps']:.0f} GB/s, MFMA {cases[-1]['sub']['mfma_utilisation']:.1%})")
    return {"n": n, "basis_columns": j, "fp64_mfma_peak_TFps": FP64_MFMA_PEAK_TFS,
            "bound": "hbm (V^T W at s = 16: 2 s j / ((j + s) 8) = 2.7 flop/B, below the "
                     "FP64 ridge of 9.8)",
            "cases": cases}


def receptivity_leg(steps=3, order=12, nsub=64, T=1.0, n_w=32, widths=(4, 8, 16, 32)):
    """C4: the action T_12(L_ T/64)^64 on the extended receptivity operator
    L_ = [L repeat(B,1,32); 0 kron(diag(-i w), I)] (LINEARIZE_L_.jl:58-67) of
    a 500 x 250-cell, b = 4 operator with 32 wavenumbers w_k = 0.1 *
    100^((k-1)/31), in extended actions per second:
      * assembled L_, FP64 values, one complex vector per sweep (the
        reference's form);
      * assembled L_, 4, 8, 16 and 32 complex vectors per sweep (8-64 doubles
        per row; 32 is the widest sweep, two of which apply a 64-vector
        batch), FP32 and FP64 values -- the batches a block eigensolver
        over the extended space applies;
      * factored: L_ built from its parts ([L B; 0 0] and the wavenumbers,
        hymor_op_create_receptivity; B kept once, never repeated), FP64
        values, one complex vector and 16 complex vectors per sweep;
      * per-wavenumber: the 32 per-wavenumber systems [L B; 0 -i w_k I] of
        ONE extended vector as 32 complex columns of one sweep.  Every flow
        SpMV is paid 32 times, against once in the other forms; what it adds
        is each wavenumber's flow response.
    Vectors resident on the device; errors against the assembled FP64
    single-vector action."""
    from hymor_amd import _lib
    from hymor_amd import operators as ops
    from hymor_amd.core import (_FORMS, DeviceMatrix, DeviceOperator, DeviceVector,
                                expm_taylor_batched_dev, expm_taylor_dev)
    from hymor_amd.receptivity import FactoredReceptivity, receptivity_operator, split_extended

    base = ops.config_operator("C4")
    w = 0.1 * 100.0 ** (np.arange(n_w) / (n_w - 1.0))
    L_ = ops.receptivity_extension(base, w)
    A_ = L_.to_scipy_varmajor()
    n_L, p = base.n, 4 * base.nchi
    L, B = split_extended(A_, n_L, p, w)
    n_ext = A_.shape[0]
    rng = np.random.default_rng(44)
    nmax = max(widths)
    X = (rng.standard_normal((n_ext, nmax)) + 1j * rng.standard_normal((n_ext, nmax)))
    x = X[:, 0].copy()
    dt = T / nsub
    out = {"workload": (f"C4: 500x250 cells, b=4, {n_w} wavenumbers, N_ext={n_ext} complex; "
                        f"action T_{order}(L_ T/{nsub})^{nsub} x, T={T}, {FORM} form"),
           "unit": "extended actions/s", "steps": steps}

    def timed(run, sync):
        run()
        sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
        sync()
        return (time.perf_counter() - t0) / steps

    log(f"receptivity leg: assembled L_ (n_ext={n_ext}), FP64 values, one vector")
    with DeviceOperator.from_blocks(L_, value_prec="f64", with_adjoint=False) as dev:
        dev.set_graphs(True)
        dx = DeviceVector.from_host(x)
        dy = DeviceVector(n_ext, np.complex128)
        t1 = timed(lambda: expm_taylor_dev(dev, dx, dy, order=order, nsub=nsub, dt=dt,
                                           form=FORM), dev.synchronize)
        y_asm = dy.download()
        dx.free()
        dy.free()
        out["assembled_fp64_single"] = {"value": 1.0 / t1, "seconds_per_sweep": t1}
        for prec, nrhs in [(p_, w_) for w_ in widths for p_ in ("f64", "f32")]:
            dev.set_value_prec(prec)
            log(f"receptivity leg: assembled L_, {prec} values, {nrhs} vectors per sweep")
            dX = DeviceMatrix.from_host(np.asfortranarray(X[:, :nrhs]))
            dY = DeviceMatrix(n_ext, nrhs, np.complex128)
            tb = timed(lambda: expm_taylor_batched_dev(dev, dX, dY, order=order, nsub=nsub,
                                                       dt=dt, form=FORM), dev.synchronize)
            y0 = dY.download()[:, 0]
            dX.free()
            dY.free()
            out[f"assembled_{prec}_values_x{nrhs}"] = {
                "value": nrhs / tb, "seconds_per_sweep": tb,
                "speedup_vs_assembled_fp64_single": t1 * nrhs / tb,
                "rel_err_vs_assembled_fp64_single": float(np.linalg.norm(y0 - y_asm) /
                                                          np.linalg.norm(y_asm))}
            log(f"receptivity leg: {prec} x{nrhs}: {nrhs / tb:.1f} actions/s "
                f"({t1 * nrhs / tb:.2f}x)")
    log("receptivity leg: factored L_ (from [L B; 0 0] and the wavenumbers), FP64 values")
    with receptivity_operator(L, B, w, nchi=base.nchi, neta=base.neta, vars_per_cell=4) as dev:
        dev.set_graphs(True)
        dx = DeviceVector.from_host(x)
        dy = DeviceVector(n_ext, np.complex128)
        t_p = timed(lambda: expm_taylor_dev(dev, dx, dy, order=order, nsub=nsub, dt=dt,
                                            form=FORM), dev.synchronize)
        y_p = dy.download()
        dx.free()
        dy.free()
        nb = 16
        dX = DeviceMatrix.from_host(np.asfortranarray(X[:, :nb]))
        dY = DeviceMatrix(n_ext, nb, np.complex128)
        t_pb = timed(lambda: expm_taylor_batched_dev(dev, dX, dY, order=order, nsub=nsub,
                                                     dt=dt, form=FORM), dev.synchronize)
        y_pb = dY.download()[:, 0]
        dX.free()
        dY.free()
        out["factored_fp64_single"] = {
            "value": 1.0 / t_p, "seconds_per_sweep": t_p,
            "speedup_vs_assembled_fp64_single": t1 / t_p,
            "rel_err_vs_assembled_fp64_single": float(np.linalg.norm(y_p - y_asm) /
                                                      np.linalg.norm(y_asm)),
            "device_bytes": dev.device_bytes}
        out[f"factored_f64_values_x{nb}"] = {
            "value": nb / t_pb, "seconds_per_sweep": t_pb,
            "speedup_vs_assembled_fp64_single": t1 * nb / t_pb,
            "rel_err_vs_assembled_fp64_single": float(np.linalg.norm(y_pb - y_asm) /
                                                      np.linalg.norm(y_asm))}
        log(f"receptivity leg: factored: {1.0 / t_p:.1f} actions/s single, "
            f"{nb / t_pb:.1f} at {nb} per sweep")
    log("receptivity leg: per-wavenumber (32 systems of one vector), FP32 values")
    fr = FactoredReceptivity(L, B, w, nchi=base.nchi, neta=base.neta, vars_per_cell=4,
                             scale=1.0, value_prec="f32", per_wavenumber=True)
    try:
        Xf = np.zeros((fr.n_sys, n_w), dtype=np.complex128, order="F")
        Xf[n_L:, :] = x[n_L:].reshape(n_w, p).T
        Xf[:n_L, 0] = x[:n_L]
        dX = DeviceMatrix.from_host(Xf)
        dY = DeviceMatrix(fr.n_sys, n_w, np.complex128)
        lib = _lib.load()

        def run():
            _lib.check(lib.hymor_expm_taylor_batched_coldiag_dev(
                fr.dev.handle, 0, order, nsub, dt, _FORMS[FORM], n_w, fr.diag.ctypes.data,
                fr.n_L, fr.n_sys, dX.ptr, fr.n_sys, dY.ptr, fr.n_sys))

        tf = timed(run, fr.dev.synchronize)
        Y = dY.download()
        y = np.empty(n_ext, dtype=np.complex128)
        y[:n_L] = Y[:n_L, :].sum(axis=1)
        y[n_L:] = Y[n_L:, :].T.reshape(-1)
        dX.free()
        dY.free()
    finally:
        fr.close()
    out["per_wavenumber_f32_values"] = {
        "value": 1.0 / tf, "seconds_per_sweep": tf,
        "speedup_vs_assembled_fp64_single": t1 / tf,
        "rel_err_vs_assembled_fp64_single": float(np.linalg.norm(y - y_asm) /
                                                  np.linalg.norm(y_asm))}
    log(f"receptivity leg: per-wavenumber: {1.0 / tf:.2f} actions/s ({t1 / tf:.2f}x)")
    # FREESTREAM_RECEPTIVITY on this operator: 4 gains at T = 1 (Chu weights
    # of a synthetic base flow of the grid), the extended operator from its
    # parts, device Krylov-Schur one vector and blocks of 4
    from hymor_amd import stability as st
    from hymor_amd.baseflow import synthetic_solution
    s = st.CFL_TIMESTEP(synthetic_solution(500, 250, seed=11, perturb_shock=False, cfl=100.0,
                                           max_dt=1.0 / 64))
    solves = {}
    for block in (1, 4):
        t0 = time.perf_counter()
        _, G, _ = st.FREESTREAM_RECEPTIVITY(A_, s, 4, 1.0, w, tol=1e-8, krylov="device", ncv=24,
                                            maxiter=400, factored=True, block=block)
        tk = s["timings"]["receptivity_krylov"]
        solves[f"block{block}"] = {"seconds": time.perf_counter() - t0,
                                   "reduce_seconds": tk["reduce_seconds"],
                                   "krylov_seconds": tk["solve_seconds"],
                                   "applications": tk.get("applications"),
                                   "gains": [float(g) for g in G]}
    out["solve"] = dict(solves, note=("FREESTREAM_RECEPTIVITY(factored=True, krylov='device'), "
                                      "k=4, T=1, tol 1e-8"))
    log(f"receptivity leg: FREESTREAM_RECEPTIVITY {solves['block1']['seconds']:.2f} s one "
        f"vector, {solves['block4']['seconds']:.2f} s blocks of 4")
    return out


COPY_MIB = 512
# the library's default launches: k_term_lds (k_term_split_p with 3 waves per
# slice at one vector) and the tile-order matrix-core batch kernel
PLAIN_TUNING = (3, 1, 1)
BATCH_KERNEL = 6
# every leg evaluates POWER_METHOD!'s polynomial by Horner's rule
# (w = v + (A dt/j) w, j = order..1): three vector passes per term instead of
# the powers recursion's four; equal to rounding (tests/test_gpu_parity.py)
FORM = "horner"


def pmc_passes(config, timeout_s=150):
    """HBM traffic per kernel launch from two rocprofv3 --pmc passes
    (FETCH_SIZE, then WRITE_SIZE; they do not fit one pass) over this script
    in --kernel-only plain mode, which also launches the streaming-copy
    kernel k_copy on COPY_MIB MiB.  Runs as child processes before this
    process touches the GPU.  Returns ({counter: {kernel name: mean KiB per
    dispatch}}, {counter: exit status of the pass}) or (None, the reason).

    A pass counts only if the profiled program printed its result line (its
    last action) and the counter CSV exists.  The exit status is reported
    either way: rocprofv3 itself can end with SIGSEGV after the program has
    finished and the counters are written, and a nonzero status is then
    shown, not hidden."""
    out = {}
    codes = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix=f"hymor_pmc_{counter}_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", str(timeout_s), "rocprofv3", "--pmc", counter,
               "--output-format", "csv", "-d", d, "-o", "run", "--", sys.executable, "-u",
               str(Path(__file__).resolve()), "--config", config, "--kernel-only", "plain",
               "--kernel-reps", "5"]
        log(f"traffic: rocprofv3 --pmc {counter} pass")
        logf = Path(d) / "pass.log"
        try:
            with open(logf, "w") as lf:
                r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=lf, stderr=subprocess.STDOUT,
                                   timeout=timeout_s + 30)
        except (OSError, subprocess.TimeoutExpired) as e:
            return None, f"{counter} pass did not run: {e}"
        files = list(Path(d).rglob("*counter_collection.csv"))
        text = logf.read_text(errors="replace")
        finished = '"kernel_only"' in text      # the program's result line
        codes[counter] = r.returncode
        if r.returncode != 0:
            log(f"traffic: {counter} pass exited with status {r.returncode} "
                f"(program {'finished' if finished else 'did not finish'})")
        if not files or not finished:
            tail = [ln for ln in text.splitlines()
                    if ln.strip() and not ln.startswith(("W20", "I20"))][-4:]
            log(f"traffic: {counter} pass failed (exit {r.returncode}): " + " | ".join(tail))
            return None, f"{counter} pass failed (exit {r.returncode}): " + " | ".join(tail)[-400:]
        acc = {}
        with open(files[0]) as f:
            for row in csv.DictReader(f):
                if row["Counter_Name"] != counter:
                    continue
                name = row["Kernel_Name"]
                acc.setdefault(name, []).append(float(row["Counter_Value"]))
        out[counter] = {k: sum(v) / len(v) for k, v in acc.items()}
    return out, codes


def traffic_from_passes(passes, kernel):
    """Bytes per launch of `kernel` (a name like "k_term_split<8, double, 1,
    4, true>"), each counter scaled by bytes / counter-reading of k_copy."""
    copy_bytes = COPY_MIB * 1024 * 1024

    def mean_kib(counter, name):
        vals = [v for k, v in passes[counter].items() if f"hymor::{name}(" in k]
        return vals[0] if vals else None

    total = 0.0
    factors = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        c_kib, k_kib = mean_kib(counter, "k_copy"), mean_kib(counter, kernel)
        if not c_kib or k_kib is None:
            return None, None
        factors[counter] = copy_bytes / (c_kib * 1024.0)
        factors[counter + "_KiB_per_dispatch"] = k_kib
        total += k_kib * 1024.0 * factors[counter]
    return total, factors


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="C5")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--nrhs", type=int, default=16, help="vectors per batched sweep")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-eigen", action="store_true", help="skip the eigenvalue leg")
    ap.add_argument("--no-receptivity", action="store_true",
                    help="skip the C4 receptivity leg (assembled vs factored)")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the launch-bound 100x80-cell leg (eager / graph / no-alloc)")
    ap.add_argument("--no-growth", action="store_true",
                    help="skip the 400x300-cell transient-growth leg")
    ap.add_argument("--no-shock", action="store_true",
                    help="skip the 200x300-cell shock-band leg (action, 20-mode eigensolve)")
    ap.add_argument("--kernel-only",
                    choices=["plain", "batched", "block4", "receptivity", "receptivity16",
                             "krylov", "shock_batch"],
                    help="time only the Taylor-term kernel of one form (for profiler runs): "
                         "plain FP64 one vector, batched FP32 --nrhs vectors, block4 FP32 "
                         "four vectors (the eigen leg's FP32 phase), receptivity: the C4 "
                         "operator extended by 32 wavenumbers, FP64 values, one complex "
                         "vector (block kernel + extra rows); receptivity16: the same "
                         "operator on 16 complex vectors (its 32 copies of B kept once); "
                         "krylov: the block Gram-Schmidt projection kernels at N=1.44M, "
                         "32 basis vectors, blocks of 4 and 16; shock_batch: the 400x300-cell "
                         "operator with 400 shock DOFs, FP64 values, --nrhs vectors (the "
                         "shock rows in the batch kernel's launch)")
    ap.add_argument("--batch-kernel", type=int, default=None,
                    help="--kernel-only batched: use this batch kernel (0..8) instead of "
                         "the default (6: the cell-pair kernel at b=8 and 16+ vectors)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes that measure HBM traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the engine has no "
              "multi-GPU path", file=sys.stderr)
        sys.exit(2)

    passes, traffic_note, pass_codes = None, "skipped (--no-traffic)", None
    if not args.kernel_only and not args.no_traffic:
        passes, info = pmc_passes(args.config)
        if passes is None:
            traffic_note = info
        else:
            pass_codes = info

    if args.kernel_only == "krylov":
        from hymor_amd.calibrate import copy_bandwidth
        kp = krylov_projection_leg(reps=args.kernel_reps)
        cp = copy_bandwidth(COPY_MIB, 3)
        print(json.dumps({"kernel_only": "krylov", "krylov_projection": kp, "copy": cp}),
              flush=True)
        return

    from hymor_amd import operators as ops
    from hymor_amd.core import (DeviceMatrix, DeviceOperator, DeviceVector,
                                expm_taylor_batched_dev, expm_taylor_dev)

    order, nsub, T = 12, 64, 1.0
    cfg = ops.CONFIGS[args.config]
    log(f"generating {args.config}: {cfg}")
    op = ops.config_operator(args.config)
    if args.kernel_only:
        prec, nv = {"plain": ("f64", 1), "batched": ("f32", args.nrhs),
                    "block4": ("f32", 4), "receptivity": ("f64", 2),
                    "receptivity16": ("f64", 32),
                    "shock_batch": ("f64", args.nrhs)}[args.kernel_only]
        if args.kernel_only == "shock_batch":
            op = ops.config_operator("C3")
        if args.kernel_only.startswith("receptivity"):
            # the receptivity leg's extended operator, default launch tuning
            op = ops.receptivity_extension(ops.config_operator("C4"),
                                           0.1 * 100.0 ** (np.arange(32) / 31.0))
        devk = DeviceOperator.from_blocks(op, value_prec=prec, with_adjoint=False)
        if args.kernel_only.startswith("receptivity"):
            pass  # default launch, as in the receptivity leg
        elif nv < 8:
            devk.set_tuning(*PLAIN_TUNING)
        else:
            devk.set_batch_kernel(args.batch_kernel if args.batch_kernel is not None
                                  else BATCH_KERNEL)
        ms, by = devk.bench_term_kernel(reps=args.kernel_reps, nv=nv,
                                        form=FORM)
        from hymor_amd.calibrate import copy_bandwidth
        cp = copy_bandwidth(COPY_MIB, 3)
        print(json.dumps({"kernel_only": args.kernel_only, "nv": nv, "ms": ms, "bytes": by,
                          "GBps": by / (ms * 1e-3) / 1e9, "tuning": devk.tuning if nv == 1
                          else None, "copy": cp}), flush=True)
        return
    log(f"n={op.n} nnzb={op.nnzb}; uploading FP64 blocks")
    dev = DeviceOperator.from_blocks(op, value_prec="f64", with_adjoint=False)
    # the autotuner's pick is reported; the timed legs (and the profiler and
    # counter runs) use the library's default launch, so every process runs
    # the same kernel (the variants are within about 2 % of each other)
    tune_ms, tune_pick = dev.autotune(reps=10)
    log(f"autotuner's pick: split={tune_pick[0]} waves={tune_pick[1]} nt={tune_pick[2]} "
        f"({tune_ms:.4f} ms)")
    dev.set_tuning(*PLAIN_TUNING)
    tuning = dev.tuning
    x = np.random.default_rng(cfg["seed"] - 1000).standard_normal(op.n)
    dx = DeviceVector.from_host(x)
    dy = DeviceVector(op.n)

    def step():
        expm_taylor_dev(dev, dx, dy, order=order, nsub=nsub, dt=T / nsub, form=FORM)

    log(f"warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    dev.synchronize()
    log(f"timing {args.steps} actions")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dev.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    actions_per_s = args.steps / elapsed
    log(f"{ms_per_step:.3f} ms/action")

    log(f"timing the Taylor-term kernel alone ({args.kernel_reps} launches, HIP events)")
    k_ms, k_bytes = dev.bench_term_kernel(reps=args.kernel_reps, form=FORM)
    achieved = k_bytes / (k_ms * 1e-3) / 1e9
    from hymor_amd.calibrate import read_bandwidth
    read_ceiling = read_bandwidth(max(1, int(k_bytes) >> 20), 20)
    log(f"streaming read of the same byte count: {read_ceiling['GBps']:.0f} GB/s")
    y_plain = dy.download()
    dev.close()
    nt_s = str(bool(tuning[2])).lower()
    # split 3 (x staged through LDS) runs k_term_split_p at one double per row,
    # with 9 waves per slice if so tuned, else 3 (its last template argument:
    # no shock bands on this operator)
    nw_p = 9 if tuning[1] == 9 else 3
    kname = {0: f"k_term<{cfg['b']}, double, 1, {nt_s}>",
             1: f"k_term_split<{cfg['b']}, double, 1, {tuning[1]}, {nt_s}>",
             2: f"k_term_split_p<{cfg['b']}, double, 1, {nw_p}, {9 // nw_p}, {nt_s}, false>"
             }[min(tuning[0], 2)]
    traffic, factors = None, None
    if passes is not None:
        traffic, factors = traffic_from_passes(passes, kname)
        traffic_note = ("FETCH_SIZE + WRITE_SIZE per launch, each scaled by bytes / counter "
                        "reading of k_copy" if traffic is not None else
                        f"{kname} not found in the pmc passes")

    batched = None
    if not args.no_batched:
        nb = args.nrhs
        log(f"batched leg: FP32-stored values, {nb} vectors per sweep")
        devb = DeviceOperator.from_blocks(op, value_prec="f32", with_adjoint=False)
        bkern_pick = None
        if nb in (16, 32, 64):
            bkern_pick, _ = devb.autotune_batch(nv=nb, reps=5)
        bkern = args.batch_kernel if args.batch_kernel is not None else BATCH_KERNEL
        devb.set_batch_kernel(bkern)
        X = np.random.default_rng(7).standard_normal((op.n, nb))
        X[:, 0] = x
        dX = DeviceMatrix.from_host(X)
        dY = DeviceMatrix(op.n, nb)

        def bstep():
            expm_taylor_batched_dev(devb, dX, dY, order=order, nsub=nsub, dt=T / nsub,
                                    form=FORM)

        bstep()
        devb.synchronize()
        ksteps = max(1, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(ksteps):
            bstep()
        devb.synchronize()
        tb = time.perf_counter() - t0
        y0 = dY.download()[:, 0]
        err = float(np.linalg.norm(y0 - y_plain) / np.linalg.norm(y_plain))
        bk_ms, bk_bytes = devb.bench_term_kernel(reps=max(5, args.kernel_reps // 5), nv=nb,
                                                 form=FORM)
        devb.close()
        # the same sweep with FP64 values: the batched form at FP64 accuracy
        log(f"batched leg: FP64 values, {nb} vectors per sweep")
        dev64 = DeviceOperator.from_blocks(op, value_prec="f64", with_adjoint=False)
        dev64.set_batch_kernel(bkern)

        def bstep64():
            expm_taylor_batched_dev(dev64, dX, dY, order=order, nsub=nsub, dt=T / nsub,
                                    form=FORM)

        bstep64()
        dev64.synchronize()
        t0 = time.perf_counter()
        for _ in range(ksteps):
            bstep64()
        dev64.synchronize()
        tb64 = time.perf_counter() - t0
        y64 = dY.download()[:, 0]
        err64 = float(np.linalg.norm(y64 - y_plain) / np.linalg.norm(y_plain))
        dev64.close()
        b64_aps = nb * ksteps / tb64
        b_aps = nb * ksteps / tb
        b_gbs = bk_bytes / (bk_ms * 1e-3) / 1e9
        # a Horner term writes its nb new vectors (n x nb doubles) beside what
        # it reads: the streaming ceiling at that read/write mix
        from hymor_amd.calibrate import mixed_bandwidth
        b_write = float(op.n) * nb * 8
        mix = mixed_bandwidth(bk_bytes - b_write, b_write, 20)
        log(f"batched term: {b_gbs:.0f} GB/s; streaming reads + writes at its write share "
            f"{b_write / bk_bytes:.1%}: {mix['GBps']:.0f} GB/s")
        batched = {
            "value": b_aps,
            "unit": "exp(AT)v actions/s",
            "nrhs": nb,
            "sweeps": ksteps,
            "ms_per_sweep": tb / ksteps * 1e3,
            "speedup_vs_plain_fp64": b_aps / actions_per_s,
            "value_storage": "f32",
            "accumulation": "f64",
            "taylor_form": FORM,
            "rel_err_vs_plain_fp64_action": err,
            "kernel": {0: f"k_spmm<{cfg['b']}, float, {nb}>",
                       1: f"k_spmm_mfma<{cfg['b']}, float, {nb}>",
                       2: f"k_spmm_mfma_mc<{cfg['b']}, float, {nb}, U, {2 if nb == 16 else 1}>",
                       3: f"k_spmm_mfma_mc<{cfg['b']}, float, {nb}, U, 1>",
                       4: f"k_spmm_mfma_w<{cfg['b']}, float, {nb}, "
                          f"{ {16: 9, 32: 5, 64: 3}.get(nb)}>",
                       5: f"k_spmm_mfma_w<{cfg['b']}, float, {nb}, "
                          f"{ {16: 5, 32: 3, 64: 2}.get(nb)}>",
                       6: (f"k_spmm_pair<float, {nb}, { {16: 12, 32: 6, 64: 4}.get(nb)}>"
                           if nb >= 16 else f"k_spmm_tile<{cfg['b']}, float, {nb}, 9>"),
                       8: f"k_spmm_tile<{cfg['b']}, float, {nb}, "
                          f"{ {8: 9, 16: 9, 32: 3, 64: 3}.get(nb)}>",
                       7: f"k_spmm_quad<float, {nb}, { {8: 9, 16: 9, 32: 5, 64: 3}.get(nb)}>"
                       }.get(bkern, "k_spmm") if cfg["b"] == 8 else {
                       # b = 4 under batch kernels 6 and 7
                       6: f"k_spmm_quad4c<float, {nb}, { {8: 9, 16: 9, 32: 3, 64: 3}.get(nb)}>",
                       7: f"k_spmm_quad4<float, {nb}, { {8: 9, 16: 9, 32: 9, 64: 5}.get(nb)}>"
                       }.get(bkern, f"batch kernel {bkern}"),
            "batch_kernel_id": bkern,
            "autotuner_pick": bkern_pick,
            "kernel_ms": bk_ms,
            "algorithmic_bytes_per_launch": bk_bytes,
            "achieved_GBps": b_gbs,
            "frac_of_hbm_peak": b_gbs / HBM_PEAK_GBS,
            "same_mix_stream": {"kernel": "k_stream_mix", "write_share": b_write / bk_bytes,
                                "GBps": mix["GBps"], "frac": b_gbs / mix["GBps"]},
            "f64_values": {"value": b64_aps, "ms_per_sweep": tb64 / ksteps * 1e3,
                           "speedup_vs_plain_fp64": b64_aps / actions_per_s,
                           "rel_err_vs_plain_fp64_action": err64},
        }
        log(f"batched: {b_aps:.2f} actions/s ({b_aps / actions_per_s:.2f}x plain), err {err:.2e}; "
            f"FP64 values {b64_aps:.2f} ({b64_aps / actions_per_s:.2f}x), err {err64:.2e}")

    result = {
        "metric": METRIC,
        "value": actions_per_s,
        "unit": "exp(AT)v actions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.config}: {cfg['nchi']}x{cfg['neta']} cells, {cfg['b']} vars/cell "
                         f"({cfg['b']}x{cfg['b']} blocks, 3x3 stencil), N={op.n}, "
                         f"nnz={op.nnzb * cfg['b'] ** 2}; action T_{order}(A T/{nsub})^{nsub} v, "
                         f"T={T}; FP64 values, FP64 vectors resident in HBM"),
            "n": op.n,
            "nnzb": op.nnzb,
            "taylor_order": order,
            "substeps": nsub,
            "taylor_form": FORM + " (POWER_METHOD!'s polynomial)",
            "kernels_per_action": order * nsub,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_note": traffic_note,
            "traffic_pass_exit_status": pass_codes,
            "traffic_counter_scale": factors,
            "kernel": kname,
            "launch": {"split": tuning[0], "waves": tuning[1], "nontemporal": tuning[2]},
            "autotuner_pick": {"split": tune_pick[0], "waves": tune_pick[1],
                               "nontemporal": tune_pick[2], "ms": tune_ms,
                               "ms_note": "timed on a powers-form term (t and v written: "
                                          "four vector passes), kernel_ms on the Horner "
                                          "step (three)"},
            "kernel_ms": k_ms,
            "algorithmic_bytes_per_launch": k_bytes,
            "read_stream_GBps": read_ceiling["GBps"],
            "frac_of_read_stream": achieved / read_ceiling["GBps"],
        },
        "batched": batched,
    }
    if not args.no_small:
        result["small_config"] = small_config_leg(cpu=not args.no_cpu)
    if not args.no_growth:
        result["transient_growth"] = transient_growth_leg()
    if not args.no_shock:
        result["shock_config"] = shock_config_leg(cpu=not args.no_cpu)
    if not args.no_receptivity:
        result["batched_receptivity"] = receptivity_leg()
    if not args.no_eigen:
        log("eigen leg: plain FP64 Krylov-Schur vs batched FP32 + FP64 refinement")
        result["batched_eigen"] = eigen_leg()
        log(f"eigen leg: {result['batched_eigen']}")
        result["batched_eigen"]["krylov_projection"] = krylov_projection_leg()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(op, order, nsub)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
