#!/usr/bin/env python3
"""Throughput of the fv3d hydro step on one MI355X.

Workload (default): 3D Sedov-style blast, 256^3 cells on the unit cube,
gamma 5/3, background rho=1 p=0.1, p=10 inside r<0.1, PLM-minmod + HLLC,
SSP-RK2, CFL 0.8, reflecting walls, FP64.  One step = one full RK2 step
(two fused stage launches, the second also reducing the next dt) over the
whole grid, with the state already resident in HBM.

Prints one JSON line:
  value      Gcell-updates/s (interior cells x steps / s)
  roofline   the stage kernel: algorithmic bytes per launch (200 B per
             cell-update over 2 launches) / mean launch time from HIP events
             recorded on the launch stream, against 8 TB/s
  traffic    measured HBM bytes per stage launch from the rocprofv3 PMC
             summary in profiles/ (FETCH_SIZE + WRITE_SIZE), if present
  cpu_baseline  the CPU restatement of the reference (oracle/, FP64, OpenMP)
             on the host cores, timed on a bounded window of the same grid
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s & %HBM roofline at 256^3 FP64/FP32; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
ALG_BYTES_PER_CELL_UPDATE = {"f64": 200.0, "f32": 100.0}
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "stage_kernel_blast256_fp64_traffic.md")


def blast_params(n):
    import fv3d_amd as F
    from fv3d_amd import Params
    return Params.defaults().replace(
        Nx=n, Ny=n, Nz=n, gamma0=5.0 / 3.0, reconstruction=F.PLM, riemann_solver=F.HLLC,
        time_stepping=F.TS_RK2, CFL=0.8, boundary_x=F.BC_REFLECTING, boundary_y=F.BC_REFLECTING,
        boundary_z=F.BC_REFLECTING, tend=0.1, save_freq=0.1)


def blast_ic(p):
    from fv3d_amd import problems
    return problems.blast(p, radius=0.1, rho_in=1.0, p_in=10.0, rho_out=1.0, p_out=0.1)


def read_traffic():
    """HBM bytes per stage launch from the committed rocprofv3 summary."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            txt = f.read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


# FP64 VALU lane-instructions MI355X issues per second: 256 CUs x 4 SIMDs x
# 16 lanes x 2.4 GHz (one wave64 FP64 instruction per SIMD every 4 cycles)
VALU_PEAK_TINST = 256 * 4 * 16 * 2.4e9 / 1e12


def read_valu_per_cell():
    """VALU lane-instructions per cell and stage of the FP64 blast stage
    kernel, from the SQ_INSTS_VALU table of the committed rocprofv3 summary."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            txt = f.read()
    except OSError:
        return None
    sq = txt.split("## SQ counters", 1)[-1]
    m = re.search(r"\| `fv3d_hip::stage_kernel<double, 8, 2, 1, false, false, 0, 2, 64>` \| ([0-9]+) \|", sq)
    return float(m.group(1)) if m else None


def cpu_model():
    """The host CPU's model name and logical CPU count (/proc/cpuinfo)."""
    name, n = "unknown", 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and name == "unknown":
                    name = line.split(":", 1)[1].strip()
                if line.startswith("processor"):
                    n += 1
    except OSError:
        pass
    return name, n


def cpu_baseline(n, budget_s):
    """The oracle (CPU restatement, reference pass structure) on the same
    grid: one warm-up step, then whole steps until budget_s has elapsed.
    OpenMP threads are bound one per core, packed (OMP_PLACES=cores,
    OMP_PROC_BIND=close), before the oracle's OpenMP runtime starts."""
    os.environ.setdefault("OMP_PLACES", "cores")
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle.oracle import Oracle, lib as olib
    import ctypes
    p = blast_params(n)
    o = Oracle(p)
    Q0 = blast_ic(p)
    U = o.prim_to_cons(Q0)
    Q = o.zeros()
    Unew = o.zeros()
    dts = np.zeros(64)
    t = ctypes.c_double()
    from oracle.oracle import _p
    L = olib()
    L.orc_run(ctypes.byref(o.p), _p(U), _p(Q), _p(Unew), _p(o.work), 1, 0, _p(dts), ctypes.byref(t))
    steps = 0
    t0 = time.perf_counter()
    while True:
        L.orc_run(ctypes.byref(o.p), _p(U), _p(Q), _p(Unew), _p(o.work), 1, 0, _p(dts), ctypes.byref(t))
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 20:
            break
    cells = p.Nx * p.Ny * p.Nz
    model, logical = cpu_model()
    return {
        "value": cells * steps / el / 1e9,
        "unit": "Gcell-updates/s",
        "cores": int(L.orc_num_threads()),
        "kind": "port",
        "cpu": model,
        "logical_cpus_visible": logical,
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PLACES", "OMP_PROC_BIND")},
        "sample": f"blast {n}^3 FP64 RK2, {steps} full steps after 1 warm-up step "
                  f"({el:.1f} s), reference pass structure, OpenMP",
    }


def stream_copy(nbytes=1 << 30, reps=20):
    """Measured device copy bandwidth: the library's plain copy kernel over
    nbytes, timed with HIP events (read + write bytes).  Also the
    calibration kernel for the rocprofv3 byte counters."""
    import fv3d_amd as F
    return F.stream_copy_gbs(nbytes, reps)


def measure(F, p, Q0, precision, steps, warmup):
    """Timed steps of one precision: HIP events on the launch stream around
    every stage launch.  The reference scheme itself develops negative
    pressure on the 256^3 blast after ~90 RK2 steps (the oracle reproduces it
    bit for bit), so steps are timed in windows of at most CHUNK steps, each
    started (untimed) from the initial condition plus the warm-up steps."""
    CHUNK = 60
    sim = F.Simulation(p, precision)
    total_ms, stage_ms, wall = 0.0, [], 0.0
    left = steps
    while left > 0:
        n = min(CHUNK, left)
        sim.set_prim(Q0, with_ghosts=True)
        if warmup > 0:
            sim.step(warmup)
        sim.synchronize()
        dt0, _ = sim.compute_dt(strict=True)
        if not np.isfinite(dt0) or dt0 <= 0:
            raise RuntimeError("non-finite dt after warm-up")
        t0 = time.perf_counter()
        ms, st = sim.time_steps(n, per_stage=True)
        sim.synchronize()
        wall += time.perf_counter() - t0
        sim.compute_dt(strict=True)  # raises if the window produced non-finite values
        total_ms += ms
        stage_ms.extend(st.tolist())
        left -= n
    sim.close()
    return total_ms, stage_ms, wall


def per_stage(stage_ms, cells, stage_bytes):
    """Roofline of each RK2 stage launch on its own: stage_ms alternates
    stage 1, stage 2; stage_bytes are their algorithmic bytes per cell."""
    out = {}
    for i, (name, b) in enumerate(zip(("stage1", "stage2"), stage_bytes)):
        ms = float(np.mean(stage_ms[i::2]))
        gbs = b * cells / (ms * 1e-3) / 1e9
        out[name] = {"alg_bytes_per_cell": b, "mean_launch_ms": round(ms, 5), "achieved_GBs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


# algorithmic bytes per cell of stage 1 and stage 2 (R U^n + W U*; R U*, U^n
# (+ compensation) + W U^{n+1} (+ compensation))
STAGE_BYTES = {"f64": (80.0, 120.0), "fp32_kahan": (40.0, 100.0), "fp32_plain": (40.0, 60.0)}


def form_line(F, p, Q0, precision, steps, warmup, alg_bytes, stage_bytes):
    cells = p.Nx * p.Ny * p.Nz
    total_ms, stage_ms, _ = measure(F, p, Q0, precision, steps, warmup)
    value = cells * steps / (total_ms * 1e-3) / 1e9
    achieved = alg_bytes * cells / 2.0 / (float(np.mean(stage_ms)) * 1e-3) / 1e9
    return {"value": round(value, 4), "unit": "Gcell-updates/s", "ms_per_step": round(total_ms / steps, 4),
            "alg_bytes_per_cell_update": alg_bytes, "roofline_achieved_GBs": round(achieved, 1),
            "roofline_frac": round(achieved / HBM_PEAK_GBS, 4),
            "per_stage": per_stage(stage_ms, cells, stage_bytes)}


def fp32_error_vs_fp64(F, p, Q0, steps):
    """Per-field error of the FP32 forms against FP64 on the same workload:
    the FP64 run's dt sequence drives every run (fv3d_step_dt), so all
    three states are at the same time; L1 and max norms relative to the
    FP64 field's mean absolute value (velocities: to the FP64 speed's)."""
    with F.Simulation(p, F.FP64) as sim:
        sim.set_prim(Q0)
        dts = []
        for _ in range(steps):
            dts.append(sim.compute_dt(strict=True)[0])
            sim.step()
        ref = sim.get_prim(with_ghosts=False)
    speed = float(np.mean(np.sqrt((ref[..., 1:4] ** 2).sum(-1))))
    scale = [float(np.mean(np.abs(ref[..., 0]))), speed, speed, speed, float(np.mean(np.abs(ref[..., 4])))]
    out = {"steps": steps, "dt": "FP64 run's sequence for every run",
           "norm": "|FP32 - FP64| / mean|FP64 field| (u, v, w: / mean FP64 speed)"}
    for name, prec in (("fp32_kahan_rk2", F.FP32), ("fp32_plain_rk2", F.FP32_PLAIN)):
        with F.Simulation(p, prec) as sim:
            sim.set_prim(Q0)
            for dt in dts:
                sim.step_dt(dt)
            q = sim.get_prim(with_ghosts=False)
        d = np.abs(q - ref)
        out[name] = {f: {"l1": float(np.mean(d[..., i]) / scale[i]), "linf": float(np.max(d[..., i]) / scale[i])}
                     for i, f in enumerate(["rho", "u", "v", "w", "prs"])}
    return out


def ensemble_form(F, nbox, steps, tune_cache=None, loop=True):
    """The batched ensemble: nbox Rayleigh-Taylor boxes (sides drawn from
    {32,48,64,96} per axis, gamma ~ U[1.3,1.7], g=-0.1, periodic x/y,
    reflecting z), PLM + HLLC, SSP-RK2, CFL 0.8, FP32 with Kahan-compensated
    RK2.  Times `steps` steps from the initial condition with the persistent
    kernels replayed from a hipGraph (launch shape from the autotuner), and
    the same steps as a per-box launch loop; then runs the FP64 ensemble for
    the same steps and reports L1(rho) of FP32 against FP64 per box."""
    from fv3d_amd import problems
    base = F.Params.defaults().replace(reconstruction=F.PLM, riemann_solver=F.HLLC,
                                       time_stepping=F.TS_RK2, CFL=0.8)
    ens = problems.rt_ensemble(base, nbox)
    ps = [q for q, _ in ens]

    def fresh(precision):
        b = F.Batch(ps, precision)
        for i, (_, Q0) in enumerate(ens):
            b.set_prim(i, Q0)
        return b

    b32 = fresh(F.FP32)
    cells = b32.cells
    nzc, grid = b32.tune(nsteps=3, cache=tune_cache)
    wx = b32.tile_width
    print(f"ensemble: tuned tile width={wx} nzc={nzc} grid={grid}", file=sys.stderr, flush=True)
    ms_graph = b32.time(steps, F.BATCH_GRAPH)
    b32.compute_dt()  # raises on non-finite state
    print(f"ensemble: graph {ms_graph:.1f} ms / {steps} steps", file=sys.stderr, flush=True)
    ms_loop = None
    if loop:
        bl = fresh(F.FP32)
        ms_loop = bl.time(steps, F.BATCH_LOOP)
        bl.close()
        print(f"ensemble: loop {ms_loop:.1f} ms / {steps} steps", file=sys.stderr, flush=True)
    b64 = fresh(F.FP64)
    b64.step(b32.step_count, F.BATCH_GRAPH)  # the graph run's capture step is one more
    b64.compute_dt()
    b32_steps = b32.step_count
    l1 = []
    for i in range(nbox):
        r32 = b32.get_prim(i, with_ghosts=False)[..., 0]
        r64 = b64.get_prim(i, with_ghosts=False)[..., 0]
        l1.append(float(np.mean(np.abs(r32 - r64))))
    b32.close()
    b64.close()
    g = cells * steps / (ms_graph * 1e-3) / 1e9
    lo = cells * steps / (ms_loop * 1e-3) / 1e9 if ms_loop else None
    return {"boxes": nbox, "cells": cells, "steps": steps, "l1_after_steps": b32_steps, "tuned_tile_width": wx, "tuned_nzc": nzc, "tuned_grid": grid,
            "graph_Gcells": round(g, 4), "loop_Gcells": round(lo, 4) if lo else None,
            "speedup_vs_loop": round(g / lo, 3) if lo else None,
            "graph_ms_per_step": round(ms_graph / steps, 4),
            "loop_ms_per_step": round(ms_loop / steps, 4) if ms_loop else None,
            "alg_bytes_per_cell_update": 140.0,
            "roofline_frac_graph": round(140.0 * g / HBM_PEAK_GBS, 4),
            "l1_rho_fp32_vs_fp64_max": max(l1), "l1_rho_fp32_vs_fp64_mean": float(np.mean(l1))}


def stratified_form(F, steps=2000, dev_window=200):
    """Stratified convection with conduction and viscosity: 256x256x128 on
    [0,2]x[0,2]x[0,1] (dx = 1/128), polytrope m=1 under g=1 (T = 1 + z/2,
    rho = T, p = T^2), gamma 5/3, kappa 0.07, mu 0.005, velocity noise 1e-3,
    HLL + PLM, SSP-RK2, periodic x/y, reflecting z with fixed temperatures
    1.0 / 1.5 at the walls.  CFL 0.4: at 0.8 the reference's conduction dt
    is unstable and the FP64 run itself blows up.  Runs FP64 and FP32
    (compensated RK2) for `steps` steps, reading the dt triple of every step,
    and reports their throughput and the L1 error of FP32 against FP64 per
    field."""
    p = F.Params.defaults().replace(
        Nx=256, Ny=256, Nz=128, xmax=2.0, ymax=2.0, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC,
        boundary_z=F.BC_REFLECTING, reconstruction=F.PLM, riemann_solver=F.HLL, time_stepping=F.TS_RK2,
        gamma0=5.0 / 3.0, gravity=1, g=1.0, CFL=0.4, thermal_conductivity_active=1, kappa=0.07,
        bctc_zmin=F.BCTC_FIXED_TEMPERATURE, bctc_zmin_value=1.0, bctc_zmax=F.BCTC_FIXED_TEMPERATURE,
        bctc_zmax_value=1.5, viscosity_active=1, mu=0.005)
    from fv3d_amd import problems
    Q0 = problems.polytrope_convection(p)
    cells = p.Nx * p.Ny * p.Nz
    out = {"grid": [p.Nx, p.Ny, p.Nz], "steps": steps, "CFL": p.CFL}
    res = {}
    for name, prec in (("fp64", F.FP64), ("fp32_kahan", F.FP32)):
        with F.Simulation(p, prec) as sim:
            sim.set_prim(Q0)
            trip = np.zeros((steps, 3))
            t0 = time.perf_counter()
            for s_ in range(steps):
                dt, inv = sim.compute_dt(strict=True)
                trip[s_] = [1.0 / inv[0], 1.0 / inv[1], 1.0 / inv[2]]
                sim.step()
            sim.synchronize()
            wall = time.perf_counter() - t0
            res[name] = (sim.get_prim(with_ghosts=False), trip)
            tot, _ = sim.time_steps(dev_window)  # device-only rate, continuing from the final state
        out[name] = {"Gcells_incl_dt_readback": round(cells * steps / wall / 1e9, 4),
                     "Gcells_device": round(cells * dev_window / (tot * 1e-3) / 1e9, 4),
                     "device_window_steps": dev_window,
                     "dt_hyp_first_last": [trip[0, 0], trip[-1, 0]], "dt_TC": trip[0, 1], "dt_visc": trip[0, 2]}
        print(f"stratified: {name} done ({wall:.1f} s)", file=sys.stderr, flush=True)
    q64, t64 = res["fp64"]
    q32, t32 = res["fp32_kahan"]
    out["l1_fp32_vs_fp64"] = {f: float(np.mean(np.abs(q32[..., i] - q64[..., i])))
                              for i, f in enumerate(["rho", "u", "v", "w", "prs"])}
    out["dt_hyp_max_rel_diff_fp32_vs_fp64"] = float(np.max(np.abs(t32[:, 0] / t64[:, 0] - 1.0)))
    return out


def adjoint_form(F, n=128, steps=200, K=20, ndirs=16, precision=None, fd=True, ref=None):
    """The discrete adjoint workload: n^3 periodic box, gamma 5/3, PLM +
    HLLC, SSP-RK2, CFL 0.4, conduction kappa=0.02 and viscosity mu=0.01;
    rho = 1 + 0.05 GRF (P(k) ~ k^-2), p = 1, solenoidal velocity of rms 0.1.
    J = sum((rho_T - rho_target)^2) after `steps` steps with the dt sequence
    of the ordinary run frozen; the target is the run from the next seed.
    Reports the adjoint's device time (checkpointed forward pass + backward
    pass) over the frozen forward run's, and central differences
    (eps = 1e-6) of J along `ndirs` unit-norm random directions of the
    initial density field (Gaussian random fields with the density's own
    k^-2 spectrum) against the adjoint's dJ/drho0, plus dJ/dkappa and dJ/dmu
    against their own central differences.  With `ref` (the FP64 call's
    returned `_ref`), an FP32 call reuses its dt sequence and target and
    reports its gradients' errors against the FP64 adjoint's."""
    from fv3d_amd import problems
    precision = F.FP64 if precision is None else precision
    p = F.Params.defaults().replace(
        Nx=n, Ny=n, Nz=n, gamma0=5.0 / 3.0, reconstruction=F.PLM, riemann_solver=F.HLLC,
        time_stepping=F.TS_RK2, CFL=0.4, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC,
        boundary_z=F.BC_PERIODIC, thermal_conductivity_active=1, kappa=0.02, viscosity_active=1,
        mu=0.01)
    Q0 = problems.turbulent_box(p, seed=12345)
    sim = F.Simulation(p, precision)
    if ref is None:
        sim.set_prim(Q0)
        dts = []
        for _ in range(steps):
            dts.append(sim.compute_dt(strict=True)[0])
            sim.step()
        dts = np.array(dts)
        sim.set_prim(problems.turbulent_box(p, seed=12346))
        sim.forward_frozen(dts)
        target = sim.get_prim(with_ghosts=False)[..., 0].copy()
    else:
        dts, target = ref["dts"], ref["target"]
    d0 = Q0[p.kbeg:p.kend, p.jbeg:p.jend, p.ibeg:p.iend]
    vel = d0[..., 1:4]

    def J_of(rho0, q=p):
        prim = d0.copy()
        prim[..., 0] = rho0
        s2 = sim if q is p else F.Simulation(q, precision)
        s2.set_prim(np.ascontiguousarray(prim), with_ghosts=False)
        ms = s2.forward_frozen(dts)
        r = s2.get_prim(with_ghosts=False)[..., 0]
        if s2 is not sim:
            s2.close()
        # summed in extended precision: the float64 sum's rounding (about
        # 1e-15 J over 2M cells) would otherwise dominate the central
        # differences at eps = 1e-6
        return float(np.sum((r - target) ** 2, dtype=np.longdouble)), r, ms

    J0, rT, _ = J_of(d0[..., 0])
    fwd_ms = []
    for _ in range(2):
        sim.set_prim(Q0)
        fwd_ms.append(sim.forward_frozen(dts))
    sim.set_prim(Q0)
    lamT = np.zeros(p.shape_domain)
    lamT[..., 0] = 2.0 * (rT - target)
    lam0, dk, dm = sim.adjoint(dts, lamT, K)
    t_fwd_pass, t_bwd, t_adj = sim.adjoint_times()
    # the same run with every segment kept in HBM from the forward pass
    # (no recomputation): 10 segments of 41 states at 128^3 FP64 = 34 GB
    sim.set_prim(Q0)
    real_bytes = 8 if precision == F.FP64 else 4
    store = (2 * K + 1) * 5 * n ** 3 * real_bytes * ((steps + K - 1) // K)
    lam0s, dks, dms = sim.adjoint(dts, lamT, K, store_bytes=store)
    ts_fwd, ts_bwd, ts_adj = sim.adjoint_times()
    stored_rel = float(np.abs(lam0s - lam0).max() / np.abs(lam0).max())
    g = lam0[..., 0] + (lam0[..., 1:4] * vel).sum(-1) + lam0[..., 4] * 0.5 * (vel * vel).sum(-1)
    print(f"adjoint: forward {min(fwd_ms):.1f} ms, adjoint {t_adj:.1f} ms", file=sys.stderr, flush=True)
    out = {"grid": [n, n, n], "steps": steps, "checkpoint_every": K,
           "dtype": "f64" if precision == F.FP64 else "f32",
           "forward_ms": round(min(fwd_ms), 3), "adjoint_ms": round(t_adj, 3),
           "adjoint_forward_pass_ms": round(t_fwd_pass, 3), "adjoint_backward_ms": round(t_bwd, 3),
           "adjoint_over_forward": round(t_adj / min(fwd_ms), 3),
           "stored_segments": {"store_gb": round(store / 1e9, 2), "adjoint_ms": round(ts_adj, 3),
                               "forward_pass_ms": round(ts_fwd, 3), "backward_ms": round(ts_bwd, 3),
                               "adjoint_over_forward": round(ts_adj / min(fwd_ms), 3),
                               "max_rel_diff_vs_recomputed": stored_rel},
           "J": J0,
           "dJ_dkappa": dk, "dJ_dmu": dm}
    if ref is not None:
        out["vs_fp64_adjoint"] = {
            "dJ_drho0_rel_l2": float(np.linalg.norm(g - ref["g"]) / np.linalg.norm(ref["g"])),
            "dJ_dkappa_rel": abs(dk - ref["dk"]) / abs(ref["dk"]),
            "dJ_dmu_rel": abs(dm - ref["dm"]) / abs(ref["dm"])}
    if fd and ndirs > 0:
        # Central differences at three step sizes per direction.  How the
        # error scales with eps tells its source: a smooth truncation error
        # falls as eps^2, crossings of the scheme's switch points (minmod
        # selections, Riemann wave regions: kinks of J along the line) make
        # an error proportional to eps (the number of kinks crossed grows
        # with the interval), and rounding grows as 1/eps.  order > 0 from
        # 1e-5 to 1e-6 means the error falls with eps.
        rng = np.random.default_rng(7)
        eps_list = (1e-5, 1e-6, 1e-7)
        rows = []
        for i in range(ndirs):
            # perturbations of the same class as the density field (P(k) ~ k^-2), unit norm
            d = problems.spectral_field(rng, g.shape, -2.0)
            d /= np.linalg.norm(d)
            ad = float(np.sum(g * d))
            errs = []
            for eps in eps_list:
                Jp = J_of(d0[..., 0] + eps * d)[0]
                Jm = J_of(d0[..., 0] - eps * d)[0]
                errs.append(abs((Jp - Jm) / (2 * eps) - ad) / abs(ad))
            # the error's order in eps between 1e-5 and 1e-6
            order = float(np.log10(max(errs[0], 1e-300) / max(errs[1], 1e-300)))
            rows.append({"adjoint": ad, "rel_err": dict(zip(["1e-5", "1e-6", "1e-7"], errs)),
                         "order_1e5_to_1e6": round(order, 2)})
        rel = [r["rel_err"]["1e-6"] for r in rows]
        best = [min(r["rel_err"].values()) for r in rows]
        print(f"adjoint: FD rho0 rel errors at eps 1e-6: max {max(rel):.2e}; best over eps: max "
              f"{max(best):.2e}", file=sys.stderr, flush=True)
        out["fd_rho0_rel_err"] = {
            "eps": 1e-6, "directions": ndirs, "max": max(rel), "median": float(np.median(rel)),
            "over_1e-6": int(sum(r > 1e-6 for r in rel)),
            "best_eps_max": max(best),
            "per_direction": rows,
            "order_note": "order ~2: smooth truncation; ~1: switch-point crossings; ~-1: rounding"}
        for key, val, ad in [("kappa", p.kappa, dk), ("mu", p.mu, dm)]:
            h = eps * val
            Jp = J_of(d0[..., 0], p.replace(**{key: val + h}))[0]
            Jm = J_of(d0[..., 0], p.replace(**{key: val - h}))[0]
            out[f"fd_{key}_rel_err"] = abs((Jp - Jm) / (2 * h) - ad) / abs(ad)
    sim.close()
    out["_ref"] = {"dts": dts, "target": target, "g": g, "dk": dk, "dm": dm}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="cells per axis")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-baseline timing (0: skip)")
    ap.add_argument("--cpu-n", type=int, default=256, help="cells per axis of the CPU-baseline sample")
    ap.add_argument("--stream", type=int, default=1, help="also measure the device copy bandwidth")
    ap.add_argument("--forms", type=int, default=1, help="also time the FP32 forms")
    ap.add_argument("--ensemble", type=int, default=512, help="boxes of the batched ensemble (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--ensemble-loop", type=int, default=1,
                    help="also time the per-box launch loop the ensemble replaces")
    ap.add_argument("--stratified", type=int, default=2000, help="steps of the stratified-convection form (0: skip)")
    ap.add_argument("--stratified-window", type=int, default=200,
                    help="steps of its device-only timing window")
    ap.add_argument("--adjoint", type=int, default=128, help="grid of the adjoint workload (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-dirs", type=int, default=16)
    ap.add_argument("--copy-only", action="store_true",
                    help="run only the 1 GiB copy (calibration run for the rocprofv3 byte counters)")
    args = ap.parse_args()

    if args.copy_only:
        print(json.dumps({"stream_copy_GBs": round(stream_copy(reps=5), 1), "bytes_read_per_launch": 1 << 30,
                          "bytes_written_per_launch": 1 << 30}))
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import fv3d_amd as F
    p = blast_params(args.n)
    cells = p.Nx * p.Ny * p.Nz
    Q0 = blast_ic(p)
    total_ms, stage_ms, wall = measure(F, p, Q0, F.FP64, args.steps, args.warmup)
    ms_per_step = total_ms / args.steps
    value = cells * args.steps / (total_ms * 1e-3) / 1e9
    mean_launch_ms = float(np.mean(stage_ms))
    bytes_per_launch = ALG_BYTES_PER_CELL_UPDATE["f64"] * cells / 2.0
    achieved = bytes_per_launch / (mean_launch_ms * 1e-3) / 1e9
    traffic = read_traffic()
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D Sedov-style blast {args.n}^3, gamma 5/3, PLM-minmod + HLLC, SSP-RK2, "
                        "CFL 0.8, reflecting walls, FP64",
            "grid": [p.Nx, p.Ny, p.Nz],
            "step": "one RK2 step = 2 fused stage launches, dt reduced on device",
            "wall_s_timed_window": round(wall, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "fv3d_hip::stage_kernel<double,8,PLM,HLLC,false,false,STAGE_PLAIN,STAGE_VAR_LEAN,64>",
            "alg_bytes_per_launch": bytes_per_launch,
            "mean_launch_ms": round(mean_launch_ms, 5),
            "per_stage": per_stage(stage_ms, cells, STAGE_BYTES["f64"]),
        },
    }
    valu = read_valu_per_cell()
    if valu:
        # the kernel is VALU-bound: its issue rate beside the HBM roofline
        rate = valu * cells / (mean_launch_ms * 1e-3) / 1e12  # one launch = one stage over every cell
        out["valu"] = {"lane_instr_per_cell_stage": valu, "achieved_Tinst_s": round(rate, 2),
                       "peak_Tinst_s": round(VALU_PEAK_TINST, 2), "frac": round(rate / VALU_PEAK_TINST, 4),
                       "source": "SQ_INSTS_VALU in profiles/stage_kernel_blast256_fp64_traffic.md"}
    if args.stream:
        out["roofline"]["stream_copy_GBs"] = round(stream_copy(), 1)
    if args.forms:
        # the FP32 forms on the same workload (error against FP64 in tests/)
        out["forms"] = {
            "fp32_kahan_rk2": form_line(F, p, Q0, F.FP32, args.steps, args.warmup, 140.0,
                                        STAGE_BYTES["fp32_kahan"]),
            "fp32_plain_rk2": form_line(F, p, Q0, F.FP32_PLAIN, args.steps, args.warmup, 100.0,
                                        STAGE_BYTES["fp32_plain"]),
            "fp32_error_vs_fp64": fp32_error_vs_fp64(F, p, Q0, min(60, args.steps + args.warmup)),
        }
    if args.ensemble > 0:
        out.setdefault("forms", {})["fp32_ensemble_rt"] = ensemble_form(F, args.ensemble, args.ensemble_steps,
                                                                        loop=bool(args.ensemble_loop))
    if args.stratified > 0:
        out.setdefault("forms", {})["stratified_tc_visc"] = stratified_form(F, args.stratified,
                                                                             args.stratified_window)
    if args.adjoint > 0:
        a64 = adjoint_form(F, args.adjoint, args.adjoint_steps, ndirs=args.adjoint_dirs)
        a32 = adjoint_form(F, args.adjoint, args.adjoint_steps, precision=F.FP32_PLAIN, fd=False,
                           ref=a64["_ref"])
        a64.pop("_ref")
        a32.pop("_ref")
        out.setdefault("forms", {})["adjoint_fp64"] = a64
        out["forms"]["adjoint_fp32"] = a32
    if args.cpu_budget > 0:
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
