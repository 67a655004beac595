"""Shared test helpers: parameter sets of the parity cases and the
main.cpp time loop driven through the GPU session."""
from __future__ import annotations

import numpy as np

import fv3d_amd as F
from fv3d_amd import Params


def sod_c1(**kw):
    """Sod along x, 256x4x4, gamma 1.4, PLM, HLLC, RK2, CFL 0.8, absorbing x,
    periodic y/z, t_end 0.2 (the reference's CPU-runnable configuration)."""
    base = dict(Nx=256, Ny=4, Nz=4, gamma0=1.4, reconstruction=F.PLM, riemann_solver=F.HLLC,
                time_stepping=F.TS_RK2, CFL=0.8, boundary_x=F.BC_ABSORBING,
                boundary_y=F.BC_PERIODIC, boundary_z=F.BC_PERIODIC, tend=0.2, save_freq=0.02)
    base.update(kw)
    return Params.defaults().replace(**base)


def blast_small(N=32, **kw):
    base = dict(Nx=N, Ny=N, Nz=N, gamma0=5.0 / 3.0, reconstruction=F.PLM, riemann_solver=F.HLLC,
                time_stepping=F.TS_RK2, CFL=0.8, boundary_x=F.BC_REFLECTING,
                boundary_y=F.BC_REFLECTING, boundary_z=F.BC_REFLECTING, tend=0.02, save_freq=0.01)
    base.update(kw)
    return Params.defaults().replace(**base)


def gpu_main_loop(params, Q0, max_steps=None, precision=F.FP64):
    """main.cpp:43-76 with the device session: returns (steps, snaps) in the
    format of oracle.MainLoop.run."""
    p = params
    with F.Simulation(p, precision) as sim:
        sim.set_prim(Q0, with_ghosts=True)
        t, ite, next_save, next_log = 0.0, 0, 0.0, 0
        steps, snaps = [], []
        while t + p.epsilon < p.tend and (max_steps is None or len(steps) < max_steps):
            save_needed = t + p.epsilon > next_save
            dt, inv = sim.compute_dt()
            logged = next_log == 0
            next_log = p.log_frequency if next_log == 0 else next_log - 1
            if save_needed:
                snaps.append((ite, t, sim.get_prim(with_ghosts=True)))
                ite += 1
                next_save += p.save_freq
            sim.step()
            steps.append((t, dt, inv, logged))
            t += dt
        snaps.append((ite, t, sim.get_stage_prim(with_ghosts=True)))
    return steps, snaps


def dom(p, Q):
    return Q[p.kbeg:p.kend, p.jbeg:p.jend, p.ibeg:p.iend]


def max_rel(a, b):
    """max |a-b| / max(|b|, tiny) over all entries."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / den)) if a.size else 0.0
