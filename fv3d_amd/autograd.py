"""torch.autograd wrapper of the frozen-dt run and its discrete adjoint, for
checking gradients only (``torch.autograd.gradcheck`` and the like).

``FrozenRun.apply(rho0, kappa, mu, setup)`` returns the final density field
of a run of ``len(setup.dts)`` steps with the dt sequence frozen, starting
from primitives (rho0, setup.velocity, setup.pressure).  Its backward calls
``fv3d_adjoint_run`` (include/fv3d_hip.h) and returns dJ/drho0, dJ/dkappa and
dJ/dmu.  Tensors are CPU float64: the GPU work runs inside libfv3d_hip, and
torch's own device runtime is never touched.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

import fv3d_amd as F


@dataclass
class RunSetup:
    params: F.Params          # numerics; kappa and mu come from the inputs
    velocity: np.ndarray      # (Nz, Ny, Nx, 3) initial velocity
    pressure: np.ndarray      # (Nz, Ny, Nx) initial pressure
    dts: np.ndarray           # frozen dt sequence
    precision: int = F.FP64
    K: int = 20               # checkpoint interval


def _prims(setup: RunSetup, rho0: np.ndarray) -> np.ndarray:
    Q = np.empty(setup.params.shape_domain)
    Q[..., 0] = rho0
    Q[..., 1:4] = setup.velocity
    Q[..., 4] = setup.pressure
    return Q


class FrozenRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho0, kappa, mu, setup: RunSetup):
        p = setup.params.replace(kappa=float(kappa), mu=float(mu))
        with F.Simulation(p, setup.precision) as sim:
            sim.set_prim(_prims(setup, rho0.detach().numpy()), with_ghosts=False)
            sim.forward_frozen(setup.dts)
            rhoT = sim.get_prim(with_ghosts=False)[..., 0].copy()
        ctx.save_for_backward(rho0, kappa, mu)
        ctx.setup = setup
        return torch.from_numpy(rhoT)

    @staticmethod
    def backward(ctx, grad_rhoT):
        rho0, kappa, mu = ctx.saved_tensors
        setup = ctx.setup
        p = setup.params.replace(kappa=float(kappa), mu=float(mu))
        lamT = np.zeros(p.shape_domain)
        lamT[..., 0] = grad_rhoT.detach().numpy()  # rho_T is the conserved density
        with F.Simulation(p, setup.precision) as sim:
            sim.set_prim(_prims(setup, rho0.detach().numpy()), with_ghosts=False)
            lam0, dk, dm = sim.adjoint(setup.dts, lamT, setup.K)
        # U0 = primToCons(rho0, v, p): d(rho u)/drho = u, dE/drho = |v|^2/2
        v = setup.velocity
        g = lam0[..., 0] + (lam0[..., 1:4] * v).sum(-1) + lam0[..., 4] * 0.5 * (v * v).sum(-1)
        return (torch.from_numpy(g), torch.tensor(dk, dtype=torch.float64),
                torch.tensor(dm, dtype=torch.float64), None)
