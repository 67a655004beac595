"""The FP32 forms against the FP64 path (which is bit-equal to the oracle).

FP32 state and fluxes; FV3D_FP32 accumulates the RK2 step with Kahan
compensation, FV3D_FP32_PLAIN does not.  Errors are relative to the FP64
field: L1 = mean|a-b| / mean|b|, Linf = max|a-b| / max|b|.  Bounds are a few
times the errors observed on MI355X (1e-7 .. 1e-6).
"""
import numpy as np
import pytest

import fv3d_amd as F
from fv3d_amd import Params
from oracle.oracle import Oracle
from tests.helpers import sod_c1, blast_small, dom

pytestmark = pytest.mark.gpu


def _run(p, Q0, prec, n):
    with F.Simulation(p, prec) as s:
        s.set_prim(Q0)
        s.step(n)
        s.synchronize()
        return s.get_prim(with_ghosts=False), s.get_cons(with_ghosts=False)


def _errs(a, b):
    l1 = [np.mean(np.abs(a[..., f] - b[..., f])) / max(np.mean(np.abs(b[..., f])), 1e-30) for f in range(5)]
    li = [np.max(np.abs(a[..., f] - b[..., f])) / max(np.max(np.abs(b[..., f])), 1e-30) for f in range(5)]
    return np.array(l1), np.array(li)


def _smooth(p):
    Q = np.zeros(p.shape_ghosted)
    d = dom(p, Q)
    z, y, x = np.meshgrid(np.arange(p.Nz), np.arange(p.Ny), np.arange(p.Nx), indexing="ij")
    d[..., 0] = 1 + 0.2 * np.sin(2 * np.pi * x / p.Nx) * np.cos(2 * np.pi * y / p.Ny)
    d[..., 1] = 0.3 * np.sin(2 * np.pi * z / p.Nz)
    d[..., 2] = 0.2
    d[..., 3] = 0.1 * np.cos(2 * np.pi * x / p.Nx)
    d[..., 4] = 1.0
    return Q


@pytest.mark.parametrize("prec", [F.FP32, F.FP32_PLAIN])
def test_sod_fp32_against_fp64(gpu_available, prec):
    p = sod_c1()
    Q0 = Oracle(p).init("sod_x")
    q64, _ = _run(p, Q0, F.FP64, 140)
    q32, _ = _run(p, Q0, prec, 140)
    l1, li = _errs(q32, q64)
    assert l1[[0, 1, 4]].max() < 2e-6 and li[[0, 1, 4]].max() < 1e-5, (l1, li)


def test_blast_fp32_against_fp64(gpu_available):
    p = blast_small(48, tend=1.0)
    Q0 = F.problems.blast(p, radius=0.1, rho_in=1.0, p_in=10.0, rho_out=1.0, p_out=0.1)
    q64, _ = _run(p, Q0, F.FP64, 40)
    q32, _ = _run(p, Q0, F.FP32, 40)
    l1, li = _errs(q32, q64)
    assert l1.max() < 5e-6 and li.max() < 5e-5, (l1, li)


def test_kahan_no_worse_than_plain_and_conserves(gpu_available):
    p = Params.defaults().replace(Nx=48, Ny=48, Nz=48, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC,
                                  boundary_z=F.BC_PERIODIC, reconstruction=F.PLM, time_stepping=F.TS_RK2)
    Q0 = _smooth(p)
    q64, u64 = _run(p, Q0, F.FP64, 300)
    qk, uk = _run(p, Q0, F.FP32, 300)
    qp, up = _run(p, Q0, F.FP32_PLAIN, 300)
    lk, _ = _errs(qk, q64)
    lp, _ = _errs(qp, q64)
    assert lk.max() < 2e-6
    assert lk.sum() <= 1.1 * lp.sum(), (lk, lp)
    for u in (uk, up):
        drift = abs(u[..., 0].sum(dtype=np.float64) - u64[..., 0].sum()) / u64[..., 0].sum()
        assert drift < 1e-6


def test_stratified_conduction_viscosity_fp32(gpu_available):
    """The conduction/viscosity workload family at its resolution (dx = 1/128)
    on a narrow column, with CFL 0.4 (at the reference's 0.8 its conduction
    time step is unstable and the FP64 run itself goes to NaN)."""
    p = Params.defaults().replace(Nx=32, Ny=32, Nz=128, xmax=0.25, ymax=0.25, boundary_x=F.BC_PERIODIC,
                                  boundary_y=F.BC_PERIODIC, boundary_z=F.BC_REFLECTING, reconstruction=F.PLM,
                                  riemann_solver=F.HLL, time_stepping=F.TS_RK2, gravity=1, g=1.0, CFL=0.4,
                                  thermal_conductivity_active=1, kappa=0.07, bctc_zmin=F.BCTC_FIXED_TEMPERATURE,
                                  bctc_zmin_value=1.0, bctc_zmax=F.BCTC_FIXED_TEMPERATURE, bctc_zmax_value=1.5,
                                  viscosity_active=1, mu=0.005)
    Q0 = F.problems.polytrope_convection(p)
    q64, _ = _run(p, Q0, F.FP64, 150)
    q32, _ = _run(p, Q0, F.FP32, 150)
    l1, li = _errs(q32, q64)
    assert l1[[0, 4]].max() < 2e-6, l1
    # velocities are O(1e-3) noise on a hydrostatic layer: bound their
    # absolute error against the sound speed (about 1.3)
    assert np.max(np.abs(q32[..., 1:4] - q64[..., 1:4])) < 1e-4


def test_fp32_stage_prim_and_dt(gpu_available):
    """The final-snapshot source (prim(U*)) and the FP64 dt of the FP32 form."""
    p = sod_c1()
    Q0 = Oracle(p).init("sod_x")
    with F.Simulation(p, F.FP64) as a, F.Simulation(p, F.FP32) as b:
        a.set_prim(Q0)
        b.set_prim(Q0)
        for _ in range(10):
            da, _ = a.compute_dt()
            db, _ = b.compute_dt()
            assert abs(da - db) / da < 1e-6
            a.step()
            b.step()
        sa = dom(p, a.get_stage_prim())
        sb = dom(p, b.get_stage_prim())
    assert np.max(np.abs(sa - sb)) < 1e-5
