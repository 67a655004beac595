"""GPU parity of the functor-level entry points: each reference pass called
on its own through the C ABI on device arrays (fv3d_array), against the
same pass of the oracle on the same arrays, bit for bit.

Covers BoundaryManager::fillBoundaries, consToPrim/primToCons over whole
arrays, UpdateFunctor::computeSlopes, computeFluxesAndUpdate, euler_step
and update, ThermalConductionFunctor::applyThermalConduction,
ViscosityFunctor::applyViscosity, ComputeDtFunctor::computeDt, and the
fused step with a caller-supplied dt (UpdateFunctor::update(Q, Unew, dt)).
"""
import numpy as np
import pytest

import fv3d_amd as F
from fv3d_amd import Params
from oracle.oracle import Oracle
from tests.helpers import dom, max_rel
from tests.test_gpu_parity import smooth_random, c91_params

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    if not np.array_equal(a, b, equal_nan=True):
        bad = np.argwhere((a != b) & ~(np.isnan(a) & np.isnan(b)))
        raise AssertionError(f"{what}: {len(bad)} entries differ, first at {bad[0].tolist()}, "
                             f"max rel {max_rel(a, b):.3e}")


def _cfg(name):
    base = dict(Nx=12, Ny=10, Nz=8, tend=10.0, kappa=0.03, mu=0.02)
    if name == "blast_reflecting":
        kw = dict(boundary_x=F.BC_REFLECTING, boundary_y=F.BC_REFLECTING, boundary_z=F.BC_REFLECTING,
                  reconstruction=F.PLM, riemann_solver=F.HLLC, time_stepping=F.TS_RK2)
    elif name == "periodic_tc_visc":
        kw = dict(boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC, boundary_z=F.BC_PERIODIC,
                  reconstruction=F.PLM, riemann_solver=F.HLLC, time_stepping=F.TS_RK2,
                  thermal_conductivity_active=1, viscosity_active=1)
    elif name == "hll_fixed_t_gravity":
        kw = dict(xmax=2.0, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC,
                  boundary_z=F.BC_REFLECTING, reconstruction=F.PLM, riemann_solver=F.HLL,
                  time_stepping=F.TS_RK2, gravity=1, g=1.0, thermal_conductivity_active=1,
                  bctc_zmin=F.BCTC_FIXED_TEMPERATURE, bctc_zmin_value=1.0,
                  bctc_zmax=F.BCTC_FIXED_TEMPERATURE, bctc_zmax_value=1.5, viscosity_active=1)
    elif name == "pcm_wb_gravity":
        kw = dict(boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC, boundary_z=F.BC_REFLECTING,
                  reconstruction=F.PCM_WB, well_balanced_flux_at_z_bc=1, gravity=1, g=-1.0,
                  time_stepping=F.TS_EULER)
    elif name == "absorbing_fixed_gradient_pcm":
        kw = dict(xmax=0.7, ymax=1.3, boundary_x=F.BC_ABSORBING, boundary_y=F.BC_REFLECTING,
                  boundary_z=F.BC_ABSORBING, reconstruction=F.PCM, riemann_solver=F.HLLC,
                  time_stepping=F.TS_RK2, thermal_conductivity_active=1,
                  bctc_zmin=F.BCTC_FIXED_GRADIENT, bctc_zmin_value=0.3,
                  bctc_zmax=F.BCTC_FIXED_GRADIENT, bctc_zmax_value=-0.2)
    elif name == "b02":
        kw = dict(boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC, boundary_z=F.BC_REFLECTING,
                  reconstruction=F.PLM, time_stepping=F.TS_RK2, thermal_conductivity_active=1,
                  thermal_conductivity_mode=F.TCM_B02, kappa=0.05, b02_zmid=0.5, b02_kappa1=1.0,
                  b02_kappa2=0.1, b02_thickness=0.1, bctc_zmin=F.BCTC_FIXED_TEMPERATURE,
                  bctc_zmin_value=1.2, bctc_zmax=F.BCTC_FIXED_TEMPERATURE, bctc_zmax_value=0.8)
    elif name == "c91":
        # Nz = 16: with 8 levels this C91 setting drives pressures negative
        # within three steps (in the oracle too)
        return c91_params(Nx=12, Ny=10, Nz=16, time_stepping=F.TS_RK2)
    else:
        raise KeyError(name)
    base.update(kw)
    return Params.defaults().replace(**base)


CFGS = ["blast_reflecting", "periodic_tc_visc", "hll_fixed_t_gravity", "pcm_wb_gravity",
        "absorbing_fixed_gradient_pcm", "b02", "c91"]


def _initial(p, name):
    Q0 = F.problems.c91(p) if name == "c91" else smooth_random(p, 11, amp_u=0.2)
    return Oracle(p).fill_boundaries(Q0)


@pytest.mark.parametrize("name", CFGS)
def test_functor_passes_bitwise(gpu_available, name):
    """Each pass on its own: fill, prim<->cons, slopes, flux update,
    conduction, viscosity and dt, against the oracle's pass."""
    p = _cfg(name)
    o = Oracle(p)
    Q0 = _initial(p, name)
    with F.Simulation(p) as sim:
        # fillBoundaries on a state whose ghosts hold other values
        Qraw = Q0.copy()
        Qraw[:p.kbeg] = 7.0
        Qraw[:, :, -1] = -3.0
        Qa = sim.array(Qraw)
        sim.fill_boundaries_array(Qa)
        Q = o.fill_boundaries(Qraw.copy())
        _same(Qa.download(), Q, "fillBoundaries")
        # primToCons / consToPrim over range_tot
        Ua = sim.array()
        sim.prim_to_cons_array(Qa, Ua)
        U = o.prim_to_cons(Q)
        _same(Ua.download(), U, "primToCons")
        Qb = sim.array()
        sim.cons_to_prim_array(Ua, Qb)
        _same(Qb.download(), o.cons_to_prim(U), "consToPrim")
        # computeDt
        dt_g, inv_g = sim.compute_dt_array(Qa)
        dt_o, inv_o = o.compute_dt(Q)
        assert dt_g == dt_o and inv_g == inv_o
        # computeSlopes
        sx, sy, sz = sim.array(), sim.array(), sim.array()
        sim.compute_slopes(Qa, sx, sy, sz)
        so = o.compute_slopes(Q)
        for a, b, ax in zip((sx, sy, sz), so, "xyz"):
            _same(a.download(), b, f"slopes{ax}")
        # computeFluxesAndUpdate, then conduction and viscosity on the same Unew
        Unew = U.copy()
        o.fluxes_and_update(Q, *so, Unew, dt_o)
        sim.fluxes_and_update(Qa, sx, sy, sz, Ua, dt_o)
        _same(Ua.download(), Unew, "computeFluxesAndUpdate")
        o.thermal_conduction(Q, Unew, dt_o)
        sim.apply_tc(Qa, Ua, dt_o)
        _same(Ua.download(), Unew, "applyThermalConduction")
        o.viscosity(Q, Unew, dt_o)
        sim.apply_visc(Qa, Ua, dt_o)
        _same(Ua.download(), Unew, "applyViscosity")


@pytest.mark.parametrize("name", CFGS)
def test_functor_update_loop_bitwise(gpu_available, name):
    """The reference's time loop driven through the array functors
    (consToPrim, computeDt, UpdateFunctor::update, U = Unew) for four steps,
    against the oracle's passes: Q (which update mutates) and Unew after
    every step.  One euler_step call is compared too."""
    p = _cfg(name)
    o = Oracle(p)
    Q0 = _initial(p, name)
    U = o.prim_to_cons(Q0)
    Q = o.zeros()
    with F.Simulation(p) as sim:
        Ua, Qa, Ua_new = sim.array(U), sim.array(), sim.array()
        for step in range(4):
            Unew = U.copy()
            Ua_new.copy_from(Ua)
            Q = o.cons_to_prim(U)
            sim.cons_to_prim_array(Ua, Qa)
            dt_o, _ = o.compute_dt(Q)
            dt_g, _ = sim.compute_dt_array(Qa)
            assert dt_g == dt_o, step
            o.update(Q, Unew, dt_o)
            sim.update(Qa, Ua_new, dt_g)
            _same(Qa.download(), Q, f"Q after update, step {step}")
            _same(Ua_new.download(), Unew, f"Unew after update, step {step}")
            U = Unew
            Ua.copy_from(Ua_new)
        # one Euler stage on its own
        Q = o.cons_to_prim(U)
        sim.cons_to_prim_array(Ua, Qa)
        dt_o, _ = o.compute_dt(Q)
        Unew = U.copy()
        Ua_new.copy_from(Ua)
        o.euler_step(Q, Unew, dt_o)
        sim.euler_step(Qa, Ua_new, dt_o)
        _same(Qa.download(), Q, "Q after euler_step")
        _same(Ua_new.download(), Unew, "Unew after euler_step")


@pytest.mark.parametrize("name", ["blast_reflecting", "periodic_tc_visc", "pcm_wb_gravity"])
@pytest.mark.parametrize("precision", [F.FP64, F.FP32])
def test_step_with_caller_dt(gpu_available, name, precision):
    """fv3d_step_dt: the fused step with the caller's dt.  FP64: a run with a
    chosen dt sequence (90% of the CFL dt) equals the oracle's update with
    that dt; the step after it reads dt from the state again.  Both
    precisions: step_dt(dt of the state) is bit-equal to step()."""
    p = _cfg(name)
    o = Oracle(p)
    Q0 = _initial(p, name)
    with F.Simulation(p, precision) as a, F.Simulation(p, precision) as b:
        a.set_prim(Q0)
        b.set_prim(Q0)
        for _ in range(3):
            dt, _ = a.compute_dt()
            a.step()
            b.step_dt(dt)
        _same(b.get_cons(), a.get_cons(), "step_dt(dt) vs step()")
    if precision != F.FP64:
        return
    U = o.prim_to_cons(Q0)
    with F.Simulation(p) as sim:
        sim.set_prim(Q0)
        for s in range(3):
            Q = o.cons_to_prim(U)
            dt, _ = o.compute_dt(Q)
            dt_s = 0.9 * dt
            Unew = U.copy()
            o.update(Q, Unew, dt_s)
            U = Unew
            sim.step_dt(dt_s)
            _same(dom(p, sim.get_cons()), dom(p, U), f"step_dt, step {s}")
        # the next computeDt and plain step follow the state again
        Q = o.cons_to_prim(U)
        dt, inv = o.compute_dt(Q)
        dt_g, inv_g = sim.compute_dt()
        assert dt_g == dt and inv_g == inv
        Unew = U.copy()
        o.update(Q, Unew, dt)
        sim.step()
        _same(dom(p, sim.get_cons()), dom(p, Unew), "step after step_dt")


@pytest.mark.parametrize("precision", [F.FP64, F.FP32])
def test_state_array_bridge(gpu_available, precision):
    """fv3d_get_cons_array / fv3d_set_cons_array move the fused path's state
    to and from a device array without the host; set_cons_array restarts
    the step count and dt like set_cons."""
    p = _cfg("periodic_tc_visc")
    Q0 = _initial(p, "periodic_tc_visc")
    with F.Simulation(p, precision) as sim, F.Simulation(p, precision) as ref:
        sim.set_prim(Q0)
        sim.step(2)
        Ua = sim.array()
        sim.get_cons_array(Ua)
        host = Ua.download()
        U2 = sim.get_cons()
        _same(dom(p, host), dom(p, U2), "get_cons_array")
        assert np.all(host[:p.kbeg] == 0.0)  # ghosts untouched
        # the same state through the host: set_cons and set_cons_array agree
        ref.set_cons(U2)
        sim.step(3)
        sim.set_cons_array(Ua)
        assert sim.step_count == 0
        _same(dom(p, sim.get_cons()), dom(p, ref.get_cons()), "set_cons_array")
        assert sim.compute_dt() == ref.compute_dt()
        sim.step(2)
        ref.step(2)
        _same(dom(p, sim.get_cons()), dom(p, ref.get_cons()), "steps after set_cons_array")


def test_functor_errors(gpu_available):
    """Arrays of another context and non-finite dt are rejected with
    FV3D_EINVAL and a message; nothing is launched."""
    p = _cfg("blast_reflecting")
    with F.Simulation(p) as a, F.Simulation(p) as b:
        Qa, Ub = a.array(), b.array()
        with pytest.raises(F.FV3DError, match="EINVAL"):
            a.apply_tc(Qa, Ub, 1e-3)
        with pytest.raises(F.FV3DError, match="EINVAL"):
            a.apply_visc(Qa, a.array(), float("nan"))
        a.set_prim(_initial(p, "blast_reflecting"))
        with pytest.raises(F.FV3DError, match="EINVAL"):
            a.step_dt(float("inf"))
