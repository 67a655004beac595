"""The benchmarked workloads at their full sizes, through checks that do not
depend on size: bit equality with the oracle over the first steps of the
256^3 blast, exact mirror symmetry of the blast, conservation of mass,
momentum and energy to round-off on a periodic 256^3 box, and a box whose
state arrays pass 2^31 elements."""
import numpy as np
import pytest

import fv3d_amd as F
from oracle.oracle import Oracle
from tests.helpers import blast_small, dom

pytestmark = pytest.mark.gpu


def blast256():
    p = blast_small(256, tend=0.1)
    return p, F.problems.blast(p, radius=0.1, rho_in=1.0, p_in=10.0, rho_out=1.0, p_out=0.1)


def test_blast_256_first_steps_match_oracle(gpu_available):
    """3 RK2 steps of the 256^3 blast: state and dt sequence bit-equal to the
    oracle (the reference's pass structure) on the full grid."""
    p, Q0 = blast256()
    print("blast 256^3: GPU", flush=True)
    with F.Simulation(p, F.FP64) as sim:
        sim.set_prim(Q0)
        dts = []
        for _ in range(3):
            dts.append(sim.compute_dt()[0])
            sim.step()
        Qg = sim.get_prim(with_ghosts=False)
    print("blast 256^3: oracle", flush=True)
    o = Oracle(p)
    U, _, dts_o, _ = o.run(Q0, 3, stop_at_tend=False)
    Qo = dom(p, o.cons_to_prim(U))
    assert np.array_equal(np.array(dts), dts_o)
    assert np.array_equal(Qg, Qo)


def test_blast_256_mirror_symmetry(gpu_available):
    """The blast is symmetric under each axis reflection, and the scheme's
    arithmetic is too (a reflected face solves the mirrored problem with
    negated velocities): 40 steps keep every mirror pair bit-equal.  Axis
    exchange is not a symmetry of the reference's scheme: its HLLC kinetic
    energy drops the second transverse velocity, which is w in the x and y
    sweeps but u in the z sweep (RiemannSolvers.h:52-55), so after one
    step x<->y transposition already differs in energy, and only there."""
    p, Q0 = blast256()
    with F.Simulation(p, F.FP64) as sim:
        sim.set_prim(Q0)
        sim.step(1)
        Q1 = sim.get_cons(with_ghosts=False)
        sim.step(39)
        Q = sim.get_prim(with_ghosts=False)
    for ax in (0, 1, 2):
        m = np.flip(Q, axis=ax).copy()
        m[..., 3 - ax] = -m[..., 3 - ax]  # the normal velocity changes sign
        assert np.array_equal(Q, m), ax
    T = np.transpose(Q1, (0, 2, 1, 3))[..., [0, 2, 1, 3, 4]]
    asym = np.abs(Q1 - T).reshape(-1, 5).max(axis=0)
    assert asym[:4].max() <= 1e-14 and asym[4] > 1e-6, asym


@pytest.mark.parametrize("precision", [F.FP64, F.FP32])
def test_periodic_256_conservation(gpu_available, precision):
    """Mass, momentum and energy on a periodic 256^3 box (hydro, PLM +
    HLLC, RK2): the flux form conserves them to round-off (FP64 sums)."""
    p = blast_small(256, tend=1.0, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC,
                    boundary_z=F.BC_PERIODIC)
    Q0 = F.problems.blast(p, radius=0.15, rho_in=1.0, p_in=10.0, rho_out=1.0, p_out=0.1)
    Q0[p.kbeg:p.kend, p.jbeg:p.jend, p.ibeg:p.iend, 1] = 0.3  # uniform drift
    with F.Simulation(p, precision) as sim:
        sim.set_prim(Q0)
        U0 = sim.get_cons(with_ghosts=False).reshape(-1, 5).sum(axis=0)
        sim.step(20)
        U1 = sim.get_cons(with_ghosts=False).reshape(-1, 5).sum(axis=0)
    # round-off of 2^24 cell updates per stage: the oracle drifts by
    # 1.4e-14 (32^3) and 1.8e-13 (64^3) relative on the same setup
    tol = 1e-10 if precision == F.FP64 else 2e-6
    for n, name in enumerate(["mass", "x-momentum", "y-momentum", "z-momentum", "energy"]):
        ref = max(abs(U0[n]), abs(U0[0]))
        assert abs(U1[n] - U0[n]) <= tol * ref, (name, U0[n], U1[n])


@pytest.mark.parametrize("precision", [F.FP64, F.FP32])
def test_stratified_full_size_conserves_mass(gpu_available, precision):
    """The stratified-convection workload at its full size (256 x 256 x 128,
    HLL, conduction with fixed-temperature walls, viscosity, gravity,
    periodic x/y, reflecting z), 20 steps: gravity, conduction and the wall
    terms change momentum and energy, but no term moves mass across the
    reflecting walls, so the total mass stays to round-off (FP64 sums).
    Momentum is not checked: the reference's viscous z branch (Viscosity.h
    :122-133) does not evaluate a z face's stress alike from its two cells."""
    p = F.Params.defaults().replace(
        Nx=256, Ny=256, Nz=128, xmax=2.0, ymax=2.0, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC,
        boundary_z=F.BC_REFLECTING, reconstruction=F.PLM, riemann_solver=F.HLL, time_stepping=F.TS_RK2,
        gamma0=5.0 / 3.0, gravity=1, g=1.0, CFL=0.4, thermal_conductivity_active=1, kappa=0.07,
        bctc_zmin=F.BCTC_FIXED_TEMPERATURE, bctc_zmin_value=1.0, bctc_zmax=F.BCTC_FIXED_TEMPERATURE,
        bctc_zmax_value=1.5, viscosity_active=1, mu=0.005, tend=100.0)
    Q0 = F.problems.polytrope_convection(p)
    with F.Simulation(p, precision) as sim:
        sim.set_prim(Q0)
        m0 = sim.get_cons(with_ghosts=False)[..., 0].sum()
        sim.step(20)
        U1 = sim.get_cons(with_ghosts=False)
    assert np.isfinite(U1).all()
    tol = 1e-12 if precision == F.FP64 else 2e-6
    assert abs(U1[..., 0].sum() - m0) <= tol * m0, (m0, U1[..., 0].sum())


def _pulse(shape, at, seed=7):
    """Background rho = 1, p = 1 at rest, plus an 8^3 block of random
    density, pressure and velocity whose low corner is `at` (z, y, x)."""
    rng = np.random.default_rng(seed)
    Q = np.empty(shape + (5,))
    Q[...] = np.array([1.0, 0.0, 0.0, 0.0, 1.0])
    b = Q[at[0]:at[0] + 8, at[1]:at[1] + 8, at[2]:at[2] + 8]
    b[..., 0] = rng.uniform(0.5, 2.0, (8, 8, 8))
    b[..., 1:4] = rng.normal(0.0, 0.5, (8, 8, 8, 3))
    b[..., 4] = rng.uniform(0.5, 4.0, (8, 8, 8))
    return Q


@pytest.mark.parametrize("precision", [F.FP64, F.FP32])
def test_state_past_32bit_offsets(gpu_available, precision):
    """A 768 x 768 x 736 periodic box (434 M cells): each state array holds
    2.17e9 values, so the last planes' offsets pass 2^31 elements (and
    2^34 bytes in FP64).  A pulse placed in those planes, after two steps,
    matches bit for bit the same pulse in a 32^3 box: the update of a cell
    depends on its neighbourhood only, and the pulse sets dt in both."""
    base = dict(gamma0=5.0 / 3.0, reconstruction=F.PLM, riemann_solver=F.HLLC, time_stepping=F.TS_RK2,
                CFL=0.8, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC, boundary_z=F.BC_PERIODIC,
                tend=1.0)
    # the same cell size, 1/32, in both boxes
    small = F.Params.defaults().replace(Nx=32, Ny=32, Nz=32, **base)
    big = F.Params.defaults().replace(Nx=768, Ny=768, Nz=736, xmax=24.0, ymax=24.0, zmax=23.0, **base)
    assert (small.dx, small.dy, small.dz) == (big.dx, big.dy, big.dz)
    assert 5 * big.Nx * big.Ny * big.Nz > 2 ** 31
    at_s, at_b = (12, 12, 12), (big.Nz - 20, big.Ny - 20, big.Nx - 20)

    def run(p, at):
        Q0 = _pulse((p.Nz, p.Ny, p.Nx), at)
        with F.Simulation(p, precision) as sim:
            sim.set_prim(Q0, with_ghosts=False)
            del Q0
            dts = []
            for _ in range(2):
                dts.append(sim.compute_dt()[0])
                sim.step()
            Q = sim.get_prim(with_ghosts=False)
        w = Q[at[0] - 10:at[0] + 18, at[1] - 10:at[1] + 18, at[2] - 10:at[2] + 18].copy()
        return dts, w

    print(f"32-bit offsets: small box, precision {precision}", flush=True)
    dts_s, w_s = run(small, at_s)
    print("32-bit offsets: 768x768x736 box", flush=True)
    dts_b, w_b = run(big, at_b)
    assert dts_s == dts_b
    assert np.array_equal(w_s, w_b)
    assert not np.array_equal(w_s[0], w_s[10])  # the pulse moved into the compared window


@pytest.mark.timeout(1200)
def test_blast_256_full_run_matches_oracle(gpu_available, capsys):
    """The whole 256^3 blast run to t = 0.1 through the reference's time
    loop (main.cpp:51-76) in lockstep with the oracle's: every dt and every
    snapshot (save_freq 0.01, plus the final stale one) bit-equal,
    NaN-aware.  The run passes through the reference scheme's own negative
    pressures (from step 91) and NaN cells; the reference's Max reducer
    leaves those out of dt, and so do both sides here."""
    p, Q0 = blast256()
    p = p.replace(save_freq=0.01)
    o = Oracle(p)
    U = o.prim_to_cons(Q0)
    Unew = o.zeros()
    t, ite, next_save, steps = 0.0, 0, 0.0, 0

    def same(a, b, what):
        assert np.array_equal(a, b, equal_nan=True), what

    with F.Simulation(p, F.FP64) as sim, pytest.warns(F.NonFiniteWarning):
        sim.set_prim(Q0)
        while t + p.epsilon < p.tend:
            Unew[...] = U
            save = t + p.epsilon > next_save
            Q = o.cons_to_prim(U)
            dt_o, inv_o = o.compute_dt(Q)
            dt_g, inv_g = sim.compute_dt()
            assert dt_g == dt_o and inv_g == inv_o, (steps, dt_g, dt_o)
            if save:
                same(sim.get_prim(with_ghosts=False), dom(p, Q), f"snapshot ite_{ite} at step {steps}")
                ite += 1
                next_save += p.save_freq
            o.update(Q, Unew, dt_o)
            U, Unew = Unew, U
            sim.step()
            t += dt_o
            steps += 1
            if steps % 20 == 0:
                with capsys.disabled():
                    print(f"\nblast 256^3 full run: step {steps}, t={t:.5f}, {ite} snapshots equal",
                          flush=True)
        # the final snapshot: the stale Q of main.cpp:76, prim(U*) of the last step
        same(sim.get_stage_prim(with_ghosts=False), dom(p, Q), "final stale snapshot")
    # saves at t = 0, 0.01, ..., 0.09 inside the loop, then the final one
    assert 250 <= steps <= 280 and ite == 10, (steps, ite)
