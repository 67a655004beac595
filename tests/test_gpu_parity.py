"""GPU parity: libfv3d_hip through its C ABI against the oracle (the CPU
restatement of the reference), on the same inputs.

The FP64 bar is bit equality: the kernels evaluate the reference's
expression trees with -ffp-contract=off and correctly rounded division and
square root, so states, dt sequences and snapshots must match exactly.
(The contract allows 1e-12 relative; equality is what is observed and is
what these tests assert.)
"""
import numpy as np
import pytest

import fv3d_amd as F
from fv3d_amd import Params
from oracle.oracle import Oracle, MainLoop
from tests.helpers import sod_c1, blast_small, gpu_main_loop, dom, max_rel

pytestmark = pytest.mark.gpu


def _assert_same(a, b, what):
    if not np.array_equal(a, b, equal_nan=True):
        diff = np.argwhere((a != b) & ~(np.isnan(a) & np.isnan(b)))
        raise AssertionError(f"{what}: {len(diff)} entries differ, first at {diff[0].tolist()}, "
                             f"max rel {max_rel(a, b):.3e}")


def _compare_runs(p, Q0, max_steps=None):
    steps_g, snaps_g = gpu_main_loop(p, Q0, max_steps)
    steps_o, snaps_o = MainLoop(p, Q0).run(max_steps)
    assert len(steps_g) == len(steps_o)
    dts_g = np.array([s[1] for s in steps_g])
    dts_o = np.array([s[1] for s in steps_o])
    _assert_same(dts_g, dts_o, "dt sequence")
    assert len(snaps_g) == len(snaps_o)
    for (ig, tg, Qg), (io, to, Qo) in zip(snaps_g, snaps_o):
        assert ig == io and tg == to
        _assert_same(dom(p, Qg), dom(p, Qo), f"snapshot ite_{ig}")
    return steps_o, snaps_o


def smooth_random(p, seed=0, amp_u=0.3):
    rng = np.random.default_rng(seed)
    Q = np.zeros(p.shape_ghosted)
    z, y, x = np.meshgrid(np.arange(p.Nz), np.arange(p.Ny), np.arange(p.Nx), indexing="ij")
    d = Q[p.kbeg:p.kend, p.jbeg:p.jend, p.ibeg:p.iend]
    ph = rng.uniform(0, 2 * np.pi, 8)
    kx, ky, kz = 2 * np.pi / p.Nx, 2 * np.pi / p.Ny, 2 * np.pi / p.Nz
    d[..., 0] = 1.0 + 0.2 * np.sin(kx * x + ph[0]) * np.cos(ky * y + ph[1]) + 0.1 * np.sin(kz * z + ph[2])
    d[..., 1] = amp_u * np.sin(ky * y + ph[3]) + 0.05 * rng.standard_normal(x.shape)
    d[..., 2] = amp_u * np.sin(kz * z + ph[4]) + 0.05 * rng.standard_normal(x.shape)
    d[..., 3] = amp_u * np.sin(kx * x + ph[5]) + 0.05 * rng.standard_normal(x.shape)
    d[..., 4] = 1.0 + 0.1 * np.cos(kx * x + ky * y + ph[6]) + 0.02 * rng.random(x.shape)
    return Q


# ---------------------------------------------------------------------------
# face level
# ---------------------------------------------------------------------------
def _riemann_states(n, seed):
    rng = np.random.default_rng(seed)
    qL = np.empty((n, 5))
    qR = np.empty((n, 5))
    for q in (qL, qR):
        q[:, 0] = rng.uniform(0.05, 3.0, n)
        q[:, 1:4] = rng.normal(0.0, 1.0, (n, 3))
        q[:, 4] = rng.uniform(0.01, 5.0, n)
    # strongly supersonic pairs in both directions, equal states and zero speeds
    m = n // 8
    qL[:m, 1] += 20.0
    qR[:m, 1] += 20.0
    qL[m:2 * m, 1] -= 20.0
    qR[m:2 * m, 1] -= 20.0
    qR[2 * m:3 * m] = qL[2 * m:3 * m]
    qL[3 * m:4 * m, 1:4] = 0.0
    qR[3 * m:4 * m, 1:4] = 0.0
    return qL, qR


@pytest.mark.parametrize("solver", [F.HLL, F.HLLC])
@pytest.mark.parametrize("gamma", [1.4, 5.0 / 3.0])
def test_riemann_face_parity(gpu_available, solver, gamma):
    qL, qR = _riemann_states(40000, seed=7 + solver)
    p = Params.defaults().replace(gamma0=gamma)
    f_o, p_o = Oracle(p).riemann(solver, qL, qR)
    f_g, p_g = F.riemann(solver, gamma, qL, qR)
    _assert_same(f_g, f_o, "flux")
    _assert_same(p_g, p_o, "pout")


BCS = [F.BC_ABSORBING, F.BC_REFLECTING, F.BC_PERIODIC]


@pytest.mark.parametrize("bcz", BCS + [F.BC_C91])
@pytest.mark.parametrize("bcx,bcy", [(a, b) for a in BCS for b in BCS] + [(F.BC_C91, F.BC_C91)])
def test_fill_boundaries_parity(gpu_available, bcx, bcy, bcz):
    p = Params.defaults().replace(Nx=6, Ny=5, Nz=4, boundary_x=bcx, boundary_y=bcy, boundary_z=bcz,
                                  gravity=1, g=-0.7, theta1=2.0)
    rng = np.random.default_rng(bcx * 16 + bcy * 4 + bcz)
    Q = rng.uniform(0.5, 2.0, p.shape_ghosted)
    Qo = Oracle(p).fill_boundaries(Q.copy())
    Qg = F.fill_boundaries(p, Q)
    _assert_same(Qg, Qo, "ghost fill")


# ---------------------------------------------------------------------------
# whole runs through the main loop
# ---------------------------------------------------------------------------
def test_sod_c1_full_run(gpu_available):
    """The reference's CPU configuration run to t=0.2, every snapshot and
    every dt bit-equal."""
    p = sod_c1()
    Q0 = Oracle(p).init("sod_x")
    steps, snaps = _compare_runs(p, Q0)
    assert len(steps) > 100 and len(snaps) == 11


@pytest.mark.usefixtures("nzc")
def test_blast_reflecting_bitwise(gpu_available):
    p = blast_small(32, tend=1.0)
    Q0 = F.problems.blast(p, radius=0.1, rho_in=1.0, p_in=10.0, rho_out=1.0, p_out=0.1)
    _compare_runs(p, Q0, max_steps=12)


def test_blast_negative_pressure_run_goes_on(gpu_available):
    """A blast strong enough (pressure ratio 1e6) that minmod-PLM drives a
    cell's pressure negative before t=0.1.  Its sound speed is NaN, the
    reference's Max reducer leaves it out of dt and the run goes on
    (ComputeDt.h:51-58); the session warns once and keeps stepping, and
    every dt and snapshot stays bit-equal, negative pressures included."""
    p = blast_small(32, tend=0.1, save_freq=0.05)
    Q0 = F.problems.blast(p, radius=0.1, rho_in=1.0, p_in=1000.0, rho_out=1.0, p_out=1e-3)
    with pytest.warns(F.NonFiniteWarning):
        steps, snaps = _compare_runs(p, Q0)
    assert len(steps) > 150
    assert np.nanmin(dom(p, snaps[-1][2])[..., 4]) < 0.0


@pytest.mark.usefixtures("nzc")
@pytest.mark.parametrize("recon", [F.PCM, F.PLM])
@pytest.mark.parametrize("solver", [F.HLL, F.HLLC])
@pytest.mark.parametrize("ts", [F.TS_EULER, F.TS_RK2])
def test_periodic_tc_visc(gpu_available, recon, solver, ts):
    p = Params.defaults().replace(Nx=16, Ny=16, Nz=16, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC,
                                  boundary_z=F.BC_PERIODIC, reconstruction=recon, riemann_solver=solver,
                                  time_stepping=ts, thermal_conductivity_active=1, kappa=0.02,
                                  viscosity_active=1, mu=0.01, tend=10.0)
    _compare_runs(p, smooth_random(p, 1), max_steps=6)


@pytest.mark.usefixtures("nzc")
@pytest.mark.parametrize("shape", [(24, 2, 3), (2, 3, 24), (3, 24, 2)])
@pytest.mark.parametrize("bc", [F.BC_PERIODIC, F.BC_REFLECTING, F.BC_ABSORBING])
def test_smallest_extents(gpu_available, shape, bc):
    """Extents of 2 and 3 cells (the library takes >= 2): ghosts wrap or
    mirror a 2-deep shell onto 2 cells, and tiles are mostly overhang."""
    p = Params.defaults().replace(Nx=shape[0], Ny=shape[1], Nz=shape[2], boundary_x=bc, boundary_y=bc,
                                  boundary_z=bc, reconstruction=F.PLM, riemann_solver=F.HLLC,
                                  time_stepping=F.TS_RK2, thermal_conductivity_active=1, kappa=0.02,
                                  viscosity_active=1, mu=0.01, tend=10.0)
    _compare_runs(p, smooth_random(p, 5), max_steps=5)


@pytest.mark.usefixtures("nzc")
def test_hll_fixed_temperature_visc_gravity(gpu_available):
    """Partial tiles (16 < 64 columns), HLL, conduction with fixed-temperature
    walls, viscosity and gravity: the configuration family of the
    stratified-convection workload."""
    p = Params.defaults().replace(Nx=16, Ny=16, Nz=8, xmax=2.0, ymax=2.0, boundary_x=F.BC_PERIODIC,
                                  boundary_y=F.BC_PERIODIC, boundary_z=F.BC_REFLECTING,
                                  reconstruction=F.PLM, riemann_solver=F.HLL, time_stepping=F.TS_RK2,
                                  gravity=1, g=1.0, thermal_conductivity_active=1, kappa=0.07,
                                  bctc_zmin=F.BCTC_FIXED_TEMPERATURE, bctc_zmin_value=1.0,
                                  bctc_zmax=F.BCTC_FIXED_TEMPERATURE, bctc_zmax_value=1.5,
                                  viscosity_active=1, mu=0.005, tend=10.0)
    Q0 = F.problems.polytrope_convection(p)
    _compare_runs(p, Q0, max_steps=8)


def _fast_range_state(p, case, seed):
    """Primitive states for the HLL FAST-quotient cases.  "wide": every
    cell in the range the kernels test (rho and p in [2^-40, 2^41), momenta
    zero or of magnitude in [2^-40, 2^41)), spanning decades, a fifth of the
    velocities exactly zero.  "edges": a smooth state with planes whose
    cells sit just outside that range (a density of 2^-41, a pressure of
    2^41, a momentum of 1e-13), so those planes and their neighbours take the
    correctly rounded forms."""
    rng = np.random.default_rng(seed)
    Q = smooth_random(p, seed, amp_u=0.3)
    d = dom(p, Q)
    if case == "wide":
        shape = d.shape[:3]
        d[..., 0] = 2.0 ** rng.uniform(-12, 12, shape)
        d[..., 4] = d[..., 0] * 2.0 ** rng.uniform(-6, 6, shape)
        for v in (1, 2, 3):
            speed = np.sqrt(d[..., 4] / d[..., 0]) * 2.0 ** rng.uniform(-14, 1, shape)
            d[..., v] = rng.choice([-1.0, 1.0], shape) * speed * (rng.random(shape) > 0.2)
    else:
        d[3, 2:5, 4:9, 0] = 2.0 ** -41
        d[3, 2:5, 4:9, 4] = 2.0 ** -40
        d[7, :, 5, 4] = 2.0 ** 41
        d[11, 6, :, 1] = 1e-13
        d[12, 1, 2, 2] = 0.0
    return Q


@pytest.mark.usefixtures("nzc")
@pytest.mark.parametrize("case", ["wide", "edges"])
def test_hll_fast_quotients_bitwise(gpu_available, case):
    """FP64 PLM-HLL faces take their quotients and square roots without the
    range-scaling steps on planes whose cells are all in a tested range
    (cell_in_fast_range), and the correctly rounded forms elsewhere; both
    give the reference's bits, including planes that switch between them."""
    p = Params.defaults().replace(Nx=24, Ny=10, Nz=16, boundary_x=F.BC_PERIODIC,
                                  boundary_y=F.BC_REFLECTING, boundary_z=F.BC_ABSORBING,
                                  reconstruction=F.PLM, riemann_solver=F.HLL,
                                  time_stepping=F.TS_RK2, CFL=0.4, tend=10.0)
    _compare_runs(p, _fast_range_state(p, case, 21), max_steps=5)


@pytest.mark.usefixtures("nzc")
def test_fixed_gradient_and_absorbing(gpu_available):
    p = Params.defaults().replace(Nx=20, Ny=12, Nz=10, boundary_x=F.BC_ABSORBING, boundary_y=F.BC_REFLECTING,
                                  boundary_z=F.BC_ABSORBING, reconstruction=F.PLM, time_stepping=F.TS_RK2,
                                  thermal_conductivity_active=1, kappa=0.05,
                                  bctc_zmin=F.BCTC_FIXED_GRADIENT, bctc_zmin_value=0.3,
                                  bctc_zmax=F.BCTC_FIXED_GRADIENT, bctc_zmax_value=-0.2, tend=10.0)
    _compare_runs(p, smooth_random(p, 3), max_steps=6)


@pytest.mark.usefixtures("nzc")
def test_pcm_wb_well_balanced_gravity(gpu_available):
    p = Params.defaults().replace(Nx=8, Ny=8, Nz=24, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC,
                                  boundary_z=F.BC_REFLECTING, reconstruction=F.PCM_WB,
                                  well_balanced_flux_at_z_bc=1, gravity=1, g=-1.0,
                                  time_stepping=F.TS_EULER, tend=10.0)
    _compare_runs(p, smooth_random(p, 4, amp_u=0.05), max_steps=6)


@pytest.mark.usefixtures("nzc")
def test_partial_tiles_non_pow2(gpu_available):
    """Sizes that fill no tile, and cell sizes that are not powers of two
    (true divisions in the update)."""
    p = Params.defaults().replace(Nx=100, Ny=13, Nz=7, xmax=0.7, ymax=1.3, zmax=0.9,
                                  boundary_x=F.BC_PERIODIC, boundary_y=F.BC_ABSORBING,
                                  boundary_z=F.BC_REFLECTING, reconstruction=F.PLM, time_stepping=F.TS_RK2,
                                  tend=10.0)
    _compare_runs(p, smooth_random(p, 5), max_steps=5)


def test_sod_axes_transposed(gpu_available):
    """sod_y on the GPU equals the transpose of sod_x bit for bit."""
    px = sod_c1(Nx=64, Ny=4, Nz=4, tend=0.05)
    py = sod_c1(Nx=4, Ny=64, Nz=4, tend=0.05, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_ABSORBING)
    sx, nx = gpu_main_loop(px, Oracle(px).init("sod_x"))
    sy, ny = gpu_main_loop(py, Oracle(py).init("sod_y"))
    assert [s[1] for s in sx] == [s[1] for s in sy]
    qx = dom(px, nx[-1][2])
    qy = dom(py, ny[-1][2])
    qy_t = np.transpose(qy, (0, 2, 1, 3))[..., [0, 2, 1, 3, 4]]
    _assert_same(qx, qy_t, "sod_x vs transposed sod_y")


def c91_params(**kw):
    """settings/C91.ini of the reference (convection under fillC91 z
    boundaries, conduction with a fixed-temperature bottom and
    fixed-gradient top, viscosity, gravity g=40), on a small grid."""
    base = dict(Nx=16, Ny=16, Nz=16, xmax=4.0, ymax=4.0, boundary_x=F.BC_PERIODIC,
                boundary_y=F.BC_PERIODIC, boundary_z=F.BC_C91, riemann_solver=F.HLLC,
                reconstruction=F.PLM, CFL=0.1, gamma0=1.66666667, gravity=1, g=40.0, c91_pert=1.0e-3,
                theta1=20.0, m1=1.0, thermal_conductivity_active=1, kappa=0.07,
                bctc_zmin=F.BCTC_FIXED_TEMPERATURE, bctc_zmin_value=1.0,
                bctc_zmax=F.BCTC_FIXED_GRADIENT, bctc_zmax_value=20.0,
                viscosity_active=1, mu=0.0028, tend=200.0)
    base.update(kw)
    return Params.defaults().replace(**base)


@pytest.mark.usefixtures("nzc")
@pytest.mark.parametrize("ts", [F.TS_EULER, F.TS_RK2])
def test_c91_z_boundaries_bitwise(gpu_available, ts):
    """boundaries_z=C91 in the fused stage (ghost planes from fillC91,
    BoundaryConditions.h:86-134, and the C91 z-wall flux of
    Update.h:138-143) against the oracle, Euler as in C91.ini and RK2."""
    p = c91_params(time_stepping=ts)
    _compare_runs(p, F.problems.c91(p), max_steps=8)


@pytest.mark.usefixtures("nzc")
def test_c91_z_with_reflecting_x_hll_pcm(gpu_available):
    """C91 z ghosts of reflected x/y ghost columns (edges and corners), HLL,
    PCM, no diffusion."""
    p = c91_params(Nx=12, Ny=20, Nz=10, boundary_x=F.BC_REFLECTING, boundary_y=F.BC_ABSORBING,
                   riemann_solver=F.HLL, reconstruction=F.PCM, thermal_conductivity_active=0,
                   viscosity_active=0, time_stepping=F.TS_RK2)
    _compare_runs(p, F.problems.c91(p), max_steps=8)


@pytest.mark.usefixtures("nzc")
@pytest.mark.parametrize("walls", ["fixed_temperature", "fixed_gradient"])
def test_b02_conductivity_bitwise(gpu_available, walls):
    """conductivity_mode=B02: the tanh profile of computeKappa
    (ThermalConduction.h:7-22) at the face positions, and ComputeDt's
    integer-index evaluation of it (ComputeDt.h:39-41)."""
    bc = F.BCTC_FIXED_TEMPERATURE if walls == "fixed_temperature" else F.BCTC_FIXED_GRADIENT
    p = Params.defaults().replace(Nx=16, Ny=8, Nz=24, boundary_x=F.BC_PERIODIC, boundary_y=F.BC_PERIODIC,
                                  boundary_z=F.BC_REFLECTING, reconstruction=F.PLM, time_stepping=F.TS_RK2,
                                  thermal_conductivity_active=1, thermal_conductivity_mode=F.TCM_B02,
                                  kappa=0.05, b02_zmid=0.5, b02_kappa1=1.0, b02_kappa2=0.1,
                                  b02_thickness=0.1, bctc_zmin=bc, bctc_zmin_value=1.2,
                                  bctc_zmax=bc, bctc_zmax_value=0.8, tend=10.0)
    _compare_runs(p, smooth_random(p, 6, amp_u=0.1), max_steps=8)
