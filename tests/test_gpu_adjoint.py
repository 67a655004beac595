"""Discrete adjoint of the frozen-dt time loop (fv3d_adjoint_run): gradients
of J = sum((rho_T - rho_target)^2) with respect to the initial state, the
initial density field, kappa and mu, checked against central finite
differences of the forward run itself (eps = 1e-6 along unit-norm random
directions), FP32 against FP64, and through the torch.autograd wrapper.

The scheme is only piecewise smooth (minmod selections, Riemann wave
regions): the adjoint is the derivative inside the piece the forward run
sits in.  Unit-norm directions keep the per-cell perturbation far below the
spacing of switch points, and J is summed in extended precision (its
float64 summation error, ~1e-15 J, would otherwise be the central
difference's largest error at eps = 1e-6), so every direction agrees to
1e-6.  A uniform field is itself a switch point of every minmod
(minmod(0, 0)), so the state-gradient cases give the pressure a random
component; the density gradient at uniform pressure is checked through the
torch wrapper, whose directions leave pressure alone."""
import numpy as np
import pytest

import fv3d_amd as F
from fv3d_amd import problems

EPS = 1e-6


def base_params(N=16, **kw):
    base = dict(Nx=N, Ny=N, Nz=N, gamma0=5.0 / 3.0, reconstruction=F.PLM, riemann_solver=F.HLLC,
                time_stepping=F.TS_RK2, CFL=0.4, boundary_x=F.BC_PERIODIC,
                boundary_y=F.BC_PERIODIC, boundary_z=F.BC_PERIODIC,
                thermal_conductivity_active=1, kappa=0.02, viscosity_active=1, mu=0.01)
    base.update(kw)
    return F.Params.defaults().replace(**base)


CONFIGS = {
    "periodic_hllc_plm_rk2_tc_visc": {},
    "walls_hll_plm_rk2_fixedT": dict(riemann_solver=F.HLL, boundary_x=F.BC_REFLECTING,
                                     boundary_y=F.BC_ABSORBING, boundary_z=F.BC_REFLECTING,
                                     bctc_zmin=F.BCTC_FIXED_TEMPERATURE, bctc_zmin_value=1.1,
                                     bctc_zmax=F.BCTC_FIXED_GRADIENT, bctc_zmax_value=0.3),
    "pcm_hllc_euler_visc": dict(reconstruction=F.PCM, time_stepping=F.TS_EULER,
                                thermal_conductivity_active=0, CFL=0.3),
    "periodic_hllc_plm_rk2_hydro": dict(thermal_conductivity_active=0, viscosity_active=0),
    # the smallest extents the library takes (2 cells) along y, then along x
    "thin_y_periodic_tc_visc": dict(Nx=20, Ny=2, Nz=3),
    "thin_x_walls_hll_tc_visc": dict(Nx=2, Ny=17, Nz=5, riemann_solver=F.HLL,
                                     boundary_x=F.BC_REFLECTING, boundary_y=F.BC_ABSORBING,
                                     boundary_z=F.BC_REFLECTING),
}


def dom(Q):
    return Q[2:-2, 2:-2, 2:-2]


def J_sum(r, target):
    """sum((r - target)^2), accumulated in extended precision so that its
    rounding stays far below the central differences' resolution."""
    return float(np.sum((r - target) ** 2, dtype=np.longdouble))


def record_dts(p, Q0, n):
    """dt sequence of the ordinary run (ComputeDt each step)."""
    with F.Simulation(p, F.FP64) as sim:
        sim.set_prim(Q0)
        dts = []
        for _ in range(n):
            dts.append(sim.compute_dt()[0])
            sim.step()
    return np.array(dts)


class Problem:
    def __init__(self, p, nsteps, seed=7, precision=F.FP64):
        self.p = p
        self.Q0 = problems.turbulent_box(p, seed=seed, p_amp=0.05)
        self.dts = record_dts(p, self.Q0, nsteps)
        self.target = self.rho_T(p, dom(problems.turbulent_box(p, seed=seed + 1)))
        self.precision = precision

    def rho_T(self, p, prims, precision=F.FP64):
        with F.Simulation(p, precision) as sim:
            sim.set_prim(np.ascontiguousarray(prims), with_ghosts=False)
            sim.forward_frozen(self.dts)
            return sim.get_prim(with_ghosts=False)[..., 0]

    def J_cons(self, p, U0):
        with F.Simulation(p, F.FP64) as sim:
            sim.set_cons(np.ascontiguousarray(U0), with_ghosts=False)
            sim.forward_frozen(self.dts)
            r = sim.get_prim(with_ghosts=False)[..., 0]
        return J_sum(r, self.target)

    def adjoint(self, p=None, precision=F.FP64, K=4, store_bytes=0):
        p = p or self.p
        with F.Simulation(p, precision) as sim:
            sim.set_prim(self.Q0)
            U0 = sim.get_cons(with_ghosts=False)
            rT = self.rho_T(p, dom(self.Q0), precision)
            lamT = np.zeros(p.shape_domain)
            lamT[..., 0] = 2.0 * (rT - self.target)
            g, dk, dm = sim.adjoint(self.dts, lamT, K, store_bytes=store_bytes)
            assert np.array_equal(sim.get_cons(with_ghosts=False), U0)  # state untouched
        return U0, g, dk, dm


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CONFIGS))
def test_adjoint_state_gradient_matches_central_differences(gpu_available, name):
    print(f"adjoint FD: {name}", flush=True)
    pr = Problem(base_params(**CONFIGS[name]), nsteps=10)
    U0, g, dk, dm = pr.adjoint(K=4)  # 10 steps: segments of 4, 4 and 2
    rng = np.random.default_rng(3)
    rel = []
    for trial in range(5):
        d = rng.standard_normal(U0.shape) * np.abs(U0).mean(axis=(0, 1, 2))
        d /= np.linalg.norm(d)
        fd = (pr.J_cons(pr.p, U0 + EPS * d) - pr.J_cons(pr.p, U0 - EPS * d)) / (2 * EPS)
        ad = float(np.sum(g * d))
        rel.append(abs(fd - ad) / abs(ad))
        print(f"  direction {trial}: fd={fd:.12e} adjoint={ad:.12e} rel={rel[-1]:.2e}")
    assert max(rel) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["periodic_hllc_plm_rk2_tc_visc", "walls_hll_plm_rk2_fixedT"])
def test_adjoint_parameter_gradients_match_central_differences(gpu_available, name):
    pr = Problem(base_params(**CONFIGS[name]), nsteps=10)
    U0, g, dk, dm = pr.adjoint(K=3)
    for key, val, ad in [("kappa", pr.p.kappa, dk), ("mu", pr.p.mu, dm)]:
        h = EPS * val
        Jp = pr.J_cons(pr.p.replace(**{key: val + h}), U0)
        Jm = pr.J_cons(pr.p.replace(**{key: val - h}), U0)
        fd = (Jp - Jm) / (2 * h)
        print(f"  d/d{key}: fd={fd:.12e} adjoint={ad:.12e}")
        assert abs(fd - ad) <= 1e-6 * abs(ad), key


@pytest.mark.gpu
def test_adjoint_checkpoint_interval_does_not_change_result(gpu_available):
    pr = Problem(base_params(), nsteps=9)
    _, g1, k1, m1 = pr.adjoint(K=1)
    _, g9, k9, m9 = pr.adjoint(K=9)
    _, g4, k4, m4 = pr.adjoint(K=4)
    scale = np.abs(g1).max()
    for g in (g9, g4):
        assert np.abs(g - g1).max() <= 1e-12 * scale  # only atomic summation order differs
    assert k4 == pytest.approx(k1, rel=1e-12) and m9 == pytest.approx(m1, rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["periodic_hllc_plm_rk2_tc_visc", "pcm_hllc_euler_visc"])
def test_adjoint_segment_store_does_not_change_result(gpu_available, name):
    """Segments kept from the forward pass (fv3d_set_adjoint_store) against
    recomputed ones: 9 steps in segments of 4, 4 and 1; a budget of two
    segments keeps the last two, a large one keeps all three."""
    p = base_params(**CONFIGS[name])
    pr = Problem(p, nsteps=9)
    rk2 = p.time_stepping == F.TS_RK2
    seg = (2 * 4 + 1 if rk2 else 4 + 1) * 5 * p.Nx * p.Ny * p.Nz * 8
    _, g0, k0, m0 = pr.adjoint(K=4)
    for store in (seg - 1, 2 * seg, 1 << 34):
        _, g, k, m = pr.adjoint(K=4, store_bytes=store)
        assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max(), store
        assert k == pytest.approx(k0, rel=1e-12, abs=1e-300) and m == pytest.approx(m0, rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CONFIGS))
def test_tiled_adjoint_equals_per_face_form(gpu_available, name, monkeypatch):
    """The LDS-tiled kernels against the one-thread-per-face/cell kernels
    (FV3D_ADJ_SIMPLE): the same sums, associated differently.  The default
    path finalizes most cells in the z-face kernel; FV3D_ADJ_UNFUSED leaves
    them all to the finalize kernel."""
    pr = Problem(base_params(N=20, **CONFIGS[name]), nsteps=5)
    _, gt, kt, mt = pr.adjoint(K=2)
    monkeypatch.setenv("FV3D_ADJ_UNFUSED", "1")
    _, gu, ku, mu_ = pr.adjoint(K=2)
    monkeypatch.delenv("FV3D_ADJ_UNFUSED")
    monkeypatch.setenv("FV3D_ADJ_COLORED", "1")  # colour passes instead of atomic flushes
    _, gc, kc, mc = pr.adjoint(K=2)
    monkeypatch.delenv("FV3D_ADJ_COLORED")
    monkeypatch.setenv("FV3D_ADJ_SIMPLE", "1")
    _, gs, ks, ms = pr.adjoint(K=2)
    for g, k, m in [(gt, kt, mt), (gu, ku, mu_), (gc, kc, mc)]:
        assert np.abs(g - gs).max() <= 1e-12 * np.abs(gs).max()
        assert k == pytest.approx(ks, rel=1e-12, abs=1e-300) and m == pytest.approx(ms, rel=1e-12, abs=1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["periodic_hllc_plm_rk2_tc_visc", "walls_hll_plm_rk2_fixedT"])
def test_adjoint_independent_of_tile_width(gpu_available, name, monkeypatch):
    """The forward passes inside the adjoint run at 64- or 32-column stage
    tiles (the same bits); the gradient differs only by the order of its
    atomic sums."""
    out = {}
    for wx in (64, 32):
        monkeypatch.setenv("FV3D_TILE_WX", str(wx))
        pr = Problem(base_params(N=20, **CONFIGS[name]), nsteps=6)
        out[wx] = pr.adjoint(K=3)
    (_, g64, k64, m64), (_, g32, k32, m32) = out[64], out[32]
    assert np.abs(g32 - g64).max() <= 1e-12 * np.abs(g64).max()
    assert k32 == pytest.approx(k64, rel=1e-12, abs=1e-300) and m32 == pytest.approx(m64, rel=1e-12, abs=1e-300)


@pytest.mark.gpu
def test_adjoint_fp32_against_fp64(gpu_available):
    pr = Problem(base_params(N=24), nsteps=12)
    _, g64, k64, m64 = pr.adjoint(precision=F.FP64)
    _, g32, k32, m32 = pr.adjoint(precision=F.FP32_PLAIN)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"FP32 vs FP64 adjoint: rel L2 {rel:.2e}, kappa {k32:.6e}/{k64:.6e}, mu {m32:.6e}/{m64:.6e}")
    assert rel < 1e-3
    assert k32 == pytest.approx(k64, rel=1e-2) and m32 == pytest.approx(m64, rel=1e-2)


@pytest.mark.gpu
def test_autograd_function_density_gradient(gpu_available):
    torch = pytest.importorskip("torch")
    from fv3d_amd.autograd import FrozenRun, RunSetup
    p = base_params(N=12)
    Q0 = dom(problems.turbulent_box(p, seed=5))
    dts = record_dts(p, problems.turbulent_box(p, seed=5), 6)
    setup = RunSetup(params=p, velocity=Q0[..., 1:4].copy(), pressure=Q0[..., 4].copy(), dts=dts, K=4)
    target = torch.from_numpy(dom(problems.turbulent_box(p, seed=6))[..., 0].copy())
    rho0 = torch.from_numpy(Q0[..., 0].copy()).requires_grad_(True)
    kappa = torch.tensor(p.kappa, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(p.mu, dtype=torch.float64, requires_grad=True)

    def J(r, k, m):
        return ((FrozenRun.apply(r, k, m, setup) - target) ** 2).sum()

    J(rho0, kappa, mu).backward()
    rng = np.random.default_rng(1)
    d = rng.standard_normal(rho0.shape)
    d = torch.from_numpy(d / np.linalg.norm(d))
    with torch.no_grad():
        fd = (J(rho0 + EPS * d, kappa, mu) - J(rho0 - EPS * d, kappa, mu)) / (2 * EPS)
        fk = (J(rho0, kappa * (1 + EPS), mu) - J(rho0, kappa * (1 - EPS), mu)) / (2 * EPS * kappa)
    ad = float((rho0.grad * d).sum())
    print(f"  rho0: fd={float(fd):.12e} adjoint={ad:.12e}; kappa: fd={float(fk):.12e} "
          f"adjoint={float(kappa.grad):.12e}")
    assert abs(float(fd) - ad) <= 1e-6 * abs(ad)
    assert abs(float(fk) - float(kappa.grad)) <= 1e-6 * abs(float(kappa.grad))


@pytest.mark.gpu
def test_adjoint_rejects_gravity(gpu_available):
    p = base_params(gravity=1, g=-1.0)
    Q0 = problems.turbulent_box(p, seed=2)
    with F.Simulation(p, F.FP64) as sim:
        sim.set_prim(Q0)
        with pytest.raises(F.FV3DError) as e:
            sim.adjoint(np.full(2, 1e-3), np.zeros(p.shape_domain), 1)
        assert e.value.code == 1


@pytest.mark.gpu
def test_adjoint_rejects_active_density_floor(gpu_available):
    """A run in which the density floor fmax(1e-6, rho) (Update.h:160) acts is
    not differentiated as if it did not: fv3d_adjoint_run fails with
    FV3D_EFLOOR and leaves the state alone.  The turbulent box scaled to
    rho ~ 5e-7 (p scaled alike: the same temperature and velocities) has
    every density below the floor; the unscaled box runs as usual."""
    p = base_params(N=12)
    lamT = np.ones(p.shape_domain)
    dts = np.full(3, 1e-3)
    for scale, expect in ((5e-7, 5), (1.0, 0)):
        Q0 = problems.turbulent_box(p, seed=4)
        Q0[..., 0] *= scale
        Q0[..., 4] *= scale
        with F.Simulation(p, F.FP64) as sim:
            sim.set_prim(Q0)
            U0 = sim.get_cons(with_ghosts=False)
            if expect:
                with pytest.raises(F.FV3DError) as e:
                    sim.adjoint(dts, lamT, 2)
                assert e.value.code == expect and "floor" in str(e.value)
            else:
                sim.adjoint(dts, lamT, 2)
            assert np.array_equal(sim.get_cons(with_ghosts=False), U0)
