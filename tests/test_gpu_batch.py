"""Ensembles of independent boxes (fv3d_batch_*): every box of a batched run
must be bit-equal to a single-box context with the same parameters and
precision, in each of the three step modes (persistent kernel replayed from
a hipGraph, persistent kernel issued from the host, per-box launch loop);
FP64 boxes are additionally checked against the oracle, and FP32 ensembles
against the FP64 ensemble in L1."""
import numpy as np
import pytest

import fv3d_amd as F
from fv3d_amd import problems
from oracle.oracle import Oracle
from tests.helpers import dom

SMALL_SIDES = (16, 24, 32, 40)


def rt_base(**kw):
    base = dict(reconstruction=F.PLM, riemann_solver=F.HLLC, time_stepping=F.TS_RK2, CFL=0.8)
    base.update(kw)
    return F.Params.defaults().replace(**base)


def single_run(p, Q0, precision, nsteps):
    with F.Simulation(p, precision) as sim:
        sim.set_prim(Q0)
        dts = []
        for _ in range(nsteps):
            dts.append(sim.compute_dt()[0])
            sim.step()
        return sim.get_prim(with_ghosts=True), dts


def batch_run(ens, precision, nsteps, mode, configure=None):
    with F.Batch([p for p, _ in ens], precision) as b:
        if configure:
            b.set_tile_width(configure[2])
            b.configure(*configure[:2])
            assert b.tile_width == configure[2]
        for i, (_, Q0) in enumerate(ens):
            b.set_prim(i, Q0)
        dts = []
        for _ in range(nsteps):
            dts.append(b.compute_dt())
            b.step(1, mode)
        assert b.step_count == nsteps
        return [b.get_prim(i, with_ghosts=True) for i in range(len(ens))], np.array(dts)


def test_batch_rejects_mixed_numerics():
    """No GPU needed: parameter checks come before the device is touched."""
    a = rt_base()
    b = rt_base(riemann_solver=F.HLL)
    with pytest.raises(F.FV3DError) as e:
        F.Batch([a, b], F.FP32)
    assert e.value.code == 1
    with pytest.raises(F.FV3DError) as e:
        F.Batch([a, a.replace(boundary_z=F.BC_C91)], F.FP32)
    assert e.value.code == 1


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [F.FP64, F.FP32, F.FP32_PLAIN])
@pytest.mark.parametrize("mode", [F.BATCH_GRAPH, F.BATCH_PERSISTENT, F.BATCH_LOOP])
def test_batch_equals_single_contexts(gpu_available, precision, mode, tile_wx):
    print(f"batch vs single: precision={precision} mode={mode} wx={tile_wx}", flush=True)
    ens = problems.rt_ensemble(rt_base(), 6, seed=7, sides=SMALL_SIDES)
    nsteps = 6
    Qb, dtb = batch_run(ens, precision, nsteps, mode)
    for i, (p, Q0) in enumerate(ens):
        Qs, dts = single_run(p, Q0, precision, nsteps)
        assert np.array_equal(dtb[:, i], np.array(dts)), i
        assert np.array_equal(dom(p, Qb[i]), dom(p, Qs)), (i, p.Nx, p.Ny, p.Nz)
        assert np.array_equal(Qb[i], Qs), i  # ghost shell from the same fill


@pytest.mark.gpu
@pytest.mark.parametrize("configure", [(8, 0, 64), (64, 1024, 32), (1, 300, 64), (3, 0, 32)])
def test_batch_launch_shape_does_not_change_results(gpu_available, configure):
    print(f"batch launch shape {configure}", flush=True)
    ens = problems.rt_ensemble(rt_base(), 5, seed=11, sides=SMALL_SIDES)
    Qa, da = batch_run(ens, F.FP32, 4, F.BATCH_GRAPH)
    Qb, db = batch_run(ens, F.FP32, 4, F.BATCH_GRAPH, configure=configure)
    assert np.array_equal(da, db)
    for a, b in zip(Qa, Qb):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_batch_default_tile_width_follows_cover(gpu_available, monkeypatch):
    """Without a setting the ensemble takes 32-column tiles when 64-column
    tiles would cover at least 10% more cells (overhang does no work)."""
    monkeypatch.delenv("FV3D_TILE_WX", raising=False)
    for sides, wx in (((32, 32), 32), ((96, 96), 32), ((64, 64), 64), ((48, 48), 64),
                      ((256, 8), 64)):
        ps = [rt_base(Nx=sides[0], Ny=sides[1], Nz=8)] * 2
        with F.Batch(ps, F.FP32) as b:
            assert b.tile_width == wx, sides
            b.set_tile_width(64 if wx == 32 else 32)
            assert b.tile_width != wx
            b.set_tile_width(0)
            assert b.tile_width == wx
            with pytest.raises(F.FV3DError) as e:
                b.set_tile_width(48)
            assert e.value.code == 1


@pytest.mark.gpu
def test_batch_fp64_boxes_match_oracle(gpu_available, tile_wx):
    ens = problems.rt_ensemble(rt_base(), 3, seed=3, sides=(16, 24))
    nsteps = 5
    Qb, dtb = batch_run(ens, F.FP64, nsteps, F.BATCH_GRAPH)
    for i, (p, Q0) in enumerate(ens):
        o = Oracle(p)
        U, _, dts, _ = o.run(Q0, nsteps, stop_at_tend=False)
        assert np.array_equal(dtb[:, i], dts), i
        Qo = o.cons_to_prim(U)
        assert np.array_equal(dom(p, Qb[i]), dom(p, Qo)), i


@pytest.mark.gpu
@pytest.mark.parametrize("ts", [F.TS_EULER, F.TS_RK2])
def test_batch_sweep_kappa_mu_gamma(gpu_available, ts, tile_wx):
    """Per-box conduction and viscosity coefficients and gamma (a parameter
    sweep), with fixed-temperature walls; Euler and RK2."""
    print(f"batch sweep ts={ts}", flush=True)
    base = rt_base(time_stepping=ts, riemann_solver=F.HLL, thermal_conductivity_active=1,
                   viscosity_active=1, CFL=0.3, bctc_zmin=F.BCTC_FIXED_TEMPERATURE,
                   bctc_zmax=F.BCTC_FIXED_TEMPERATURE, bctc_zmin_value=2.5, bctc_zmax_value=1.2)
    ens = []
    for b, (kappa, mu) in enumerate([(0.0, 0.0), (1e-3, 2e-3), (5e-3, 1e-3), (2e-3, 8e-3)]):
        p = problems.rt_box_params(base, b, seed=5, sides=(16, 24)).replace(kappa=kappa, mu=mu)
        ens.append((p, problems.rt_box(p, b, seed=5)))
    nsteps = 5
    for precision in (F.FP64, F.FP32):
        Qb, dtb = batch_run(ens, precision, nsteps, F.BATCH_GRAPH)
        for i, (p, Q0) in enumerate(ens):
            Qs, dts = single_run(p, Q0, precision, nsteps)
            assert np.array_equal(dtb[:, i], np.array(dts)), (precision, i)
            assert np.array_equal(Qb[i], Qs), (precision, i)


@pytest.mark.gpu
def test_batch_fp32_ensemble_l1_against_fp64(gpu_available):
    """L1(rho) of the FP32 (Kahan RK2) ensemble against the FP64 ensemble,
    per box, after 100 steps: <= 1e-5."""
    ens = problems.rt_ensemble(rt_base(), 8, seed=12345, sides=(32, 48))
    nsteps = 100
    ps = [p for p, _ in ens]
    out = {}
    for precision in (F.FP64, F.FP32):
        with F.Batch(ps, precision) as b:
            for i, (_, Q0) in enumerate(ens):
                b.set_prim(i, Q0)
            b.step(nsteps, F.BATCH_GRAPH)
            out[precision] = [b.get_prim(i, with_ghosts=False) for i in range(len(ens))]
    for i in range(len(ens)):
        l1 = np.mean(np.abs(out[F.FP32][i][..., 0] - out[F.FP64][i][..., 0]))
        print(f"box {i}: L1(rho) = {l1:.3e}")
        assert l1 <= 1e-5, (i, l1)


@pytest.mark.gpu
def test_batch_tune_restores_state(gpu_available, tmp_path):
    ens = problems.rt_ensemble(rt_base(), 4, seed=9, sides=SMALL_SIDES)
    ps = [p for p, _ in ens]
    cache = str(tmp_path / "tune.txt")
    with F.Batch(ps, F.FP32) as b:
        for i, (_, Q0) in enumerate(ens):
            b.set_prim(i, Q0)
        b.step(2)
        before = [b.get_prim(i) for i in range(len(ens))]
        dt0 = b.compute_dt()
        nzc, grid = b.tune(nsteps=2, cache=cache)
        wx = b.tile_width
        assert nzc in (8, 16, 32, 64, 128) and grid > 0 and wx in (32, 64)
        assert b.step_count == 2
        assert np.array_equal(b.compute_dt(), dt0)
        for i in range(len(ens)):
            assert np.array_equal(b.get_prim(i), before[i])
        b.step(3)
        after = [b.get_prim(i) for i in range(len(ens))]
    # a cache hit configures the same shape without a search, and the run
    # continues identically
    with F.Batch(ps, F.FP32) as b:
        for i, (_, Q0) in enumerate(ens):
            b.set_prim(i, Q0)
        b.step(2)
        assert b.tune(nsteps=2, cache=cache) == (nzc, grid)
        assert b.tile_width == wx
        b.step(3)
        for i in range(len(ens)):
            assert np.array_equal(b.get_prim(i), after[i])
    assert open(cache).read().count("\n") == 1


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_full_rt_ensemble_512_boxes(gpu_available, capsys):
    """The benchmarked ensemble at its full size: 512 Rayleigh-Taylor boxes
    with sides drawn from {32, 48, 64, 96} per axis (110.9 M cells), FP32
    with the compensated RK2, 100 steps replayed from the hipGraph with the
    autotuned launch shape.  Every box stays within L1(rho) <= 1e-5 of the
    FP64 ensemble run alike, and four sampled boxes are bit-equal to
    single-box contexts of the same parameters."""
    base = rt_base()
    ens = problems.rt_ensemble(base, 512)
    nsteps = 100

    def run(precision, tune):
        b = F.Batch([p for p, _ in ens], precision)
        for i, (_, Q0) in enumerate(ens):
            b.set_prim(i, Q0)
        if tune:
            b.tune(nsteps=2)
        b.step(nsteps, F.BATCH_GRAPH)
        b.compute_dt()  # raises if a box went non-finite
        return b

    b32 = run(F.FP32, True)
    b64 = run(F.FP64, False)
    assert b32.cells == sum(p.Nx * p.Ny * p.Nz for p, _ in ens) > 100_000_000
    worst = 0.0
    for i in range(len(ens)):
        r32 = b32.get_prim(i, with_ghosts=False)[..., 0]
        r64 = b64.get_prim(i, with_ghosts=False)[..., 0]
        worst = max(worst, float(np.mean(np.abs(r32 - r64))))
    with capsys.disabled():
        print(f"\n512-box ensemble: max per-box L1(rho) FP32 vs FP64 after {nsteps} steps {worst:.2e}",
              flush=True)
    assert worst <= 1e-5
    for i in (0, 137, 300, 511):
        p, Q0 = ens[i]
        q_single, _ = single_run(p, Q0, F.FP32, nsteps)
        assert np.array_equal(b32.get_prim(i, with_ghosts=True)[2:-2, 2:-2, 2:-2],
                              q_single[2:-2, 2:-2, 2:-2]), i
    b32.close()
    b64.close()
