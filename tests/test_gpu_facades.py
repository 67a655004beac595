"""The C++ facades of the reference's functors (fv3d_amd/driver/functors.h:
BoundaryManager, primToCons/consToPrim of arrays and of a State, hll/hllc,
ComputeDtFunctor, UpdateFunctor, ThermalConductionFunctor and
ViscosityFunctor) driven by a C++ program (driver/functor_check.cpp)
through the C ABI, every output against the oracle's, bit for bit."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import fv3d_amd as F
from oracle.oracle import Oracle, lib as olib, _p
from tests.helpers import dom
from tests.test_gpu_functors import _cfg, _initial, _same

pytestmark = pytest.mark.gpu

CHECK = os.path.join(os.path.dirname(F.__file__), "fv3d_functor_check")


@pytest.mark.parametrize("name", ["periodic_tc_visc", "hll_fixed_t_gravity", "b02", "pcm_wb_gravity"])
def test_cpp_facades_match_oracle(gpu_available, tmp_path, name):
    p = _cfg(name)
    o = Oracle(p)
    Q = _initial(p, name)
    U = o.prim_to_cons(Q)
    dt = 0.5 * o.compute_dt(Q)[0]
    qL = np.array([1.2, 0.3, -0.4, 0.5, 0.9])
    qR = np.array([0.7, -0.2, 0.1, -0.6, 1.3])
    with open(tmp_path / "in.bin", "wb") as f:
        f.write(bytes(p))
        f.write(np.float64(dt).tobytes())
        f.write(Q.tobytes())
        f.write(U.tobytes())
        f.write(qL.tobytes())
        f.write(qR.tobytes())
    r = subprocess.run([CHECK, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.float64)
    n = int(np.prod(p.shape_ghosted))
    pos = 0

    def take(k, shape=None):
        nonlocal pos
        a = raw[pos:pos + k]
        pos += k
        return a.reshape(shape) if shape else a

    arr = lambda: take(n, p.shape_ghosted)
    Qf = o.fill_boundaries(Q.copy())
    _same(arr(), Qf, "BoundaryManager::fillBoundaries")
    U2 = o.prim_to_cons(Qf)
    _same(arr(), U2, "primToCons(Q, U)")
    _same(arr(), o.cons_to_prim(U2), "consToPrim(U, Q)")
    for s, ax in zip(o.compute_slopes(Qf), "XYZ"):
        _same(arr(), s, f"slopes{ax}")
    sx, sy, sz = o.compute_slopes(Qf)
    Unew = U.copy()
    o.fluxes_and_update(Qf, sx, sy, sz, Unew, dt)
    _same(arr(), Unew, "computeFluxesAndUpdate")
    o.thermal_conduction(Qf, Unew, dt)
    _same(arr(), Unew, "applyThermalConduction")
    o.viscosity(Qf, Unew, dt)
    _same(arr(), Unew, "applyViscosity")
    assert take(1)[0] == o.compute_dt(Qf)[0]
    uL, qb = np.zeros(5), np.zeros(5)
    olib().orc_state_prim_to_cons(ctypes.byref(o.p), _p(qL), _p(uL))
    olib().orc_state_cons_to_prim(ctypes.byref(o.p), _p(uL), _p(qb))
    _same(take(5), uL, "State primToCons")
    _same(take(5), qb, "State consToPrim")
    for solver, what in ((F.HLL, "hll"), (F.HLLC, "hllc")):
        f_o, p_o = o.riemann(solver, qL, qR)
        _same(take(5), f_o[0], what)
        assert take(1)[0] == p_o[0], what
    Q3, U3 = Qf.copy(), U.copy()
    o.update(Q3, U3, dt)
    _same(arr(), Q3, "UpdateFunctor::update Q")
    _same(arr(), U3, "UpdateFunctor::update Unew")
    Us = o.prim_to_cons(Q)
    Qs = o.cons_to_prim(Us)
    Unew = Us.copy()
    o.update(Qs, Unew, dt)
    _same(dom(p, arr()), dom(p, Unew), "UpdateFunctor::update(Device, dt)")
    assert pos == raw.size
