"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol its headers declare, the ctypes mirror of fv3d_params has the
C layout, and read_inifile follows readInifile (SimInfo.h:169-305) with
inih semantics."""
import ctypes
import os
import re

import pytest

import fv3d_amd as F
from fv3d_amd import Params, read_inifile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            for m in re.finditer(r"^\s*([A-Za-z_][\w\s\*]*?)\b(fv3d_\w+)\s*\(", txt, flags=re.M):
                if "static" not in m.group(1):  # header-only inline helpers are not exported
                    names.add(m.group(2))
    return names


def test_library_exports_every_declared_symbol():
    L = F.lib()
    syms = _declared_symbols()
    assert {"fv3d_create", "fv3d_step", "fv3d_compute_dt", "fv3d_riemann"} <= syms
    missing = [s for s in sorted(syms) if not hasattr(L, s)]
    assert not missing, missing


def test_params_layout_matches_library():
    assert F.lib().fv3d_params_size() == ctypes.sizeof(Params)
    assert F.lib().fv3d_abi_version() >= 1


def test_create_without_gpu_fails_loudly():
    """No silent fallback: with no device visible the session refuses."""
    import subprocess, sys
    code = ("import fv3d_amd as F\n"
            "try:\n    F.Simulation(F.Params.defaults())\nexcept F.FV3DError as e:\n    print('raised', e.code)\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, cwd=ROOT,
                       timeout=120)
    assert "raised" in r.stdout, r.stdout + r.stderr


INI = """
[mesh]
Nx=64
Ny = 16
nz=8
xmin=0.0
xmax=2.0

[run]
tend=0.25
save_freq=0.01 ; inline comment
boundaries_x=periodic
boundaries_y=absorbing
boundaries_z=C91

[solvers]
riemann_solver=hll
reconstruction=plm
time_stepping=RK2
CFL=0.4

[physics]
gamma0=1.4
gravity=yes
g=-0.1
problem=rayleigh-taylor

[thermal_conduction]
active=true
kappa=0.07
bc_zmin=fixed_temperature
bc_zmax=fixed_gradient
bc_zmax_value=0.5

[viscosity]
active=1
mu=5e-3

[misc]
log_frequency=0x10
"""


def test_read_inifile_values_and_defaults():
    cfg = read_inifile(text=INI)
    p = cfg.params
    assert (p.Nx, p.Ny, p.Nz, p.Ng) == (64, 16, 8, 2)
    assert (p.Ntx, p.ibeg, p.iend) == (68, 2, 66)
    assert p.dx == 2.0 / 64 and p.dy == 1.0 / 16
    assert p.tend == 0.25 and p.save_freq == 0.01
    assert (p.boundary_x, p.boundary_y, p.boundary_z) == (F.BC_PERIODIC, F.BC_ABSORBING, F.BC_C91)
    assert (p.riemann_solver, p.reconstruction, p.time_stepping) == (F.HLL, F.PLM, F.TS_RK2)
    assert p.CFL == 0.4 and p.gamma0 == 1.4 and p.gravity == 1 and p.g == -0.1
    assert cfg.problem == "rayleigh-taylor" and cfg.output_filename == "run"
    assert p.thermal_conductivity_active == 1 and p.kappa == 0.07
    assert (p.bctc_zmin, p.bctc_zmax) == (F.BCTC_FIXED_TEMPERATURE, F.BCTC_FIXED_GRADIENT)
    assert p.bctc_zmin_value == 1.0 and p.bctc_zmax_value == 0.5
    assert p.viscosity_active == 1 and p.mu == 5e-3
    assert p.log_frequency == 16 and p.seed == 12345 and p.epsilon == 1e-6


def test_read_inifile_reference_quirks():
    # unknown enum strings map to enumerator 0; "rk2" is not "RK2"
    cfg = read_inifile(text="[solvers]\ntime_stepping=rk2\nreconstruction=weno\n[run]\nboundaries_x=open\n")
    p = cfg.params
    assert p.time_stepping == F.TS_EULER and p.reconstruction == F.PCM
    assert p.boundary_x == F.BC_ABSORBING
    assert len(cfg.warnings) == 3
    # a missing file gives every default, like INIReader's unchecked ParseError
    cfg = read_inifile("/nonexistent/file.ini")
    p = cfg.params
    assert (p.Nx, p.riemann_solver, p.reconstruction, p.boundary_x) == (32, F.HLLC, F.PCM, F.BC_REFLECTING)
    assert cfg.problem == "blast" and p.CFL == 0.8 and p.gamma0 == 5.0 / 3.0
    # a duplicated key is joined with a newline, so the number falls back
    cfg = read_inifile(text="[mesh]\nNx=10\nNx=20\n")
    assert cfg.params.Nx == 32
    # malformed numbers fall back to the default
    cfg = read_inifile(text="[mesh]\nNx=12abc\n[solvers]\nCFL=0.5x\n")
    assert cfg.params.Nx == 32 and cfg.params.CFL == 0.8


def test_problem_generators_match_oracle():
    """Sod and blast initial conditions are the reference's, bit for bit."""
    import numpy as np
    from oracle.oracle import Oracle
    p = Params.defaults().replace(Nx=20, Ny=12, Nz=10)
    o = Oracle(p)
    d = (slice(p.kbeg, p.kend), slice(p.jbeg, p.jend), slice(p.ibeg, p.iend))
    for name, fn in [("sod_x", lambda: F.problems.sod(p, 0)), ("sod_y", lambda: F.problems.sod(p, 1)),
                     ("sod_z", lambda: F.problems.sod(p, 2)), ("blast", lambda: F.problems.blast(p)),
                     ("diffusion", lambda: F.problems.diffusion(p))]:
        assert np.array_equal(fn()[d], o.init(name)[d]), name
