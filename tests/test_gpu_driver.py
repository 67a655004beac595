"""The fv3d driver executable (C++ host loop on libfv3d_hip) against the
oracle's restatement of main.cpp: every snapshot of the HDF5 file, the
file's attributes and vertex coordinates, the XDMF side file and the log
cadence."""
import os
import re
import shutil
import subprocess
import xml.dom.minidom

import numpy as np
import pytest

import fv3d_amd as F
from fv3d_amd import read_inifile
from oracle.oracle import MainLoop, Oracle
from tests.h5read import H5File

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "fv3d_amd", "fv3d")
SOD_INI = os.path.join(ROOT, "tests", "data", "sod_c1.ini")


C91_INI = os.path.join(ROOT, "tests", "data", "c91_small.ini")


def _run_driver(ini, cwd, *extra):
    return subprocess.run([DRIVER, ini, *extra], cwd=cwd, capture_output=True, text=True, timeout=300)


def test_driver_usage_and_unknown_problem(tmp_path):
    """No GPU needed: argument and problem errors are reported before the device is touched."""
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    ini = tmp_path / "bad.ini"
    ini.write_text("[physics]\nproblem=kelvin-helmholtz\n")
    r = _run_driver(str(ini), str(tmp_path))
    assert r.returncode == 1 and "Error unknown problem kelvin-helmholtz" in r.stderr


@pytest.mark.gpu
def test_driver_sod_snapshots_match_oracle(gpu_available, tmp_path):
    shutil.copy(SOD_INI, tmp_path / "sod.ini")
    r = _run_driver("sod.ini", str(tmp_path))
    assert r.returncode == 0, r.stderr
    cfg = read_inifile(SOD_INI)
    p = cfg.params
    Q0 = Oracle(p).init("sod_x")
    steps, snaps = MainLoop(p, Q0).run()

    # log cadence: step 0, then every log_frequency+1 steps (main.cpp:58-61)
    logs = re.findall(r"Computing dts at \(t=(\S+)\) : dt_hyp=(\S+); dt_TC=(\S+); dt_visc=(\S+)", r.stdout)
    logged = [s for s in steps if s[3]]
    assert len(logs) == len(logged)
    for (t, h, tc, vi), (ts, dt, inv, _) in zip(logs, logged):
        assert float(t) == pytest.approx(ts, rel=1e-5, abs=1e-12)
        assert float(h) == pytest.approx(1.0 / inv[0], rel=1e-5)
        assert float(tc) == pytest.approx(1e6) and float(vi) == pytest.approx(1e6)
    assert r.stdout.count(" - Saving at time") == len(snaps) - 1

    with H5File(str(tmp_path / "sod.h5")) as f:
        for name, v in [("Ntx", p.Ntx), ("Nty", p.Nty), ("Ntz", p.Ntz), ("Nx", p.Nx), ("Ny", p.Ny),
                        ("Nz", p.Nz), ("ibeg", p.ibeg), ("iend", p.iend), ("jbeg", p.jbeg),
                        ("jend", p.jend), ("kbeg", p.kbeg), ("kend", p.kend)]:
            assert f.int_attr(name) == v, name
        x = f.dataset("x")
        k, j, i = np.meshgrid(np.arange(p.Nz + 1), np.arange(p.Ny + 1), np.arange(p.Nx + 1), indexing="ij")
        assert np.array_equal(x, (i * p.dx).ravel())
        assert np.array_equal(f.dataset("z"), (k * p.dz).ravel())
        for ite, t, Q in snaps:
            g = f"ite_{ite}"
            assert f.group_double_attr(g, "time") == t
            d = Q[p.kbeg:p.kend, p.jbeg:p.jend, p.ibeg:p.iend]
            for field, slot in [("rho", 0), ("u", 1), ("v", 2), ("w", 3), ("prs", 4)]:
                assert np.array_equal(f.dataset(f"{g}/{field}"), d[..., slot]), (g, field)
        assert not f.exists(f"ite_{len(snaps)}")

    doc = xml.dom.minidom.parse(str(tmp_path / "sod.xdmf"))
    grids = [g for g in doc.getElementsByTagName("Grid") if g.getAttribute("Name") == "Cells"]
    assert len(grids) == len(snaps)
    first = grids[0].getElementsByTagName("Attribute")[0].getElementsByTagName("DataItem")[0]
    assert first.firstChild.data.strip() == "sod.h5:/ite_0/rho"


@pytest.mark.gpu
def test_driver_fp32_runs(gpu_available, tmp_path):
    shutil.copy(SOD_INI, tmp_path / "sod.ini")
    r = _run_driver("sod.ini", str(tmp_path), "--precision", "fp32")
    assert r.returncode == 0, r.stderr
    with H5File(str(tmp_path / "sod.h5")) as f:
        rho = f.dataset("ite_10/rho")
    assert np.all(np.isfinite(rho)) and rho.min() > 0.1


@pytest.mark.usefixtures("nzc")
@pytest.mark.gpu
def test_driver_c91_settings_match_oracle(gpu_available, tmp_path):
    """The reference's settings/C91.ini setup (C91 z boundaries, conduction
    with fixed-temperature / fixed-gradient walls, viscosity, gravity) on a
    16^3 grid: every snapshot of the driver's file against the oracle."""
    shutil.copy(C91_INI, tmp_path / "c91.ini")
    r = _run_driver("c91.ini", str(tmp_path))
    assert r.returncode == 0, r.stderr
    cfg = read_inifile(C91_INI)
    p = cfg.params
    Q0 = F.problems.initial_condition(cfg)
    steps, snaps = MainLoop(p, Q0).run()
    assert len(snaps) >= 3
    # settings/C91.ini names no output file: the default name applies
    with H5File(str(tmp_path / (cfg.output_filename + ".h5"))) as f:
        for ite, t, Q in snaps:
            g = f"ite_{ite}"
            assert f.group_double_attr(g, "time") == t
            d = Q[p.kbeg:p.kend, p.jbeg:p.jend, p.ibeg:p.iend]
            for field, slot in [("rho", 0), ("u", 1), ("v", 2), ("w", 3), ("prs", 4)]:
                assert np.array_equal(f.dataset(f"{g}/{field}"), d[..., slot]), (g, field)


STRONG_BLAST_INI = """; a blast strong enough (pressure ratio 1e6) to drive a cell's pressure negative
[mesh]
Nx=32
Ny=32
Nz=32

[run]
tend=0.1
save_freq=0.05
boundaries_x=reflecting
boundaries_y=reflecting
boundaries_z=reflecting
output_filename=strong

[solvers]
riemann_solver=hllc
reconstruction=plm
time_stepping=RK2
CFL=0.8

[physics]
gamma0=1.6666666666666667
problem=blast

[blast]
radius=0.1
rho_in=1.0
p_in=1000.0
rho_out=1.0
p_out=0.001
"""


@pytest.mark.gpu
def test_driver_runs_through_negative_pressure(gpu_available, tmp_path):
    """The reference leaves cells with a NaN sound speed out of dt and goes
    on (ComputeDt.h:51-58).  The driver warns once, finishes the run, and
    its snapshots match the oracle's bit for bit, negative pressures
    included."""
    (tmp_path / "strong.ini").write_text(STRONG_BLAST_INI)
    r = _run_driver("strong.ini", str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert r.stderr.count("fv3d: warning") == 1 and "non-finite" in r.stderr, r.stderr
    cfg = read_inifile(str(tmp_path / "strong.ini"))
    p = cfg.params
    steps, snaps = MainLoop(p, F.problems.initial_condition(cfg)).run()
    with H5File(str(tmp_path / "strong.h5")) as f:
        for ite, t, Q in snaps:
            g = f"ite_{ite}"
            assert f.group_double_attr(g, "time") == t
            d = Q[p.kbeg:p.kend, p.jbeg:p.jend, p.ibeg:p.iend]
            for field, slot in [("rho", 0), ("u", 1), ("v", 2), ("w", 3), ("prs", 4)]:
                assert np.array_equal(f.dataset(f"{g}/{field}"), d[..., slot]), (g, field)
    assert np.nanmin(snaps[-1][2][p.kbeg:p.kend, p.jbeg:p.jend, p.ibeg:p.iend, 4]) < 0.0
