"""readInifile (SimInfo.h:169-305) with inih INIReader semantics.

INIReader behaviour kept here:
  * section and key lookups are case-insensitive; values are stripped;
  * ``;`` and ``#`` start a comment line, `` ;`` starts an inline comment;
  * ``key=value`` and ``key: value`` are both accepted;
  * a key given twice has its values joined by a newline, which makes the
    numeric getters fall back to the default;
  * GetInteger parses with strtol(base 0) and GetFloat with strtod, both
    requiring the whole value to parse, else the default is returned;
  * GetBoolean accepts true/yes/on/1 and false/no/off/0 (case-insensitive).
An unreadable file yields every default, as INIReader's unchecked
ParseError does in the reference (SimInfo.h:170); a warning is printed.

Enum-valued keys go through std::map::operator[] in the reference
(SimInfo.h:207-283): an unknown string silently maps to the enumerator 0.
That is kept, with a warning.

Keys that the reference does not have (marked "extension") carry the
workload parameters of the benchmarked configurations; their defaults leave
the reference's behaviour unchanged.
"""
from __future__ import annotations

import math
import re
import sys
from dataclasses import dataclass, field
from typing import Dict, Optional


class IniFile:
    def __init__(self, path: Optional[str] = None, text: Optional[str] = None):
        self.values: Dict[str, str] = {}
        self.error = 0
        if text is None:
            try:
                with open(path, "r") as f:
                    text = f.read()
            except OSError:
                self.error = -1
                return
        self._parse(text)

    def _parse(self, text: str):
        section = ""
        prev_key = None
        for lineno, raw in enumerate(text.splitlines(), 1):
            line = raw
            if prev_key is not None and line[:1] in (" ", "\t") and line.strip() and \
                    line.strip()[0] not in ";#":
                # multi-line continuation (INI_ALLOW_MULTILINE)
                self.values[prev_key] += "\n" + _strip_inline(line.strip())
                continue
            s = line.strip()
            if not s or s[0] in ";#":
                continue
            if s[0] == "[":
                end = s.find("]")
                if end < 0:
                    if not self.error:
                        self.error = lineno
                    continue
                section = s[1:end].strip()
                prev_key = None
                continue
            m = re.search(r"[=:]", s)
            if not m:
                if not self.error:
                    self.error = lineno
                continue
            name = s[:m.start()].strip()
            value = _strip_inline(s[m.end():].strip())
            key = (section.lower() + "=" + name.lower())
            if key in self.values and self.values[key]:
                self.values[key] += "\n" + value
            else:
                self.values[key] = value
            prev_key = key

    def get(self, section: str, name: str, default: str) -> str:
        return self.values.get(section.lower() + "=" + name.lower(), default)

    def get_integer(self, section, name, default: int) -> int:
        v = self.get(section, name, None)
        if v is None:
            return default
        return _strtol(v, default)

    def get_float(self, section, name, default: float) -> float:
        v = self.get(section, name, None)
        if v is None:
            return default
        return _strtod(v, default)

    def get_boolean(self, section, name, default: bool) -> bool:
        v = self.get(section, name, None)
        if v is None:
            return default
        lv = v.lower()
        if lv in ("true", "yes", "on", "1"):
            return True
        if lv in ("false", "no", "off", "0"):
            return False
        return default


def _strip_inline(v: str) -> str:
    m = re.search(r"\s;", v)
    return v[:m.start()].rstrip() if m else v


def _strtol(v: str, default: int) -> int:
    s = v.strip()
    try:
        neg = s.startswith("-")
        body = s[1:] if s[:1] in "+-" else s
        if re.fullmatch(r"0[xX][0-9a-fA-F]+", body):
            r = int(body, 16)
        elif re.fullmatch(r"0[0-7]*", body):
            r = int(body, 8) if len(body) > 1 else 0
        elif re.fullmatch(r"[1-9][0-9]*", body):
            r = int(body, 10)
        else:
            return default
        return -r if neg else r
    except ValueError:
        return default


def _strtod(v: str, default: float) -> float:
    s = v.strip()
    if not s:
        return default
    try:
        if re.fullmatch(r"[+-]?0[xX][0-9a-fA-F]*\.?[0-9a-fA-F]*([pP][+-]?\d+)?", s):
            return float.fromhex(s)
        if re.fullmatch(r"[+-]?(inf|infinity|nan)", s, flags=re.I):
            return float(s)
        if re.fullmatch(r"[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?", s):
            return float(s)
    except ValueError:
        pass
    return default


@dataclass
class RunConfig:
    """Params plus the string members kept on the host side of the C ABI."""

    params: "object"
    problem: str = "blast"
    output_filename: str = "run"
    # extension: blast initial condition (reference values by default, Init.h:88-95)
    blast_radius: float = 0.2
    blast_rho_in: float = 1.0
    blast_p_in: float = 10.0
    blast_rho_out: float = 1.2
    blast_p_out: float = 0.1
    # extension: stop after this many steps (0: run to tend only)
    max_steps: int = 0
    warnings: list = field(default_factory=list)


_BC = {"reflecting": 1, "absorbing": 0, "periodic": 2, "C91": 3}
_RECON = {"pcm": 0, "pcm_wb": 1, "plm": 2}
_RIEMANN = {"hll": 0, "hllc": 1}
_TS = {"euler": 0, "RK2": 1}
_TCM = {"constant": 0, "B02": 1}
_BCTC = {"none": 0, "fixed_temperature": 1, "fixed_gradient": 2}
_VSC = {"constant": 0}


def _enum(table, value, what, warnings):
    if value not in table:
        warnings.append(f"unknown {what} '{value}': the reference maps it to 0")
        return 0
    return table[value]


def read_inifile(path: Optional[str] = None, text: Optional[str] = None) -> RunConfig:
    """Params from an .ini file, as readInifile (SimInfo.h:169-305) builds them."""
    from . import Params  # local import: Params lives in the package root

    r = IniFile(path, text)
    w: list = []
    if r.error:
        w.append(f"ini parse error {r.error} in {path!r}: using defaults for what was not read")
    p = Params()
    p.Nx = r.get_integer("mesh", "Nx", 32)
    p.Ny = r.get_integer("mesh", "Ny", 32)
    p.Nz = r.get_integer("mesh", "Nz", 32)
    p.Ng = r.get_integer("mesh", "Nghosts", 2)
    p.xmin = r.get_float("mesh", "xmin", 0.0)
    p.xmax = r.get_float("mesh", "xmax", 1.0)
    p.ymin = r.get_float("mesh", "ymin", 0.0)
    p.ymax = r.get_float("mesh", "ymax", 1.0)
    p.zmin = r.get_float("mesh", "zmin", 0.0)
    p.zmax = r.get_float("mesh", "zmax", 1.0)
    p.tend = r.get_float("run", "tend", 1.0)
    p.save_freq = r.get_float("run", "save_freq", 1.0e-1)
    out = r.get("run", "output_filename", "run")
    p.boundary_x = _enum(_BC, r.get("run", "boundaries_x", "reflecting"), "boundaries_x", w)
    p.boundary_y = _enum(_BC, r.get("run", "boundaries_y", "reflecting"), "boundaries_y", w)
    p.boundary_z = _enum(_BC, r.get("run", "boundaries_z", "reflecting"), "boundaries_z", w)
    p.reconstruction = _enum(_RECON, r.get("solvers", "reconstruction", "pcm"), "reconstruction", w)
    p.riemann_solver = _enum(_RIEMANN, r.get("solvers", "riemann_solver", "hllc"), "riemann_solver", w)
    p.time_stepping = _enum(_TS, r.get("solvers", "time_stepping", "euler"), "time_stepping", w)
    p.CFL = r.get_float("solvers", "CFL", 0.8)
    p.epsilon = r.get_float("misc", "epsilon", 1.0e-6)
    p.gamma0 = r.get_float("physics", "gamma0", 5.0 / 3.0)
    p.gravity = int(r.get_boolean("physics", "gravity", False))
    p.g = r.get_float("physics", "g", 0.0)
    p.m1 = r.get_float("polytrope", "m1", 1.0)
    p.theta1 = r.get_float("polytrope", "theta1", 10.0)
    p.m2 = r.get_float("polytrope", "m2", 1.0)
    p.theta2 = r.get_float("polytrope", "theta2", 10.0)
    problem = r.get("physics", "problem", "blast")
    p.well_balanced_flux_at_z_bc = int(r.get_boolean("physics", "well_balanced_flux_at_z_bc", False))
    p.thermal_conductivity_active = int(r.get_boolean("thermal_conduction", "active", False))
    p.thermal_conductivity_mode = _enum(
        _TCM, r.get("thermal_conduction", "conductivity_mode", "constant"), "conductivity_mode", w)
    p.kappa = r.get_float("thermal_conduction", "kappa", 0.0)
    p.bctc_zmin = _enum(_BCTC, r.get("thermal_conduction", "bc_zmin", "none"), "bc_zmin", w)
    p.bctc_zmax = _enum(_BCTC, r.get("thermal_conduction", "bc_zmax", "none"), "bc_zmax", w)
    p.bctc_zmin_value = r.get_float("thermal_conduction", "bc_zmin_value", 1.0)
    p.bctc_zmax_value = r.get_float("thermal_conduction", "bc_zmax_value", 1.0)
    p.viscosity_active = int(r.get_boolean("viscosity", "active", False))
    p.viscosity_mode = _enum(_VSC, r.get("viscosity", "viscosity_mode", "constant"), "viscosity_mode", w)
    p.mu = r.get_float("viscosity", "mu", 0.0)
    p.h84_pert = r.get_float("H84", "perturbation", 1.0e-4)
    p.c91_pert = r.get_float("C91", "perturbation", 1.0e-3)
    p.seed = r.get_integer("misc", "seed", 12345)
    p.log_frequency = r.get_integer("misc", "log_frequency", 10)
    p.finalize()
    cfg = RunConfig(params=p, problem=problem, output_filename=out, warnings=w)
    cfg.blast_radius = r.get_float("blast", "radius", 0.2)
    cfg.blast_rho_in = r.get_float("blast", "rho_in", 1.0)
    cfg.blast_p_in = r.get_float("blast", "p_in", 10.0)
    cfg.blast_rho_out = r.get_float("blast", "rho_out", 1.2)
    cfg.blast_p_out = r.get_float("blast", "p_out", 0.1)
    cfg.max_steps = r.get_integer("run", "max_steps", 0)
    for msg in w:
        print("fv3d warning:", msg, file=sys.stderr)
    return cfg
