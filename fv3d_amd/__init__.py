"""fv3d_amd -- MI355X-native per-timestep finite-volume hydro path of fv3d.

The compute path is libfv3d_hip.so (hand-written HIP for gfx950 behind the
C ABI in include/fv3d_hip.h).  This module binds that ABI with ctypes and
mirrors the reference's host-side interface: ``Params`` (SimInfo.h:72-159),
``read_inifile`` (SimInfo.h:169-305) and a ``Simulation`` session that owns
the device state the reference's main() allocates (main.cpp:28-30).

There is no CPU fallback: if the library is missing or no GPU is visible,
the calls raise.
"""
from __future__ import annotations

import ctypes
import os
import sys
import warnings
import weakref
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfv3d_hip.so")

# enumerator values of SimInfo.h:17-69
IR, IU, IV, IW, IP, IE = 0, 1, 2, 3, 4, 4
HLL, HLLC = 0, 1
BC_ABSORBING, BC_REFLECTING, BC_PERIODIC, BC_C91 = 0, 1, 2, 3
TS_EULER, TS_RK2 = 0, 1
PCM, PCM_WB, PLM = 0, 1, 2
TCM_CONSTANT, TCM_B02 = 0, 1
BCTC_NONE, BCTC_FIXED_TEMPERATURE, BCTC_FIXED_GRADIENT = 0, 1, 2
VSC_CONSTANT = 0

FP64, FP32, FP32_PLAIN = 0, 1, 2

ERRORS = {1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "ENONFINITE", 5: "EFLOOR"}
ENONFINITE = 4


class NonFiniteWarning(RuntimeWarning):
    """The state holds a cell with a non-finite inverse time step."""


class Params(ctypes.Structure):
    """ctypes mirror of ``fv3d_params`` (include/fv3d_params.h)."""

    _fields_ = [
        ("Nx", ctypes.c_int), ("Ny", ctypes.c_int), ("Nz", ctypes.c_int), ("Ng", ctypes.c_int),
        ("Ntx", ctypes.c_int), ("Nty", ctypes.c_int), ("Ntz", ctypes.c_int),
        ("ibeg", ctypes.c_int), ("iend", ctypes.c_int), ("jbeg", ctypes.c_int),
        ("jend", ctypes.c_int), ("kbeg", ctypes.c_int), ("kend", ctypes.c_int),
        ("xmin", ctypes.c_double), ("xmax", ctypes.c_double), ("ymin", ctypes.c_double),
        ("ymax", ctypes.c_double), ("zmin", ctypes.c_double), ("zmax", ctypes.c_double),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("save_freq", ctypes.c_double), ("tend", ctypes.c_double),
        ("boundary_x", ctypes.c_int), ("boundary_y", ctypes.c_int), ("boundary_z", ctypes.c_int),
        ("reconstruction", ctypes.c_int), ("riemann_solver", ctypes.c_int),
        ("time_stepping", ctypes.c_int), ("CFL", ctypes.c_double),
        ("epsilon", ctypes.c_double), ("gamma0", ctypes.c_double),
        ("gravity", ctypes.c_int), ("g", ctypes.c_double),
        ("well_balanced_flux_at_z_bc", ctypes.c_int), ("well_balanced", ctypes.c_int),
        ("thermal_conductivity_active", ctypes.c_int), ("thermal_conductivity_mode", ctypes.c_int),
        ("kappa", ctypes.c_double),
        ("bctc_zmin", ctypes.c_int), ("bctc_zmax", ctypes.c_int),
        ("bctc_zmin_value", ctypes.c_double), ("bctc_zmax_value", ctypes.c_double),
        ("viscosity_active", ctypes.c_int), ("viscosity_mode", ctypes.c_int),
        ("mu", ctypes.c_double),
        ("m1", ctypes.c_double), ("theta1", ctypes.c_double), ("m2", ctypes.c_double),
        ("theta2", ctypes.c_double), ("h84_pert", ctypes.c_double), ("c91_pert", ctypes.c_double),
        ("b02_zmid", ctypes.c_double), ("b02_kappa1", ctypes.c_double),
        ("b02_kappa2", ctypes.c_double), ("b02_thickness", ctypes.c_double),
        ("seed", ctypes.c_int), ("log_frequency", ctypes.c_int),
    ]

    @classmethod
    def defaults(cls) -> "Params":
        """The values readInifile gives an empty .ini (SimInfo.h:175-294)."""
        p = cls()
        p.Nx = p.Ny = p.Nz = 32
        p.Ng = 2
        p.xmin, p.xmax, p.ymin, p.ymax, p.zmin, p.zmax = 0.0, 1.0, 0.0, 1.0, 0.0, 1.0
        p.tend, p.save_freq = 1.0, 1.0e-1
        p.boundary_x = p.boundary_y = p.boundary_z = BC_REFLECTING
        p.reconstruction = PCM
        p.riemann_solver = HLLC
        p.time_stepping = TS_EULER
        p.CFL = 0.8
        p.epsilon = 1.0e-6
        p.gamma0 = 5.0 / 3.0
        p.gravity, p.g = 0, 0.0
        p.m1, p.theta1, p.m2, p.theta2 = 1.0, 10.0, 1.0, 10.0
        p.thermal_conductivity_mode = TCM_CONSTANT
        p.bctc_zmin = p.bctc_zmax = BCTC_NONE
        p.bctc_zmin_value = p.bctc_zmax_value = 1.0
        p.viscosity_mode = VSC_CONSTANT
        p.h84_pert, p.c91_pert = 1.0e-4, 1.0e-3
        p.seed, p.log_frequency = 12345, 10
        p.finalize()
        return p

    def finalize(self) -> "Params":
        """Derived mesh members (SimInfo.h:186-198)."""
        self.Ntx = self.Nx + 2 * self.Ng
        self.Nty = self.Ny + 2 * self.Ng
        self.Ntz = self.Nz + 2 * self.Ng
        self.ibeg, self.iend = self.Ng, self.Ng + self.Nx
        self.jbeg, self.jend = self.Ng, self.Ng + self.Ny
        self.kbeg, self.kend = self.Ng, self.Ng + self.Nz
        self.dx = (self.xmax - self.xmin) / self.Nx
        self.dy = (self.ymax - self.ymin) / self.Ny
        self.dz = (self.zmax - self.zmin) / self.Nz
        return self

    def copy(self) -> "Params":
        q = Params()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(self), ctypes.sizeof(Params))
        return q

    def replace(self, **kw) -> "Params":
        q = self.copy()
        for k, v in kw.items():
            if not hasattr(q, k):
                raise AttributeError(k)
            setattr(q, k, v)
        return q.finalize()

    @property
    def shape_ghosted(self):
        return (self.Ntz, self.Nty, self.Ntx, 5)

    @property
    def shape_domain(self):
        return (self.Nz, self.Ny, self.Nx, 5)


# ---------------------------------------------------------------------------
# library loading
# ---------------------------------------------------------------------------
_lib_handle: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """The loaded libfv3d_hip.so; raises if it has not been built."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m fv3d_amd.build` first")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    dp = P(ctypes.c_double)
    vp = ctypes.c_void_p
    sig = {
        "fv3d_abi_version": (ctypes.c_int, []),
        "fv3d_params_size": (ctypes.c_size_t, []),
        "fv3d_last_error": (ctypes.c_char_p, [vp]),
        "fv3d_create": (ctypes.c_int, [P(Params), ctypes.c_int, ctypes.c_int, P(vp)]),
        "fv3d_destroy": (ctypes.c_int, [vp]),
        "fv3d_set_prim": (ctypes.c_int, [vp, dp, ctypes.c_int]),
        "fv3d_get_prim": (ctypes.c_int, [vp, dp, ctypes.c_int]),
        "fv3d_get_stage_prim": (ctypes.c_int, [vp, dp, ctypes.c_int]),
        "fv3d_set_cons": (ctypes.c_int, [vp, dp, ctypes.c_int]),
        "fv3d_get_cons": (ctypes.c_int, [vp, dp, ctypes.c_int]),
        "fv3d_compute_dt": (ctypes.c_int, [vp, dp, dp]),
        "fv3d_step": (ctypes.c_int, [vp]),
        "fv3d_step_dt": (ctypes.c_int, [vp, ctypes.c_double]),
        "fv3d_steps": (ctypes.c_int, [vp, ctypes.c_int]),
        "fv3d_array_create": (ctypes.c_int, [vp, P(vp)]),
        "fv3d_array_destroy": (ctypes.c_int, [vp]),
        "fv3d_array_data": (ctypes.c_void_p, [vp]),
        "fv3d_array_upload": (ctypes.c_int, [vp, dp]),
        "fv3d_array_download": (ctypes.c_int, [vp, dp]),
        "fv3d_array_copy": (ctypes.c_int, [vp, vp]),
        "fv3d_cons_to_prim_array": (ctypes.c_int, [vp, vp, vp]),
        "fv3d_prim_to_cons_array": (ctypes.c_int, [vp, vp, vp]),
        "fv3d_fill_boundaries_array": (ctypes.c_int, [vp, vp]),
        "fv3d_compute_slopes": (ctypes.c_int, [vp, vp, vp, vp, vp]),
        "fv3d_fluxes_and_update": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, ctypes.c_double]),
        "fv3d_apply_tc": (ctypes.c_int, [vp, vp, vp, ctypes.c_double]),
        "fv3d_apply_visc": (ctypes.c_int, [vp, vp, vp, ctypes.c_double]),
        "fv3d_euler_step": (ctypes.c_int, [vp, vp, vp, ctypes.c_double]),
        "fv3d_update": (ctypes.c_int, [vp, vp, vp, ctypes.c_double]),
        "fv3d_compute_dt_array": (ctypes.c_int, [vp, vp, dp, dp]),
        "fv3d_get_cons_array": (ctypes.c_int, [vp, vp]),
        "fv3d_set_cons_array": (ctypes.c_int, [vp, vp]),
        "fv3d_step_count": (ctypes.c_long, [vp]),
        "fv3d_synchronize": (ctypes.c_int, [vp]),
        "fv3d_time_steps": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_float), P(ctypes.c_float)]),
        "fv3d_fill_boundaries": (ctypes.c_int, [P(Params), dp]),
        "fv3d_riemann": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_long, dp, dp, dp, dp]),
        "fv3d_stream_copy": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, P(ctypes.c_float)]),
        "fv3d_forward_frozen": (ctypes.c_int, [vp, ctypes.c_int, dp, P(ctypes.c_float)]),
        "fv3d_adjoint_run": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp, dp]),
        "fv3d_adjoint_times": (ctypes.c_int, [vp, P(ctypes.c_float)]),
        "fv3d_set_adjoint_store": (ctypes.c_int, [vp, ctypes.c_size_t]),
        "fv3d_batch_create": (ctypes.c_int, [ctypes.c_int, P(Params), ctypes.c_int, ctypes.c_int, P(vp)]),
        "fv3d_batch_destroy": (ctypes.c_int, [vp]),
        "fv3d_batch_last_error": (ctypes.c_char_p, [vp]),
        "fv3d_batch_set_prim": (ctypes.c_int, [vp, ctypes.c_int, dp, ctypes.c_int]),
        "fv3d_batch_get_prim": (ctypes.c_int, [vp, ctypes.c_int, dp, ctypes.c_int]),
        "fv3d_batch_compute_dt": (ctypes.c_int, [vp, dp]),
        "fv3d_batch_step": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "fv3d_batch_time": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_float)]),
        "fv3d_batch_step_count": (ctypes.c_long, [vp]),
        "fv3d_batch_cells": (ctypes.c_longlong, [vp]),
        "fv3d_batch_configure": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
        "fv3d_batch_set_tile_width": (ctypes.c_int, [vp, ctypes.c_int]),
        "fv3d_batch_tile_width": (ctypes.c_int, [vp]),
        "fv3d_batch_tune_shape": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_char_p, P(ctypes.c_int),
                                                 P(ctypes.c_int), P(ctypes.c_int)]),
        "fv3d_batch_tune": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_char_p, P(ctypes.c_int),
                                           P(ctypes.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.fv3d_params_size() != ctypes.sizeof(Params):
        raise RuntimeError("fv3d_params layout mismatch between libfv3d_hip.so and fv3d_amd.Params")
    _lib_handle = L
    return L


class FV3DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fv3d error {ERRORS.get(code, code)}: {msg}")
        self.code = code


def _check(rc: int, ctx=None):
    if rc != 0:
        msg = lib().fv3d_last_error(ctx)
        raise FV3DError(rc, msg.decode() if msg else "")


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ---------------------------------------------------------------------------
# session
# ---------------------------------------------------------------------------
class Simulation:
    """Device state and functors of one box (main.cpp:28-41).

    ``set_prim(Q)`` is init + primToCons; ``compute_dt()`` is
    ComputeDtFunctor::computeDt on prim(U); ``step()`` is UpdateFunctor::update
    followed by the U <- Unew copy of main.cpp:71.
    """

    def __init__(self, params: Params, precision: int = FP64, device: int = 0):
        self.params = params.copy()
        self.nonfinite = False
        self._ctx = ctypes.c_void_p()
        self._arrays = weakref.WeakSet()
        _check(lib().fv3d_create(ctypes.byref(self.params), precision, device, ctypes.byref(self._ctx)))

    def close(self):
        if self._ctx:
            for a in list(self._arrays):  # arrays die before their context
                a.close()
            lib().fv3d_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _shape(self, with_ghosts):
        return self.params.shape_ghosted if with_ghosts else self.params.shape_domain

    def set_prim(self, Q: np.ndarray, with_ghosts: bool = True):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        if Q.shape != self._shape(with_ghosts):
            raise ValueError(f"Q has shape {Q.shape}, expected {self._shape(with_ghosts)}")
        _check(lib().fv3d_set_prim(self._ctx, _dptr(Q), int(with_ghosts)), self._ctx)
        self.nonfinite = False

    def set_cons(self, U: np.ndarray, with_ghosts: bool = True):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if U.shape != self._shape(with_ghosts):
            raise ValueError(f"U has shape {U.shape}, expected {self._shape(with_ghosts)}")
        _check(lib().fv3d_set_cons(self._ctx, _dptr(U), int(with_ghosts)), self._ctx)
        self.nonfinite = False

    def get_prim(self, with_ghosts: bool = True) -> np.ndarray:
        out = np.zeros(self._shape(with_ghosts))
        _check(lib().fv3d_get_prim(self._ctx, _dptr(out), int(with_ghosts)), self._ctx)
        return out

    def get_stage_prim(self, with_ghosts: bool = True) -> np.ndarray:
        out = np.zeros(self._shape(with_ghosts))
        _check(lib().fv3d_get_stage_prim(self._ctx, _dptr(out), int(with_ghosts)), self._ctx)
        return out

    def get_cons(self, with_ghosts: bool = True) -> np.ndarray:
        out = np.zeros(self._shape(with_ghosts))
        _check(lib().fv3d_get_cons(self._ctx, _dptr(out), int(with_ghosts)), self._ctx)
        return out

    def compute_dt(self, strict: bool = False):
        """Returns (dt, (inv_dt_hyp, inv_dt_par_tc, inv_dt_par_visc)).

        A cell whose inverse time step is not finite (negative pressure or
        zero density) is left out of the maximum, as the reference's Max
        reducer leaves out NaN (ComputeDt.h:51-58), and its run goes on.  The
        session then sets ``nonfinite`` and warns once; with ``strict`` it
        raises FV3DError (ENONFINITE) instead."""
        inv = np.zeros(3)
        dt = ctypes.c_double()
        rc = lib().fv3d_compute_dt(self._ctx, _dptr(inv), ctypes.byref(dt))
        if rc == ENONFINITE and not strict:
            if not self.nonfinite:
                warnings.warn("fv3d: a cell has a non-finite inverse time step (negative pressure or "
                              "zero density); it is left out of dt, as in the reference",
                              NonFiniteWarning, stacklevel=2)
            self.nonfinite = True
        else:
            _check(rc, self._ctx)
        return dt.value, tuple(inv)

    def step(self, n: int = 1):
        _check(lib().fv3d_steps(self._ctx, int(n)), self._ctx)

    def step_dt(self, dt: float):
        """UpdateFunctor::update(Q, Unew, dt) with the caller's dt (Update.h:180)."""
        _check(lib().fv3d_step_dt(self._ctx, float(dt)), self._ctx)

    # ---- the reference's functors one pass at a time, on device arrays ----
    def array(self, host: Optional[np.ndarray] = None) -> "Array":
        """A new device Array of this box (zero, or a copy of `host`)."""
        a = Array(self)
        if host is not None:
            a.upload(host)
        return a

    def _fn(self, name, *args):
        _check(getattr(lib(), name)(self._ctx, *args), self._ctx)

    def cons_to_prim_array(self, U: "Array", Q: "Array"):
        self._fn("fv3d_cons_to_prim_array", U._a, Q._a)

    def prim_to_cons_array(self, Q: "Array", U: "Array"):
        self._fn("fv3d_prim_to_cons_array", Q._a, U._a)

    def fill_boundaries_array(self, Q: "Array"):
        self._fn("fv3d_fill_boundaries_array", Q._a)

    def compute_slopes(self, Q: "Array", sx: "Array", sy: "Array", sz: "Array"):
        self._fn("fv3d_compute_slopes", Q._a, sx._a, sy._a, sz._a)

    def fluxes_and_update(self, Q, sx, sy, sz, Unew, dt: float):
        self._fn("fv3d_fluxes_and_update", Q._a, sx._a, sy._a, sz._a, Unew._a, float(dt))

    def apply_tc(self, Q: "Array", Unew: "Array", dt: float):
        self._fn("fv3d_apply_tc", Q._a, Unew._a, float(dt))

    def apply_visc(self, Q: "Array", Unew: "Array", dt: float):
        self._fn("fv3d_apply_visc", Q._a, Unew._a, float(dt))

    def euler_step(self, Q: "Array", Unew: "Array", dt: float):
        self._fn("fv3d_euler_step", Q._a, Unew._a, float(dt))

    def update(self, Q: "Array", Unew: "Array", dt: float):
        self._fn("fv3d_update", Q._a, Unew._a, float(dt))

    def compute_dt_array(self, Q: "Array", strict: bool = False):
        """(dt, (inv_dt_hyp, inv_dt_par_tc, inv_dt_par_visc)) of a primitive
        array.  Non-finite cells are left out of the maximum with a
        NonFiniteWarning, or raise with ``strict``, as in compute_dt."""
        inv = np.zeros(3)
        dt = ctypes.c_double()
        rc = lib().fv3d_compute_dt_array(self._ctx, Q._a, _dptr(inv), ctypes.byref(dt))
        if rc == ENONFINITE and not strict:
            warnings.warn("fv3d: a cell of the array has a non-finite inverse time step; it is "
                          "left out of dt, as in the reference", NonFiniteWarning, stacklevel=2)
        else:
            _check(rc, self._ctx)
        return dt.value, tuple(inv)

    def get_cons_array(self, U: "Array"):
        self._fn("fv3d_get_cons_array", U._a)

    def set_cons_array(self, U: "Array"):
        self._fn("fv3d_set_cons_array", U._a)
        self.nonfinite = False

    @property
    def step_count(self) -> int:
        return int(lib().fv3d_step_count(self._ctx))

    def synchronize(self):
        _check(lib().fv3d_synchronize(self._ctx), self._ctx)

    def time_steps(self, n: int, per_stage: bool = False):
        """HIP-event timing of n steps on the session's stream.

        Returns (total_ms, per-launch stage ms array or None)."""
        tot = ctypes.c_float()
        nst = 2 if self.params.time_stepping == TS_RK2 else 1
        st = np.zeros(n * nst, dtype=np.float32) if per_stage else None
        _check(lib().fv3d_time_steps(
            self._ctx, int(n), ctypes.byref(tot),
            st.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if per_stage else None), self._ctx)
        return tot.value, st

    # ---- frozen-dt runs and the discrete adjoint ----
    def forward_frozen(self, dts) -> float:
        """Advance the state by len(dts) steps with the given dt sequence;
        returns the device time in ms."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        ms = ctypes.c_float()
        _check(lib().fv3d_forward_frozen(self._ctx, len(dts), _dptr(dts), ctypes.byref(ms)), self._ctx)
        return ms.value

    def adjoint(self, dts, dJ_dUT: np.ndarray, K: int = 20, store_bytes: int = 0):
        """dJ/dU0 (Nz, Ny, Nx, 5), dJ/dkappa, dJ/dmu for J(U_T) after the
        frozen-dt run of len(dts) steps from the current state, given
        dJ/dU_T (Nz, Ny, Nx, 5).  Checkpoints every K steps; segments that
        fit in store_bytes of device memory are kept from the forward pass
        instead of being recomputed."""
        _check(lib().fv3d_set_adjoint_store(self._ctx, int(store_bytes)), self._ctx)
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        lamT = np.ascontiguousarray(dJ_dUT, dtype=np.float64)
        if lamT.shape != self.params.shape_domain:
            raise ValueError(f"dJ_dUT has shape {lamT.shape}, expected {self.params.shape_domain}")
        out = np.zeros(self.params.shape_domain)
        dk, dm = ctypes.c_double(), ctypes.c_double()
        _check(lib().fv3d_adjoint_run(self._ctx, len(dts), int(K), _dptr(dts), _dptr(lamT), _dptr(out),
                                      ctypes.byref(dk), ctypes.byref(dm)), self._ctx)
        return out, dk.value, dm.value

    def adjoint_times(self):
        """(forward pass, backward pass, total) device ms of the last adjoint run."""
        ms = (ctypes.c_float * 3)()
        _check(lib().fv3d_adjoint_times(self._ctx, ms), self._ctx)
        return tuple(ms)


class Array:
    """One Array of the reference (Kokkos::View<double****> of extents
    (Ntz, Nty, Ntx, 5), SimInfo.h:14) in device memory, owned by a
    Simulation; upload/download use that host layout."""

    def __init__(self, sim: Simulation):
        self.sim = sim
        self._a = ctypes.c_void_p()
        _check(lib().fv3d_array_create(sim._ctx, ctypes.byref(self._a)), sim._ctx)
        sim._arrays.add(self)

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, dtype=np.float64)
        if host.shape != self.sim.params.shape_ghosted:
            raise ValueError(f"shape {host.shape}, expected {self.sim.params.shape_ghosted}")
        _check(lib().fv3d_array_upload(self._a, _dptr(host)), self.sim._ctx)

    def download(self) -> np.ndarray:
        out = np.zeros(self.sim.params.shape_ghosted)
        _check(lib().fv3d_array_download(self._a, _dptr(out)), self.sim._ctx)
        return out

    def copy_from(self, other: "Array"):
        _check(lib().fv3d_array_copy(self._a, other._a), self.sim._ctx)

    def close(self):
        if self._a:
            lib().fv3d_array_destroy(self._a)
            self._a = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_GRAPH, BATCH_PERSISTENT, BATCH_LOOP = 0, 1, 2


class Batch:
    """An ensemble of independent boxes advanced together (fv3d_batch_* in
    include/fv3d_hip.h).  Each box is one reference run (main.cpp:51-76)
    with its own Params and dt; the boxes must share reconstruction, solver,
    time stepping and the conduction / viscosity switches."""

    def __init__(self, params_list, precision: int = FP32, device: int = 0):
        self.params = [q.copy() for q in params_list]
        arr = (Params * len(self.params))(*self.params)
        self._b = ctypes.c_void_p()
        _check(lib().fv3d_batch_create(len(self.params), arr, precision, device, ctypes.byref(self._b)))

    def _check(self, rc):
        if rc != 0:
            msg = lib().fv3d_batch_last_error(self._b)
            raise FV3DError(rc, msg.decode() if msg else "")

    def close(self):
        if self._b:
            lib().fv3d_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __len__(self):
        return len(self.params)

    def set_prim(self, box: int, Q: np.ndarray, with_ghosts: bool = True):
        p = self.params[box]
        shape = p.shape_ghosted if with_ghosts else p.shape_domain
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        if Q.shape != shape:
            raise ValueError(f"Q has shape {Q.shape}, expected {shape}")
        self._check(lib().fv3d_batch_set_prim(self._b, box, _dptr(Q), int(with_ghosts)))

    def get_prim(self, box: int, with_ghosts: bool = True) -> np.ndarray:
        p = self.params[box]
        out = np.zeros(p.shape_ghosted if with_ghosts else p.shape_domain)
        self._check(lib().fv3d_batch_get_prim(self._b, box, _dptr(out), int(with_ghosts)))
        return out

    def compute_dt(self) -> np.ndarray:
        """The next dt of every box.  Raises FV3DError (ENONFINITE) when a
        box holds a cell with a non-finite inverse time step: unlike a single
        run, which goes on as the reference does, an ensemble has no
        counterpart to follow, and one broken box among hundreds would
        otherwise pass unnoticed."""
        dts = np.zeros(len(self.params))
        self._check(lib().fv3d_batch_compute_dt(self._b, _dptr(dts)))
        return dts

    def step(self, n: int = 1, mode: int = BATCH_GRAPH):
        self._check(lib().fv3d_batch_step(self._b, int(n), int(mode)))

    def time(self, n: int, mode: int = BATCH_GRAPH) -> float:
        ms = ctypes.c_float()
        self._check(lib().fv3d_batch_time(self._b, int(n), int(mode), ctypes.byref(ms)))
        return ms.value

    @property
    def step_count(self) -> int:
        return int(lib().fv3d_batch_step_count(self._b))

    @property
    def cells(self) -> int:
        return int(lib().fv3d_batch_cells(self._b))

    def configure(self, nzc_cap: int = 0, grid: int = 0):
        self._check(lib().fv3d_batch_configure(self._b, int(nzc_cap), int(grid)))

    @property
    def tile_width(self) -> int:
        """Stage tile width (64 or 32 columns) of the persistent kernels."""
        return int(lib().fv3d_batch_tile_width(self._b))

    def set_tile_width(self, wx: int = 0):
        """64 or 32 columns, or 0 for the default (by tile cover)."""
        self._check(lib().fv3d_batch_set_tile_width(self._b, int(wx)))

    def tune(self, nsteps: int = 3, cache: Optional[str] = None):
        """Autotune the persistent launch shape (tile width, z-chunk cap,
        grid); returns (nzc_cap, grid); the width is `tile_width` after."""
        nzc, grid, wx = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(lib().fv3d_batch_tune_shape(self._b, int(nsteps), cache.encode() if cache else None,
                                                ctypes.byref(nzc), ctypes.byref(grid), ctypes.byref(wx)))
        return nzc.value, grid.value


def fill_boundaries(params: Params, Q: np.ndarray) -> np.ndarray:
    """BoundaryManager::fillBoundaries on a host primitive array (returns a copy)."""
    Q = np.array(Q, dtype=np.float64, order="C", copy=True)
    if Q.shape != params.shape_ghosted:
        raise ValueError("Q must include ghosts")
    _check(lib().fv3d_fill_boundaries(ctypes.byref(params), _dptr(Q)))
    return Q


def stream_copy_gbs(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Measured device copy bandwidth (GB/s, read + write bytes)."""
    ms = ctypes.c_float()
    _check(lib().fv3d_stream_copy(int(nbytes), int(reps), ctypes.byref(ms)))
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9


def riemann(solver: int, gamma0: float, qL: np.ndarray, qR: np.ndarray):
    """hll / hllc on [n, 5] face-frame states; returns (flux [n, 5], pout [n])."""
    qL = np.ascontiguousarray(qL, dtype=np.float64).reshape(-1, 5)
    qR = np.ascontiguousarray(qR, dtype=np.float64).reshape(-1, 5)
    n = qL.shape[0]
    flux = np.zeros((n, 5))
    pout = np.zeros(n)
    _check(lib().fv3d_riemann(int(solver), float(gamma0), n, _dptr(qL), _dptr(qR), _dptr(flux), _dptr(pout)))
    return flux, pout


from .config import read_inifile, RunConfig  # noqa: E402  (pure-Python, no GPU needed)
from . import problems  # noqa: E402
