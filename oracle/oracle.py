"""ctypes binding of liboracle.so -- the CPU restatement of fv3d's hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's CPU-baseline leg.  Arrays are numpy float64 in
the reference's host layout (Ntz, Nty, Ntx, 5).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PROBLEM_IDS = {"sod_x": 0, "sod_y": 1, "sod_z": 2, "blast": 3, "rayleigh-taylor": 4,
               "diffusion": 5, "H84": 6, "C91": 7}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `python -m oracle.build`")
        from fv3d_amd import Params
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        pp = P(Params)
        dp = P(ctypes.c_double)
        sig = {
            "orc_state_prim_to_cons": (None, [pp, dp, dp]),
            "orc_state_cons_to_prim": (None, [pp, dp, dp]),
            "orc_compute_flux": (None, [pp, dp, dp]),
            "orc_riemann_batch": (None, [pp, ctypes.c_int, ctypes.c_long, dp, dp, dp, dp]),
            "orc_cons_to_prim": (None, [pp, dp, dp]),
            "orc_prim_to_cons": (None, [pp, dp, dp]),
            "orc_fill_boundaries": (None, [pp, dp]),
            "orc_compute_slopes": (None, [pp, dp, dp, dp, dp]),
            "orc_fluxes_and_update": (None, [pp, dp, dp, dp, dp, dp, ctypes.c_double]),
            "orc_thermal_conduction": (None, [pp, dp, dp, ctypes.c_double]),
            "orc_viscosity": (None, [pp, dp, dp, ctypes.c_double]),
            "orc_work_size": (ctypes.c_long, [pp]),
            "orc_euler_step": (None, [pp, dp, dp, ctypes.c_double, dp]),
            "orc_update": (None, [pp, dp, dp, ctypes.c_double, dp]),
            "orc_compute_dt": (ctypes.c_double, [pp, dp, dp]),
            "orc_init": (ctypes.c_int, [pp, ctypes.c_int, dp]),
            "orc_run": (ctypes.c_long, [pp, dp, dp, dp, dp, ctypes.c_long, ctypes.c_int, dp, dp]),
            "orc_num_threads": (ctypes.c_int, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need C-contiguous float64"
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """The reference's functors, on host arrays with ghosts."""

    def __init__(self, params):
        self.p = params.copy()
        self.shape = (params.Ntz, params.Nty, params.Ntx, 5)
        self.work = np.zeros(lib().orc_work_size(ctypes.byref(self.p)))

    def zeros(self):
        return np.zeros(self.shape)

    def init(self, problem):
        Q = self.zeros()
        pid = PROBLEM_IDS[problem] if isinstance(problem, str) else int(problem)
        if lib().orc_init(ctypes.byref(self.p), pid, _p(Q)) != 0:
            raise RuntimeError("unknown problem")
        return Q

    def cons_to_prim(self, U):
        Q = self.zeros()
        lib().orc_cons_to_prim(ctypes.byref(self.p), _p(U), _p(Q))
        return Q

    def prim_to_cons(self, Q):
        U = self.zeros()
        lib().orc_prim_to_cons(ctypes.byref(self.p), _p(Q), _p(U))
        return U

    def fill_boundaries(self, Q):
        lib().orc_fill_boundaries(ctypes.byref(self.p), _p(Q))
        return Q

    def compute_slopes(self, Q):
        sx, sy, sz = self.zeros(), self.zeros(), self.zeros()
        lib().orc_compute_slopes(ctypes.byref(self.p), _p(Q), _p(sx), _p(sy), _p(sz))
        return sx, sy, sz

    def fluxes_and_update(self, Q, sx, sy, sz, Unew, dt):
        lib().orc_fluxes_and_update(ctypes.byref(self.p), _p(Q), _p(sx), _p(sy), _p(sz), _p(Unew), dt)

    def thermal_conduction(self, Q, Unew, dt):
        lib().orc_thermal_conduction(ctypes.byref(self.p), _p(Q), _p(Unew), dt)

    def viscosity(self, Q, Unew, dt):
        lib().orc_viscosity(ctypes.byref(self.p), _p(Q), _p(Unew), dt)

    def compute_dt(self, Q):
        inv = np.zeros(3)
        dt = lib().orc_compute_dt(ctypes.byref(self.p), _p(Q), _p(inv))
        return dt, tuple(inv)

    def update(self, Q, Unew, dt):
        """UpdateFunctor::update: mutates Q and Unew in place."""
        lib().orc_update(ctypes.byref(self.p), _p(Q), _p(Unew), dt, _p(self.work))

    def euler_step(self, Q, Unew, dt):
        """UpdateFunctor::euler_step: fills Q's ghosts and updates Unew in place."""
        lib().orc_euler_step(ctypes.byref(self.p), _p(Q), _p(Unew), dt, _p(self.work))

    def riemann(self, solver, qL, qR):
        qL = np.ascontiguousarray(qL, dtype=np.float64).reshape(-1, 5)
        qR = np.ascontiguousarray(qR, dtype=np.float64).reshape(-1, 5)
        n = qL.shape[0]
        flux = np.zeros((n, 5))
        pout = np.zeros(n)
        lib().orc_riemann_batch(ctypes.byref(self.p), int(solver), n, _p(qL), _p(qR), _p(flux), _p(pout))
        return flux, pout

    def run(self, Q0, max_steps, stop_at_tend=True):
        """main.cpp:43-74 from the initial primitive array Q0: returns
        (U, Q, dts, t) after the loop (no snapshots)."""
        U = self.prim_to_cons(np.ascontiguousarray(Q0))
        Q = self.zeros()
        Unew = self.zeros()
        dts = np.zeros(max_steps)
        t = ctypes.c_double()
        n = lib().orc_run(ctypes.byref(self.p), _p(U), _p(Q), _p(Unew), _p(self.work), int(max_steps),
                          int(stop_at_tend), _p(dts), ctypes.byref(t))
        return U, Q, dts[:n], t.value


class MainLoop:
    """main.cpp:51-76 step by step on the oracle, recording what the
    reference would print and save: (t, dt, inv) per step and the snapshots
    (ite, t, Q) taken when t+eps > next_save, plus the final stale one."""

    def __init__(self, params, Q0):
        self.o = Oracle(params)
        self.p = params
        self.U = self.o.prim_to_cons(np.ascontiguousarray(Q0))
        self.Q = self.o.zeros()
        self.Unew = self.o.zeros()

    def run(self, max_steps=None):
        p = self.p
        t, ite, next_save, next_log = 0.0, 0, 0.0, 0
        steps, snaps = [], []
        dt = 0.0
        while t + p.epsilon < p.tend and (max_steps is None or len(steps) < max_steps):
            self.Unew[...] = self.U
            save_needed = t + p.epsilon > next_save
            lib().orc_cons_to_prim(ctypes.byref(self.o.p), _p(self.U), _p(self.Q))
            dt, inv = self.o.compute_dt(self.Q)
            logged = next_log == 0
            next_log = p.log_frequency if next_log == 0 else next_log - 1
            if save_needed:
                snaps.append((ite, t, self.Q.copy()))
                ite += 1
                next_save += p.save_freq
            self.o.update(self.Q, self.Unew, dt)
            self.U[...] = self.Unew
            steps.append((t, dt, inv, logged))
            t += dt
        snaps.append((ite, t, self.Q.copy()))
        return steps, snaps
