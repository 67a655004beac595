"""The allocation-free Riemann entry point (fv3d_riemann_noalloc) through
the C ABI: bit-equal to fv3d_riemann on two calls in a row, the second with
fewer faces than the first, at the capacity limit, and rejecting one face
more than the limit."""
import ctypes

import numpy as np
import pytest

import fv3d_amd as F
from tests.test_gpu_parity import _riemann_states

pytestmark = pytest.mark.gpu

MAX_FACES = 65536  # FV3D_NOALLOC_MAX_FACES


def _noalloc(solver, gamma, qL, qR):
    L = F.lib()
    f = L.fv3d_riemann_noalloc
    dp = ctypes.POINTER(ctypes.c_double)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_long, dp, dp, dp, dp]
    n = qL.shape[0]
    flux = np.full((n, 5), np.nan)
    pout = np.full(n, np.nan)
    rc = f(solver, gamma, n, F._dptr(qL), F._dptr(qR), F._dptr(flux), F._dptr(pout))
    return rc, flux, pout


@pytest.mark.parametrize("solver", [F.HLL, F.HLLC])
def test_riemann_noalloc_matches_allocating_entry(gpu_available, solver):
    gamma = 5.0 / 3.0
    for n, seed in ((MAX_FACES, 3), (1000, 4), (MAX_FACES, 5), (7, 6)):
        qL, qR = _riemann_states(n, seed=seed + 10 * solver)
        rc, f_na, p_na = _noalloc(solver, gamma, qL, qR)
        assert rc == 0, F.lib().fv3d_last_error(None)
        f_ref, p_ref = F.riemann(solver, gamma, qL, qR)
        assert np.array_equal(f_na, f_ref), n
        assert np.array_equal(p_na, p_ref), n


def test_riemann_noalloc_rejects_beyond_capacity(gpu_available):
    qL, qR = _riemann_states(MAX_FACES + 1, seed=1)
    rc, _, _ = _noalloc(F.HLLC, 1.4, qL, qR)
    assert rc == 1  # FV3D_EINVAL
    assert b"exceeds" in F.lib().fv3d_last_error(None)
    rc, _, _ = _noalloc(F.HLLC, 1.4, qL[:0], qR[:0])
    assert rc == 0
