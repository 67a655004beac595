"""Minimal HDF5 reader over the HDF5 C library (ctypes), for checking the
driver's snapshot files (h5py is not available)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_L = None
HID = ctypes.c_int64


def lib():
    global _L
    if _L is None:
        root = os.environ.get("FV3D_HDF5_ROOT", "/opt/conda")
        L = ctypes.CDLL(os.path.join(root, "lib", "libhdf5.so"))
        L.H5open()
        for n, r, a in [("H5Fopen", HID, [ctypes.c_char_p, ctypes.c_uint, HID]),
                        ("H5Fclose", ctypes.c_int, [HID]),
                        ("H5Dopen2", HID, [HID, ctypes.c_char_p, HID]),
                        ("H5Dclose", ctypes.c_int, [HID]),
                        ("H5Dget_space", HID, [HID]),
                        ("H5Sget_simple_extent_ndims", ctypes.c_int, [HID]),
                        ("H5Sget_simple_extent_dims", ctypes.c_int, [HID, ctypes.c_void_p, ctypes.c_void_p]),
                        ("H5Sclose", ctypes.c_int, [HID]),
                        ("H5Dread", ctypes.c_int, [HID, HID, HID, HID, HID, ctypes.c_void_p]),
                        ("H5Aopen", HID, [HID, ctypes.c_char_p, HID]),
                        ("H5Aread", ctypes.c_int, [HID, HID, ctypes.c_void_p]),
                        ("H5Aclose", ctypes.c_int, [HID]),
                        ("H5Lexists", ctypes.c_int, [HID, ctypes.c_char_p, HID]),
                        ("H5Gopen2", HID, [HID, ctypes.c_char_p, HID]),
                        ("H5Gclose", ctypes.c_int, [HID])]:
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _L = L
    return _L


def _type(name):
    return HID.in_dll(lib(), name).value


class H5File:
    def __init__(self, path):
        self.f = lib().H5Fopen(path.encode(), 0, 0)
        if self.f < 0:
            raise OSError(path)

    def close(self):
        lib().H5Fclose(self.f)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def exists(self, name):
        return lib().H5Lexists(self.f, name.encode(), 0) > 0

    def dataset(self, name):
        d = lib().H5Dopen2(self.f, name.encode(), 0)
        if d < 0:
            raise KeyError(name)
        sp = lib().H5Dget_space(d)
        nd = lib().H5Sget_simple_extent_ndims(sp)
        dims = (ctypes.c_uint64 * nd)()
        lib().H5Sget_simple_extent_dims(sp, dims, None)
        out = np.empty(tuple(dims), dtype=np.float64)
        lib().H5Dread(d, _type("H5T_NATIVE_DOUBLE_g"), 0, 0, 0, out.ctypes.data)
        lib().H5Sclose(sp)
        lib().H5Dclose(d)
        return out

    def _attr(self, obj, name, tname, ctype):
        a = lib().H5Aopen(obj, name.encode(), 0)
        if a < 0:
            raise KeyError(name)
        v = ctype()
        lib().H5Aread(a, _type(tname), ctypes.byref(v))
        lib().H5Aclose(a)
        return v.value

    def int_attr(self, name):
        return self._attr(self.f, name, "H5T_NATIVE_INT_g", ctypes.c_int)

    def group_double_attr(self, group, name):
        g = lib().H5Gopen2(self.f, group.encode(), 0)
        try:
            return self._attr(g, name, "H5T_NATIVE_DOUBLE_g", ctypes.c_double)
        finally:
            lib().H5Gclose(g)
