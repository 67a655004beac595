"""Build step for libfv3d_hip.so (hipcc, gfx950) and the fv3d driver.

Run as ``python -m fv3d_amd.build`` or through ``__graft_entry__.build()``.
The library is built in-tree next to this file so that it travels with the
source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfv3d_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FV3D_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no FMA contraction, so FP64 results keep the reference's
# rounding.  Packed FP32 instructions are disabled for every HIP object.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}",
]

SOURCES = ["fv3d_stage.hip", "fv3d_stage_w32.hip", "fv3d_stage_f32.hip", "fv3d_stage_f32_w32.hip",
           "fv3d_batch.hip", "fv3d_batch_f32.hip", "fv3d_batch_f32_w32.hip", "fv3d_aux.hip", "fv3d_api.hip", "fv3d_batch_api.hip", "fv3d_adjoint.hip",
           "fv3d_adjoint_f32.hip", "fv3d_adjoint_api.hip", "fv3d_functors.hip",
           "fv3d_noalloc.hip"]
# per-source extra flags.  The FP32 kernels flush FP32 denormals (FP64 keeps
# them): state values sit far above 1e-38, and without denormal support the
# FP32 square root and division need no range scaling.  They also contract
# a*b + c within one expression into an FMA (-ffp-contract=on, decided per
# source expression, so the batched and single-box FP32 kernels stay
# bit-equal); the FP64 translation units keep -ffp-contract=off.
F32 = ["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero",
       "-ffp-contract=on"]
EXTRA = {
    "fv3d_stage_f32.hip": F32,
    "fv3d_stage_f32_w32.hip": F32,
    "fv3d_batch_f32.hip": F32,
    "fv3d_batch_f32_w32.hip": F32,
    "fv3d_adjoint_f32.hip": F32,
}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build_library(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(len(cmds), 8)) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


DRIVER = os.path.join(HERE, "fv3d")
FUNCTOR_CHECK = os.path.join(HERE, "fv3d_functor_check")
DRIVER_SOURCES = ["main.cpp", "sim_info.cpp", "init.cpp", "io_manager.cpp"]
HDF5_ROOT = os.environ.get("FV3D_HDF5_ROOT", "/opt/conda")


def build_driver(verbose: bool = False) -> str:
    """The fv3d executable: the reference's driver loop on libfv3d_hip, with
    HDF5 snapshot output (HDF5 C library from FV3D_HDF5_ROOT)."""
    hdf5_lib = os.path.join(HDF5_ROOT, "lib", "libhdf5.so")
    # Only libhdf5 is taken from HDF5_ROOT: a private directory holds a link
    # to it, so the rest of that prefix (an older libstdc++) stays off the
    # executable's search path.
    libdir = os.path.join(HERE, "_hdf5")
    os.makedirs(libdir, exist_ok=True)
    real = os.path.realpath(hdf5_lib)
    link = os.path.join(libdir, os.path.basename(real).split(".so")[0] + ".so.103")
    for name in os.listdir(libdir):
        os.unlink(os.path.join(libdir, name))
    os.symlink(os.path.join(HDF5_ROOT, "lib", os.path.basename(link)), link)
    src = [os.path.join(HERE, "driver", f) for f in DRIVER_SOURCES]
    tmp = DRIVER + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}",
           f"-I{os.path.join(HERE, 'driver')}", f"-I{os.path.join(HDF5_ROOT, 'include')}", *src,
           "-o", tmp, LIB, hdf5_lib, "-Wl,-rpath,$ORIGIN:$ORIGIN/_hdf5"]
    _run(cmd)
    os.replace(tmp, DRIVER)
    # the functor facades' check program (no HDF5)
    tmp = FUNCTOR_CHECK + ".tmp"
    _run([os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}",
          f"-I{os.path.join(HERE, 'driver')}", os.path.join(HERE, "driver", "functor_check.cpp"),
          "-o", tmp, LIB, "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, FUNCTOR_CHECK)
    if verbose:
        print("built", DRIVER, "and", FUNCTOR_CHECK)
    return DRIVER


if __name__ == "__main__":
    build_library(verbose=True)
    build_driver(verbose=True)
