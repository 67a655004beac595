"""Builds liboracle.so, the CPU restatement used as the checker (test
infrastructure only).

Flags follow the reference's parity build: -O2 -ffp-contract=off
-fno-fast-math and no -march, so no FMA is formed and every operation rounds
as the reference's expression trees do.  OpenMP parallelises the passes.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liboracle.so")

CFLAGS = ["-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC", "-shared",
          "-std=c11", "-D_DEFAULT_SOURCE", f"-I{os.path.join(ROOT, 'include')}", f"-I{HERE}"]


def build_oracle(verbose: bool = False) -> str:
    cc = os.environ.get("CC", "gcc")
    tmp = LIB + ".tmp"
    cmd = [cc, *CFLAGS, os.path.join(HERE, "fv3d_oracle.c"), "-o", tmp, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build_oracle(verbose=True)
