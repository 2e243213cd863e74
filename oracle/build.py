"""Build recipe for the CPU oracle (test infrastructure only).

Compiles ``mas_oracle.c`` (a plain-C restatement of the reference MAS path) and
``jacobi_cg_omp.c`` (the reference's Jacobi-CG CPU solver, OpenMP) into ``oracle/liborc.so`` with
gcc.  The reference itself (CUDA C++ with CCCL, cuSPARSE, KaMinPar,
Eigen, jse fetched at configure time) cannot be compiled in this image, so there is no
``oracle/_ref`` build.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build() -> Path:
    out = HERE / "liborc.so"
    tmp = HERE / "liborc.so.tmp"
    cc = shutil.which("gcc") or "gcc"
    cmd = [cc, "-O2", "-std=c99", "-fPIC", "-shared", "-Wall", "-fopenmp", str(HERE / "mas_oracle.c"),
           str(HERE / "jacobi_cg_omp.c"), "-o", str(tmp), "-lm"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{proc.stdout}")
    tmp.replace(out)
    return out


if __name__ == "__main__":
    print(build())
