"""Build the native libraries of polysolve_amd in-tree.

* ``libpsmas.so``     — the MAS solver: HIP kernels for gfx950 + C++ host + the C ABI of
  ``include/psmas.h``; compiled with hipcc.
* ``libpsproblems.so`` — the Q1-hex elasticity assembler used to manufacture benchmark and test
  systems; plain C++ (g++).

Run ``python -m polysolve_amd.build`` or call :func:`build`.  Every call recompiles.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
BUILD_DIR = HERE.parent / "_build" / "psmas"
ARCH = os.environ.get("PSMAS_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["spmv_vec.hip", "mas_kernels.hip", "coarse_solve.hip", "mlp.hip", "lcc.hip", "sell_build.hip", "aggregate.hip", "dirichlet.hip", "solver.hip"]
CXX_SOURCES = ["host_setup.cpp", "params.cpp", "capi.cpp", "mlp_capi.cpp"]

# Packed-FP32 VALU forms are disabled for every HIP translation unit (hipcc warns that the host
# pass does not know the feature; that warning is harmless).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the MAS solver needs the ROCm HIP compiler")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "unknown target feature" not in line and "packed-fp32-ops" not in line
                    and line.strip() != "1 warning generated.")
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build_psmas(jobs: int = 6) -> Path:
    BUILD_DIR.mkdir(parents=True, exist_ok=True)
    hipcc = _hipcc()
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        obj = BUILD_DIR / (Path(src).stem + ".hip.o")
        cmds.append([hipcc, *HIP_FLAGS, f"-I{INCLUDE}", "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = BUILD_DIR / (Path(src).stem + ".o")
        cmds.append([hipcc, *CXX_FLAGS, f"-I{INCLUDE}", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = HERE / "libpsmas.so"
    tmp = HERE / "libpsmas.so.tmp"
    _run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), "-lpthread"])
    tmp.replace(out)
    return out


def build_problems() -> Path:
    out = HERE / "libpsproblems.so"
    tmp = HERE / "libpsproblems.so.tmp"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, *CXX_FLAGS, "-fopenmp", "-shared", str(CSRC / "problems.cpp"), "-o", str(tmp)])
    tmp.replace(out)
    return out


def build(jobs: int = 6) -> list[Path]:
    return [build_problems(), build_psmas(jobs)]


if __name__ == "__main__":
    for p in build():
        print(p)
