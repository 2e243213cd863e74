"""Host builder of the symmetric matrix form (build_symmetric_sell), compiled with the host
compiler and checked by tests/cpp/sym_sell_check.cpp: the form evaluated as the GPU kernel does
equals the full matrix's product on irregular block patterns (bd 1, 2, 3); a structurally
asymmetric pattern is refused."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
ROCM = Path("/opt/rocm")


def test_symmetric_form_reproduces_the_full_product(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None or not (ROCM / "include" / "hip" / "hip_runtime.h").exists():
        pytest.skip("needs g++ and the ROCm headers")
    exe = tmp_path / "sym_sell_check"
    csrc = ROOT / "polysolve_amd" / "csrc"
    build = subprocess.run([cxx, "-O2", "-std=c++17", "-pthread", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}",
                            f"-I{csrc}", str(ROOT / "tests" / "cpp" / "sym_sell_check.cpp"), str(csrc / "host_setup.cpp"),
                            "-o", str(exe), f"-L{ROCM / 'lib'}", "-lamdhip64", "-w"],
                           capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120,
                         env={"LD_LIBRARY_PATH": str(ROCM / "lib")})
    print(run.stdout)
    assert run.returncode == 0 and run.stdout.strip().endswith("OK"), run.stdout + run.stderr
