"""Launch guards (CPU): tests/launch_guard_check.cpp calls launchers of libpsmas.so with a null
device pointer and expects each to refuse before any HIP call.  A kernel that dereferences a null
pointer faults the GPU; the guard turns that fault into an error on the host."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB_DIR = ROOT / "polysolve_amd"


def test_null_device_pointers_are_refused(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    if not (LIB_DIR / "libpsmas.so").exists():
        pytest.fail("libpsmas.so is not built (run __graft_entry__.build())")
    exe = tmp_path / "launch_guard_check"
    cmd = [hipcc, "-std=c++17", "-O1", str(ROOT / "tests" / "launch_guard_check.cpp"), "-o", str(exe),
           f"-L{LIB_DIR}", "-lpsmas", f"-Wl,-rpath,{LIB_DIR}"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert proc.returncode == 0, proc.stdout
    run = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert run.returncode == 0, run.stdout
    assert "all launches refused" in run.stdout
    assert run.stdout.count("refused ") == 7
