"""A C consumer of the drop-in boundary: tests/capi_consumer.c is compiled with gcc against
include/psmas.h alone (C99, -pedantic -Werror) and linked to polysolve_amd/libpsmas.so, then run.

On the CPU the program must build and report that no GPU is visible (exit 77, before any compute
call); on the GPU it runs create / set_parameters / analyze_pattern / factorize_csc / solve /
get_info (coarsest-level Cholesky, then a Dirichlet factorisation: info texts above 1 KB) /
destroy and checks the residual, the fixed dofs and the size-query contract of get_info.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB_DIR = ROOT / "polysolve_amd"


def _build(tmp_path: Path) -> Path:
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    if not (LIB_DIR / "libpsmas.so").exists():
        pytest.fail("libpsmas.so is not built (run __graft_entry__.build())")
    exe = tmp_path / "capi_consumer"
    cmd = [gcc, "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O1", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "capi_consumer.c"), "-o", str(exe), f"-L{LIB_DIR}", "-lpsmas",
           f"-Wl,-rpath,{LIB_DIR}", "-Wl,--allow-shlib-undefined", "-lm"]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert proc.returncode == 0, proc.stdout
    return exe


def test_consumer_builds_and_sees_no_device_on_cpu(tmp_path):
    from polysolve_amd import linear

    exe = _build(tmp_path)
    if linear.device_count() > 0:
        pytest.skip("a GPU is visible: the GPU test runs the program")
    proc = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert proc.returncode == 77, proc.stdout


@pytest.mark.gpu
def test_consumer_runs_on_gpu(gpu, tmp_path):
    exe = _build(tmp_path)
    proc = subprocess.run([str(exe)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(proc.stdout)
    assert proc.returncode == 0, proc.stdout
    assert "consumer ok" in proc.stdout
    # The Dirichlet factorisation's info text is above the 1 KB a fixed buffer would hold.
    fields = dict(kv.split("=") for kv in proc.stdout.split(":", 1)[1].split())
    assert int(fields["dirichlet_info_bytes"]) > 1024
