"""The C ABI library: loads, exports what include/psmas.h declares, parameter contract (CPU)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

from polysolve_amd import MASError, Solver, _native, linear, parse_parameters

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "psmas.h").read_text()
    return sorted(set(re.findall(r"\b(psmas_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_native.SIGNATURES), "ctypes table and header disagree"


def test_name_is_MAS():
    assert _native.lib().psmas_name() == b"MAS"


def test_default_parameters_are_reference_in_code_defaults():
    p = parse_parameters({})
    # MASSolver.cu:137-143
    assert p["block_dim"] == 1 and p["max_iter"] == 5000 and p["relative_tolerance"] == 1e-5
    assert p["absolute_tolerance"] == 1e-8 and p["lazy_partitioning"] is False
    assert p["use_preconditioned_residual_norm"] is False and p["true_residual_period"] == 4
    assert p["precision"] == "fp64" and p["division_guard"] == 0.0


def test_mas_member_is_read():
    p = parse_parameters({"MAS": {"block_dim": 3, "max_iter": 77, "relative_tolerance": 0.0,
                                  "absolute_tolerance": 1e-9, "lazy_partitioning": True,
                                  "use_preconditioned_residual_norm": True}, "solver": "MAS"})
    assert (p["block_dim"], p["max_iter"], p["relative_tolerance"], p["absolute_tolerance"]) == (3, 77, 0.0, 1e-9)
    assert p["lazy_partitioning"] is True and p["use_preconditioned_residual_norm"] is True


def test_without_mas_member_parameters_are_ignored():
    # MASSolver::set_parameters returns when params has no "MAS" key (MASSolver.cu:605-614).
    p = parse_parameters({"Eigen::ConjugateGradient": {"max_iter": 3, "tolerance": 1e-3}})
    assert p["max_iter"] == 5000


def test_inner_object_extension():
    p = parse_parameters({"max_iter": 12, "precision": "mixed"})
    assert p["max_iter"] == 12 and p["precision"] == "mixed"


@pytest.mark.parametrize("key,default", [("spmv_fuse_direction", False), ("single_reduction", False),
                                         ("apply_wave", True), ("spmv_symmetric", True)])
def test_kernel_variant_switches(key, default):
    """Boolean switches between measured variants of the iteration's kernels: the default, the
    other value, and a non-boolean refused with the key named."""
    assert parse_parameters({})[key] is default
    assert parse_parameters({"MAS": {key: not default}})[key] is (not default)
    with pytest.raises(MASError, match=key):
        parse_parameters({"MAS": {key: 0}})


@pytest.mark.parametrize("bad", [
    '{"MAS": {"block_dim": 4}}',
    '{"MAS": {"precision": "bf8"}}',
    '{"MAS": {"max_iter": 1.5}}',
    '{"MAS": {"lazy_partitioning": 1}}',
    '{"MAS": {"reliable_update_delta": 2}}',
    '{"MAS": {"block_dim": 3',
    '[1, 2]',
    '{"MAS": {"relative_tolerance": 1e999}}',
])
def test_invalid_parameters_raise(bad):
    with pytest.raises(MASError, match=r"\[MAS\]"):
        parse_parameters(bad)


def test_create_without_device_raises_runtime_error():
    if linear.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(MASError, match="No AMD GPU"):
        Solver.create("MAS", "")


def test_unknown_solver_name():
    with pytest.raises(MASError):
        Solver.create("Eigen::SimplicialLDLT", "")
    assert Solver.available_solvers() == ["MAS"]


def test_null_handle_calls_fail_cleanly():
    lib = _native.lib()
    assert lib.psmas_solve(None, 3, None, None) == _native.PSMAS_ERROR_INVALID_ARGUMENT
    assert lib.psmas_factorize_csc(None, 3, 0, None, None, None) == _native.PSMAS_ERROR_INVALID_ARGUMENT
    assert lib.psmas_destroy(None) == _native.PSMAS_OK


def test_parameter_json_is_strict_json():
    """Every number the library writes is valid JSON (no nan / inf tokens): parsed with a decoder
    that refuses the non-standard constants."""
    import json

    def refuse(tok):
        raise ValueError(f"non-standard JSON constant {tok}")

    for prm in ({}, {"MAS": {"relative_tolerance": 1e-300, "division_guard": 1e-20, "learned_output_scale": -0.5}}):
        text = json.dumps(prm)
        rc, out = _native.read_string(_native.lib().psmas_parse_parameters, text.encode())
        assert rc == _native.PSMAS_OK
        json.loads(out, parse_constant=refuse)
