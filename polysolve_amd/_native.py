"""ctypes bindings of ``libpsmas.so`` (the C ABI declared in ``include/psmas.h``).

The library is loaded from the package directory.  There is no fallback: when the native
library is missing the import fails loudly with instructions to build it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libpsmas.so"
PROBLEMS_PATH = _HERE / "libpsproblems.so"

PSMAS_OK = 0
PSMAS_ERROR_RUNTIME = 1
PSMAS_ERROR_OUT_OF_MEMORY = 2
PSMAS_ERROR_INVALID_ARGUMENT = 3
PSMAS_ERROR_NO_DEVICE = 4
PSMAS_ERROR_BUFFER_TOO_SMALL = 5

_lib = None
_plib = None

_i = C.c_int
_ll = C.c_longlong
_sz = C.c_size_t
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_vp = C.c_void_p
_cp = C.c_char_p

# name -> (restype, argtypes)
SIGNATURES = {
    "psmas_create": (_i, [C.POINTER(_vp)]),
    "psmas_destroy": (_i, [_vp]),
    "psmas_name": (_cp, []),
    "psmas_set_parameters": (_i, [_vp, _cp]),
    "psmas_get_parameters": (_i, [_vp, C.c_char_p, _sz, C.POINTER(_sz)]),
    "psmas_get_info": (_i, [_vp, C.c_char_p, _sz, C.POINTER(_sz)]),
    "psmas_analyze_pattern": (_i, [_vp, _i, _pi, _pi, _i]),
    "psmas_factorize_csc": (_i, [_vp, _i, _ll, _pi, _pi, _pd]),
    "psmas_solve": (_i, [_vp, _i, _pd, _pd]),
    "psmas_last_error": (_cp, [_vp]),
    "psmas_solve_device": (_i, [_vp, _i, _vp, _vp]),
    "psmas_solve_device_noalloc": (_i, [_vp, _i, _vp, _vp]),
    "psmas_factorize_values_device": (_i, [_vp, _ll, _vp]),
    "psmas_factorize_csc_device": (_i, [_vp, _i, _ll, _vp, _vp, _vp]),
    "psmas_factorize_multi_csc": (_i, [_vp, _i, _i, _ll, _pi, _pi, _pd]),
    "psmas_factorize_multi_csc_device": (_i, [_vp, _i, _i, _ll, _vp, _vp, _vp]),
    "psmas_solve_multi": (_i, [_vp, _i, _i, _pd, _pd]),
    "psmas_solve_multi_device": (_i, [_vp, _i, _i, _vp, _vp]),
    "psmas_dirichlet_factorize_csc": (_i, [_vp, _i, _ll, _pi, _pi, _pd, _i, _pi]),
    "psmas_dirichlet_factorize_csc_device": (_i, [_vp, _i, _ll, _vp, _vp, _vp, _i, _pi]),
    "psmas_dirichlet_solve": (_i, [_vp, _i, _pd, _pd]),
    "psmas_dirichlet_solve_device": (_i, [_vp, _i, _vp, _vp]),
    "psmas_set_coordinates": (_i, [_vp, _i, _i, _pd]),
    "psmas_batch_factorize_csc": (_i, [_vp, _i, _pi, C.POINTER(_ll), C.POINTER(_pi), C.POINTER(_pi),
                                       C.POINTER(_pd)]),
    "psmas_batch_solve": (_i, [_vp, _i, C.POINTER(_pd), C.POINTER(_pd), _i]),
    "psmas_batch_get_info": (_i, [_vp, _i, C.c_char_p, _sz, C.POINTER(_sz)]),
    "psmas_batch_set_coordinates": (_i, [_vp, _i, _i, _i, _pd]),
    "psmas_autotune": (_i, [_vp, _i, C.c_char_p, _sz, C.POINTER(_sz)]),
    "psmas_lattice_size": (_i, [_vp, _pi]),
    "psmas_lattice_restrict": (_i, [_vp, _i, _pd, _pd]),
    "psmas_lattice_prolong": (_i, [_vp, _i, _pd, _pd]),
    "psmas_learned_train": (_i, [_vp, _i, _i, C.c_float, C.c_uint, _pd]),
    "psmas_learned_reset": (_i, [_vp, _i]),
    "psmas_learned_exact": (_i, [_vp, C.POINTER(C.c_double)]),
    "psmas_device_count": (_i, []),
    "psmas_device_alloc": (_i, [_sz, C.POINTER(_vp)]),
    "psmas_device_free": (_i, [_vp]),
    "psmas_memcpy": (_i, [_vp, _vp, _sz, _i]),
    "psmas_device_synchronize": (_i, []),
    "psmas_read_bandwidth": (_i, [_ll, _i, C.POINTER(C.c_double)]),
    "psmas_mlp_create": (_i, [_i, _i, _i, _i, C.c_uint, C.POINTER(_vp)]),
    "psmas_mlp_destroy": (_i, [_vp]),
    "psmas_mlp_last_error": (_cp, [_vp]),
    "psmas_mlp_num_parameters": (_ll, [_vp]),
    "psmas_mlp_set_parameters": (_i, [_vp, C.POINTER(C.c_float)]),
    "psmas_mlp_get_parameters": (_i, [_vp, C.POINTER(C.c_float)]),
    "psmas_mlp_forward": (_i, [_vp, _i, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "psmas_mlp_loss_grad": (_i, [_vp, _i, C.POINTER(C.c_float), C.POINTER(C.c_float), _pd,
                                 C.POINTER(C.c_float)]),
    "psmas_mlp_train": (_i, [_vp, _i, C.POINTER(C.c_float), C.POINTER(C.c_float), _i, _i, C.c_float, C.c_uint,
                             _pd]),
    "psmas_gemm_bf16": (_i, [_i, _i, _i, _i, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "psmas_parse_parameters": (_i, [_cp, C.c_char_p, _sz, C.POINTER(_sz)]),
    "psmas_block_graph": (_i, [_i, _ll, _pi, _pi, _pd, _i, _pi, _i, _pi, _pd, _ll, C.POINTER(_ll)]),
    "psmas_get_level_info": (_i, [_vp, _pi, _i]),
    "psmas_get_partition": (_i, [_vp, _pi, _pi, _i, _pi, _i]),
    "psmas_get_coarse_map": (_i, [_vp, _pi, _ll]),
    "psmas_get_bank_inverses": (_i, [_vp, _pd, _ll]),
    "psmas_apply_preconditioner": (_i, [_vp, _i, _pd, _pd, _i]),
    "psmas_spmv": (_i, [_vp, _i, _pd, _pd, _i]),
    "psmas_set_kernel_timing": (_i, [_vp, _i]),
    "psmas_get_kernel_timing": (_i, [_vp, C.c_char_p, _sz, C.POINTER(_sz)]),
}


def lib() -> C.CDLL:
    """Load libpsmas.so (once)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -m polysolve_amd.build, or __graft_entry__.build()). "
                "The MAS solver has no CPU fallback.")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def problems_lib() -> C.CDLL:
    global _plib
    if _plib is None:
        if not PROBLEMS_PATH.exists():
            raise ImportError(f"{PROBLEMS_PATH} is missing: run python -m polysolve_amd.build")
        L = C.CDLL(str(PROBLEMS_PATH))
        L.psgen_hex_assemble.restype = _vp
        L.psgen_hex_assemble.argtypes = [_i, _pd, C.c_double, C.c_double, _i, _pi, _pi, _pd, _pd, _i,
                                         _pi, C.POINTER(_ll)]
        L.psgen_fetch.restype = _i
        L.psgen_fetch.argtypes = [_vp, _pi, _pi, _pd, _pd]
        L.psgen_free.restype = None
        L.psgen_free.argtypes = [_vp]
        L.psgen_unit_element_stiffness.restype = None
        L.psgen_unit_element_stiffness.argtypes = [C.c_double, _pd]
        _plib = L
    return _plib


def read_string(fn, *args) -> str:
    """Call a psmas_* writer (buf, cap, needed) and return the string."""
    needed = _sz(0)
    cap = 4096
    while True:
        buf = C.create_string_buffer(cap)
        rc = fn(*args, buf, cap, C.byref(needed))
        if rc == PSMAS_ERROR_BUFFER_TOO_SMALL and needed.value > cap:
            cap = needed.value
            continue
        return rc, buf.value.decode()
