"""ctypes binding of libvbh.so (the C ABI declared in include/vbh.h).

The HIP library is the product path: if it is missing this module raises
instead of falling back to any CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvbh.so")

VBH_OK, VBH_ERR_ARG, VBH_ERR_NO_DEVICE, VBH_ERR_SOLVER, VBH_ERR_HIP, VBH_ERR_STATE = range(6)
VBH_PREC_FP64, VBH_PREC_MIXED = 0, 1


class vbh_solve_stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("refinements", C.c_int),
        ("rel_res", C.c_double),
        ("solve_ms", C.c_double),
        ("spmv_ms_avg", C.c_double),
        ("format", C.c_int),
    ]


class vbh_assembly_stats(C.Structure):
    _fields_ = [
        ("nnz_topological", C.c_longlong),
        ("nnz", C.c_longlong),
        ("pattern_ms", C.c_double),
        ("element_ms", C.c_double),
        ("mpc_ms", C.c_double),
    ]


class vbh_block_config(C.Structure):
    _fields_ = [
        ("col_tile", C.c_int),
        ("cols_per_lane", C.c_int),
        ("lanes_per_group", C.c_int),
        ("row_blocks", C.c_int),
        ("xcd_rows", C.c_int),
        ("value_path", C.c_int),
        ("ldk", C.c_int),
        ("spmm_ms", C.c_double),
    ]


class vbh_deck_sizes(C.Structure):
    _fields_ = [
        ("nnodes", C.c_int),
        ("nelem", C.c_int),
        ("nconn", C.c_int),
        ("nsubcase", C.c_int),
        ("solid_form", C.c_int),
        ("shell_form", C.c_int),
        ("nunsupported", C.c_int),
        ("nrigid_mpc", C.c_int),
        ("nrigid_terms", C.c_int),
    ]


class vbh_deck_subcase_info(C.Structure):
    _fields_ = [
        ("id", C.c_int),
        ("load_set", C.c_int),
        ("spc_set", C.c_int),
        ("nspc", C.c_int),
        ("npload4", C.c_int),
        ("disp_print", C.c_int),
        ("stress_print", C.c_int),
        ("grav", C.c_double * 3),
        ("mpc_set", C.c_int),
        ("nmpc", C.c_int),
        ("nmpc_terms", C.c_int),
    ]


class vbh_mpc_stats(C.Structure):
    _fields_ = [
        ("n_full", C.c_int),
        ("n_reduced", C.c_int),
        ("n_eliminations", C.c_int),
        ("n_skipped", C.c_int),
        ("t_nnz", C.c_int),
    ]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_LL = C.c_longlong
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); every symbol declared in include/vbh.h.
SIGNATURES = {
    "vbh_last_error": (C.c_char_p, []),
    "vbh_device_count": (_I, []),
    "vbh_ctx_create": (_I, [_I, _D, _I, C.POINTER(_P)]),
    "vbh_ctx_destroy": (None, [_P]),
    "vbh_device_name": (C.c_char_p, [_P]),
    "vbh_set_precision": (_I, [_P, _I]),
    "vbh_set_tolerance": (_I, [_P, _D, _I]),
    "vbh_set_format": (_I, [_P, _I]),
    "vbh_set_csr": (_I, [_P, _I, _I, _IP, _IP, _DP, _I]),
    "vbh_get_csr_size": (_I, [_P, _IP, _IP]),
    "vbh_get_csr": (_I, [_P, _IP, _IP, _DP]),
    "vbh_dof_map": (_I, [_P, _I, _IP, _I, _IP, _IP, _IP, _I, _IP, _IP, _IP, _IP]),
    "vbh_assemble": (_I, [_P, _I, _I, _DP, _IP, _I, _IP, _IP, _IP, _DP, _DP, _DP, _I, _I]),
    "vbh_last_assembly_stats": (_I, [_P, C.POINTER(vbh_assembly_stats)]),
    "vbh_set_mpc": (_I, [_P, _I, _IP, _I, _I, _IP, _IP, _IP, _DP, _IP]),
    "vbh_mpc_info": (_I, [_P, C.POINTER(vbh_mpc_stats)]),
    "vbh_mpc_transform": (_I, [_P, _IP, _IP, _DP]),
    "vbh_mpc_expand": (_I, [_P, _DP, _DP]),
    "vbh_assemble_rhs": (_I, [_P, _DP, _I, _IP, _DP, _DP, _DP, _DP]),
    "vbh_assemble_rhs_device": (_I, [_P, _DP, _I, _IP, _DP, _DP, _DP, _P]),
    "vbh_recover_stress": (_I, [_P, _DP, _DP, _DP]),
    "vbh_recover_stress_device": (_I, [_P, _P, _P, _P]),
    "vbh_recover_displacements": (_I, [_P, _DP, _DP]),
    "vbh_recover_displacements_device": (_I, [_P, _P, _P]),
    "vbh_write_node_csv": (_I, [C.c_char_p, _I, _I, _I, _IP, _DP]),
    "vbh_write_elem_csv": (_I, [C.c_char_p, _I, _I, _I, _IP, _IP, _DP]),
    "vbh_element_forces": (_I, [_P, _DP, _DP]),
    "vbh_element_forces_device": (_I, [_P, _P, _P]),
    "vbh_reactions": (_I, [_P, _DP, _DP, _I, _IP, _DP, _DP, _DP, _DP]),
    "vbh_deck_read": (_I, [C.c_char_p, C.POINTER(_P)]),
    "vbh_deck_parse": (_I, [C.c_char_p, _LL, C.POINTER(_P)]),
    "vbh_deck_free": (None, [_P]),
    "vbh_deck_info": (_I, [_P, C.POINTER(vbh_deck_sizes)]),
    "vbh_deck_mesh": (_I, [_P, _IP, _DP, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _DP]),
    "vbh_deck_subcase": (_I, [_P, _I, C.POINTER(vbh_deck_subcase_info)]),
    "vbh_deck_subcase_data": (_I, [_P, _I, _IP, _IP, _DP, _IP, _IP, _IP, _DP]),
    "vbh_deck_subcase_mpc": (_I, [_P, _I, _IP, _IP, _IP, _DP]),
    "vbh_deck_frames": (_I, [_P, _IP, _DP]),
    "vbh_deck_rigid": (_I, [_P, _IP, _IP, _IP, _IP, _DP]),
    "vbh_deck_unsupported": (C.c_char_p, [_P, _I]),
    "vbh_element_sensitivity": (_I, [_P, _I, _DP, _DP, _DP]),
    "vbh_element_sensitivity_device": (_I, [_P, _I, _P, _P, _P]),
    "vbh_solve": (_I, [_P, _DP, _DP, _IP, _DP]),
    "vbh_solve_device": (_I, [_P, _P, _P, _IP, _DP]),
    "vbh_solve_device_noalloc": (_I, [_P, _P, _P, _IP, _DP]),
    "vbh_solve_block": (_I, [_P, _I, _DP, _DP, _IP, _DP]),
    "vbh_solve_block_device": (_I, [_P, _I, _P, _P, _IP, _DP]),
    "vbh_last_block_config": (_I, [_P, C.POINTER(vbh_block_config)]),
    "vbh_last_stats": (_I, [_P, C.POINTER(vbh_solve_stats)]),
    "vbh_spmv": (_I, [_P, _DP, _DP]),
    "vbh_bench_spmv": (_I, [_P, _I, _I, _DP]),
    "vbh_operator_stream_bytes": (_I, [_P, _I, C.POINTER(_LL)]),
    "vbh_device_alloc": (_I, [_P, _LL, C.POINTER(_P)]),
    "vbh_device_free": (_I, [_P, _P]),
    "vbh_memcpy_h2d": (_I, [_P, _P, _P, _LL]),
    "vbh_memcpy_d2h": (_I, [_P, _P, _P, _LL]),
}

_lib = None


def load(required_symbols=True):
    """Load libvbh.so. Raises RuntimeError when the extension is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libvbh.so not found at {LIB_PATH}: build the HIP extension first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        if not hasattr(lib, name):
            if required_symbols:
                raise RuntimeError(f"libvbh.so lacks symbol {name}")
            continue
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().vbh_last_error().decode(errors="replace")


def ip(a):
    return a.ctypes.data_as(_IP)


def dp(a):
    return a.ctypes.data_as(_DP)
