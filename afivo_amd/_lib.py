"""ctypes binding of libafivo_amd.so (the C ABI of include/afivo_amd.h).

The library must have been built (``python -m afivo_amd.build``); loading
fails loudly otherwise -- there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libafivo_amd.so")

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
F64 = ctypes.c_double
I32P = ctypes.POINTER(ctypes.c_int32)
I64P = ctypes.POINTER(ctypes.c_int64)
F64P = ctypes.POINTER(ctypes.c_double)

AFG_OP_LPL, AFG_OP_AUTO, AFG_OP_LSF = 0, 1, 2
AFG_RSTR_LPL, AFG_RSTR_AUTO, AFG_RSTR_LSF = 0, 1, 2
AFG_BC_DIRICHLET, AFG_BC_NEUMANN = -1, -2


class MgOpts(ctypes.Structure):
    """afg_mg_opts: the fields of mg$D_t (reference src/m_aX_mg.f90:20-52)."""
    _fields_ = [("i_phi", I32), ("i_rhs", I32), ("i_tmp", I32), ("i_eps", I32),
                ("n_cycle_down", I32), ("n_cycle_up", I32), ("n_cycle_base", I32),
                ("op", I32), ("initialized", I32), ("i_lsf", I32), ("i_bval", I32), ("rstr", I32),
                ("reserved", I32 * 4)]

    def __init__(self, i_phi=-1, i_rhs=-1, i_tmp=-1, i_eps=-1, n_cycle_down=-1,
                 n_cycle_up=-1, n_cycle_base=-1, op=AFG_OP_LPL, i_lsf=-1, i_bval=-1, rstr=AFG_RSTR_LPL):
        super().__init__(i_phi, i_rhs, i_tmp, i_eps, n_cycle_down, n_cycle_up, n_cycle_base,
                         op, 0, i_lsf, i_bval, rstr)


# name -> (restype, argtypes); every symbol of include/afivo_amd.h
SIGNATURES = {
    "afg_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(P)]),
    "afg_ctx_destroy": (ctypes.c_int, [P]),
    "afg_last_error": (ctypes.c_char_p, [P]),
    "afg_sync": (ctypes.c_int, [P]),
    "afg_set_tree": (ctypes.c_int, [P, ctypes.c_int, I32P, F64]),
    "afg_set_forest": (ctypes.c_int, [P, ctypes.c_int, I32P, F64, I32P]),
    "afg_gc_tree": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "afg_gc_level": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "afg_restrict_tree": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "afg_gc2_boxes": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I32P, F64P]),
    "afg_prolong_level": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "afg_add_cc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "afg_sub_cc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "afg_times_cc": (ctypes.c_int, [P, F64, ctypes.c_int]),
    "afg_lincomb_cc": (ctypes.c_int, [P, F64, ctypes.c_int, F64, ctypes.c_int]),
    "afg_upload_fc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I32P, F64P]),
    "afg_download_fc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I32P, F64P]),
    "afg_consistent_fluxes": (ctypes.c_int, [P, ctypes.c_int, I32P]),
    "afg_set_coord": (ctypes.c_int, [P, ctypes.c_int, F64]),
    "afg_update_tree": (ctypes.c_int, [P, ctypes.c_int, I32P, F64, I32P, ctypes.c_int, ctypes.c_int, I32P]),
    "afg_update_forest": (ctypes.c_int, [P, ctypes.c_int, I32P, F64, I32P, I32P]),
    "afg_mixed_first_fmg": (ctypes.c_int, [P, ctypes.c_int]),
    "afg_mixed_fp64_base": (ctypes.c_int, [P, ctypes.c_int]),
    "afg_mixed_fp64_tau": (ctypes.c_int, [P, ctypes.c_int]),
    "afg_forest_trees": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int), I32P]),
    "afg_forest_max_abs": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, F64P]),
    "afg_set_bc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, F64P]),
    "afg_packed_size": (ctypes.c_int, [P, ctypes.c_int, I32P, I64P]),
    "afg_upload_cc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I32P, F64P]),
    "afg_download_cc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I32P, F64P]),
    "afg_copy_cc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "afg_set_cc": (ctypes.c_int, [P, ctypes.c_int, F64]),
    "afg_mg_init": (ctypes.c_int, [ctypes.POINTER(MgOpts)]),
    "afg_fas_fmg": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int]),
    "afg_fas_vcycle": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int]),
    "afg_levels": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "afg_level_ids": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I32P, ctypes.POINTER(ctypes.c_int)]),
    "afg_tree_max_cc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, F64P]),
    "afg_tree_min_cc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, F64P]),
    "afg_tree_extreme_noalloc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "afg_fill_gc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "afg_gsrb_level": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int]),
    "afg_residual_level": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int]),
    "afg_restrict_level": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "afg_prolong2_level": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "afg_correct_children": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int]),
    "afg_update_coarse": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int]),
    "afg_fp32_from_cc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "afg_fp32_to_cc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "afg_fas_fmg_fp32": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int]),
    "afg_fas_vcycle_fp32": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int]),
    "afg_set_pass_pairs": (ctypes.c_int, [P, ctypes.c_int]),
    "afg_get_pass_pairs": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "afg_set_level_pass_pairs": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "afg_set_tile": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "afg_get_tile": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "afg_get_bottom_cells": (ctypes.c_int, [P, ctypes.c_int, I64P]),
    "afg_autotune": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int, ctypes.c_int, F64P]),
    "afg_mixed_begin": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int]),
    "afg_mixed_correct": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int]),
    "afg_mixed_solve": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int, F64, ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_int), F64P]),
    "afg_set_bottom_cells": (ctypes.c_int, [P, I64, I64]),
    "afg_set_persist": (ctypes.c_int, [P, ctypes.c_int, I64]),
    "afg_set_split_passes": (ctypes.c_int, [P, ctypes.c_int]),
    "afg_persist_trace": (ctypes.c_int, [P, ctypes.c_int, I64P, ctypes.POINTER(ctypes.c_int)]),
    "afg_get_persist": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int), I64P, ctypes.POINTER(ctypes.c_int)]),
    "afg_set_graphs": (ctypes.c_int, [P, ctypes.c_int]),
    "afg_set_timing": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int]),
    "afg_timing": (ctypes.c_int, [P, I64P, F64P, F64P]),
    "afg_vcycle_cell_updates": (ctypes.c_int, [P, ctypes.POINTER(MgOpts), ctypes.c_int, F64P]),
    "afg_device_bytes": (ctypes.c_int, [P, I64P]),
    "afg_upload_boxes": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I32P, ctypes.POINTER(F64P)]),
    "afg_download_boxes": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, I32P, ctypes.POINTER(F64P)]),
    # per-box operators on the host
    "afg_box_op": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, F64, ctypes.c_int, ctypes.c_int,
                                  F64P, F64P, F64P, F64P, F64P]),
    "afg_box_gsrb": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, F64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    F64P, F64P, F64P, F64P, F64P]),
    "afg_box_corr": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, I32P,
                                    F64P, F64P, F64P, F64P, F64P]),
    "afg_box_rstr": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, I32P,
                                    F64P, F64P, F64P, F64, F64, F64P]),
    "afg_box_sides_rb": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, I32P,
                                        F64P, F64P]),
}
AFG_BOX_LPL, AFG_BOX_LPLD, AFG_BOX_LSF, AFG_BOX_CLPL, AFG_BOX_CLPLD = 0, 1, 2, 3, 4

_lib = None


class AfgError(RuntimeError):
    pass


def load():
    """Load libafivo_amd.so, raising if it is missing (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AfgError(f"{LIB_PATH} is missing: build it with `python -m afivo_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, ctx=None):
    if rc != 0:
        msg = load().afg_last_error(ctx)
        raise AfgError(f"afivo_amd error {rc}: {msg.decode() if msg else ''}")
