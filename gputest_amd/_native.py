"""ctypes binding of libftmi.so (the C ABI in ``include/ftmi.h``).

There is no fallback: if the library is missing or was built for another ABI, ``lib()``
raises, and every operator that needs it fails loudly.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libftmi.so")
ABI_VERSION = 107

P = ctypes.c_void_p
I32, I64, F64 = ctypes.c_int, ctypes.c_int64, ctypes.c_double


class FtmiMesh(ctypes.Structure):
    _fields_ = [("nc", I64), ("nf", I64), ("nb", I64), ("dim", I64),
                ("P", P), ("N", P), ("w", P), ("alpha", P), ("wl", P),
                ("S", P), ("d", P), ("Tvec", P), ("dP_f", P), ("dN_f", P), ("dxm", P),
                ("Pb", P), ("Sb", P), ("db", P), ("db_t", P), ("nb_unit", P),
                ("alpha_b", P), ("wlb", P), ("Sb_mag", P),
                ("vol", P), ("Ginv", P),
                ("cptr", P), ("cface", P), ("cnbr", P), ("bptr", P), ("bface", P)]


class FtmiBc(ctypes.Structure):
    _fields_ = [("kind", P), ("tkind", P), ("ub_fixed", P), ("pb_fixed", P), ("tval", P)]


class FtmiRed(ctypes.Structure):
    _fields_ = [("partials", P), ("counter", P)]


class FtmiAssembleArgs(ctypes.Structure):
    _fields_ = [("F", P), ("Fb", P), ("gamma", F64), ("energy", I32), ("nfields", I32),
                ("phi", P * 3), ("grad", P * 3), ("phib", P * 3), ("gp", P),
                ("drive", F64 * 3), ("has_drive", I32), ("relax", F64), ("second_order", I32),
                ("simplec", I32), ("cp", F64),
                ("diag", P), ("diag_r", P), ("off", P), ("b", P * 3), ("dP", P), ("resid", P),
                ("inv_r", P), ("x0", P * 3), ("out_fp32", I32),
                ("diag32", P), ("diag_r32", P), ("off32", P), ("b32", P * 3), ("inv_r32", P), ("ex32", P),
                ("gam_f", P), ("gam_b", P), ("extra", P * 3), ("cell_resid", P)]


class FtmiSstCoef(ctypes.Structure):
    _fields_ = [("rho", F64), ("mu", F64), ("two_rho_sw2", F64), ("four_rho_sw2", F64), ("nu500", F64),
                ("ten_bstar_rho", F64), ("bstar_rho", F64), ("rho_a1", F64), ("g1", F64), ("g2", F64)]


_SIGS = {
    "ftmi_abi_version": ([], I32),
    "ftmi_sst_bvals": ([P] * 10 + [P], I32),
    "ftmi_sst_cell": ([P] * 17, I32),
    "ftmi_sst_work_doubles": ([], I32),
    "ftmi_sst_assemble": ([P] * 11 + [F64, F64] + [P] * 10, I32),
    "ftmi_sst_hold": ([I64] + [P] * 5 + [F64, F64] + [P] * 6, I32),
    "ftmi_sst_update": ([I64, P, P, F64, F64] + [P] * 7, I32),
    "ftmi_sst_face_gamma": ([P, P, P, F64, I32, P, P, P], I32),
    "ftmi_sst_wall": ([P, P, P, P, F64, F64, P, P, P, P], I32),
    "ftmi_sst_tstress": ([P] * 5, I32),
    "ftmi_last_error": ([], ctypes.c_char_p),
    "ftmi_device_count": ([], I32),
    "ftmi_csr_spmm": ([I64, I64, P, P, P, P, P, P], I32),
    "ftmi_face_interp": ([I64, I64, P, P, P, P, P, P], I32),
    "ftmi_ls_gradient": ([I64, I64] + [P] * 15, I32),
    "ftmi_gg_gradient": ([I64, I64] + [P] * 17, I32),
    "ftmi_assemble": ([I64, I64, I64] + [P] * 9, I32),
    "ftmi_assemble_plan_host": ([I64, I64, P, P, P, P, P], I32),
    "ftmi_capture_mode": ([I32], I32),
    "ftmi_capture_plan": ([P, I32, P, I32], I32),
    "ftmi_capture_slots_used": ([], I32),
    "ftmi_set_tpr_bias": ([I32], I32),
    "ftmi_amg_retune": ([P], I32),
    "ftmi_set_grid_cap": ([I32], I32),
    "ftmi_timer_enable": ([I32], I32),
    "ftmi_timer_read": ([P, P], I32),
    "ftmi_bicgstab_work_bytes": ([I64], I64),
    "ftmi_bicgstab_jacobi": ([I64, P, P, P, P, P, P, F64, I64, I32, P, P, P, P], I32),
    "ftmi_bicgstab_noalloc_max_rows": ([], I64),
    "ftmi_bicgstab_jacobi_noalloc": ([I64, P, P, P, P, P, P, F64, I64, I32, P, P, P], I32),
    "ftmi_bicgstab_ldu_sliced": ([I64, P, P, P, P, P, P, P, P, F64, I64, I32, P, P, P, P], I32),
    "ftmi_ldu_slice": ([I64, P, P, P, P], I32),
    "ftmi_ldu_slice_f32": ([I64, P, P, P, P], I32),
    "ftmi_bicgstab_ldu": ([I64, P, P, P, P, P, P, P, F64, I64, I32, P, P, P, P], I32),
    "ftmi_pcg_work_bytes": ([I64], I64),
    "ftmi_pcg_ldu": ([I64, P, P, P, P, P, P, F64, I64, P, I32, P, P, P, P], I32),
    "ftmi_pcg_ldu_sliced": ([I64, P, P, P, P, P, P, P, F64, I64, P, I32, P, P, P, P], I32),
    "ftmi_pcg_csr": ([I64, P, P, P, P, P, F64, I64, P, I32, P, P, P, P], I32),
    "ftmi_amg_create": ([F64, I32], P),
    "ftmi_amg_destroy": ([P], None),
    "ftmi_amg_add_level": ([P, I64, I64, I64] + [P] * 14, I32),
    "ftmi_amg_set_coarse": ([P, I64, P, P, P], I32),
    "ftmi_amg_set_coarse_iter": ([P, I64, I64, P, P, P, P, I32, F64, P, P, P], I32),
    "ftmi_amg_set_coarse_iter_f32": ([P, I64, I64, P, P, P, P, I32, F64, P, P, P], I32),
    "ftmi_amg_vcycle": ([P, P, P, P], I32),
    "ftmi_amg_create_f32": ([F64, I32], P),
    "ftmi_amg_add_level_f32": ([P, I64, I64, I64] + [P] * 14, I32),
    "ftmi_amg_set_coarse_f32": ([P, I64, P, P, P], I32),
    "ftmi_amg_vcycle_f32": ([P, P, P, P], I32),
    "ftmi_pcg_ldu_f32": ([I64, P, P, P, P, P, P, F64, I64, P, I32, P, P, P, P], I32),
    "ftmi_bicgstab_ldu_mixed": ([I64, P, P, P, P, P, P, P, F64, I64, I32, I32, P, P, P, P], I32),
    "ftmi_bicgstab_ldu_mixed_sliced": ([I64, P, P, P, P, P, P, P, P, F64, I64, I32, P, P, P, P], I32),
    "ftmi_pcg_ldu_mixed": ([I64, P, P, P, P, P, P, P, F64, I64, P, I32, I32, P, P, P, P, P], I32),
    "ftmi_cast_f64_f32": ([I64, P, P, P], I32),
    "ftmi_gather_f64": ([I64, P, P, P, P], I32),
    "ftmi_amg_aggregate_host": ([I64, P, P, P], I64),
    "ftmi_amg_aggregate_mask_host": ([I64, P, P, P, P, P], I64),
    "ftmi_amg_aggregate_csr_host": ([I64, P, P, P, P], I64),
    "ftmi_amgs_strong_count": ([I64, P, P, P, P], I32),
    "ftmi_amgs_strong_fill": ([I64, P, P, P, P, P, P], I32),
    "ftmi_diag_rows": ([I64, P, P, P, P, P, P, F64, P, P, P, P, P], I32),
    "ftmi_diag_cell_flux": ([I64, P, P, P, P, P, P, P, P, P], I32),
    "ftmi_diag_work_doubles": ([], I64),
    "ftmi_diag_rowset": ([I64, I32, P, P, F64, P, P, P], I32),
    "ftmi_diag_vstats": ([I64, I32, I64, P, P, P, P, P], I32),
    "ftmi_parse_hex": ([P, I64, P, I64], I64),
    "ftmi_parse_f64": ([P, I64, P, I64], I64),
    "ftmi_fluent_index": ([P, I64, P, I64, P], I64),
    "ftmi_segment_sums": ([I64, P, I64, I64, P, I32, P, P], I32),
    "ftmi_fluent_walk": ([P, I64, I64, P], I64),
    "ftmi_amgs_ldu_to_csr": ([I64] + [P] * 8, I32),
    "ftmi_amgs_diag": ([I64, I64] + [P] * 6, I32),
    "ftmi_amgs_strength": ([I64, I64, P, P, P, P, F64, P, P], I32),
    "ftmi_amgs_strength_cmp": ([I64, I64, P, P, P, P, F64, P, P, P, I32, P], I32),
    "ftmi_differs_u64": ([I64, P, P, P, I32, P], I32),
    "ftmi_amgs_tentative": ([I64] + [P] * 5, I32),
    "ftmi_amgs_power": ([I64, I64, P, P, P, P, P, P, I32, P, P, P], I32),
    "ftmi_amgs_spgemm": ([I32, I64, I64] + [P] * 10 + [F64] + [P] * 8, I32),
    "ftmi_scan_i32": ([I64, P, P, P, P], I32),
    "ftmi_amgs_count_cols": ([I64, P, P, P], I32),
    "ftmi_amgs_transpose_fill": ([I64, I64] + [P] * 8, I32),
    "ftmi_simplec_bvals": ([P] * 9, I32),
    "ftmi_energy_bvals": ([P] * 7, I32),
    "ftmi_ls_gradients": ([P, I32, P, P, P, P, P], I32),
    "ftmi_gg_gradient_fused": ([P] * 6, I32),
    "ftmi_transport_assemble": ([P] * 5, I32),
    "ftmi_rhie_chow": ([P, P, F64] + [P] * 15, I32),
    "ftmi_correct": ([P] * 9 + [F64, I32] + [P] * 8, I32),
    "ftmi_init_flux": ([P, P, F64, P, P, P, P, P], I32),
    "ftmi_abs_sum": ([I64, P, P, P, P], I32),
}

_lib = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it is missing or stale."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libftmi.so not built ({LIB_PATH}); run `python -m gputest_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes, fn.restype = args, res
        if L.ftmi_abi_version() != ABI_VERSION:
            raise RuntimeError("libftmi.so ABI mismatch: rebuild it")
        _lib = L
    return _lib


def check(status: int, what: str = "") -> int:
    if status < 0 or (status > 1 and what != "krylov"):
        msg = lib().ftmi_last_error().decode(errors="replace")
        raise RuntimeError(f"libftmi {what} failed ({status}): {msg}")
    return status


def ptr(t) -> int:
    """Device address of a torch tensor (None -> NULL)."""
    return 0 if t is None else t.data_ptr()


def stream_ptr() -> int:
    import torch

    return torch.cuda.current_stream().cuda_stream
