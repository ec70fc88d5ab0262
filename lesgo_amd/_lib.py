"""ctypes binding of liblesgo_amd.so (the C ABI declared in include/lesgo_amd.h).

The library is loaded from this package directory only; if it is missing
the import fails loudly -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblesgo_amd.so")

FIELDS = ("u", "v", "w", "dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz",
          "dwdx", "dwdy", "dwdz", "RHSx", "RHSy", "RHSz", "RHSx_f", "RHSy_f",
          "RHSz_f", "dpdx", "dpdy", "dpdz", "txx", "txy", "tyy", "txz", "tyz",
          "tzz", "divtx", "divty", "divtz", "p", "fxa", "fya", "fza",
          "S11", "S12", "S13", "S22", "S23", "S33", "Nu_t", "Cs_opt2",
          "F_LM", "F_MM", "F_QN", "F_NN", "Beta", "Tn_all")
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

LG_FP64, LG_MIXED = 0, 1


class LgParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("Nz", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
                ("dt", C.c_double), ("zo", C.c_double), ("vonk", C.c_double),
                ("cs_count", C.c_int), ("use_mean_p_force", C.c_int),
                ("mean_p_force_x", C.c_double), ("mean_p_force_y", C.c_double),
                ("nu", C.c_double), ("precision", C.c_int), ("initu", C.c_int),
                ("dyn_init", C.c_int), ("use_cfl_dt", C.c_int), ("cfl", C.c_double)]


# name -> (restype, argtypes)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
SIGNATURES = {
    "lg_create": (C.c_int, [C.POINTER(LgParams), C.POINTER(C.c_void_p)]),
    "lg_destroy": (C.c_int, [_P]),
    "lg_last_error": (C.c_char_p, []),
    "lg_sizes": (C.c_int, [_P, _I, _I, _I]),
    "lg_upload": (C.c_int, [_P, C.c_int, _D]),
    "lg_download": (C.c_int, [_P, C.c_int, _D]),
    "lg_seed_state": (C.c_int, [_P, C.c_long, C.c_double]),
    "lg_set_counters": (C.c_int, [_P, C.c_int, C.c_int]),
    "lg_get_counters": (C.c_int, [_P, _I, _I]),
    "lg_synchronize": (C.c_int, [_P]),
    "lg_filt_da": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "lg_ddx": (C.c_int, [_P, C.c_int, C.c_int]),
    "lg_ddy": (C.c_int, [_P, C.c_int, C.c_int]),
    "lg_ddxy": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "lg_ddz_uv": (C.c_int, [_P, C.c_int, C.c_int]),
    "lg_ddz_w": (C.c_int, [_P, C.c_int, C.c_int]),
    "lg_wallstress": (C.c_int, [_P]),
    "lg_sgs_stag": (C.c_int, [_P]),
    "lg_lagrange_sdep": (C.c_int, [_P]),
    "lg_interpolag_sdep": (C.c_int, [_P]),
    "lg_divstress_uv": (C.c_int, [_P]),
    "lg_divstress_w": (C.c_int, [_P]),
    "lg_convec": (C.c_int, [_P]),
    "lg_press_stag_array": (C.c_int, [_P]),
    "lg_project": (C.c_int, [_P]),
    "lg_copy_field": (C.c_int, [_P, C.c_int, C.c_int]),
    "lg_test_filter": (C.c_int, [_P, C.c_int, _D]),
    "lg_tridag_array": (C.c_int, [C.c_int, C.c_int, C.c_int, _D, _D, _D, _D, _D]),
    "lg_tridag_array_device": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lg_tridag_array_device_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p,
                                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p]),
    "lg_rhs_ab2": (C.c_int, [_P]),
    "lg_rhs_gradp": (C.c_int, [_P]),
    "lg_step": (C.c_int, [_P, C.c_int]),
    "lg_rmsdiv": (C.c_int, [_P, _D]),
    "lg_energy": (C.c_int, [_P, _D]),
    "lg_max_cfl": (C.c_int, [_P, _D]),
    "lg_tau_wall_bot": (C.c_int, [_P, _D]),
    "lg_turbines_init": (C.c_int, [_P, C.c_int, _D, C.c_double, C.c_double, C.c_double, C.c_int]),
    "lg_turbines_forcing": (C.c_int, [_P]),
    "lg_turbines_get": (C.c_int, [_P, _D, _D, C.c_int]),
    "lg_timing_enable": (C.c_int, [_P, C.c_int]),
    "lg_timing_report": (C.c_int, [_P, C.c_int, _D, _D, _D]),
    "lg_calibrate_copy": (C.c_int, [C.c_long, C.c_int, _D]),
    "lg_create_batch": (C.c_int, [C.POINTER(LgParams), C.c_int, C.POINTER(C.c_void_p)]),
    "lg_batch_info": (C.c_int, [_P, _I, _I]),
    "lg_set_active": (C.c_int, [_P, C.c_int]),
    "lg_member_params": (C.c_int, [_P, C.c_int, C.c_double, C.c_double, C.c_double]),
    "lg_upload_member": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "lg_download_member": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "lg_seed_member": (C.c_int, [_P, C.c_int, C.c_long, C.c_double]),
    "lg_copy_member": (C.c_int, [_P, C.c_int, C.c_int]),
    "lg_rmsdiv_member": (C.c_int, [_P, C.c_int, _D]),
    "lg_energy_member": (C.c_int, [_P, C.c_int, _D]),
    "lg_max_cfl_member": (C.c_int, [_P, C.c_int, _D]),
    "lg_tau_wall_bot_member": (C.c_int, [_P, C.c_int, _D]),
    "lg_clock_member": (C.c_int, [_P, C.c_int, _D]),
    "lg_advance_clock": (C.c_int, [_P]),
    "lg_resync": (C.c_int, [_P]),
    "lg_set_clock_member": (C.c_int, [_P, C.c_int, C.c_double, C.c_double, C.c_int]),
    "lg_turbines_init_member": (C.c_int, [_P, C.c_int, C.c_int, _D, C.c_double, C.c_double,
                                          C.c_double, C.c_int]),
    "lg_turbines_get_member": (C.c_int, [_P, C.c_int, _D, _D, C.c_int]),
    "lg_turbines_history_member": (C.c_int, [_P, C.c_int, _D, C.c_int]),
    "lg_turbines_time_filter": (C.c_int, [_P, C.c_double, C.c_double, C.c_double]),
    "lg_turbines_set_ct_member": (C.c_int, [_P, C.c_int, _D, C.c_int]),
    "lg_turbines_set_udt_member": (C.c_int, [_P, C.c_int, _D, C.c_int]),
    "lg_step_graph": (C.c_int, [_P, C.c_int]),
    "lg_autotune": (C.c_int, [_P, C.c_int, _I]),
    "lg_tune_get": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _I, _I, _I, _I]),
    "lg_tune_set": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_int]),
    "lg_tune_clear": (C.c_int, []),
    "lg_tune_count": (C.c_int, [_I]),
    "lg_tune_entry": (C.c_int, [C.c_int, _I, _I, _I, _I, _I, _I, _I, _I]),
    "lg_case_bytes": (C.c_int, [C.POINTER(LgParams), C.c_int, _D]),
    "lg_stats_enable": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "lg_stats_info": (C.c_int, [_P, _D, _I]),
    "lg_stats_time": (C.c_int, [_P, C.c_int, _D]),
    "lg_stats_mean": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "lg_stats_sums": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "lg_stats_restore": (C.c_int, [_P, C.c_int, C.c_double, _D]),
    "lg_stats_spectrum": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "lg_diff_l2": (C.c_int, [_P, C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int, _D, _D]),
    "lg_mem_info": (C.c_int, [_D, _D]),
    "lg_server_create": (C.c_int, [C.c_double, C.c_int, C.POINTER(C.c_void_p)]),
    "lg_server_destroy": (C.c_int, [_P]),
    "lg_server_now": (C.c_int, [_P, _D]),
    "lg_server_submit": (C.c_int, [_P, C.POINTER(LgParams), C.c_long, C.c_double, C.c_int,
                                   C.c_int, _D, C.c_double, C.c_int, _I]),
    "lg_server_tick": (C.c_int, [_P, _D, _I, _I]),
    "lg_server_job": (C.c_int, [_P, C.c_int, _I, _I, _D, _D, _D]),
    "lg_server_result": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "lg_server_stats": (C.c_int, [_P, _D, C.POINTER(C.c_long), _D, _I]),
}

TAVG_NAMES = ("u", "v", "w", "w_uv", "u_w", "v_w", "u2", "v2", "w2", "uv", "uw", "vw",
              "txx", "tyy", "tzz", "txy", "txz", "tyz", "p", "fx", "fy", "fz", "cs_opt2",
              "vortx", "vorty", "vortz")

PASS_KINDS = ("r2c", "c2r", "cols_fwd", "cols_inv", "tridag", "cols_unpad", "cols_pad")

# timing classes of lg_timing_report (KernelKind in csrc/timer.hpp): the row
# and column passes on the nx x ny grid, the 3/2 grid's passes, the pressure
# tridiagonal and the Lagrangian model's fused row passes
KERNEL_KINDS = ("rows_r2c", "rows_c2r", "cols_fwd", "cols_inv", "tridag", "rows_r2c_pad",
                "rows_c2r_pad", "cols_unpad", "cols_pad", "lag_rows", "lag_prep_rows",
                "rows_r2c_prod")


class LesgoError(RuntimeError):
    pass


_lib = None


def lib():
    """Load liblesgo_amd.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not found: build it with `python -m lesgo_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().lg_last_error().decode(errors="replace")
        raise LesgoError(f"{what} failed (code {rc}): {msg}")
