"""ctypes binding of libe2edp (include/e2edp.h).

The library is loaded from the package directory (built in-tree by
``python -m e2e_gpu_dp_amd.build``). A missing library is an error, never a fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("E2EDP_LIB", PKG / "libe2edp.so"))

STATUS = {
    0: "E2E_OK", 1: "E2E_ERR_INVALID", 2: "E2E_ERR_HIP", 3: "E2E_ERR_NOMEM",
    4: "E2E_ERR_NOT_CONVERGED", 5: "E2E_ERR_TIMEOUT", 6: "E2E_ERR_IO",
    7: "E2E_ERR_CAPACITY", 8: "E2E_ERR_FIELD",
}


class E2EError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("gsize", C.c_int32), ("nt", C.c_int32), ("nrzns", C.c_int32), ("nmodes", C.c_int32),
        ("num_ac_speeds", C.c_int32), ("num_ac_angles", C.c_int32), ("i_term", C.c_int32),
        ("j_term", C.c_int32), ("term_subgrid_size", C.c_int32), ("reward_type", C.c_int32),
        ("neighb_gsize", C.c_int32), ("is_stationary", C.c_int32),
        ("F", C.c_float), ("dt", C.c_float), ("dx", C.c_float), ("dy", C.c_float),
        ("x0", C.c_float), ("y0", C.c_float), ("r_outbound", C.c_float),
        ("r_terminal", C.c_float), ("alpha", C.c_float), ("z", C.c_float),
    ]


class Fields(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("u", "v", "ui", "vi", "yi", "s", "mask")]


class BuildOpts(C.Structure):
    _fields_ = [
        ("fields_on_device", C.c_int32), ("keep_rsum", C.c_int32), ("slot_k", C.c_int32),
        ("radix_bits", C.c_int32), ("slices_per_launch", C.c_int32), ("block_threads", C.c_int32),
        ("stream", C.c_void_p), ("count_kernel", C.c_int32), ("paired_rounds", C.c_int32),
        ("n_alpha", C.c_int32), ("lane_moves", C.c_int32), ("alphas", C.c_void_p), ("comp_span", C.c_int32),
    ]


class ModelInfo(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "ns", "na", "nnz", "nc", "nt", "lanes", "n_out_of_window", "max_row_nnz",
        "n_overflow_rows", "n_multipass_rows", "device_bytes")] + \
        [(n, C.c_double) for n in ("build_ms", "move_kernel_ms", "scan_ms", "csr_ms")] + \
        [("has_rsum", C.c_int32), ("move_launches", C.c_int32), ("block_threads", C.c_int32),
         ("unit_cells", C.c_int32), ("wave_blocks", C.c_int32), ("n_deferred_rows", C.c_int32),
         ("paired_rounds", C.c_int32), ("counted_cells", C.c_int32)]


class SolveOpts(C.Structure):
    _fields_ = [
        ("thresh", C.c_float), ("discount", C.c_float), ("max_sweeps", C.c_int32),
        ("sweeps_per_check", C.c_int32), ("use_graph", C.c_int32), ("loads_in_flight", C.c_int32),
        ("timeout_ms", C.c_double), ("kernel_variant", C.c_int32), ("blocks_per_cu", C.c_int32),
        ("comp_span", C.c_int32), ("full_sweeps", C.c_int32),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("sweeps", C.c_int32), ("converged", C.c_int32), ("final_residual", C.c_float),
        ("loads_in_flight", C.c_int32), ("solve_ms", C.c_double), ("sweep_kernel_ms", C.c_double),
        ("sweep_kernel_ms_per_sweep", C.c_double), ("bytes_per_sweep", C.c_double),
        ("moved_bytes_per_sweep", C.c_double),
        ("launches", C.c_int32), ("kernel_variant", C.c_int32), ("blocks", C.c_int32),
        ("compressed_windows", C.c_int32),
    ]


class BatchOpts(C.Structure):
    _fields_ = [("variant", C.c_int32), ("use_graph", C.c_int32), ("reserved", C.c_int32 * 2)]


class BatchStats(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("sweeps", "converged", "variant", "cols_per_lane",
                                          "states_per_wave", "col_block")] + \
        [(n, C.c_double) for n in ("solve_ms", "sweep_kernel_ms", "tune_ms", "bytes_per_sweep")] + \
        [("launches", C.c_int32), ("reserved", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_cols", C.c_int32),
                ("sweeps_per_step", C.c_int32), ("keep_full_results", C.c_int32),
                ("serial", C.c_int32), ("thresh", C.c_float), ("exact_pass", C.c_int32),
                ("max_models", C.c_int32)]


class RequestResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("sweeps", C.c_int32), ("value_at_start", C.c_float),
                ("action_at_start", C.c_int32), ("latency_ms", C.c_double), ("queue_ms", C.c_double),
                ("cache_hit", C.c_int32), ("reserved", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("submitted", "completed", "rejected", "models_built",
                                          "model_evictions", "cache_hits", "peak_reserved_bytes",
                                          "sweeps_launched")] + \
        [(n, C.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "wall_ms", "solves_per_s")]


class RolloutOpts(C.Structure):
    _fields_ = [("interp", C.c_int32), ("fields_on_device", C.c_int32), ("cf", C.c_double),
                ("cr", C.c_double), ("stream", C.c_void_p), ("reserved", C.c_int32 * 4)]


class RolloutStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("trajectories", "reached", "rejected", "moves")] + \
        [("kernel_ms", C.c_double)]


class EnvParams(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("gsize", "nt", "n_wsamples", "n_modes")] + \
        [(n, C.c_double) for n in ("dt", "dxy", "A", "eps", "wx", "wy", "w_lo", "w_hi", "A_sc")] + \
        [(n, C.c_int32) for n in ("obstacle_mode", "obs_upper_row", "obs_left_col", "obs_width")] + \
        [("obs_speed_x", C.c_double), ("obs_speed_y", C.c_double), ("reserved", C.c_int32 * 4)]


class EnvOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("u", "v", "ui", "vi", "yi", "s", "mask", "phi", "phi_i", "phi_yi",
                                           "singular_values")]


class EnvStats(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("modes_ms", C.c_double), ("slices_per_pass", C.c_int32),
                ("max_iterations", C.c_int32)]


STRUCTS = [Params, Fields, BuildOpts, ModelInfo, SolveOpts, SolveStats, BatchOpts, BatchStats,
           ServerOpts, RequestResult, ServerStats, RolloutOpts, RolloutStats, EnvParams, EnvOut, EnvStats]

VP = C.c_void_p
P = C.POINTER

# name -> (restype, argtypes); every symbol declared in include/e2edp.h
SIGNATURES = {
    "e2e_last_error": (C.c_char_p, []),
    "e2e_version": (C.c_char_p, []),
    "e2e_device_count": (C.c_int32, []),
    "e2e_cached_device_bytes": (C.c_int64, []),
    "e2e_release_cached_memory": (None, []),
    "e2e_struct_size": (C.c_int32, [C.c_int32]),
    "e2e_params_pack": (C.c_int, [P(Params), VP]),
    "e2e_actions": (C.c_int, [P(Params), VP, VP]),
    "e2e_model_build": (C.c_int, [P(Params), P(Fields), P(BuildOpts), P(VP)]),
    "e2e_model_import_coo": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, VP, VP, VP, VP, VP, P(VP)]),
    "e2e_model_info_get": (C.c_int, [VP, P(ModelInfo)]),
    "e2e_model_destroy": (C.c_int, [VP]),
    "e2e_model_export_coo": (C.c_int, [VP, VP, VP, VP, VP]),
    "e2e_model_export_coo_i32": (C.c_int, [VP, VP, VP, VP, VP]),
    "e2e_model_export_csr": (C.c_int, [VP, VP, VP, VP, VP]),
    "e2e_vi_solve": (C.c_int, [VP, P(SolveOpts), VP, VP, P(SolveStats)]),
    "e2e_vi_solve_noalloc": (C.c_int, [VP, P(SolveOpts), VP, VP, P(SolveStats)]),
    "e2e_model_export_R_alpha": (C.c_int, [VP, C.c_int32, VP]),
    "e2e_vi_solve_alphas": (C.c_int, [VP, P(SolveOpts), P(BatchOpts), VP, VP, VP, P(BatchStats)]),
    "e2e_vi_solve_goals": (C.c_int, [VP, VP, C.c_int32, P(SolveOpts), P(BatchOpts), VP, VP, VP,
                                     P(BatchStats)]),
    "e2e_rollout": (C.c_int, [P(Params), P(Fields), VP, C.c_int32, VP, P(RolloutOpts), VP, VP, VP, VP,
                              P(RolloutStats)]),
    "e2e_env_generate": (C.c_int, [P(EnvParams), P(EnvOut), P(EnvStats)]),
    "e2e_server_create": (C.c_int, [P(ServerOpts), P(VP)]),
    "e2e_server_add_env": (C.c_int, [VP, C.c_int64, P(Params), P(Fields)]),
    "e2e_server_submit": (C.c_int, [VP, C.c_int64, C.c_int32, C.c_int32, P(C.c_uint64)]),
    "e2e_server_poll": (C.c_int, [VP, C.c_uint64, P(RequestResult)]),
    "e2e_server_wait": (C.c_int, [VP, C.c_uint64, C.c_double, P(RequestResult)]),
    "e2e_server_result_columns": (C.c_int, [VP, C.c_uint64, VP, VP]),
    "e2e_server_drain": (C.c_int, [VP, C.c_double]),
    "e2e_server_stats_get": (C.c_int, [VP, P(ServerStats)]),
    "e2e_server_destroy": (C.c_int, [VP]),
}

_lib = None


def lib():
    """Loads libe2edp.so once; raises if it is missing."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} not found: build it with `python -m e2e_gpu_dp_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int):
    if code != 0:
        raise E2EError(code, lib().e2e_last_error().decode(errors="replace"))
