"""ctypes binding of libabc.so (the C ABI in include/abc.h).

The library is the only compute path: if it is missing or fails to load this
module raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libabc.so")

ABC_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP runtime error", -3: "device allocation failed",
          -4: "model index outside 1..5", -5: "output buffer too small",
          -6: "not allowed in this state"}

class AbcDesign(C.Structure):
    _fields_ = [
        ("cycle_h", C.c_double), ("n_bins", C.c_int32), ("n_conds", C.c_int32),
        ("pulse_h", C.c_void_p), ("chase_h", C.c_void_p), ("is_chase", C.c_void_p),
        ("cells_per_particle", C.c_int32), ("age_mode", C.c_int32),
        ("capture_mode", C.c_int32), ("beta_a", C.c_int32), ("beta_b", C.c_int32),
        ("betas", C.c_void_p), ("beta_bin", C.c_void_p), ("beta_is_chase", C.c_void_p),
        ("n_betas", C.c_int64), ("age_weights", C.c_void_p), ("burn_eps", C.c_double),
        ("burn_max", C.c_int32), ("prior_mode", C.c_int32), ("splice_log10", C.c_double),
        ("split_target", C.c_double), ("split_max", C.c_int32),
    ]


class AbcMomentCfg(C.Structure):
    _fields_ = [
        ("cycle", C.c_double), ("n_steps", C.c_int32), ("n_ages", C.c_int32),
        ("ages", C.c_void_p), ("n_conds", C.c_int32), ("pulse", C.c_void_p),
        ("chase", C.c_void_p), ("is_chase", C.c_void_p), ("t0", C.c_double),
        ("iv", C.c_double * 9), ("transient_mode", C.c_int32), ("rtol", C.c_double),
        ("atol", C.c_double), ("method", C.c_int32),
    ]


_lib = None


def load() -> C.CDLL:
    """Load libabc.so; raise if it is absent (build it with
    ``python -m abc_inference_transcription_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build the HIP library first "
                           "(python -m abc_inference_transcription_amd.build)")
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7)
    # as the system one libabc.so links against; whichever loads first serves
    # the whole process, and torch only works with its own.  Loading torch
    # first makes libabc bind to that runtime too.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    _declare(L)
    _lib = L
    return L


def _declare(L):
    V, I32, I64, U64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
    sig = {
        "abc_create": (I32, [C.POINTER(V), I32, C.POINTER(AbcDesign)]),
        "abc_destroy": (I32, [V]),
        "abc_last_error": (C.c_char_p, [V]),
        "abc_n_params": (I32, [V, I32]),
        "abc_n_stats": (I32, [V]),
        "abc_sample_prior": (I32, [V, I32, U64, I64, I64, V]),
        "abc_simulate": (I32, [V, I32, U64, I64, I64, V, V, V, V, V, V, V, V]),
        "abc_errors": (I32, [V, I64, I32, V, I64, V, V, V, I32, D, V]),
        "abc_accept": (I32, [V, I64, I32, V, I64, V, V, V, I32, D, I32, D, I64, I64, V, V, V, I64]),
        "abc_set_stream": (I32, [V, V]),
        "abc_synchronize": (I32, [V]),
        "abc_sample_prior_dev": (I32, [V, I32, U64, I64, I64, V]),
        "abc_simulate_dev": (I32, [V, I32, U64, I64, I64, V, V, V, V, V]),
        "abc_errors_dev": (I32, [V, I64, I32, V, I64, V, V, V, I32, D, V]),
        "abc_simulate_multi_dev": (I32, [V, I32, V, U64, V, V, V, V, V, V, V]),
        "abc_simulate_multi_idx_dev": (I32, [V, I32, V, U64, V, V, V, V, V, V, V]),
        "abc_last_timing": (I32, [V, V]),
        "abc_ssa_time": (I32, [V, I32, V, V, V]),
        "abc_set_ssa_blocks": (I32, [V, I32]),
        "abc_get_ssa_blocks": (I32, [V]),
        "abc_set_ssa_claim": (I32, [V, I32]),
        "abc_get_ssa_claim": (I32, [V]),
        "abc_autotune_ssa": (I32, [V, I32, V, U64, I64, I64, V, V]),
        "abc_graph_begin": (I32, [V]),
        "abc_graph_end": (I32, [V, C.POINTER(V)]),
        "abc_graph_launch": (I32, [V, V]),
        "abc_graph_destroy": (I32, [V]),
        "abc_recover_dev": (I32, [V, I32, I64, V, I64, C.POINTER(AbcMomentCfg), V, V]),
        "abc_recover": (I32, [V, I32, I64, V, I64, C.POINTER(AbcMomentCfg), V, V]),
        "abc_moment_sim_dev": (I32, [V, I32, I64, V, I64, C.POINTER(AbcMomentCfg), V, V, I32, I32,
                                     V, V, V]),
        "abc_moment_sim": (I32, [V, I32, I64, V, I64, C.POINTER(AbcMomentCfg), V, V, I32, I32, V,
                                 V, V]),
        "abc_model_probs": (I32, [V, I64, I32, V, U64, I32, D, V, V, V]),
        "abc_data_stats": (I32, [V, I64, I64, V, V, I32, V, I32, V, V, I64, V, I64, V, I32, U64,
                                 V]),
        "abc_data_stats_dev": (I32, [V, I64, I64, V, V, I32, V, I32, V, V, I64, V, I64, V, I32,
                                     U64, V]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _declare_optional(L)


def _declare_optional(L):
    """Entry points of the screening / surrogate units (declared when present)."""
    V, I32, I64, U64, D, F = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double, C.c_float
    sig = {
        "abc_accept_begin": (I32, [V, I32, I64, V, V, V, I32, D, I32, D, I64, D]),
        "abc_accept_add": (I32, [V, I64, V, I64]),
        "abc_accept_add_dev": (I32, [V, I64, V, I64]),
        "abc_accept_end": (I32, [V, V, V, V, I64, V]),
        "abc_accept_add_idx_dev": (I32, [V, I64, V, V]),
        "abc_accept_set_thresholds": (I32, [V, V]),
        "abc_simulate_idx_dev": (I32, [V, I32, U64, I64, V, V, V, V, V, V]),
        "abc_simulate_noalloc_dev": (I32, [V, I32, U64, I64, I64, V, V, V, V, V]),
        "abc_screen_errors": (I32, [V, I64, I32, V, I64, V, V, V, I32, D, V, V]),
        "abc_math_probe_dev": (I32, [V, I32, I64, V, V]),
        "abc_mlp_create": (I32, [C.POINTER(V), I32, I32, I32, I32, I32, I32, U64]),
        "abc_mlp_destroy": (I32, [V]),
        "abc_mlp_last_error": (C.c_char_p, [V]),
        "abc_mlp_n_params": (I64, [V]),
        "abc_mlp_set_count_outputs": (I32, [V, I32]),
        "abc_mlp_set_tile_rows": (I32, [V, I32]),
        "abc_mlp_set_tile_cols": (I32, [V, I32]),
        "abc_mlp_n_stats": (I32, [V]),
        "abc_mlp_set_stream": (I32, [V, V]),
        "abc_mlp_synchronize": (I32, [V]),
        "abc_mlp_get_params": (I32, [V, V]),
        "abc_mlp_set_params": (I32, [V, V]),
        "abc_mlp_forward_dev": (I32, [V, I64, V, V]),
        "abc_mlp_grad_dev": (I32, [V, I64, V, V, V, V]),
        "abc_mlp_train_step_dev": (I32, [V, I64, V, V, F, F, F, F, F, V]),
        "abc_features_dev": (I32, [I32, I64, V, I32, V, V]),
        "abc_targets_dev": (I32, [I64, V, I32, V, V, V, V]),
        "abc_screen_dev": (I32, [V, V, I64, V, V, V, D, D, V, V]),
        "abc_screen_score_dev": (I32, [V, V, I64, V, V, V, D, V]),
    }
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


def check(rc: int, ctx=None, what: str = ""):
    if rc != ABC_OK:
        msg = ""
        if ctx is not None and _lib is not None:
            m = _lib.abc_last_error(ctx)
            msg = m.decode() if m else ""
        raise RuntimeError(f"{what} failed ({rc}: {ERRORS.get(rc, '?')}): {msg}")
