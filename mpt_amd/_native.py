"""ctypes binding of ``libmptg.so`` (the C ABI in ``include/mptg.h``).

The shared object is loaded from this package directory only; there is no
fallback implementation — if it is missing, importing fails loudly.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmptg.so"

c_double_p = C.POINTER(C.c_double)
c_int32_p = C.POINTER(C.c_int32)
c_uint8_p = C.POINTER(C.c_uint8)
c_uint32_p = C.POINTER(C.c_uint32)
c_int64_p = C.POINTER(C.c_int64)


class PrrtParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("range", C.c_double),
        ("goal_bias", C.c_double),
        ("goal_radius", C.c_double),
        ("batch", C.c_int32),
        ("order", C.c_int32),
    ]


class TuneParams(C.Structure):
    _fields_ = [("knn_blocks_per_cu", C.c_int32), ("dmv_long_waves_per_cu", C.c_int32),
                ("mesh_waves_per_cu", C.c_int32), ("arm_tail_waves", C.c_int32), ("arm_cull", C.c_int32),
                ("mesh_edge_waves", C.c_int32), ("knn_select", C.c_int32), ("knn_grid", C.c_int32),
                ("dmv_edge_lanes", C.c_int32), ("knn_q8", C.c_int32)]


class TuneReport(C.Structure):
    _fields_ = [("knob", C.c_int32), ("ncand", C.c_int32), ("chosen", C.c_int32), ("_pad", C.c_int32),
                ("value", C.c_int32 * 8), ("ms", C.c_float * 8)]

    def as_dict(self) -> dict:
        n = int(self.ncand)
        return {"knob": int(self.knob), "chosen": int(self.chosen),
                "candidates": {int(self.value[i]): float(self.ms[i]) for i in range(n)}}


class PrrtStarParams(C.Structure):
    _fields_ = [("k", C.c_int32), ("neighbors", C.c_int32), ("rewire_factor", C.c_double)]


class PrrtStats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "iterations", "samples", "zero_dist", "invalid_states", "link_calls", "link_ok",
        "valid_calls", "goals", "size", "rewire_tests", "rewires", "parent_changes", "cp_links", "rw_links",
        "deferred", "graph_launches", "rw_late")]

    def as_dict(self) -> dict:
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


# (name, restype, argtypes)
_SIGNATURES = [
    ("mptg_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("mptg_destroy", None, [C.c_void_p]),
    ("mptg_last_error", C.c_char_p, [C.c_void_p]),
    ("mptg_sync", C.c_int, [C.c_void_p]),
    ("mptg_stream", C.c_void_p, [C.c_void_p]),
    ("mptg_dev_alloc", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    ("mptg_dev_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mptg_dev_upload", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("mptg_dev_download", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("mptg_dev_memset", C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64]),
    ("mptg_timer_start", C.c_int, [C.c_void_p]),
    ("mptg_timer_stop", C.c_int, [C.c_void_p, C.POINTER(C.c_float)]),
    ("mptg_space", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]),
    ("mptg_tree_reserve", C.c_int, [C.c_void_p, C.c_int64]),
    ("mptg_tree_clear", C.c_int, [C.c_void_p]),
    ("mptg_tree_insert", C.c_int, [C.c_void_p, c_double_p, c_int32_p, C.c_int64]),
    ("mptg_tree_size", C.c_int64, [C.c_void_p]),
    ("mptg_tree_read", C.c_int, [C.c_void_p, C.c_int64, C.c_int64, c_double_p, c_int32_p]),
    ("mptg_knn", C.c_int, [C.c_void_p, c_double_p, C.c_int64, C.c_int, C.c_double, c_int32_p, c_double_p]),
    ("mptg_knn_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_void_p, C.c_void_p]),
    ("mptg_rnn", C.c_int, [C.c_void_p, c_double_p, C.c_int64, C.c_double, C.c_int64, C.POINTER(C.c_int64),
                           C.POINTER(C.c_int32), c_double_p]),
    ("mptg_scene_boxes", C.c_int, [C.c_void_p, c_double_p, C.c_int]),
    ("mptg_scene_shapes", C.c_int, [C.c_void_p, c_double_p, C.c_int, c_double_p, C.c_int]),
    ("mptg_scene_host", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mptg_scene_grid", C.c_int, [C.c_void_p, c_uint8_p, C.c_int, C.c_int]),
    ("mptg_scene_arm", C.c_int, [C.c_void_p, c_double_p, C.c_int, C.c_double, c_double_p, C.c_int]),
    ("mptg_scene_mesh", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float), C.c_int]),
    ("mptg_set_link", C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    ("mptg_valid", C.c_int, [C.c_void_p, c_double_p, C.c_int64, c_uint8_p]),
    ("mptg_link", C.c_int, [C.c_void_p, c_double_p, c_double_p, C.c_int64, C.c_int, C.c_int, c_uint8_p, c_uint32_p]),
    ("mptg_link_dev", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p,
                                C.c_void_p]),
    ("mptg_link_dev_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int,
                                        C.c_void_p, C.c_void_p]),
    ("mptg_prrt_init", C.c_int, [C.c_void_p, C.POINTER(PrrtParams), c_double_p, c_double_p, c_double_p]),
    ("mptg_prrt_add_start", C.c_int, [C.c_void_p, c_double_p]),
    ("mptg_prrt_run", C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(PrrtStats)]),
    ("mptg_prrt_stats_get", C.c_int, [C.c_void_p, C.POINTER(PrrtStats)]),
    ("mptg_prrt_graph", C.c_int, [C.c_void_p, C.c_int]),
    ("mptg_tune_get", C.c_int, [C.c_void_p, C.POINTER(TuneParams)]),
    ("mptg_tune_set", C.c_int, [C.c_void_p, C.POINTER(TuneParams)]),
    ("mptg_autotune_knn", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.POINTER(TuneReport)]),
    ("mptg_autotune_link", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(TuneReport)]),
    ("mptg_autotune_prrt", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(TuneReport)]),
    ("mptg_prrt_goals", C.c_int, [C.c_void_p, c_int32_p, C.c_int64, c_int64_p]),
    ("mptg_prrt_goal_edges", C.c_int, [C.c_void_p, c_double_p, c_int32_p, C.c_int64, c_int64_p]),
    ("mptg_prrt_sampler", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mptg_prrt_star_init", C.c_int, [C.c_void_p, C.POINTER(PrrtParams), C.POINTER(PrrtStarParams), c_double_p,
                                      c_double_p, c_double_p]),
    ("mptg_prrt_costs", C.c_int, [C.c_void_p, C.c_int64, C.c_int64, c_double_p]),
    ("mptg_surrogate_nparams", C.c_int64, []),
    ("mptg_surrogate_init", C.c_int, [C.c_void_p, C.c_uint64]),
    ("mptg_surrogate_get", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    ("mptg_surrogate_set", C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int64]),
    ("mptg_surrogate_features", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    ("mptg_surrogate_samples", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64, C.c_void_p,
                                         C.c_void_p]),
    ("mptg_surrogate_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    ("mptg_surrogate_loss_grad", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_float),
                                           C.POINTER(C.c_float)]),
    ("mptg_surrogate_train", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_float,
                                       C.POINTER(C.c_float)]),
    ("mptg_surrogate_score_edges", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    ("mptg_link_phases", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_void_p,
                                   C.c_void_p, C.POINTER(C.c_float)]),
    ("mptg_diag_copy", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    ("mptg_diag_math", C.c_int, [C.c_void_p, C.c_int, c_double_p, C.c_int64, c_double_p]),
    ("mptg_link_work", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int,
                                 C.POINTER(C.c_uint64)]),
]

EXPORTED = [name for name, _, _ in _SIGNATURES]

_lib = None


def lib() -> C.CDLL:
    """Loads ``libmptg.so`` (once) and declares every entry point."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m mpt_amd.build` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in _SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
