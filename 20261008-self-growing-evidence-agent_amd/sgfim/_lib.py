"""ctypes binding of libsgfim.so (the C ABI in include/sgfim.h).

The library is loaded from this package directory. There is no fallback: if the
HIP library is missing or cannot be loaded, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgfim.so")

OK, E_INVAL, E_NOMEM, E_DEVICE, E_DEPTH, E_CANCELLED, E_TIMEOUT, E_STATE, E_CAPACITY = 0, -1, -2, -3, -4, -5, -6, -7, -8
STATUS_NAMES = {0: "OK", -1: "E_INVAL", -2: "E_NOMEM", -3: "E_DEVICE", -4: "E_DEPTH", -5: "E_CANCELLED",
                -6: "E_TIMEOUT", -7: "E_STATE", -8: "E_CAPACITY"}

ROOT_STRICT = 0x1
ROOT_TOPK = 0x2
ORDER_LEARNED = 0x4
NO_GRAPH = 0x8
FRESH_PAIR = 0x10


class SgfimError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("mem_cap_bytes", C.c_int64), ("batch_nodes", C.c_int32),
                ("chunk_words", C.c_int32), ("frontier_cap", C.c_int64), ("ring_cap", C.c_int64)]


class MineArgs(C.Structure):
    _fields_ = [("min_count", C.c_int64), ("flags", C.c_uint32), ("rel_thr", C.c_double),
                ("root_topk", C.c_int32), ("deadline_s", C.c_double)]


class RoundStats(C.Structure):
    _fields_ = [("round", C.c_int64), ("nodes", C.c_int64), ("counts", C.c_int64), ("mfis", C.c_int64),
                ("frontier", C.c_int64), ("words", C.c_int64), ("pair_counts", C.c_int64),
                ("count_launches", C.c_int64), ("count_ms", C.c_double), ("event_ms", C.c_double),
                ("elapsed_s", C.c_double),
                ("done", C.c_int32), ("status", C.c_int32), ("projections", C.c_int64), ("proj_words", C.c_int64),
                ("proj_from_rows", C.c_int64), ("proj_block_words", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# name -> (restype, argtypes)
P = C.c_void_p
PP = C.POINTER(C.c_void_p)
_SIGS = {
    "sgfim_version": (C.c_char_p, []),
    "sgfim_create": (C.c_int, [C.POINTER(Opts), PP]),
    "sgfim_destroy": (None, [P]),
    "sgfim_last_error": (C.c_char_p, [P]),
    "sgfim_load_csr": (C.c_int, [P, P, C.c_int64, P, C.c_int64]),
    "sgfim_load_windows": (C.c_int, [P, C.c_int32, P, P, C.c_int64, P, C.c_int64]),
    "sgfim_num_windows": (C.c_int, [P, C.POINTER(C.c_int32)]),
    "sgfim_window_supports": (C.c_int, [P, C.c_int32, P]),
    "sgfim_set_window_thresholds": (C.c_int, [P, P, C.c_int32]),
    "sgfim_set_thresholds": (C.c_int, [P, P, P, C.c_int32]),
    "sgfim_load_sensors": (C.c_int, [P, P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "sgfim_load_sensors_online": (C.c_int, [P, P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P]),
    "sgfim_sensor_items": (C.c_int, [P, P, C.c_int64, P]),
    "sgfim_set_tile": (C.c_int, [P, C.c_int32, C.c_int32]),
    "sgfim_set_sib_ring": (C.c_int, [P, C.c_int64]),
    "sgfim_get_tile": (C.c_int, [P, P, P]),
    "sgfim_set_projection": (C.c_int, [P, C.c_int32, C.c_int64]),
    "sgfim_stage_times": (C.c_int, [P, P, C.c_int32]),
    "sgfim_count_batch_noalloc": (C.c_int, [P, P, P, C.c_int32, P, P, P]),
    "sgfim_get_projection": (C.c_int, [P, P, P]),
    "sgfim_set_block_rule": (C.c_int, [P, C.c_int32, C.c_int32]),
    "sgfim_get_block_rule": (C.c_int, [P, P, P]),
    "sgfim_device_info": (C.c_int, [P, C.c_char_p, C.c_int32, P]),
    "sgfim_result_windows": (C.c_int, [P, PP]),
    "sgfim_load_zipf": (C.c_int, [P, C.c_int64, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                  C.c_int32, C.c_double, C.c_double, C.c_uint64]),
    "sgfim_zipf_spec": (C.c_int, [C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                  C.c_double, C.c_uint64, P, P, P, P]),
    "sgfim_info": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "sgfim_item_ids": (C.c_int, [P, P]),
    "sgfim_item_supports": (C.c_int, [P, P]),
    "sgfim_pair_matrix": (C.c_int, [P, P]),
    "sgfim_pair_matrix_valu": (C.c_int, [P, P]),
    "sgfim_pair_time": (C.c_int, [P, C.POINTER(C.c_double)]),
    "sgfim_set_pair_kernel": (C.c_int, [P, C.c_int32, C.c_int32]),
    "sgfim_set_count_kernel": (C.c_int, [P, C.c_int32]),
    "sgfim_mfma_probe": (C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "sgfim_count_batch": (C.c_int, [P, P, P, C.c_int32, P, P, P]),
    "sgfim_mine": (C.c_int, [P, C.POINTER(MineArgs), PP]),
    "sgfim_anytime_begin": (C.c_int, [P, C.POINTER(MineArgs)]),
    "sgfim_anytime_step": (C.c_int, [P, C.c_int32, C.POINTER(RoundStats)]),
    "sgfim_anytime_poll": (C.c_int, [P, C.POINTER(RoundStats)]),
    "sgfim_anytime_cancel": (C.c_int, [P]),
    "sgfim_anytime_take": (C.c_int, [P, PP]),
    "sgfim_result_get": (C.c_int, [P, PP, PP, PP, PP, C.POINTER(C.c_int64)]),
    "sgfim_result_times": (C.c_int, [P, PP]),
    "sgfim_result_free": (None, [P]),
    "sgfim_copy_probe": (C.c_int, [P, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "sgfim_last_stats": (C.c_int, [P, C.POINTER(RoundStats)]),
    "sgfim_set_scorer": (C.c_int, [P, P, P, P, C.c_float]),
    "sgfim_set_trace": (C.c_int, [P, C.c_int32, C.c_int64]),
    "sgfim_trace_get": (C.c_int, [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), P, P, P, C.c_int64]),
    "sgfim_mlp_create": (C.c_int, [C.c_int32, P, P, P, C.c_float, PP]),
    "sgfim_mlp_destroy": (None, [P]),
    "sgfim_mlp_last_error": (C.c_char_p, [P]),
    "sgfim_mlp_step": (C.c_int, [P, P, P, C.c_int32, C.c_float, P, P]),
    "sgfim_mlp_set_data": (C.c_int, [P, P, P, C.c_int64]),
    "sgfim_mlp_train": (C.c_int, [P, C.c_int32, C.c_int32, C.c_float, C.c_uint64, P]),
    "sgfim_mlp_get": (C.c_int, [P, P, P, P, P]),
    "sgfim_mlp_forward": (C.c_int, [P, P, C.c_int32, C.c_int32, P]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libsgfim.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SgfimError(E_STATE, f"{LIB_PATH} is not built; run python -m sgfim.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return {name: hasattr(L, name) for name in _SIGS}
