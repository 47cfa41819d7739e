"""ctypes binding of libjpgpu.so (include/jpgpu.h).

The library is built in-tree by j_presto_amd.build (hipcc, gfx950).  There is
no fallback: if the library or a GPU is missing, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

from .abi import EnergyLog, SysDesc

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjpgpu.so")

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p

# name -> (restype, argtypes); every symbol include/jpgpu.h declares
SIGNATURES = {
    "jp_version": (C.c_char_p, []),
    "jp_ctx_create": (C.c_int, [C.POINTER(SysDesc), C.POINTER(vp)]),
    "jp_ctx_destroy": (C.c_int, [vp]),
    "jp_set_state": (C.c_int, [vp, f64p, f64p]),
    "jp_get_state": (C.c_int, [vp, f64p, f64p]),
    "jp_build_list": (C.c_int, [vp, f64p]),
    "jp_get_list": (C.c_int, [vp, i32p, i64p, i32p, i32p, C.c_int64]),
    "jp_nonbond": (C.c_int, [vp, f64p, f64p, f64p, f64p, f64p, C.c_int]),
    "jp_bonded": (C.c_int, [vp, f64p, f64p, f64p]),
    "jp_shake": (C.c_int, [vp, f64p, f64p]),
    "jp_run_nve": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(EnergyLog)]),
    "jp_last_timing": (C.c_int, [vp, f64p, f64p, f64p, f64p]),
    "jp_count_pairs": (C.c_int, [vp, i64p, i64p, i64p]),
    "jp_count_pairs_noalloc": (C.c_int, [vp, i64p, i64p, i64p]),
    "jp_set_step": (C.c_int, [vp, C.c_int64]),
    "jp_restart_step": (C.c_int, [vp, C.c_int32, C.POINTER(EnergyLog)]),
    "jp_prepare_graphs": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "jp_batch_prepare_graphs": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "jp_get_step": (C.c_int, [vp, i64p]),
    "jp_batch_set_step": (C.c_int, [vp, C.c_int64]),
    "jp_batch_restart_step": (C.c_int, [vp, C.c_int32, C.POINTER(EnergyLog)]),
    "jp_time_pair": (C.c_int, [vp, C.c_int32, C.c_int32, f64p]),
    "jp_set_canonical": (C.c_int, [vp, C.c_int32, C.c_double, C.c_int32, C.c_double, i32p]),
    "jp_set_mcmd": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, f64p, f64p, f64p, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_double]),
    "jp_batch_set_canonical": (C.c_int, [vp, C.c_int32, C.c_double, C.c_int32, C.c_double, i32p]),
    "jp_set_alsd": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, f64p, f64p, f64p, C.c_double, C.c_double,
                              C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32]),
    "jp_batch_set_alsd": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, f64p, f64p, f64p, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32]),
    "jp_set_cmm": (C.c_int, [vp, C.c_int32, C.c_double, C.c_int32, C.c_int32]),
    "jp_set_heating": (C.c_int, [vp, C.c_int32, C.c_double]),
    "jp_set_stopce": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32]),
    "jp_batch_set_stopce": (C.c_int, [vp, C.c_int32, i32p, i32p]),
    "jp_start_step": (C.c_int, [vp, C.c_int32]),
    "jp_batch_start_step": (C.c_int, [vp, C.c_int32]),
    "jp_batch_set_heating": (C.c_int, [vp, C.c_int32, C.c_double]),
    "jp_time_cmm": (C.c_int, [vp, C.c_int32, f64p, f64p]),
    "jp_get_lambda": (C.c_int, [vp, f64p, f64p]),
    "jp_set_lambda": (C.c_int, [vp, C.c_double, C.c_double]),
    "jp_batch_get_lambda": (C.c_int, [vp, C.c_int32, f64p, f64p]),
    "jp_batch_set_lambda": (C.c_int, [vp, C.c_int32, C.c_double, C.c_double]),
    "jp_batch_create": (C.c_int, [C.POINTER(SysDesc), C.c_int32, C.POINTER(vp)]),
    "jp_batch_run_nve": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(EnergyLog)]),
    "jp_batch_set_state": (C.c_int, [vp, C.c_int32, f64p, f64p]),
    "jp_batch_get_state": (C.c_int, [vp, C.c_int32, f64p, f64p]),
    "jp_batch_last_timing": (C.c_int, [vp, f64p]),
    "jp_batch_destroy": (C.c_int, [vp]),
    "jp_set_option": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "jp_batch_set_option": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "jp_get_option": (C.c_int, [vp, C.c_int32, i32p]),
    "jp_batch_get_option": (C.c_int, [vp, C.c_int32, i32p]),
}
JP_OPT_GRAPHS, JP_OPT_AUTOTUNE, JP_OPT_PAIR_WAVES, JP_OPT_LIST_WAVES, JP_OPT_LIST_SEGMENT = 1, 2, 3, 4, 5
JP_OPT_CELL_TABLE, JP_OPT_GRAPH_CAPTURES, JP_OPT_HALF_LIST, JP_OPT_LIST_BATCH, JP_OPT_GRID_SHIFT = 6, 7, 8, 9, 10
JP_EINVAL, JP_ENOMEM, JP_ENODEV = -101, -102, -103

_lib = None


class JpError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} returned {code}")
        self.code = code


def load() -> C.CDLL:
    """Load libjpgpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(fn: str, code: int, ok=(0,)) -> int:
    """Runtime errors (<= -100) raise; md_run ier codes are returned."""
    if code <= -100:
        raise JpError(fn, code)
    return code
