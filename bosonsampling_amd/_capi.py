"""ctypes binding of libbsmpo.so (the C ABI declared in include/bsmpo.h).

The library is built in-tree by ``bosonsampling_amd.build``; importing this
module never falls back to anything else: if the library is missing,
``lib()`` raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbsmpo.so")

C128, C64, BF16X3 = 0, 1, 2

_lib = None

# (name, restype, argtypes) for every entry point of include/bsmpo.h
_vp, _i, _sz, _i32p, _i64p, _dp = C.c_void_p, C.c_int, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p
SIGNATURES = {
    "bsmpo_last_error": (C.c_char_p, []),
    "bsmpo_abi_version": (C.c_int, []),
    "bsmpo_set_eig_threshold": (C.c_int, [_i]),
    "bsmpo_set_tri_split": (C.c_int, [_i]),
    "bsmpo_tri_split": (C.c_int, []),
    "bsmpo_eig_threshold": (C.c_int, []),
    "bsmpo_set_tile_order": (C.c_int, [_i]),
    "bsmpo_tile_order": (C.c_int, []),
    "bsmpo_set_two_stage": (C.c_int, [_i]),
    "bsmpo_two_stage_threshold": (C.c_int, []),
    "bsmpo_set_tie_rel": (C.c_int, [C.c_double]),
    "bsmpo_set_drop_rel": (C.c_int, [C.c_double]),
    "bsmpo_drop_rel": (C.c_double, []),
    "bsmpo_bs_table": (C.c_int, [_i, _i, _dp, _dp, _vp]),
    "bsmpo_plan_query": (C.c_int, [_i, _i, _i32p, _i32p, _i32p, _i, _i, C.POINTER(C.c_size_t),
                                   C.POINTER(C.c_size_t)]),
    "bsmpo_plan_build": (C.c_int, [_i, _i, _i32p, _i32p, _i32p, _i, _i, _vp, _sz]),
    "bsmpo_plan_info_get": (C.c_int, [_vp, _vp]),
    "bsmpo_plan_sectors": (C.c_int, [_vp, _i32p, _i64p, _i32p, _i32p]),
    "bsmpo_site_update": (C.c_int, [_i, _i, _vp, _dp, _dp, _dp, _dp, _i32p, _dp, _i32p, _i, _i,
                                    _vp, _sz, _vp]),
    "bsmpo_contract": (C.c_int, [_i, _i, _vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _vp,
                                 _sz, _vp]),
    "bsmpo_contract_noalloc": (C.c_int, [_i, _i, _vp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i,
                                         _vp]),
    "bsmpo_observables": (C.c_int, [_i, _i, _i, _dp, _dp, _i32p, _dp, _dp, _dp, _dp, _vp]),
    "bsmpo_entanglement_entropy": (C.c_int, [_i, _i, _dp, _dp, _vp]),
    "bsmpo_group_query": (C.c_int, [_i, _vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "bsmpo_group_build": (C.c_int, [_i, _vp, _vp, _vp, _vp, _sz]),
    "bsmpo_group_run": (C.c_int, [_vp, _i, _vp, _sz, _vp]),
    "bsmpo_site_update_vjp_query": (C.c_int, [_i, _i, _vp, _i32p, C.POINTER(C.c_size_t)]),
    "bsmpo_site_update_vjp": (C.c_int, [_i, _i, _vp, _i32p, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                        _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _vp, _sz, _vp]),
    "bsmpo_bs_table_vjp": (C.c_int, [_i, _i, _dp, _dp, _dp, _vp]),
    "bsmpo_observable_vjp": (C.c_int, [_i, _i, _i, _dp, _dp, _i32p, _i32p, _dp, _dp, _dp, _dp,
                                       _vp]),
    "bsmpo_group_graph_create": (C.c_int, [_vp, _i, _vp, _sz, _vp, C.POINTER(C.c_void_p)]),
    "bsmpo_graph_launch": (C.c_int, [_vp, _vp]),
    "bsmpo_graph_destroy": (C.c_int, [_vp]),
    "bsmpo_layer_query": (C.c_int, [_vp, C.POINTER(C.c_size_t)]),
    "bsmpo_layer_create": (C.c_int, [_vp, _vp, _sz, _vp, C.POINTER(C.c_void_p)]),
    "bsmpo_layer_seek": (C.c_int, [_vp, _i, _vp]),
    "bsmpo_layer_launch": (C.c_int, [_vp, _vp]),
    "bsmpo_layer_state_read": (C.c_int, [_vp, _vp, _vp]),
    "bsmpo_layer_destroy": (C.c_int, [_vp]),
    "bsmpo_layer_needs": (C.c_int, [_i, _vp, _vp]),
    "bsmpo_timing_enable": (C.c_int, [_i]),
    "bsmpo_timing_read": (C.c_int, [_dp, _dp, _dp, _i64p, _i]),
}

STAGES = ("bs_table", "contract", "gram", "jacobi", "vectors", "wmat", "topk", "scatter",
          "observables", "tridiag", "bisect", "trivec", "backtransform")


def timing_enable(on: bool = True):
    check(lib().bsmpo_timing_enable(1 if on else 0), "timing_enable")


def timing_read(reset: bool = True):
    """{stage: (ms, flops, bytes, launches)} accumulated since the last reset."""
    import numpy as np
    n = len(STAGES)
    ms, fl, by = np.zeros(n), np.zeros(n), np.zeros(n)
    cnt = np.zeros(n, np.int64)
    check(lib().bsmpo_timing_read(ms.ctypes.data, fl.ctypes.data, by.ctypes.data,
                                  cnt.ctypes.data, 1 if reset else 0), "timing_read")
    return {s: (float(ms[i]), float(fl[i]), float(by[i]), int(cnt[i])) for i, s in enumerate(STAGES)}


class PlanInfo(C.Structure):
    _fields_ = [("n_sectors", C.c_int), ("n_items", C.c_int), ("c_elems", C.c_int64),
                ("g_elems", C.c_int64), ("sv_total", C.c_int64), ("flops_alg", C.c_double),
                ("flops_tau", C.c_double), ("bytes_alg16", C.c_double)]


class UpdateIO(C.Structure):
    _fields_ = [("Glc", C.c_void_p), ("Gcr", C.c_void_p), ("LR", C.c_void_p), ("U3", C.c_void_p),
                ("charge_c", C.c_void_p), ("lambda_c", C.c_void_p), ("num_lambda", C.c_void_p),
                ("edge", C.c_int32), ("reserved", C.c_int32)]


class LayerUpdate(C.Structure):
    _fields_ = [("CL", C.c_void_p), ("CC", C.c_void_p), ("CR", C.c_void_p), ("Glc", C.c_void_p),
                ("Gcr", C.c_void_p), ("LR", C.c_void_p), ("charge_c", C.c_void_p),
                ("lambda_c", C.c_void_p), ("num_lambda", C.c_void_p), ("edge", C.c_int32),
                ("first_replay", C.c_int32)]


class LayerCaps(C.Structure):
    _fields_ = [("var_bytes", C.c_int64), ("work_bytes", C.c_int64), ("tiles", C.c_int32),
                ("gram", C.c_int32), ("jac", C.c_int32), ("tri", C.c_int32), ("sec", C.c_int32),
                ("v", C.c_int32), ("w", C.c_int32), ("wyt", C.c_int32), ("wya", C.c_int32),
                ("status", C.c_int32)]
    INT_FIELDS = ("tiles", "gram", "jac", "tri", "sec", "v", "w", "wyt", "wya")

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class LayerSpec(C.Structure):
    _fields_ = [("d", C.c_int), ("chi", C.c_int), ("precision", C.c_int), ("n_sets", C.c_int),
                ("n_upd", C.c_int * 2), ("upd", C.c_void_p * 2), ("n_groups", C.c_int),
                ("n_streams", C.c_int), ("n_layers", C.c_int), ("d_B", C.c_void_p),
                ("caps", LayerCaps)]


class LayerState(C.Structure):
    _fields_ = [("status", C.c_int32), ("replays", C.c_int32), ("failed_at", C.c_int32),
                ("n_groups", C.c_int32)]


LAYER_STATUS = {0: "ok", 1: "over capacity", 2: "bad charges", 3: "tile too big",
                4: "sector too big", 5: "no BS matrices", 6: "skipped"}


class BsmpoError(RuntimeError):
    pass


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BsmpoError(f"{LIB_PATH} not built: run `python -m bosonsampling_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != 0:
        msg = lib().bsmpo_last_error().decode(errors="replace")
        raise BsmpoError(f"{what} failed ({rc}): {msg}")


def ptr(t) -> int:
    """Raw address of a torch tensor / numpy array (no copy)."""
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data
