"""ctypes binding of librvf_hip.so (the C ABI declared in include/rvf.h).

Loading fails loudly when the library is missing: there is no CPU fallback
in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librvf_hip.so")

RVF_WANT_F, RVF_AUDIT, RVF_TIMING, RVF_GRAPH, RVF_SELF_VF = 1, 2, 4, 8, 16
METHODS = {"quad": 0, "mc": 1, "duffy": 2}
PRECISIONS = {"fp64": 0, "mixed": 1}


class RVFError(RuntimeError):
    """Error returned through the C ABI (message from rvf_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"rvf error {code}: {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("nquad", C.c_int32), ("method", C.c_int32), ("precision", C.c_int32), ("n_samples", C.c_int32),
                ("seed", C.c_uint64), ("flags", C.c_int32), ("variant", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_active", C.c_int32), ("round_tiles", C.c_int32),
                ("direct_tiles", C.c_int32), ("reserved", C.c_int32)]


class RequestStats(C.Structure):
    _fields_ = [("t_submit_ms", C.c_double), ("t_admit_ms", C.c_double), ("t_start_ms", C.c_double),
                ("t_done_ms", C.c_double), ("footprint_bytes", C.c_int64), ("tiles", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerCounters(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("submitted", "completed", "failed", "cancelled", "rounds", "tiles",
                                          "peak_bytes", "used_bytes", "pending", "active")] + \
               [(k, C.c_double) for k in ("launch_ms", "admit_ms", "wait_ms", "retire_ms")] + \
               [("batched_tiles", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneEntry(C.Structure):
    _fields_ = [("variant", C.c_int32), ("refill_min", C.c_int32), ("tiles", C.c_int64), ("ms", C.c_double),
                ("columns", C.c_int32), ("waves", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("precompute_ms", C.c_double), ("pair_ms", C.c_double), ("duffy_ms", C.c_double),
                ("fill_ms", C.c_double), ("reduce_ms", C.c_double), ("total_ms", C.c_double),
                ("n_elems", C.c_int64), ("unique_pairs", C.c_int64), ("tiles", C.c_int64), ("rays", C.c_int64),
                ("retests", C.c_int64), ("singular_pairs", C.c_int64), ("aabb_tests", C.c_int64),
                ("tri_tests", C.c_int64), ("trav_iters", C.c_int64), ("slot_steps", C.c_int64),
                ("task_batches", C.c_int64), ("exact_batches", C.c_int64), ("sure_hits", C.c_int64),
                ("culled_rays", C.c_int64), ("culled_waves", C.c_int64), ("tile_columns", C.c_int64),
                ("waves", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"librvf_hip.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    PP = C.POINTER(C.c_void_p)
    sig = {
        "rvf_ctx_create": ([C.c_int, PP], C.c_int),
        "rvf_ctx_destroy": ([P], None),
        "rvf_last_error": ([], C.c_char_p),
        "rvf_version": ([], C.c_int),
        "rvf_mesh_upload": ([P, P, I64, I64, P, P, P, P, P, PP], C.c_int),
        "rvf_curve_mesh_upload": ([P, P, I64, I64, P, P, P, P, I64, P, I64, PP], C.c_int),
        "rvf_mesh_free": ([P], None),
        "rvf_mesh_info": ([P, P, P, P], C.c_int),
        "rvf_obs_build": ([P, P, P, P, I64, P, I32, I32, PP], C.c_int),
        "rvf_obs_free": ([P], None),
        "rvf_obs_info": ([P, P, P, P], C.c_int),
        "rvf_obs_export": ([P, I32, I64, P, P, P, P, P], C.c_int),
        "rvf_bvh_build_host": ([P, I64, I32, I64, P, P, P, P], C.c_int),
        "rvf_plan_create": ([P, P, P, P, PP], C.c_int),
        "rvf_plan_run": ([P], C.c_int),
        "rvf_plan_sync": ([P], C.c_int),
        "rvf_plan_fetch": ([P, P, P, P, P, P, P, P, P], C.c_int),
        "rvf_plan_stats": ([P, P], C.c_int),
        "rvf_host_alloc": ([C.c_size_t, PP], C.c_int),
        "rvf_host_free": ([P], C.c_int),
        "rvf_plan_device_ptr": ([P, C.c_int, PP], C.c_int),
        "rvf_plan_tune": ([P, C.c_int32, C.c_int32], C.c_int),
        "rvf_plan_tune_tiles": ([P, C.c_int32], C.c_int),
        "rvf_plan_tune_waves": ([P, C.c_int32], C.c_int),
        "rvf_plan_kernel": ([P, C.c_char_p, C.c_int32], C.c_int),
        "rvf_plan_autotune": ([P, I64, P, I32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
                              C.c_int),
        "rvf_meshfile_read": ([C.c_char_p, I32, I32, C.c_char_p, PP], C.c_int),
        "rvf_meshfile_info": ([P, P, P, P, P], C.c_int),
        "rvf_meshfile_arrays": ([P, P, P, P, P], C.c_int),
        "rvf_meshfile_group": ([P, I32, P, C.c_char_p, I32], C.c_int),
        "rvf_meshfile_free": ([P], None),
        "rvf_server_create": ([P, P, PP], C.c_int),
        "rvf_server_destroy": ([P], None),
        "rvf_server_submit": ([P, P, I64, I64, P, P, P, P, P, P, P, I64, P, I32, I32, P, P], C.c_int),
        "rvf_server_poll": ([P, I64], C.c_int),
        "rvf_server_wait": ([P, I64, I32], C.c_int),
        "rvf_server_result": ([P, I64, P, P, P, P, P, P, P], C.c_int),
        "rvf_server_cancel": ([P, I64], C.c_int),
        "rvf_server_release": ([P, I64], C.c_int),
        "rvf_server_stats": ([P, P], C.c_int),
        "rvf_plan_free": ([P], None),
        "rvf_compute_view_factors": ([P, P, P, P, P, P, P, P, P, P], C.c_int),
        "rvf_compute_view_factors_noalloc": ([P, P, P, P, P, P, P, P, P, P], C.c_int),
        "rvf_quad_rule": ([I32, I32, P, P], C.c_int),
        "rvf_ctx_event_record": ([P, I32], C.c_int),
        "rvf_ctx_event_elapsed": ([P, I32, I32, C.POINTER(C.c_float)], C.c_int),
        "rvf_ctx_sync": ([P], C.c_int),
        "rvf_copy_probe": ([P, I64, I32, C.POINTER(C.c_float)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc):
    if rc is not None and rc < 0:
        raise RVFError(rc, lib().rvf_last_error().decode())
    return rc


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
