"""Loader for libngap_amd.so (the C ABI declared in include/ngap_amd.h).

The library is built in-tree by ``make -C ngap_amd`` (``__graft_entry__.build``).
There is no fallback: a missing or stale library raises immediately.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libngap_amd.so")

NGAP_OK = 0
NGAP_ERR_INVALID = 1
NGAP_ERR_DEVICE = 2
NGAP_ERR_OOM = 3
NGAP_ERR_OVERFLOW = 4
NGAP_ERR_STATE = 5
NGAP_ERR_HOST = 6

ALGORITHMS = {  # --algorithm names (code/src/ngap/main.cu:114-130)
    "blockinggroups": 0,
    "NAPgroups": 1,
    "nonblockinggroups": 2,
    "nonblockingr1groups": 3,
    "nonblockingr2groups": 4,
    "nonblockingpcgroups": 5,
    "nonblockingallgroups": 6,
}
SCHED_PLAIN = 0
SCHED_BATCHED = 1
BUILD_DEVICE = 0   # ngap_load_groups_ex: image built by HIP kernels
BUILD_HOST = 1     # ... by host threads


class GroupDesc(C.Structure):
    _fields_ = [
        ("num_states", C.c_int32),
        ("row_offsets", C.c_void_p),
        ("col_indices", C.c_void_p),
        ("symbol_sets", C.c_void_p),
        ("node_attrs", C.c_void_p),
        ("num_always_active", C.c_int32),
        ("always_active", C.c_void_p),
        ("num_start", C.c_int32),
        ("start_states", C.c_void_p),
    ]


class Task(C.Structure):
    _fields_ = [("group", C.c_uint32), ("chunk", C.c_uint32), ("begin", C.c_uint32),
                ("end", C.c_uint32)]


class RunOpts(C.Structure):
    _fields_ = [
        ("algorithm", C.c_int32),
        ("schedule", C.c_int32),
        ("precompute_depth", C.c_int32),
        ("active_threshold", C.c_int32),
        ("unique", C.c_int32),
        ("is_dup", C.c_int32),
        ("report_off", C.c_int32),
        ("data_buffer_fetch_size", C.c_uint32),
        ("add_aas_start", C.c_uint32),
        ("add_aas_interval", C.c_uint32),
        ("result_capacity", C.c_uint64),
        ("window", C.c_uint32),
        ("spill_capacity", C.c_uint32),
        ("persistent_blocks", C.c_uint32),
        ("use_graph", C.c_int32),
        ("tasks", C.c_void_p),
        ("num_tasks", C.c_uint32),
        ("fallback", C.c_int32),
        ("cross_groups", C.c_int32),
        ("ring_capacity", C.c_uint32),
        ("task_traces", C.c_int32),
        ("root_pass", C.c_int32),
        ("root_capacity", C.c_uint32),
        ("tasks_id", C.c_uint64),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double),
        ("reports_raw", C.c_uint64),
        ("overflow_units", C.c_uint64),
        ("items", C.c_uint64),
        ("num_units", C.c_uint32),
        ("num_tasks", C.c_uint32),
        ("fallback_tasks", C.c_uint32),
        ("fallback_ms", C.c_double),
        ("tasks_id", C.c_uint64),
    ]


class ImageInfo(C.Structure):
    _fields_ = [
        ("built_on_device", C.c_int32),
        ("memo_depth", C.c_int32),
        ("num_groups", C.c_int32),
        ("num_states", C.c_uint64),
        ("bytes", C.c_uint64),
        ("build_ms", C.c_double),
    ]


class ServerInfo(C.Structure):
    _fields_ = [
        ("budget_bytes", C.c_uint64),
        ("queued", C.c_uint64),
        ("inflight_requests", C.c_uint64),
        ("inflight_bytes", C.c_uint64),
        ("held_bytes", C.c_uint64),
        ("peak_inflight_bytes", C.c_uint64),
        ("peak_held_bytes", C.c_uint64),
        ("launches", C.c_uint64),
        ("rejected", C.c_uint64),
        ("redone", C.c_uint64),
    ]


# Every symbol include/ngap_amd.h declares, with (restype, argtypes).
SIGNATURES = {
    "ngap_last_error": (C.c_char_p, []),
    "ngap_default_opts": (None, [C.POINTER(RunOpts)]),
    "ngap_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "ngap_destroy": (C.c_int, [C.c_void_p]),
    "ngap_load_groups": (C.c_int, [C.c_void_p, C.POINTER(GroupDesc), C.c_int32, C.c_int32]),
    "ngap_load_groups_ex": (C.c_int, [C.c_void_p, C.POINTER(GroupDesc), C.c_int32, C.c_int32,
                                      C.c_int32]),
    "ngap_image_info": (C.c_int, [C.c_void_p, C.POINTER(ImageInfo)]),
    "ngap_image_array": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_uint64,
                                   C.POINTER(C.c_uint64)]),
    "ngap_set_input": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32]),
    "ngap_run": (C.c_int, [C.c_void_p, C.POINTER(RunOpts), C.POINTER(RunStats)]),
    "ngap_run_noalloc": (C.c_int, [C.c_void_p, C.POINTER(RunOpts), C.c_void_p, C.c_uint64,
                                   C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(RunStats)]),
    "ngap_autotune": (C.c_int, [C.c_void_p, C.POINTER(RunOpts), C.c_int32, C.POINTER(RunOpts),
                                C.POINTER(C.c_double)]),
    "ngap_batched_grid": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32),
                                    C.POINTER(C.c_uint32)]),
    "ngap_get_reports": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "ngap_get_task_items": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "ngap_get_tasks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "ngap_server_create": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.POINTER(RunOpts),
                                     C.POINTER(C.c_void_p)]),
    "ngap_server_destroy": (C.c_int, [C.c_void_p]),
    "ngap_server_submit": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64,
                                     C.POINTER(C.c_uint64)]),
    "ngap_server_poll": (C.c_int, [C.c_void_p, C.c_int32]),
    "ngap_server_take": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32), C.c_void_p,
                                   C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "ngap_server_pending": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "ngap_server_stats": (C.c_int, [C.c_void_p, C.POINTER(ServerInfo)]),
    "ngap_mlp_create": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "ngap_mlp_destroy": (C.c_int, [C.c_void_p]),
    "ngap_mlp_num_params": (C.c_int32, []),
    "ngap_mlp_get_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ngap_mlp_set_params": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ngap_mlp_predict": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "ngap_mlp_predict_pairs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                         C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p]),
    "ngap_mlp_gradient": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                    C.POINTER(C.c_float)]),
    "ngap_mlp_train": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32,
                                 C.c_uint32, C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
    "ngap_frontend_load": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                     C.POINTER(C.c_void_p)]),
    "ngap_frontend_load_text": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "ngap_frontend_free": (C.c_int, [C.c_void_p]),
    "ngap_frontend_num_groups": (C.c_int32, [C.c_void_p]),
    "ngap_frontend_group": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(GroupDesc)]),
    "ngap_frontend_state_name": (C.c_char_p, [C.c_void_p, C.c_int32, C.c_int32]),
    "ngap_frontend_dump": (C.c_char_p, [C.c_void_p]),
    "ngap_parse_symbol_set": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint32 * 8)]),
}

_lib = None


def lib():
    """Returns the loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C ngap_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class NgapError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ngap error {code}: {msg}")
        self.code = code


def check(code):
    if code != NGAP_OK:
        raise NgapError(code, lib().ngap_last_error().decode(errors="replace"))
