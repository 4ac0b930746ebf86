"""ctypes binding of libdaggerfft_hip (the C ABI declared in include/daggerfft_hip.h).

The library is the product path: there is no fallback.  If the shared object
is missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "libdaggerfft_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG_ROOT), "include", "daggerfft_hip.h")

DFFT_OK = 0
DFFT_ERR_ARG = 1
DFFT_ERR_UNSUPPORTED_SIZE = 2
DFFT_ERR_HIP = 3
DFFT_ERR_OOM = 4
DFFT_ERR_BUDGET = 5
DFFT_ERR_STATE = 6
DFFT_ERR_NOT_READY = 7
DFFT_ERR_UNSUPPORTED = 8
DFFT_ERR_SCRATCH = 9
DFFT_ERR_WISDOM = 10
STATIC_SCRATCH_BYTES = 256 << 20

C2C, R2C, C2R, R2R = 0, 1, 2, 3
F32, F64 = 0, 1
PENCIL, SLAB = 0, 1
FLAG_UNNORMALIZED = 0x1
FLAG_NORMALIZE_R2R = 0x2
FLAG_DESTROY_INPUT = 0x4
FLAG_NO_TUNE = 0x8
FLAG_GRAPH = 0x10
FLAG_STAGE_ORDER = 0x20
FLAG_PLANE_PIPELINE = 0x40
FLAG_BATCH_GROUPED = 0x80
FLAG_BATCH_LANES = 0x100
FLAG_MEASURE = 0x200
FLAG_DCT_FULL_LENGTH = 0x400
FLAG_DCT_HALF_LENGTH = 0x800
FLAG_BATCH_PERSISTENT = 0x1000

_lib = None
_lock = threading.Lock()


class DFFTError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        msg = strerror(status) if _lib is not None else str(status)
        if _lib is not None and status in (DFFT_ERR_HIP, DFFT_ERR_SCRATCH, DFFT_ERR_WISDOM):
            msg += " [" + _lib.dfft_last_error().decode() + "]"
        super().__init__("%s: %s (status %d)" % (where or "libdaggerfft_hip", msg, status))


class dfft_batch_desc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64 * 3), ("kind", ctypes.c_int), ("sign", ctypes.c_int),
                ("precision", ctypes.c_int), ("flags", ctypes.c_uint32)]


class dfft_server_config(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_inflight", ctypes.c_int), ("lanes", ctypes.c_int),
                ("flags", ctypes.c_uint32), ("reserved", ctypes.c_int)]


class dfft_request(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64 * 3), ("kind", ctypes.c_int), ("sign", ctypes.c_int),
                ("precision", ctypes.c_int), ("decomp", ctypes.c_int), ("flags", ctypes.c_uint32),
                ("req_flags", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("d_in", ctypes.c_void_p),
                ("d_out", ctypes.c_void_p), ("t_arrival", ctypes.c_double)]


class dfft_request_info(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("status", ctypes.c_int), ("lane", ctypes.c_int),
                ("reserved", ctypes.c_int), ("t_arrival", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double), ("d_result", ctypes.c_void_p), ("result_bytes", ctypes.c_size_t),
                ("algorithmic_bytes", ctypes.c_double)]


class dfft_server_stats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("submitted", "completed", "failed", "queued", "running",
                                              "peak_running", "plans", "overtakes", "buffer_allocs",
                                              "arena_chunks", "arena_waits")] + \
               [(k, ctypes.c_size_t) for k in ("budget", "bytes_held", "bytes_busy", "peak_bytes_held")]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3
REQ_DISCARD = 0x1


def _declare(L):
    P = ctypes.c_void_p
    I = ctypes.c_int
    I64P = ctypes.POINTER(ctypes.c_int64)
    IP = ctypes.POINTER(ctypes.c_int)
    SZP = ctypes.POINTER(ctypes.c_size_t)
    sig = {
        "dfft_plan_create": [ctypes.POINTER(P), I, I64P, I, I, IP, I, I, IP, I, ctypes.c_uint32],
        "dfft_plan_workspace_size": [P, SZP],
        "dfft_plan_io_size": [P, SZP, SZP],
        "dfft_execute": [P, P, P, P, P],
        "dfft_execute_static": [P, P, P, P],
        "dfft_device_copy": [P, P, ctypes.c_size_t, P],
        "dfft_execute_pass": [P, I, P, P, P, P],
        "dfft_plan_destroy": [P],
        "dfft_plan_num_passes": [P, IP],
        "dfft_plan_describe": [P, I, ctypes.c_char_p, ctypes.c_size_t],
        "dfft_plan_algorithmic_bytes": [P, ctypes.POINTER(ctypes.c_double)],
        "dfft_plan_schedule": [P, IP, IP],
        "dfft_version": [IP, IP],
        "dfft_box_pack": [P, P, I64P, I64P, I64P, ctypes.c_int64, I, P],
        "dfft_box_unpack": [P, P, I64P, I64P, I64P, ctypes.c_int64, I, P],
        "dfft_box_copy": [P, I64P, I64P, P, I64P, I64P, I64P, I, P],
        "dfft_permute": [P, P, I64P, IP, I, P],
        "dfft_fill_uniform": [P, ctypes.c_int64, I, I, ctypes.c_uint64, P],
        "dfft_batch_create": [ctypes.POINTER(P), I, ctypes.POINTER(dfft_batch_desc), ctypes.c_uint32],
        "dfft_batch_execute": [P, ctypes.POINTER(P), ctypes.POINTER(P), P],
        "dfft_batch_workspace_size": [P, SZP],
        "dfft_batch_num_items": [P, I64P],
        "dfft_batch_destroy": [P],
        "dfft_autotune": [I, I64P, I, I, IP, I, I, IP, I, ctypes.c_uint32, I, P],
        "dfft_tune_save": [ctypes.c_char_p],
        "dfft_tune_autosave": [ctypes.c_char_p],
        "dfft_plan_export_wisdom": [P, ctypes.c_char_p],
        "dfft_tune_load": [ctypes.c_char_p],
        "dfft_tune_clear": [],
        "dfft_tune_entries": [IP],
        "dfft_batch_algorithmic_bytes": [P, ctypes.POINTER(ctypes.c_double)],
        "dfft_batch_num_launches": [P, IP],
        "dfft_server_create": [ctypes.POINTER(P), ctypes.POINTER(dfft_server_config)],
        "dfft_server_clock": [P, ctypes.POINTER(ctypes.c_double)],
        "dfft_server_prepare": [P, ctypes.POINTER(dfft_request), I],
        "dfft_server_submit": [P, ctypes.POINTER(dfft_request), I64P],
        "dfft_server_wait": [P, ctypes.c_int64, ctypes.c_double],
        "dfft_server_wait_any": [P, ctypes.c_double, I64P],
        "dfft_server_info": [P, ctypes.c_int64, ctypes.POINTER(dfft_request_info)],
        "dfft_server_copy_result": [P, ctypes.c_int64, P, ctypes.c_size_t, P],
        "dfft_server_release": [P, ctypes.c_int64],
        "dfft_server_drain": [P, ctypes.c_double],
        "dfft_server_stats_get": [P, ctypes.POINTER(dfft_server_stats)],
        "dfft_server_destroy": [P],
    }
    for name, args in sig.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = ctypes.c_int
    L.dfft_strerror.argtypes = [ctypes.c_int]
    L.dfft_strerror.restype = ctypes.c_char_p
    L.dfft_last_error.argtypes = []
    L.dfft_last_error.restype = ctypes.c_char_p


def lib():
    """Load libdaggerfft_hip.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError("libdaggerfft_hip.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            _declare(L)
            _lib = L
    return _lib


def strerror(status):
    return lib().dfft_strerror(int(status)).decode()


def check(status, where=""):
    if status != DFFT_OK:
        raise DFFTError(status, where)


def i64(vals):
    vals = list(vals)
    return (ctypes.c_int64 * max(len(vals), 1))(*vals)


def i32(vals):
    vals = list(vals)
    return (ctypes.c_int * max(len(vals), 1))(*vals)


def header_symbols(path=HEADER_PATH):
    """Names of the functions declared in include/daggerfft_hip.h."""
    import re
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dfft_[a-z0-9_]+)\s*\(", txt)))
