"""ctypes binding of libheu_paillier.so (the C ABI in include/heu_paillier.h).

Loading fails loudly when the library has not been built: there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libheu_paillier.so")

HEU_OK, HEU_EINVAL, HEU_ERANGE, HEU_ENOTINV, HEU_ENOMEM, HEU_EDEVICE, HEU_EUNSUPPORTED, HEU_EKEY = range(8)

# Every entry point declared in include/heu_paillier.h, with its ctypes signature.
P, U32, U64, SZ, I = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int
SIGNATURES = {
    "heu_last_error": (ctypes.c_char_p, []),
    "heu_ctx_create": (I, [I, ctypes.POINTER(P)]),
    "heu_ctx_destroy": (I, [P]),
    "heu_ctx_set_stream": (I, [P, P]),
    "heu_ctx_stream": (P, [P]),
    "heu_ctx_sync": (I, [P]),
    "heu_ctx_profile": (I, [P, I]),
    "heu_ctx_profile_read": (I, [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(U64)]),
    "heu_keygen": (I, [U32, ctypes.POINTER(U64), P, P, P, P]),
    "heu_key_create": (I, [P, P, P, U32, P, P, U32, U32, U32, ctypes.POINTER(P)]),
    "heu_key_destroy": (I, [P]),
    "heu_key_info": (I, [P, ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32), ctypes.POINTER(U32)]),
    "heu_key_device_bytes": (SZ, [P]),
    "heu_encrypt_i64": (I, [P, P, P, U32, SZ, P]),
    "heu_encrypt": (I, [P, P, P, U32, P, U32, SZ, P]),
    "heu_randomize": (I, [P, P, P, U32, SZ, P]),
    "heu_decrypt": (I, [P, P, SZ, P, U32, P]),
    "heu_add": (I, [P, P, P, SZ, P]),
    "heu_add_i64": (I, [P, P, P, SZ, P]),
    "heu_sub": (I, [P, P, P, SZ, P]),
    "heu_negate": (I, [P, P, SZ, P]),
    "heu_mul_i64": (I, [P, P, P, SZ, P]),
    "heu_mul": (I, [P, P, P, P, U32, SZ, P]),
    "heu_sum": (I, [P, P, P, SZ, P]),
    "heu_sum_noalloc": (I, [P, P, P, SZ, P]),
    "heu_dot_i64": (I, [P, P, P, P, SZ, P]),
    "heu_sum_gather": (I, [P, P, P, P, SZ, P]),
    "heu_dot_gather": (I, [P, P, P, P, P, SZ, P]),
    "heu_prefix_mul": (I, [P, P, P, SZ, P]),
    "heu_to_mont": (I, [P, P, SZ, U32, P]),
    "heu_from_mont": (I, [P, P, SZ, U32, P]),
    "heu_key_mform_words": (I, [P, ctypes.POINTER(U32)]),
    "heu_to_mform": (I, [P, P, SZ, P]),
    "heu_from_mform": (I, [P, P, SZ, P]),
    "heu_add_mform": (I, [P, P, P, SZ, P]),
    "heu_sum_mform": (I, [P, P, P, SZ, P]),
    "heu_batch_create": (I, [P, P, SZ, U32, ctypes.POINTER(P)]),
    "heu_batch_run": (I, [P]),
    "heu_batch_destroy": (I, [P]),
    "heu_batch_stats": (I, [P, ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(SZ)]),
    "heu_ctx_set_grid_multiplier": (I, [P, U32, I, U32]),
    "heu_key_set_param": (I, [P, I, U32]),
    "heu_key_get_param": (I, [P, I, ctypes.POINTER(U32)]),
    "heu_autotune": (I, [P, U32, ctypes.c_char_p, SZ]),
    "heu_server_create": (I, [P, P, ctypes.POINTER(P)]),
    "heu_server_submit": (I, [P, P, ctypes.POINTER(U64)]),
    "heu_server_wait": (I, [P, U64, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]),
    "heu_server_stats": (I, [P, P]),
    "heu_server_destroy": (I, [P]),
    "heu_wire_put_ciphertext": (I, [P, U32, P, SZ, ctypes.POINTER(SZ)]),
    "heu_wire_get_ciphertext": (I, [P, SZ, P, U32]),
    "heu_wire_put_matrix": (I, [P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, U32, P, SZ, ctypes.POINTER(SZ)]),
    "heu_wire_get_matrix": (I, [P, SZ, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                ctypes.POINTER(ctypes.c_int64), P, U32, SZ]),
    "heu_wire_put_matrix_ic": (I, [P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, U32, P, SZ,
                                   ctypes.POINTER(SZ)]),
    "heu_wire_get_matrix_ic": (I, [P, SZ, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                   ctypes.POINTER(ctypes.c_int64), P, U32, SZ]),
    "heu_wire_put_public_key": (I, [P, U32, P, P, SZ, ctypes.POINTER(SZ)]),
    "heu_wire_get_public_key": (I, [P, SZ, P, P, U32, ctypes.POINTER(U32)]),
    "heu_wire_put_secret_key": (I, [P, P, U32, P, SZ, ctypes.POINTER(SZ)]),
    "heu_wire_get_secret_key": (I, [P, SZ, P, P, U32, ctypes.POINTER(U32)]),
}

# heu_batch ops and the heu_task struct (include/heu_paillier.h)
(HEU_OP_ENCRYPT_I64, HEU_OP_RANDOMIZE, HEU_OP_ADD, HEU_OP_ADD_I64, HEU_OP_MUL_I64, HEU_OP_DECRYPT, HEU_OP_SUM,
 HEU_OP_DOT_I64, HEU_OP_BARRIER) = range(1, 10)
HEU_BATCH_GRAPH = 1
HEU_PARAM_TABLE_WINDOW, HEU_PARAM_DEC_LANES = 1, 2
HEU_TUNE_TABLE, HEU_TUNE_DEC_LANES, HEU_TUNE_GRID, HEU_TUNE_ALL = 1, 2, 4, 7
HEU_SCHEMA_ZPAILLIER = 4


class HeuTask(ctypes.Structure):
    _fields_ = [("key", P), ("op", ctypes.c_int32), ("aux", ctypes.c_uint32), ("a", P), ("b", P),
                ("count", SZ), ("offsets", P), ("segments", SZ), ("out", P)]

_lib = None
_lock = threading.Lock()


class HeuError(RuntimeError):
    """A failing engine call (the reference raises yacl::Exception, surfaced
    to Python as RuntimeError)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[heu error {code}] {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m heu_amd.build` "
                                   "(the engine has no CPU fallback)")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != HEU_OK:
        msg = lib().heu_last_error()
        raise HeuError(rc, msg.decode() if msg else "")


class HeuServerConfig(ctypes.Structure):
    _fields_ = [("workspace_bytes", SZ), ("max_inflight", U32), ("max_batch_tasks", U32),
                ("batch_rounds", U32), ("parts_per_round", U32), ("urgent_ms", U32), ("max_workspace_bytes", SZ)]


class HeuServerStats(ctypes.Structure):
    _fields_ = [("submitted", U64), ("completed", U64), ("failed", U64), ("batches", U64),
                ("max_batch_requests", U64), ("peak_workspace_bytes", SZ), ("workspace_bytes", SZ)]
