"""ctypes binding of librotormap_hip.so (the C ABI in include/rotormap_hip.h).

The library is required: there is no fallback path.  Importing works on a
machine without a GPU (tests check the exports there); any call that needs a
device raises RotorMapError when none is present.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librotormap_hip.so")

RM_OK, RM_ERR_ARG, RM_ERR_SHAPE, RM_ERR_NOMEM, RM_ERR_HIP, RM_ERR_STATE = range(6)
RM_F32, RM_F16, RM_E4M3 = 0, 1, 2
RM_MODE_FP8, RM_MODE_FP8_RESCORE, RM_MODE_FP16 = 0, 1, 2
RM_INDEX_HOST, RM_INDEX_FP16 = 1, 2
RM_SCHED_FIFO, RM_SCHED_SRPT = 0, 1

_CODE_NAMES = {1: "bad argument", 2: "shape", 3: "out of memory", 4: "HIP error", 5: "state"}


class RotorMapError(RuntimeError):
    """A failed library call (status != RM_OK)."""

    def __init__(self, code, msg):
        super().__init__(f"{_CODE_NAMES.get(code, code)}: {msg}")
        self.code = code


class RotorMapArgumentError(RotorMapError, ValueError):
    """Bad argument or shape (the reference raises ArgumentError here)."""


vp, i64, i32, u64, f64p = C.c_void_p, C.c_int64, C.c_int, C.c_uint64, C.c_void_p

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "rm_version": (C.c_char_p, []),
    "rm_last_error": (C.c_char_p, []),
    "rm_ctx_create": (i32, [i32, C.POINTER(vp)]),
    "rm_ctx_destroy": (i32, [vp]),
    "rm_sync": (i32, [vp, vp]),
    "rm_ctx_bytes": (i64, [vp]),
    "rm_encode_batch": (i32, [vp, vp, i64, i32, i32, i32, i32, i32, vp, i32, i64, i64, vp, i32, i64,
                              i64, vp, vp, vp]),
    "rm_phase_table": (i32, [i32, i32, i32, vp]),
    "rm_revcomp_words": (i32, [vp, vp, vp, i64, i32, vp]),
    "rm_extract_windows": (i32, [vp, vp, vp, vp, i64, i32, vp]),
    "rm_quantize_e4m3": (i32, [vp, vp, vp, i32, i64, vp]),
    "rm_dequantize_e4m3": (i32, [vp, vp, vp, i64, vp]),
    "rm_gemm_chunk": (i32, [vp, vp, i32, i64, vp, i64, i64, vp, i64, i32, i64, i64, i64, vp]),
    "rm_rowtopk_merge": (i32, [vp, vp, vp, i64, vp, i32, i64, i64, i64, i64, i32, vp]),
    "rm_index_create": (i32, [vp, vp, i64, i64, i32, C.POINTER(vp)]),
    "rm_index_create_f16": (i32, [vp, vp, i64, i64, i32, C.POINTER(vp)]),
    "rm_index_destroy": (i32, [vp]),
    "rm_index_cols": (i64, [vp]),
    "rm_index_rdim": (i64, [vp]),
    "rm_index_e4m3": (vp, [vp]),
    "rm_index_f16": (vp, [vp]),
    "rm_index_set_rescore": (i32, [vp, i32]),
    "rm_index_rescore": (i32, [vp]),
    "rm_topk_batch": (i32, [vp, vp, vp, i64, vp, i64, i64, i32, i32, i32, vp, vp, i64, vp]),
    "rm_map_batch": (i32, [vp, vp, vp, i64, i32, i32, i32, i32, i32, i32, i32, vp, vp, i64, vp, vp]),
    "rm_map_batch_noalloc": (i32, [vp, vp, vp, i64, i32, i32, i32, i32, i32, i32, i32, vp, vp, i64,
                                   vp, vp]),
    "rm_map_batches": (i32, [vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, i32, i32, vp, vp, i64,
                             vp]),
    "rm_plan_get": (i32, [vp, vp, i64, i32, i32, vp, vp, vp]),
    "rm_plan_set": (i32, [vp, i64, i32, i32, i32, i32]),
    "rm_launch_rows": (i32, [vp, vp, i64, i32, i32, vp]),
    "rm_autotune": (i32, [vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp]),
    "rm_set_persistent": (i32, [vp, i32]),
    "rm_prof_enable": (i32, [vp, i32]),
    "rm_prof_read": (i32, [vp, vp, vp, vp]),
    "rm_batch_bytes": (i64, [i64, i32, i32, i32, i32, i32]),
    "rm_device_free_bytes": (i64, [vp]),
    "rm_server_create": (i32, [vp, vp, i32, i32, i32, i32, i32, i32, i32, i64, i32, i64,
                               C.POINTER(vp)]),
    "rm_server_submit": (i32, [vp, vp, i64, C.POINTER(i64)]),
    "rm_server_submit_ref": (i32, [vp, vp, i64, C.POINTER(i64)]),
    "rm_server_now": (C.c_double, [vp]),
    "rm_server_wait": (i32, [vp, i64, i32, vp, vp, vp, vp, vp]),
    "rm_server_cancel": (i32, [vp, i64]),
    "rm_server_stats": (i32, [vp, vp, vp, vp]),
    "rm_server_set_policy": (i32, [vp, i32]),
    "rm_server_set_lowload": (i32, [vp, i64, i64]),
    "rm_server_destroy": (i32, [vp]),
    "rm_fasta_reads": (i32, [vp, vp, i64, i32, vp, i64, vp, vp, C.POINTER(i64), vp]),
    "rm_fasta_reads_host": (i32, [vp, vp, i64, i32, vp, i64, vp, vp, C.POINTER(i64)]),
    "rm_synth_record": (i32, [u64, i64, i64, C.c_double, vp]),
    "rm_mark_n": (i32, [vp, i64, i64]),
    "rm_extract_windows_host": (i32, [vp, vp, vp, i64, i32]),
    "rm_mutate_words": (i32, [vp, i64, i32, vp, u64, vp, i32]),
}

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RotorMapError(RM_ERR_STATE, f"{LIB_PATH} is missing: build it with "
                                              "qpg-rotormap_amd/build.py (no fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().rm_last_error().decode(errors="replace")


def check(rc):
    if rc != RM_OK:
        msg = last_error()
        if rc in (RM_ERR_ARG, RM_ERR_SHAPE):
            raise RotorMapArgumentError(rc, msg)
        raise RotorMapError(rc, msg)


_ctxs = {}


def ctx(device=0):
    """The process-wide context of a device (created on first use)."""
    c = _ctxs.get(device)
    if c is None:
        h = C.c_void_p()
        check(lib().rm_ctx_create(device, C.byref(h)))
        c = _ctxs[device] = h
    return c


def ptr(t):
    """Raw address of a torch tensor / numpy array / None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return C.c_void_p(t.data_ptr())
    if hasattr(t, "ctypes"):
        return C.c_void_p(t.ctypes.data)
    return C.c_void_p(t)


def stream_of(device=0):
    """torch's current HIP stream on `device`, as a raw handle."""
    import torch
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)
