"""Loader for libmojoblas_hip.so (the C-ABI declared in include/mojoblas_hip.h).

The library is built in-tree (mojoblas_amd/csrc/Makefile) and loaded from
this directory only; a missing library is an error, never a fallback.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmojoblas_hip.so")
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "mojoblas_hip.h")

_lib = None

i64 = ctypes.c_int64
f32 = ctypes.c_float
f64 = ctypes.c_double
vp = ctypes.c_void_p
ch = ctypes.c_char
i32 = ctypes.c_int

# Argument lists of every exported entry, in header order.
SIGNATURES = {
    "mbh_version": ([], ctypes.c_char_p),
    "mbh_last_error": ([], ctypes.c_char_p),
    "mbh_set_default_precision": ([i32], i32),
    "mbh_get_default_precision": ([], i32),
    "mbh_set_thread_precision": ([i32], i32),
    "mbh_device_sync": ([], i32),
    "mbh_debug_fail_allocation": ([ctypes.c_long], i32),
    "mbh_timing_enable": ([i32], i32),
    "mbh_timing_read": ([vp, i32], i32),
    "mbh_sgemm_host": ([ch, ch, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64], i32),
    "mbh_dgemm_host": ([ch, ch, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64], i32),
    "mbh_sgemm_dev": ([ch, ch, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, i32, vp], i32),
    "mbh_dgemm_dev": ([ch, ch, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64, vp], i32),
    "mbh_sgemm_dev_noalloc": ([ch, ch, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, i32, vp], i32),
    "mbh_hgemm_host": ([ch, ch, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64], i32),
    "mbh_bgemm_host": ([ch, ch, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64], i32),
    "mbh_hgemm_dev": ([ch, ch, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, vp], i32),
    "mbh_bgemm_dev": ([ch, ch, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, vp], i32),
    "mbh_strsm_host": ([ch, ch, ch, ch, i64, i64, f32, vp, i64, vp, i64], i32),
    "mbh_dtrsm_host": ([ch, ch, ch, ch, i64, i64, f64, vp, i64, vp, i64], i32),
    "mbh_strsm_dev": ([ch, ch, ch, ch, i64, i64, f32, vp, i64, vp, i64, i32, i32, vp], i32),
    "mbh_dtrsm_dev": ([ch, ch, ch, ch, i64, i64, f64, vp, i64, vp, i64, vp], i32),
    "mbh_strmm_host": ([ch, ch, ch, ch, i64, i64, f32, vp, i64, vp, i64], i32),
    "mbh_dtrmm_host": ([ch, ch, ch, ch, i64, i64, f64, vp, i64, vp, i64], i32),
    "mbh_strmm_dev": ([ch, ch, ch, ch, i64, i64, f32, vp, i64, vp, i64, i32, vp], i32),
    "mbh_dtrmm_dev": ([ch, ch, ch, ch, i64, i64, f64, vp, i64, vp, i64, vp], i32),
    "mbh_ssyrk_host": ([ch, ch, i64, i64, f32, vp, i64, f32, vp, i64], i32),
    "mbh_dsyrk_host": ([ch, ch, i64, i64, f64, vp, i64, f64, vp, i64], i32),
    "mbh_ssyrk_dev": ([ch, ch, i64, i64, f32, vp, i64, f32, vp, i64, i32, vp], i32),
    "mbh_dsyrk_dev": ([ch, ch, i64, i64, f64, vp, i64, f64, vp, i64, vp], i32),
    "mbh_ssyr2k_host": ([ch, ch, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64], i32),
    "mbh_dsyr2k_host": ([ch, ch, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64], i32),
    "mbh_ssyr2k_dev": ([ch, ch, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, i32, vp], i32),
    "mbh_dsyr2k_dev": ([ch, ch, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64, vp], i32),
    "mbh_ssymm_host": ([ch, ch, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64], i32),
    "mbh_dsymm_host": ([ch, ch, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64], i32),
    "mbh_ssymm_dev": ([ch, ch, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, i32, vp], i32),
    "mbh_dsymm_dev": ([ch, ch, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64, vp], i32),
    "mbh_sgemv_host": ([ch, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64], i32),
    "mbh_dgemv_host": ([ch, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64], i32),
    "mbh_sgemv_dev": ([ch, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, vp], i32),
    "mbh_dgemv_dev": ([ch, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64, vp], i32),
    "mbh_strsv_host": ([ch, ch, ch, i64, vp, i64, vp, i64], i32),
    "mbh_dtrsv_host": ([ch, ch, ch, i64, vp, i64, vp, i64], i32),
    "mbh_strsv_dev": ([ch, ch, ch, i64, vp, i64, vp, i64, vp], i32),
    "mbh_dtrsv_dev": ([ch, ch, ch, i64, vp, i64, vp, i64, vp], i32),
    "mbh_trsv_status": ([vp, ctypes.POINTER(i32)], i32),
    "mbh_vbatch_plan_create": ([i64, vp, vp, vp, vp, vp, vp, i32, vp], i32),
    "mbh_sgemm_vbatched_exec": ([vp, f32, vp, vp, f32, vp, vp], i32),
    "mbh_bgemm_vbatched_exec": ([vp, f32, vp, vp, f32, vp, vp], i32),
    "mbh_vbatch_plan_info": ([vp, vp], i32),
    "mbh_vbatch_plan_destroy": ([vp], i32),
    "mbh_vbatch_set_tuning": ([i32, vp, i32], i32),
    "mbh_sgemm_vbatched_dev": ([i64, vp, vp, vp, f32, vp, vp, vp, vp, f32, vp, vp, i32, vp], i32),
    "mbh_sgemm_backward": ([ch, ch, i64, i64, i64, f32, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64, i32, vp], i32),
    "mbh_strsm_backward": ([ch, ch, ch, ch, i64, i64, f32, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64, i32, vp],
                           i32),
    "mbh_dev_alloc": ([ctypes.c_size_t, ctypes.POINTER(vp)], i32),
    "mbh_dev_free": ([vp], i32),
    "mbh_host_alloc": ([ctypes.c_size_t, ctypes.POINTER(vp)], i32),
    "mbh_host_free": ([vp], i32),
    "mbh_set_matrix": ([i64, i64, i64, vp, i64, vp, i64, vp], i32),
    "mbh_get_matrix": ([i64, i64, i64, vp, i64, vp, i64, vp], i32),
    "mbh_stream_create": ([ctypes.POINTER(vp)], i32),
    "mbh_stream_destroy": ([vp], i32),
    "mbh_stream_sync": ([vp], i32),
    "mbh_workspace_release": ([vp], i32),
    "mbh_workspace_bytes": ([vp, ctypes.POINTER(ctypes.c_size_t)], i32),
}


def build(jobs=8):
    """Compile libmojoblas_hip.so for gfx950 with hipcc (no GPU needed)."""
    subprocess.check_call(["make", "-s", "-j%d" % jobs, "-C", os.path.join(HERE, "csrc")])


def lib():
    """The loaded C-ABI library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libmojoblas_hip.so is not built (%s); run mojoblas_amd._lib.build() or "
                "__graft_entry__.build() first" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def last_error():
    return lib().mbh_last_error().decode()
