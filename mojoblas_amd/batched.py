"""Variable-size batched gemm over the C-ABI (mbh_vbatch_* in
include/mojoblas_hip.h).

    plan = VBatchPlan(m, n, k, precision="fp32")      # shapes fixed here
    plan.exec(alpha, a_ptrs, b_ptrs, beta, c_ptrs)    # capturable in a graph

a_ptrs / b_ptrs / c_ptrs are int64 CUDA tensors holding one device address
per problem (see pointer_array).  Leading dimensions default to the row
counts (lda = m, ldb = k, ldc = m).
"""
import ctypes

import numpy as np

from . import _lib, PRECISIONS, HipBlasError


def _i64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int64))


def _check(rc, what):
    if rc != 0:
        raise HipBlasError("%s failed (%d): %s" % (what, rc, _lib.last_error()))


class VBatchPlan:
    def __init__(self, m, n, k, lda=None, ldb=None, ldc=None, precision="fp32"):
        self.m, self.n, self.k = _i64(m), _i64(n), _i64(k)
        count = self.m.size
        if not (self.n.size == count and self.k.size == count):
            raise ValueError("m, n, k must have the same length")
        self.lda = _i64(np.maximum(self.m, 1) if lda is None else lda)
        self.ldb = _i64(np.maximum(self.k, 1) if ldb is None else ldb)
        self.ldc = _i64(np.maximum(self.m, 1) if ldc is None else ldc)
        self.count = count
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        h = ctypes.c_void_p()
        P = lambda a: ctypes.c_void_p(a.ctypes.data)
        rc = _lib.lib().mbh_vbatch_plan_create(count, P(self.m), P(self.n), P(self.k), P(self.lda), P(self.ldb),
                                               P(self.ldc), self.precision, ctypes.byref(h))
        _check(rc, "mbh_vbatch_plan_create")
        self._h = h

    def info(self):
        out = np.zeros(8, np.float64)
        _check(_lib.lib().mbh_vbatch_plan_info(self._h, ctypes.c_void_p(out.ctypes.data)), "mbh_vbatch_plan_info")
        keys = ["count", "tiles", "workgroups", "flops", "problems_64", "problems_128", "kernel", "precision"]
        return dict(zip(keys, out.tolist()))

    def exec(self, alpha, a_ptrs, b_ptrs, beta, c_ptrs, stream=None, operands="fp32"):
        """operands "fp32": A_p / B_p are fp32 (mbh_sgemm_vbatched_exec);
        "bf16": they are bf16 (mbh_bgemm_vbatched_exec, a "bf16" plan only).
        C_p is fp32 either way."""
        for t in (a_ptrs, b_ptrs, c_ptrs):
            if not (t.is_cuda and str(t.dtype) == "torch.int64" and t.numel() == self.count):
                raise TypeError("pointer arrays must be int64 CUDA tensors with one entry per problem")
        if operands not in ("fp32", "bf16"):
            raise ValueError("operands must be 'fp32' or 'bf16'")
        if stream is None:
            import torch

            stream = torch.cuda.current_stream().cuda_stream
        name = "mbh_sgemm_vbatched_exec" if operands == "fp32" else "mbh_bgemm_vbatched_exec"
        rc = getattr(_lib.lib(), name)(self._h, alpha, ctypes.c_void_p(a_ptrs.data_ptr()),
                                       ctypes.c_void_p(b_ptrs.data_ptr()), beta, ctypes.c_void_p(c_ptrs.data_ptr()),
                                       ctypes.c_void_p(stream))
        _check(rc, name)

    def close(self):
        if getattr(self, "_h", None):
            _check(_lib.lib().mbh_vbatch_plan_destroy(self._h), "mbh_vbatch_plan_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_tuning(class_table=None, wgs_per_cu=0, precision="fp32"):
    """Override, for one precision, the tile-class table (6x6 of 0/1, indexed
    [bucket(m)][bucket(n)]) and workgroups per CU for plans created
    afterwards; None / 0 restore the compiled defaults."""
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    if class_table is None:
        rc = _lib.lib().mbh_vbatch_set_tuning(prec, None, int(wgs_per_cu))
    else:
        t = np.ascontiguousarray(np.asarray(class_table, dtype=np.int32).reshape(36))
        rc = _lib.lib().mbh_vbatch_set_tuning(prec, ctypes.c_void_p(t.ctypes.data), int(wgs_per_cu))
    _check(rc, "mbh_vbatch_set_tuning")


def pack_problems(torch, sizes_m, sizes_n, sizes_k, gen=None, device="cuda", fill="uniform"):
    """Allocate A, B, C for a problem list as three flat fp32 buffers (each
    operand 16-byte aligned, column-major with ld = rows) and return
    (buffers, offsets, pointer arrays)."""
    m, n, k = (np.asarray(x, np.int64) for x in (sizes_m, sizes_n, sizes_k))

    def layout(elems):
        padded = (elems + 3) // 4 * 4
        off = np.concatenate([[0], np.cumsum(padded)[:-1]]).astype(np.int64)
        return off, int(padded.sum())

    aoff, at = layout(m * k)
    boff, bt = layout(k * n)
    coff, ct = layout(m * n)
    if fill == "uniform":
        A = torch.rand(max(at, 1), device=device, generator=gen) * 2 - 1
        B = torch.rand(max(bt, 1), device=device, generator=gen) * 2 - 1
    else:
        A = torch.randn(max(at, 1), device=device, generator=gen)
        B = torch.randn(max(bt, 1), device=device, generator=gen)
    C = torch.zeros(max(ct, 1), device=device)
    ptr = lambda T, off: torch.from_numpy(T.data_ptr() + 4 * off).to(device)
    return (A, B, C), (aoff, boff, coff), (ptr(A, aoff), ptr(B, boff), ptr(C, coff))


def to_bf16(torch, buffers, offsets):
    """bf16 copies (nearest even) of the flat fp32 operand buffers of
    pack_problems and their pointer arrays (same element offsets, so the
    operands keep their leading dimensions)."""
    out, ptrs = [], []
    for T, off in zip(buffers, offsets):
        H = T.to(torch.bfloat16)
        out.append(H)
        ptrs.append(torch.from_numpy(H.data_ptr() + 2 * np.asarray(off, np.int64)).to(T.device))
    return out, ptrs


def config4_shapes(count=4096, seed=2024, lo=16, hi=512):
    """m, n, k = floor(exp(U[ln lo, ln hi])) per problem (log-uniform), from seed."""
    rng = np.random.default_rng(seed)
    draw = lambda: np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(np.int64)
    return draw(), draw(), draw()
