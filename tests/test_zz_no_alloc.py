"""mbh_sgemm_dev_noalloc against mbh_sgemm_dev, through the C-ABI.

The no-allocation entry runs the same kernels with the pack planes in the
library's static device array, so its output must equal mbh_sgemm_dev's bit
for bit, in every precision and transpose, on two calls in a row (the
second smaller, so stale planes from the first would show), without
allocating device memory or touching the stream's workspace, and it must
refuse a product whose planes exceed the array before queueing anything.
"""
import ctypes

import pytest

torch = pytest.importorskip("torch")
import mojoblas_amd as mb  # noqa: F401  (loads the library)
from mojoblas_amd import _lib

pytestmark = pytest.mark.gpu
PREC = {"fp32": 0, "bf16x3": 1, "fp8": 2}
E_ARG = -4
LIMIT = 64 * 1024 * 1024


def _call(fn, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, prec, stream):
    return fn(ta.encode(), tb.encode(), m, n, k, alpha, ctypes.c_void_p(A.data_ptr()), lda,
              ctypes.c_void_p(B.data_ptr()), ldb, beta, ctypes.c_void_p(C.data_ptr()), ldc, PREC[prec],
              ctypes.c_void_p(stream))


def _operands(g, ta, tb, m, n, k, pad):
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + pad, br + pad, m + pad
    A = torch.randn(lda * ac, device="cuda", generator=g)
    B = torch.randn(ldb * bc, device="cuda", generator=g)
    C0 = torch.randn(ldc * n, device="cuda", generator=g)
    return A, lda, B, ldb, C0, ldc


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp8"])
@pytest.mark.parametrize("tt", ["NN", "NT", "TN", "TT"])
def test_noalloc_matches_dev_entry_on_two_calls(prec, tt):
    L = _lib.lib()
    ta, tb = tt
    g = torch.Generator(device="cuda").manual_seed(21)
    stream = torch.cuda.current_stream().cuda_stream
    # the second product is smaller than the first: stale planes would show
    for (m, n, k, pad, alpha, beta) in [(1100, 900, 1300, 3, 1.25, 0.5), (300, 257, 190, 0, -0.75, 0.0)]:
        A, lda, B, ldb, C0, ldc = _operands(g, ta, tb, m, n, k, pad)
        want, got = C0.clone(), C0.clone()
        torch.cuda.synchronize()
        assert _call(L.mbh_sgemm_dev, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, want, ldc, prec, stream) == 0
        assert _call(L.mbh_sgemm_dev_noalloc, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, got, ldc, prec,
                     stream) == 0
        torch.cuda.synchronize()
        assert torch.isfinite(want).all()
        assert torch.equal(got, want), (prec, tt, m, n, k)


def test_noalloc_allocates_nothing_and_leaves_the_workspace_alone():
    L = _lib.lib()
    g = torch.Generator(device="cuda").manual_seed(22)
    s = mb.Stream()
    shapes = [(2048, 1536, 2048), (512, 512, 768)]  # BF16x3 planes: 29 MiB, then 5 MiB
    ops = [_operands(g, "N", "N", m, n, k, 0) for m, n, k in shapes]
    outs = [torch.empty_like(o[4]) for o in ops]
    warm = torch.empty(64 * 64, device="cuda")
    torch.cuda.synchronize()
    A0, B0 = ops[1][0], ops[1][2]
    assert _call(L.mbh_sgemm_dev_noalloc, "N", "N", 64, 64, 64, 1.0, A0, 512, B0, 768, 0.0, warm, 64, "bf16x3",
                 s.handle) == 0
    s.sync()
    free0 = torch.cuda.mem_get_info()[0]
    for rep in range(2):
        for (m, n, k), (A, lda, B, ldb, _, ldc), C in zip(shapes, ops, outs):
            for prec in ("bf16x3", "fp8", "fp32"):
                assert _call(L.mbh_sgemm_dev_noalloc, "N", "N", m, n, k, 1.0, A, lda, B, ldb, 0.0, C, ldc, prec,
                             s.handle) == 0
    s.sync()
    assert torch.cuda.mem_get_info()[0] == free0
    held = ctypes.c_size_t(1)
    assert L.mbh_workspace_bytes(ctypes.c_void_p(s.handle), ctypes.byref(held)) == 0 and held.value == 0
    s.close()


def test_noalloc_refuses_products_beyond_its_scratch():
    L = _lib.lib()
    m = n = k = 4096  # BF16x3 planes: 4 B x 2 x 4096^2 = 128 MiB
    A = torch.randn(m * k, device="cuda")
    B = torch.randn(k * n, device="cuda")
    C = torch.full((m * n,), float("nan"), device="cuda")
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    rc = _call(L.mbh_sgemm_dev_noalloc, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, "bf16x3", stream)
    assert rc == E_ARG and "pack scratch" in _lib.last_error()
    torch.cuda.synchronize()
    assert torch.isnan(C).all()  # nothing was queued
    # fp32 needs no planes: the same product runs, and equals the dev entry
    want = torch.empty_like(C)
    assert _call(L.mbh_sgemm_dev_noalloc, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, "fp32", stream) == 0
    assert _call(L.mbh_sgemm_dev, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, want, m, "fp32", stream) == 0
    torch.cuda.synchronize()
    assert torch.equal(C, want)
