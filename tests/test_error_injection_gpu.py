"""Error paths under an allocation failure forced from the host
(mbh_debug_fail_allocation: the next internal device allocation fails once,
as on an exhausted device).  Each entry returns its runtime error code with
a message, queues nothing, leaves its outputs untouched, and the library
works normally afterwards."""
import ctypes
import threading

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import mojoblas_amd as mb
from mojoblas_amd import _lib

pytestmark = pytest.mark.gpu
E_HIP = -1


def test_dev_gemm_reports_a_failed_workspace_allocation():
    L = _lib.lib()
    m = n = k = 512
    A = torch.randn(m * k, device="cuda")
    B = torch.randn(k * n, device="cuda")
    C = torch.full((m * n,), float("nan"), device="cuda")
    s = mb.Stream()  # a fresh stream: its first BF16x3 call must allocate pack planes
    torch.cuda.synchronize()
    args = (b"N", b"N", m, n, k, 1.0, ctypes.c_void_p(A.data_ptr()), m, ctypes.c_void_p(B.data_ptr()), k, 0.0,
            ctypes.c_void_p(C.data_ptr()), m, 1, ctypes.c_void_p(s.handle))
    assert L.mbh_debug_fail_allocation(0) == 0
    assert L.mbh_sgemm_dev(*args) == E_HIP
    assert "out of memory" in _lib.last_error()
    s.sync()
    assert torch.isnan(C).all()  # nothing was queued
    assert L.mbh_sgemm_dev(*args) == 0  # the failure was one-shot
    s.sync()
    assert torch.isfinite(C).all()
    s.close()


def test_dev_trsm_reports_a_failed_workspace_allocation():
    L = _lib.lib()
    m = n = 256
    a = np.tril(np.random.default_rng(1).uniform(-1, 1, (m, m)) / m) + np.eye(m)
    A = torch.from_numpy(np.asfortranarray(a.astype(np.float32)).ravel(order="F")).cuda()
    B0 = torch.randn(m * n, device="cuda")
    B = B0.clone()
    s = mb.Stream()
    torch.cuda.synchronize()
    assert L.mbh_debug_fail_allocation(0) == 0
    rc = L.mbh_strsm_dev(b"L", b"L", b"N", b"N", m, n, 1.0, ctypes.c_void_p(A.data_ptr()), m,
                         ctypes.c_void_p(B.data_ptr()), m, 0, 1, ctypes.c_void_p(s.handle))
    assert rc == E_HIP and "out of memory" in _lib.last_error()
    s.sync()
    assert torch.equal(B, B0)
    s.close()


def test_host_gemm_reports_a_failed_staging_allocation():
    L = _lib.lib()
    out = {}

    def worker():  # a new thread has no staging buffers yet
        c = np.full(64 * 64, 7.0, np.float32)
        a = np.ones(64 * 64, np.float32)
        L.mbh_debug_fail_allocation(0)
        rc = L.mbh_sgemm_host(b"N", b"N", 64, 64, 64, 1.0, ctypes.c_void_p(a.ctypes.data), 64,
                              ctypes.c_void_p(a.ctypes.data), 64, 0.0, ctypes.c_void_p(c.ctypes.data), 64)
        out["rc"], out["msg"], out["c"] = rc, _lib.last_error(), c.copy()
        out["rc2"] = L.mbh_sgemm_host(b"N", b"N", 64, 64, 64, 1.0, ctypes.c_void_p(a.ctypes.data), 64,
                                      ctypes.c_void_p(a.ctypes.data), 64, 0.0, ctypes.c_void_p(c.ctypes.data), 64)
        out["c2"] = c

    t = threading.Thread(target=worker)
    t.start()
    t.join()
    assert out["rc"] == E_HIP and "out of memory" in out["msg"]
    assert (out["c"] == 7.0).all()
    assert out["rc2"] == 0 and (out["c2"] == 64.0).all()


def test_batched_plan_reports_a_failed_allocation():
    from mojoblas_amd import batched

    L = _lib.lib()
    assert L.mbh_debug_fail_allocation(0) == 0
    with pytest.raises(mb.HipBlasError, match="out of memory"):
        batched.VBatchPlan(np.array([64, 32]), np.array([32, 64]), np.array([16, 48]), precision="fp32")
    plan = batched.VBatchPlan(np.array([64, 32]), np.array([32, 64]), np.array([16, 48]), precision="fp32")
    plan.close()
