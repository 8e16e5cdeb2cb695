"""The C-ABI boundary under concurrency, graph capture and resource release.

* host entries from several threads at once, each thread with its own
  precision (mbh_set_thread_precision through the Python layer): every
  result equals the single-threaded result of its precision bit for bit;
* device entries from two threads on one stream, with shapes that make the
  stream's workspace grow while the other thread is between calls: every
  result equals the single-threaded one (calls hold the stream's workspace
  until their last kernel is queued);
* workspace release: mbh_workspace_release and mbh_stream_destroy hand the
  pack buffers back to the device;
* graph capture: a call whose workspace would have to grow during capture is
  refused with MBH_E_ARG before queueing anything; after one eager call the
  same call captures, and its graph replays correctly, also after a larger
  eager call has grown the stream's workspace (the captured buffer is kept);
* trsv status: mbh_trsv_status reports no give-up for a normal solve.
"""
import ctypes
import threading

import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
import mojoblas_amd as mb
from mojoblas_amd import _lib

pytestmark = pytest.mark.gpu
E_ARG = -4


def _host_gemm(m, n, k, seed, prec):
    a = oracle.fill_f32(m * k, seed, "normal")
    b = oracle.fill_f32(k * n, seed + 1, "normal")
    c = np.zeros(m * n, np.float32)
    assert mb.gemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m, precision=prec) == 0
    return c


def test_host_threads_keep_their_own_precision():
    shapes = [(384, 320, 512), (200, 130, 700)]
    want = {(s, p): _host_gemm(*s, 7, p) for s in shapes for p in ("fp32", "bf16x3")}
    assert not np.array_equal(want[(shapes[0], "fp32")], want[(shapes[0], "bf16x3")])
    bad = []

    def worker(prec, reps):
        try:
            for r in range(reps):
                s = shapes[r % 2]
                if not np.array_equal(_host_gemm(*s, 7, prec), want[(s, prec)]):
                    bad.append((prec, r))
        except Exception as exc:  # noqa: BLE001 - reported below
            bad.append((prec, repr(exc)))

    threads = [threading.Thread(target=worker, args=(p, 12)) for p in ("fp32", "bf16x3", "fp32", "bf16x3")]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert bad == []
    assert _lib.lib().mbh_get_default_precision() == 0  # no thread touched the default


def test_dev_calls_from_two_threads_share_one_stream_safely():
    """Both threads issue BF16x3 gemms on torch's default stream; the larger
    shape grows the stream's workspace (synchronise, free, allocate) while
    the other thread's calls come and go."""
    g = torch.Generator(device="cuda").manual_seed(3)
    shapes = [(256, 256, 256), (1024, 768, 2048), (512, 640, 384)]
    ops = []
    for m, n, k in shapes:
        A = torch.randn(m * k, device="cuda", generator=g)
        B = torch.randn(k * n, device="cuda", generator=g)
        ops.append((m, n, k, A, B))

    def run(i, C):
        m, n, k, A, B = ops[i]
        return mb.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, precision="bf16x3")

    want = []
    for i in range(3):
        C = torch.empty(ops[i][0] * ops[i][1], device="cuda")
        assert run(i, C) == 0
        want.append(C)
    torch.cuda.synchronize()
    assert _lib.lib().mbh_workspace_release(None) == 0  # the first concurrent call grows again
    orders = {0: [0, 1, 2], 1: [2, 0, 1]}
    outs = {t: [torch.full_like(want[orders[t][r % 3]], float("nan")) for r in range(18)] for t in range(2)}
    errs = []

    def worker(t):
        try:
            order = orders[t]
            for r in range(18):
                i = order[r % 3]
                if run(i, outs[t][r]) != 0:
                    errs.append((t, r))
        except Exception as exc:  # noqa: BLE001
            errs.append((t, repr(exc)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    torch.cuda.synchronize()
    assert errs == []
    for t in range(2):
        for r in range(18):
            assert torch.equal(outs[t][r], want[orders[t][r % 3]]), (t, r)


def _ws_bytes(stream):
    out = ctypes.c_size_t()
    assert _lib.lib().mbh_workspace_bytes(ctypes.c_void_p(stream), ctypes.byref(out)) == 0
    return out.value


def test_workspace_release_and_stream_destroy_free_the_pack_buffers():
    m = n = k = 2048
    g = torch.Generator(device="cuda").manual_seed(4)
    A = torch.randn(m * k, device="cuda", generator=g)
    B = torch.randn(k * n, device="cuda", generator=g)
    C = torch.empty(m * n, device="cuda")
    torch.cuda.synchronize()
    L = _lib.lib()
    s = mb.Stream()
    assert _ws_bytes(s.handle) == 0
    assert mb.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, precision="bf16x3", stream=s.handle) == 0
    s.sync()
    held = _ws_bytes(s.handle)
    assert held >= 2 * m * k * 4  # hi and lo planes of A and B
    assert L.mbh_workspace_release(ctypes.c_void_p(s.handle)) == 0
    assert _ws_bytes(s.handle) == 0
    assert mb.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, precision="bf16x3", stream=s.handle) == 0
    s.sync()
    assert _ws_bytes(s.handle) == held
    free_before = torch.cuda.mem_get_info()[0]
    s.close()  # mbh_stream_destroy
    free_after = torch.cuda.mem_get_info()[0]
    assert free_after - free_before >= 0.9 * held, (free_before, free_after, held)


def test_graph_capture_of_a_dev_gemm():
    L = _lib.lib()
    m, n, k = 1024, 512, 768
    g = torch.Generator(device="cuda").manual_seed(5)
    A = torch.randn(m * k, device="cuda", generator=g)
    B = torch.randn(k * n, device="cuda", generator=g)
    C = torch.zeros(m * n, device="cuda")
    cs = torch.cuda.Stream()
    torch.cuda.synchronize()

    def raw(stream, mm=m, nn=n, kk=k, AA=A, BB=B, CC=C):
        return L.mbh_sgemm_dev(b"N", b"N", mm, nn, kk, 1.0, ctypes.c_void_p(AA.data_ptr()), mm,
                               ctypes.c_void_p(BB.data_ptr()), kk, 0.0, ctypes.c_void_p(CC.data_ptr()), mm, 1,
                               ctypes.c_void_p(stream.cuda_stream))

    # a fresh stream has no workspace: the call is refused during capture
    g0 = torch.cuda.CUDAGraph()
    flag = torch.zeros(1, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.graph(g0, stream=cs):
        flag.fill_(1.0)  # keeps the captured graph non-empty
        rc = raw(torch.cuda.current_stream())
    assert rc == E_ARG and "captured" in _lib.last_error()
    assert (C == 0).all()

    with torch.cuda.stream(cs):
        assert raw(cs) == 0
    cs.synchronize()
    want = C.clone()
    C.zero_()
    torch.cuda.synchronize()
    g1 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g1, stream=cs):
        assert raw(torch.cuda.current_stream()) == 0
    for _ in range(3):
        C.zero_()
        torch.cuda.synchronize()
        g1.replay()
        torch.cuda.synchronize()
        assert torch.equal(C, want)
    held = _ws_bytes(cs.cuda_stream)

    # a larger eager call grows the workspace; the graph's buffer is kept
    m2, k2 = 2048, 2048
    A2 = torch.randn(m2 * k2, device="cuda", generator=g)
    B2 = torch.randn(k2 * n, device="cuda", generator=g)
    C2 = torch.empty(m2 * n, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(cs):
        assert raw(cs, m2, n, k2, A2, B2, C2) == 0
    cs.synchronize()
    assert _ws_bytes(cs.cuda_stream) > held
    C.zero_()
    torch.cuda.synchronize()
    g1.replay()
    torch.cuda.synchronize()
    assert torch.equal(C, want)
    del g1, g0
    assert L.mbh_workspace_release(ctypes.c_void_p(cs.cuda_stream)) == 0
    assert _ws_bytes(cs.cuda_stream) == 0


def test_trsv_status_after_a_solve():
    L = _lib.lib()
    n = 1000
    rng = np.random.default_rng(6)
    a = np.tril(rng.uniform(-1, 1, (n, n)) / n) + np.diag(rng.uniform(1, 2, n))
    x = rng.uniform(-1, 1, n)
    A = torch.from_numpy(np.asfortranarray(a).ravel(order="F").astype(np.float32)).cuda()
    X = torch.from_numpy(x.astype(np.float32)).cuda()
    assert mb.trsv("L", "N", "N", n, A, n, X, 1) == 0
    failed = ctypes.c_int(-1)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.mbh_trsv_status(s, ctypes.byref(failed)) == 0 and failed.value == 0
    want = x.copy()
    assert oracle.dtrsv("L", "N", "N", n, np.asfortranarray(a.astype(np.float32).astype(np.float64)).ravel(order="F"),
                        n, want, 1) == 0
    assert np.abs(X.cpu().numpy() - want).max() < 1e-5
    other = mb.Stream()
    assert L.mbh_trsv_status(ctypes.c_void_p(other.handle), ctypes.byref(failed)) == 0 and failed.value == 0
    other.close()
