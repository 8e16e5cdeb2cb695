"""Device-resident buffers through the library's own entries
(mbh_dev_alloc / mbh_set_matrix / mbh_get_matrix / mbh_stream_*, wrapped by
mojoblas_amd.DeviceMatrix / Stream), used as operands of the _dev entries,
checked against the FP64 oracle (bounds as in test_gemm_gpu / test_level3_gpu)."""
import ctypes

import numpy as np
import pytest

import oracle
from tests import helpers as H

pytest.importorskip("torch")
import mojoblas_amd as mb
from mojoblas_amd import _lib

pytestmark = pytest.mark.gpu


def test_round_trip_with_padded_leading_dimension():
    rng = np.random.default_rng(80)
    for dt in (np.float32, np.float64):
        a = rng.uniform(-1, 1, (37, 11)).astype(dt)
        with mb.DeviceMatrix.from_numpy(a, ld=40) as A:
            assert A.ld == 40 and A.ptr != 0
            assert np.array_equal(A.to_numpy(), a)
            A.copy_from(2 * a)
            assert np.array_equal(A.to_numpy(), 2 * a)


def test_gemm_on_library_buffers_and_stream():
    rng = np.random.default_rng(81)
    m, n, k = 300, 200, 150
    a = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    c0 = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    want = c0.astype(np.float64).ravel(order="F").copy()
    assert oracle.dgemm("N", "N", m, n, k, 1.5, a.astype(np.float64).ravel(order="F"), m,
                        b.astype(np.float64).ravel(order="F"), k, -0.5, want, m) == 0
    absab = np.abs(a.astype(np.float64)) @ np.abs(b.astype(np.float64))
    with mb.Stream() as s, mb.DeviceMatrix.from_numpy(a, ld=m + 3) as A, \
            mb.DeviceMatrix.from_numpy(b) as B, mb.DeviceMatrix.from_numpy(c0, ld=m + 1) as C:
        for prec in ("fp32", "bf16x3"):
            C.copy_from(c0)
            assert mb.gemm("N", "N", m, n, k, 1.5, A, A.ld, B, B.ld, -0.5, C, C.ld, precision=prec,
                           stream=s.handle) == 0
            got = C.to_numpy(stream=s.handle).astype(np.float64)
            bound = H.gemm_bound(prec, k, 1.5, absab, -0.5, c0.astype(np.float64))
            err = np.abs(got - want.reshape(n, m).T)
            assert (err <= bound).all(), (prec, float((err / bound).max()))


def test_trsm_and_gemv_on_library_buffers():
    rng = np.random.default_rng(82)
    m, n = 260, 70
    t = rng.uniform(-1, 1, (m, m)) / m
    t[np.diag_indices(m)] = rng.uniform(1, 2, m)
    t = np.tril(t).astype(np.float32)
    b = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    want = b.astype(np.float64).ravel(order="F").copy()
    oracle.dtrsm("L", "L", "N", "N", m, n, 1.0, t.astype(np.float64).ravel(order="F"), m, want, m)
    with mb.DeviceMatrix.from_numpy(t) as T, mb.DeviceMatrix.from_numpy(b, ld=m + 5) as B:
        assert mb.trsm("L", "L", "N", "N", m, n, 1.0, T, T.ld, B, B.ld, stream=0) == 0
        got = B.to_numpy()
        x = want.reshape(n, m).T
        assert np.abs(got - x).max() / np.abs(x).max() <= 2e-6
        # gemv y := T x0 on the same buffers (x0 = the first solution column, n x 1 matrices)
        with mb.DeviceMatrix.from_numpy(got[:, :1]) as X, mb.DeviceMatrix(m, 1, np.float32) as Y:
            assert mb.gemv("N", m, m, 1.0, T, T.ld, X, 1, 0.0, Y, 1, stream=0) == 0
            y = Y.to_numpy()[:, 0].astype(np.float64)
            y64 = t.astype(np.float64) @ got[:, 0].astype(np.float64)
            bound = (m + 2) * H.U32 * (np.abs(t.astype(np.float64)) @ np.abs(got[:, 0].astype(np.float64))) + 1e-30
            assert (np.abs(y - y64) <= bound).all()


def test_pinned_host_buffer_round_trip():
    L = _lib.lib()
    rows, cols = 64, 9
    p = ctypes.c_void_p()
    assert L.mbh_host_alloc(rows * cols * 8, ctypes.byref(p)) == 0 and p.value
    try:
        host = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), shape=(rows * cols,))
        host[:] = np.arange(rows * cols, dtype=np.float64)
        with mb.DeviceMatrix(rows, cols, np.float64) as D:
            assert L.mbh_set_matrix(rows, cols, 8, p, rows, ctypes.c_void_p(D.ptr), D.ld, None) == 0
            assert L.mbh_stream_sync(None) == 0
            assert np.array_equal(D.to_numpy().ravel(order="F"), host)
    finally:
        assert L.mbh_host_free(p) == 0


@pytest.mark.parametrize("fmt", ["float16", "bfloat16"])
def test_16bit_gemm_on_library_buffers(fmt):
    """hgemm / bgemm with DeviceMatrix operands (padded ld) on a library
    stream: integer data, so every partial sum is exact in fp32 and C must
    equal the correctly rounded 16-bit result bit for bit."""
    rng = np.random.default_rng(84)
    m, n, k = 200, 136, 96
    av = rng.integers(-4, 5, (m, k)).astype(np.float64)
    bv = rng.integers(-4, 5, (k, n)).astype(np.float64)
    cv = rng.integers(-4, 5, (m, n)).astype(np.float64)
    if fmt == "float16":
        enc = lambda x: np.asarray(x, np.float16)
        bits = lambda x: np.asarray(x).view(np.uint16)
        dt = np.float16
    else:
        enc = lambda x: H.round_bf16(np.asarray(x, np.float32))
        bits = lambda x: np.asarray(x)
        dt = "bfloat16"
    want = bits(enc((-1.0 * (av @ bv) + 2.0 * cv).astype(np.float32)))
    with mb.Stream() as s, mb.DeviceMatrix.from_numpy(enc(av), ld=m + 8, dtype=dt) as A, \
            mb.DeviceMatrix.from_numpy(enc(bv), dtype=dt) as B, \
            mb.DeviceMatrix.from_numpy(enc(cv), ld=m + 3, dtype=dt) as C:
        assert mb.gemm("N", "N", m, n, k, -1.0, A, A.ld, B, B.ld, 2.0, C, C.ld, stream=s.handle) == 0
        got = bits(C.to_numpy(stream=s.handle))
    assert np.array_equal(got, want)
