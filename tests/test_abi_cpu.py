"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares, and rejects invalid arguments with the reference's Netlib
info codes before any GPU call (so these run without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from mojoblas_amd import _lib
import mojoblas_amd as mb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(mbh_\w+)\s*\(", text))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.lib()
    declared = header_symbols()
    assert declared, "no declarations found"
    for name in sorted(declared):
        assert hasattr(L, name), "missing export " + name
    # and the Python binding knows the signature of each
    assert declared <= set(_lib.SIGNATURES), sorted(declared - set(_lib.SIGNATURES))


def test_version_string():
    assert b"gfx950" in _lib.lib().mbh_version()


# (args, info) for gemm, gemm.mojo:179-213.
GEMM_INFO = [
    ((b"X", b"N", 2, 2, 2, 2, 2, 2), 1),
    ((b"N", b"\x00", 2, 2, 2, 2, 2, 2), 2),
    ((b"N", b"N", -1, 2, 2, 2, 2, 2), 3),
    ((b"N", b"N", 2, -1, 2, 2, 2, 2), 4),
    ((b"N", b"N", 2, 2, -1, 2, 2, 2), 5),
    ((b"N", b"N", 2, 2, 2, 1, 2, 2), 8),
    ((b"t", b"N", 2, 2, 3, 2, 3, 2), 8),
    ((b"N", b"N", 2, 2, 2, 2, 1, 2), 10),
    ((b"N", b"c", 2, 3, 2, 2, 2, 2), 10),
    ((b"N", b"N", 2, 2, 2, 2, 2, 1), 13),
]


@pytest.mark.parametrize("args,info", GEMM_INFO)
@pytest.mark.parametrize("entry", ["host_s", "host_d", "dev_s", "dev_d", "host_h", "host_b", "dev_h", "dev_b"])
def test_gemm_info_codes_without_gpu(args, info, entry):
    L = _lib.lib()
    ta, tb, m, n, k, lda, ldb, ldc = args
    dt = {"s": np.float32, "d": np.float64, "h": np.uint16, "b": np.uint16}[entry[-1]]
    a = np.full(16, 7, dt)
    b = np.full(16, 7, dt)
    c = np.full(16, 7, dt)
    P = lambda x: ctypes.c_void_p(x.ctypes.data)
    if entry == "host_s":
        rc = L.mbh_sgemm_host(ta, tb, m, n, k, 1.0, P(a), lda, P(b), ldb, 0.0, P(c), ldc)
    elif entry == "host_d":
        rc = L.mbh_dgemm_host(ta, tb, m, n, k, 1.0, P(a), lda, P(b), ldb, 0.0, P(c), ldc)
    elif entry == "dev_s":
        rc = L.mbh_sgemm_dev(ta, tb, m, n, k, 1.0, P(a), lda, P(b), ldb, 0.0, P(c), ldc, 0, None)
    elif entry == "dev_d":
        rc = L.mbh_dgemm_dev(ta, tb, m, n, k, 1.0, P(a), lda, P(b), ldb, 0.0, P(c), ldc, None)
    elif entry.startswith("host"):
        fn = L.mbh_hgemm_host if entry == "host_h" else L.mbh_bgemm_host
        rc = fn(ta, tb, m, n, k, 1.0, P(a), lda, P(b), ldb, 0.0, P(c), ldc)
    else:
        fn = L.mbh_hgemm_dev if entry == "dev_h" else L.mbh_bgemm_dev
        rc = fn(ta, tb, m, n, k, 1.0, P(a), lda, P(b), ldb, 0.0, P(c), ldc, None)
    assert rc == info
    assert (c == 7).all() and (a == 7).all() and (b == 7).all()


def test_python_mirror_prints_info_like_the_reference(capsys):
    a = np.zeros(4, np.float32)
    assert mb.gemm("No", "N", 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a.copy(), 2) == 1
    assert "gemm: Info 1" in capsys.readouterr().out
    assert mb.gemm("N", "N", 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a.copy(), 1) == 13
    assert "gemm: Info 13" in capsys.readouterr().out


def test_python_gemm_16bit_dtype_routing_without_gpu(capsys):
    # numpy float16 host arrays go to the hgemm entry: its info codes come
    # back before any GPU call; precision= is refused for 16-bit data.
    a = np.zeros(4, np.float16)
    assert mb.gemm("X", "N", 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a.copy(), 2) == 1
    assert "gemm: Info 1" in capsys.readouterr().out
    with pytest.raises(ValueError):
        mb.gemm("N", "N", 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a.copy(), 2, precision="bf16x3")
    with pytest.raises(TypeError):
        mb.gemm("N", "N", 2, 2, 2, 1.0, a, 2, np.zeros(4, np.float32), 2, 0.0, a.copy(), 2)
    with pytest.raises(TypeError):  # the other level-3 routines stay fp32 / fp64
        mb.syrk("L", "N", 2, 2, 1.0, a, 2, 0.0, a.copy(), 2)


def test_zero_dimension_quick_return_touches_nothing():
    # m == 0 returns before any GPU call (gemm.mojo:215-216)
    c = np.full(4, 7, np.float32)
    assert mb.gemm("N", "N", 0, 2, 2, 1.0, c, 1, c, 2, 0.0, c, 1) == 0
    assert (c == 7).all()


def test_invalid_precision_rejected():
    L = _lib.lib()
    a = np.zeros(4, np.float32)
    rc = L.mbh_sgemm_dev(b"N", b"N", 2, 2, 2, 1.0, ctypes.c_void_p(a.ctypes.data), 2,
                         ctypes.c_void_p(a.ctypes.data), 2, 0.0, ctypes.c_void_p(a.ctypes.data), 2, 99, None)
    assert rc == -3
    assert "precision" in _lib.last_error()


def test_batched_plan_validates_every_problem_without_gpu():
    from mojoblas_amd import batched

    with pytest.raises(mb.HipBlasError, match="problem 1 fails gemm argument 8"):
        batched.VBatchPlan([4, 4], [4, 4], [4, 4], lda=[4, 3])
    with pytest.raises(mb.HipBlasError, match="problem 0 fails gemm argument 3"):
        batched.VBatchPlan([-1], [4], [4])
    with pytest.raises(mb.HipBlasError, match="precision"):
        batched.VBatchPlan([4], [4], [4], precision="bf16x3")
    with pytest.raises(mb.HipBlasError, match="classes are 0 and 1"):
        batched.set_tuning([2] * 36)


# (args, info) for gemv, gemv.mojo:76-91 (trans uppercase only).
GEMV_INFO = [
    ((b"n", 2, 2, 2, 1, 1), 1),
    ((b"\x00", 2, 2, 2, 1, 1), 1),
    ((b"N", -1, 2, 2, 1, 1), 2),
    ((b"T", 2, -1, 2, 1, 1), 3),
    ((b"C", 3, 2, 2, 1, 1), 6),
    ((b"N", 2, 2, 2, 0, 1), 8),
    ((b"N", 2, 2, 2, -1, 0), 11),
]


@pytest.mark.parametrize("args,info", GEMV_INFO)
@pytest.mark.parametrize("entry", ["host_s", "host_d", "dev_s", "dev_d"])
def test_gemv_info_codes_without_gpu(args, info, entry):
    L = _lib.lib()
    trans, m, n, lda, incx, incy = args
    dt = np.float32 if entry.endswith("s") else np.float64
    a, x, y = np.full(16, 7, dt), np.full(16, 7, dt), np.full(16, 7, dt)
    P = lambda v: ctypes.c_void_p(v.ctypes.data)
    fn = getattr(L, "mbh_%sgemv_%s" % (entry[-1], entry[:-2]))
    extra = [None] if entry.startswith("dev") else []
    assert fn(trans, m, n, 1.0, P(a), lda, P(x), incx, 0.0, P(y), incy, *extra) == info
    assert (y == 7).all()


# (args, info) for trsv, trsv.mojo:57-80 (both cases accepted).
TRSV_INFO = [
    ((b"X", b"N", b"N", 2, 2, 1), 1),
    ((b"u", b"X", b"n", 2, 2, 1), 2),
    ((b"U", b"t", b"X", 2, 2, 1), 3),
    ((b"L", b"N", b"N", -1, 2, 1), 4),
    ((b"l", b"c", b"u", 3, 2, 1), 6),
    ((b"U", b"N", b"N", 2, 2, 0), 8),
]


@pytest.mark.parametrize("args,info", TRSV_INFO)
@pytest.mark.parametrize("entry", ["host_s", "host_d", "dev_s", "dev_d"])
def test_trsv_info_codes_without_gpu(args, info, entry):
    L = _lib.lib()
    uplo, trans, diag, n, lda, incx = args
    dt = np.float32 if entry.endswith("s") else np.float64
    a, x = np.full(16, 7, dt), np.full(16, 7, dt)
    P = lambda v: ctypes.c_void_p(v.ctypes.data)
    fn = getattr(L, "mbh_%strsv_%s" % (entry[-1], entry[:-2]))
    extra = [None] if entry.startswith("dev") else []
    assert fn(uplo, trans, diag, n, P(a), lda, P(x), incx, *extra) == info
    assert (x == 7).all()


def test_level2_python_mirror_prints_info(capsys):
    a = np.zeros(4, np.float32)
    assert mb.gemv("n", 2, 2, 1.0, a, 2, a, 1, 0.0, a.copy(), 1) == 1
    assert "gemv: Info 1" in capsys.readouterr().out
    assert mb.trsv("U", "N", "N", 2, a, 1, a.copy(), 1) == 6
    assert "trsv: Info 6" in capsys.readouterr().out


# ---- device-resident buffer entries: argument checks run before any HIP call

E_ARG = -4


def test_buffer_entries_validate_without_gpu():
    L = _lib.lib()
    p = ctypes.c_void_p(123)
    assert L.mbh_dev_alloc(0, ctypes.byref(p)) == 0 and p.value is None  # zero bytes: NULL, no allocation
    assert L.mbh_host_alloc(0, ctypes.byref(p)) == 0 and p.value is None
    assert L.mbh_dev_alloc(16, None) == E_ARG
    assert L.mbh_host_alloc(16, None) == E_ARG
    assert L.mbh_stream_create(None) == E_ARG
    assert L.mbh_dev_free(None) == 0 and L.mbh_host_free(None) == 0 and L.mbh_stream_destroy(None) == 0
    host = np.zeros(16, np.float32)
    hp, dp = host.ctypes.data, ctypes.c_void_p(0x1000)  # dp is never dereferenced: every case fails validation
    for fn in (L.mbh_set_matrix, L.mbh_get_matrix):
        assert fn(-1, 2, 4, hp, 4, dp, 4, None) == E_ARG          # negative rows
        assert fn(2, -1, 4, hp, 4, dp, 4, None) == E_ARG          # negative cols
        assert fn(2, 2, 0, hp, 4, dp, 4, None) == E_ARG           # element size
        assert fn(2, 2, 17, hp, 4, dp, 4, None) == E_ARG
        assert fn(4, 2, 4, hp, 3, dp, 4, None) == E_ARG           # source ld below rows
        assert fn(4, 2, 4, hp, 4, dp, 3, None) == E_ARG           # destination ld below rows
        assert fn(4, 2, 4, None, 4, dp, 4, None) == E_ARG         # null buffer
        assert fn(0, 2, 4, None, 1, None, 1, None) == 0           # empty: quick return
    assert L.mbh_set_matrix(4, 2, 4, hp, 3, dp, 4, None) == E_ARG
    assert "leading dimensions" in _lib.last_error()


def test_device_matrix_rejects_bad_shapes_without_gpu():
    with pytest.raises(TypeError):
        mb.DeviceMatrix(2, 2, np.int32)
    with pytest.raises(TypeError):
        mb.DeviceMatrix(2, 2, "float8")
    with pytest.raises(ValueError):
        mb.DeviceMatrix(4, 2, np.float32, ld=3)


def test_thread_precision_validated_and_thread_local_without_gpu():
    import threading

    L = _lib.lib()
    assert L.mbh_set_thread_precision(-1) == -1  # nothing set on this thread yet
    assert L.mbh_set_thread_precision(7) == -3  # MBH_E_PRECISION, setting unchanged
    assert L.mbh_set_thread_precision(3) == -3  # bf16 is the batched gemm's mode only
    assert L.mbh_set_thread_precision(1) == -1
    seen = []
    t = threading.Thread(target=lambda: seen.append(L.mbh_set_thread_precision(2)))
    t.start()
    t.join()
    assert seen == [-1]  # another thread starts with no setting of its own
    assert L.mbh_set_thread_precision(-1) == 1
    assert L.mbh_get_default_precision() == 0  # the library default is untouched


def test_python_rejects_bf16_precision_outside_batched_without_gpu():
    a = np.zeros(4, np.float32)
    for call in (lambda: mb.gemm("N", "N", 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a.copy(), 2, precision="bf16"),
                 lambda: mb.trsm("L", "L", "N", "N", 2, 2, 1.0, a, 2, a.copy(), 2, precision="bf16"),
                 lambda: mb.syrk("L", "N", 2, 2, 1.0, a, 2, 0.0, a.copy(), 2, precision=3),
                 lambda: mb.gemm("N", "N", 2, 2, 2, 1.0, a, 2, a, 2, 0.0, a.copy(), 2, precision="tf32")):
        with pytest.raises(ValueError):
            call()


@pytest.mark.parametrize("routine", ["gemm_a", "gemm_b", "gemm_c", "gemm_t", "trsm", "trmm", "syrk", "syr2k",
                                     "symm", "gemv_x", "gemv_y", "trsv"])
def test_python_rejects_short_buffers_without_gpu(routine):
    """A buffer shorter than the storage a call reaches is refused before the
    C-ABI sees it (host entries would otherwise copy past the array's end)."""
    f = np.float32
    ok, short = np.zeros(64, f), np.zeros(5, f)
    calls = {
        "gemm_a": lambda: mb.gemm("N", "N", 4, 4, 4, 1.0, short, 4, ok, 4, 0.0, ok.copy(), 4),
        "gemm_b": lambda: mb.gemm("N", "T", 4, 4, 4, 1.0, ok, 4, short, 4, 0.0, ok.copy(), 4),
        "gemm_c": lambda: mb.gemm("N", "N", 4, 4, 4, 1.0, ok, 4, ok, 4, 0.0, short, 4),
        "gemm_t": lambda: mb.gemm("T", "N", 3, 4, 5, 1.0, np.zeros(14, f), 5, ok, 5, 0.0, ok.copy(), 3),
        "trsm": lambda: mb.trsm("L", "L", "N", "N", 4, 4, 1.0, ok, 4, short, 4),
        "trmm": lambda: mb.trmm("R", "U", "N", "N", 2, 4, 1.0, short, 4, ok.copy(), 2),
        "syrk": lambda: mb.syrk("L", "T", 4, 4, 1.0, short, 4, 0.0, ok.copy(), 4),
        "syr2k": lambda: mb.syr2k("L", "N", 4, 4, 1.0, ok, 4, short, 4, 0.0, ok.copy(), 4),
        "symm": lambda: mb.symm("L", "U", 4, 4, 1.0, ok, 4, ok, 4, 1.0, short, 4),
        "gemv_x": lambda: mb.gemv("N", 4, 4, 1.0, ok, 4, np.zeros(4, f), 2, 0.0, ok.copy(), 1),
        "gemv_y": lambda: mb.gemv("T", 4, 4, 1.0, ok, 4, ok, 1, 0.0, np.zeros(3, f), 1),
        "trsv": lambda: mb.trsv("U", "N", "N", 4, ok, 5, np.zeros(3, f), 1),
    }
    with pytest.raises(ValueError, match="holds"):
        calls[routine]()


def test_python_length_check_spares_unread_operands_without_gpu(capsys):
    """alpha == 0 does not read A or B (gemm.mojo:215-237); invalid arguments
    are the C-ABI's to report as info codes, not the length check's."""
    c = np.full(4, 7, np.float32)
    assert mb.gemm("N", "N", 0, 2, 2, 1.0, np.zeros(1, np.float32), 1, np.zeros(1, np.float32), 2, 0.0, c, 1) == 0
    assert mb.gemm("N", "N", 2, 2, 2, 1.0, np.zeros(1, np.float32), 1, c, 2, 0.0, c, 2) == 8
    assert "gemm: Info 8" in capsys.readouterr().out
    assert (c == 7).all()


def test_noalloc_entry_validates_before_any_gpu_work():
    L = _lib.lib()
    p = ctypes.c_void_p(0x1000)  # never dereferenced: every case returns first
    args = lambda m, n, k, lda, prec: (b"N", b"N", m, n, k, 1.0, p, lda, p, k, 0.0, p, max(m, 1), prec, None)  # noqa: E731
    assert L.mbh_sgemm_dev_noalloc(*args(4, 4, 4, 3, 1)) == 8  # gemm info codes first
    assert L.mbh_sgemm_dev_noalloc(*args(4, 4, 4, 4, 9)) == -3  # then the precision
    assert L.mbh_sgemm_dev_noalloc(*args(0, 4, 4, 1, 1)) == 0  # quick return
    assert L.mbh_sgemm_dev_noalloc(*args(4096, 4096, 4096, 4096, 1)) == E_ARG  # planes beyond the scratch
    assert "pack scratch" in _lib.last_error()
