"""GPU parity of the level-2 gemv / trsv (csrc/level2.hip) through the C-ABI,
against the FP64 oracle (orc_dgemv / orc_dtrsv, pinned by
tests/test_oracle_level2.py) on seeded inputs.

gemv bound, componentwise: |y - y64| <= (lenx + 2) u |alpha| |op(A)| |x| + 2 u |beta y|
(u = 2^-24 for float32, 2^-53 for float64; the sums are in the data type).
trsv: relative max error against the oracle <= 2e-6 (float32, one fp64-residual
refinement step) / 1e-12 (float64) on well-conditioned triangles (diagonal in
[1, 2], off-diagonal U[-1,1]/n), as for trsm.
"""
import itertools
import json
import os

import numpy as np
import pytest

import oracle

torch = pytest.importorskip("torch")
import mojoblas_amd as mb

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "level2_kat.json")))
U = {np.float32: 2.0 ** -24, np.float64: 2.0 ** -53}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("case", KAT["tests"], ids=[t["name"] for t in KAT["tests"]])
@pytest.mark.parametrize("entry", ["dev", "host"])
def test_known_answers(case, entry):
    dt = np.float32 if case["dtype"] == "float32" else np.float64
    arrays = {k: np.array(v, dtype=dt) for k, v in case["arrays"].items()}
    bufs = {k: (dev(v) if entry == "dev" else v) for k, v in arrays.items()}
    args = [bufs[a["array"]] if isinstance(a, dict) else a for a in case["args"]]
    rc = getattr(mb, case["routine"])(*args)
    assert rc == 0
    torch.cuda.synchronize()
    for e in case["expects"]:
        buf = bufs[e["array"]]
        got = float(buf.cpu().numpy()[e["index"]] if entry == "dev" else buf[e["index"]])
        assert abs(got - e["value"]) <= e["atol"] + e["rtol"] * abs(e["value"]), (case["name"], e, got)


def strided(rng, length, inc, dt, gap=np.nan):
    buf = np.full(abs(inc) * max(length - 1, 0) + 1, gap)
    pos = [i * inc if inc > 0 else (length - 1 - i) * (-inc) for i in range(length)]
    buf[pos] = rng.uniform(-1, 1, length)
    return buf.astype(dt), np.array(pos, np.int64)


GEMV_SIZES = [(1, 1), (5, 3), (257, 130), (1000, 70), (40, 3000)]


@pytest.mark.parametrize("trans", ["N", "T", "C"])
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_gemv_vs_oracle(trans, dt):
    rng = np.random.default_rng(60)
    for (m, n), (incx, incy) in itertools.product(GEMV_SIZES, [(1, 1), (2, -1), (-3, 2)]):
        alpha, beta = [(1.0, 0.0), (1.5, 0.5), (-0.75, 1.0), (2.0, -1.25)][(m + n + incx) % 4]
        lda = m + 3
        abuf = np.full(lda * n, np.nan)
        A = rng.uniform(-1, 1, (m, n))
        abuf.reshape(n, lda).T[:m] = A
        abuf = abuf.astype(dt)
        lx, ly = (n, m) if trans == "N" else (m, n)
        xbuf, xpos = strided(rng, lx, incx, dt)
        ybuf, ypos = strided(rng, ly, incy, dt)
        if beta == 0.0:
            ybuf[ypos] = np.nan  # must not be read
        want = np.nan_to_num(ybuf.astype(np.float64), nan=0.0) if beta == 0.0 else ybuf.astype(np.float64)
        want = want.copy()
        assert oracle.dgemv(trans, m, n, alpha, abuf.astype(np.float64), lda, xbuf.astype(np.float64), incx, beta,
                            want, incy) == 0
        Y = dev(ybuf)
        assert mb.gemv(trans, m, n, alpha, dev(abuf), lda, dev(xbuf), incx, beta, Y, incy) == 0
        got = Y.cpu().numpy().astype(np.float64)
        Aop = np.abs(A if trans == "N" else A.T)
        absax = Aop @ np.abs(xbuf[xpos].astype(np.float64))
        yin = 0.0 if beta == 0.0 else np.abs(ybuf[ypos].astype(np.float64))
        bound = (lx + 2) * U[dt] * abs(alpha) * absax + 2 * U[dt] * abs(beta) * yin + 1e-300
        err = np.abs(got[ypos] - want[ypos])
        assert (err <= bound).all(), (trans, m, n, incx, incy, float((err / bound).max()))
        gaps = np.setdiff1d(np.arange(ybuf.size), ypos)
        assert np.isnan(got[gaps]).all(), "strided y gaps were written"


def test_gemv_host_entry_matches_device():
    rng = np.random.default_rng(61)
    m, n = 300, 200
    A = rng.uniform(-1, 1, m * n).astype(np.float32)
    x = rng.uniform(-1, 1, 2 * n).astype(np.float32)
    y0 = rng.uniform(-1, 1, m).astype(np.float32)
    yh = y0.copy()
    assert mb.gemv("N", m, n, 1.25, A, m, x, 2, 0.5, yh, 1) == 0
    Y = dev(y0)
    assert mb.gemv("N", m, n, 1.25, dev(A), m, dev(x), 2, 0.5, Y, 1) == 0
    assert np.array_equal(Y.cpu().numpy(), yh)  # same kernels, deterministic


def test_gemv_alpha_zero_and_quick_returns():
    y = dev(np.full(8, np.nan, np.float32))
    a = dev(np.ones(16, np.float32))
    assert mb.gemv("N", 4, 4, 0.0, a, 4, a, 1, 0.0, y, 2) == 0  # y := 0 on the 4 strided slots
    got = y.cpu().numpy()
    assert (got[[0, 2, 4, 6]] == 0).all() and np.isnan(got[[1, 3, 5, 7]]).all()
    y2 = dev(np.full(4, np.nan, np.float32))
    assert mb.gemv("T", 0, 4, 1.0, a, 1, a, 1, 0.0, y2, 1) == 0  # m == 0: untouched
    assert np.isnan(y2.cpu().numpy()).all()


def tri_operand(rng, n, upper, dt, unit_junk=False):
    A = rng.uniform(-1, 1, (n, n)) / n
    A[np.diag_indices(n)] = rng.uniform(1, 2, n)
    store = np.triu(A) if upper else np.tril(A)
    store[np.tril_indices(n, -1) if upper else np.triu_indices(n, 1)] = np.nan
    if unit_junk:
        store[np.diag_indices(n)] = 1e30
    return store.astype(dt)


@pytest.mark.parametrize("uplo,trans,diag", list(itertools.product("UL", "NT", "NU")))
@pytest.mark.parametrize("dt", [np.float32, np.float64], ids=["f32", "f64"])
def test_trsv_vs_oracle(uplo, trans, diag, dt):
    rng = np.random.default_rng(62)
    for n, incx in [(1, 1), (130, 1), (300, -2), (1000, 3)]:
        store = tri_operand(rng, n, uplo == "U", dt, unit_junk=(diag == "U"))
        a = np.asfortranarray(store).ravel(order="F")
        xbuf, xpos = strided(rng, n, incx, dt)
        want = xbuf.astype(np.float64).copy()
        a64 = a.astype(np.float64)
        assert oracle.dtrsv(uplo, trans, diag, n, a64, n, want, incx) == 0
        X = dev(xbuf)
        assert mb.trsv(uplo, trans, diag, n, dev(a), n, X, incx) == 0
        got = X.cpu().numpy().astype(np.float64)
        tol = 2e-6 if dt == np.float32 else 1e-12
        err = np.abs(got[xpos] - want[xpos]).max() / max(np.abs(want[xpos]).max(), 1e-300)
        assert err <= tol, (uplo, trans, diag, n, incx, err)
        gaps = np.setdiff1d(np.arange(xbuf.size), xpos)
        assert np.isnan(got[gaps]).all()


def test_trsv_host_entry():
    rng = np.random.default_rng(63)
    n = 257
    a = np.asfortranarray(tri_operand(rng, n, False, np.float32)).ravel(order="F")
    x = rng.uniform(-1, 1, n).astype(np.float32)
    want = x.astype(np.float64).copy()
    oracle.dtrsv("L", "N", "N", n, a.astype(np.float64), n, want, 1)
    got = x.copy()
    assert mb.trsv("L", "N", "N", n, a, n, got, 1) == 0
    assert np.abs(got - want).max() / np.abs(want).max() <= 2e-6
