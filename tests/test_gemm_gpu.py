"""GPU parity of gemm (all four transposes, alpha/beta, lda/ldb/ldc, edge
tiles) against the fp64 oracle, through the C-ABI (device and host entries).

Bound, componentwise: |C - C_oracle| <= (e_rep + (2k+2) u)|alpha| |A||B| + 4u|beta C|
with u = 2^-24 and e_rep = 0 (fp32) or 3*2^-16 (bf16x3, see helpers.rep_error).
FP8 (e4m3, per-128x128-tile scales): bit-exact on representable inputs and
within the per-element rounding model |x - qs| <= 2^-4|x| + 2^-10 s otherwise.
"""
import itertools
import json
import os

import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
import mojoblas_amd as mb

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "level3_kat.json")))


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def run_dev(ta, tb, m, n, k, alpha, abuf, lda, bbuf, ldb, beta, cbuf, ldc, prec):
    A, B, C = dev(abuf), dev(bbuf), dev(cbuf)
    rc = mb.gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, precision=prec)
    torch.cuda.synchronize()
    assert rc == 0
    return C.cpu().numpy()


@pytest.mark.parametrize("case", [t for t in KAT["tests"] if t["routine"] == "gemm"], ids=lambda t: t["name"])
@pytest.mark.parametrize("entry", ["dev", "host"])
def test_known_answers(case, entry):
    dt = np.float32 if case["dtype"] == "float32" else np.float64
    arrays = {k: np.array(v, dtype=dt) for k, v in case["arrays"].items()}
    if entry == "dev":
        arrays = {k: dev(v) for k, v in arrays.items()}
    args = [arrays[a["array"]] if isinstance(a, dict) else a for a in case["args"]]
    assert mb.gemm(*args) == 0
    torch.cuda.synchronize()
    for e in case["expects"]:
        got = float(arrays[e["array"]][e["index"]])
        assert abs(got - e["value"]) <= e["atol"] + e["rtol"] * abs(e["value"]), (case["name"], e, got)


SHAPES = [(1, 1, 1), (7, 13, 5), (128, 128, 16), (129, 65, 33), (64, 300, 257), (300, 257, 190)]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_gemm_vs_oracle(ta, tb, prec):
    rng = np.random.default_rng(11)
    for idx, (m, n, k) in enumerate(SHAPES):
        pad = idx % 3  # exercises odd leading dimensions (scalar load path)
        alpha, beta = [(1.0, 0.0), (1.5, 0.5), (-0.75, 1.0)][idx % 3]
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        lda, ldb, ldc = ar + pad, br + pad, m + pad
        abuf, amat = H.col_major(rng, ar, ac, lda)
        bbuf, bmat = H.col_major(rng, br, bc, ldb)
        cbuf, cmat = H.col_major(rng, m, n, ldc, fill=np.nan if beta == 0 else None)
        got = run_dev(ta, tb, m, n, k, alpha, abuf, lda, bbuf, ldb, beta, cbuf, ldc, prec)
        want = H.oracle_gemm(ta, tb, m, n, k, alpha, abuf, lda, bbuf, ldb, beta, np.nan_to_num(cbuf), ldc)
        bound = H.gemm_bound(prec, k, alpha, H.abs_product(ta, tb, amat, bmat), beta, np.nan_to_num(cmat))
        err = np.abs(H.logical(got, m, n, ldc).astype(np.float64) - want)
        assert (err <= bound).all(), (m, n, k, ta, tb, prec, float(err.max()), float((err / bound).max()))
        # padding rows of C untouched
        full = got[: ldc * n].reshape(n, ldc).T
        assert np.isnan(full[m:]).all()


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_config1_sgemm_512(prec):
    """SGEMM N,N 512^3, lda=ldb=ldc=512, alpha=1.5, beta=0.5, A/B/C U[-1,1]."""
    m = n = k = 512
    a = oracle.fill_f32(m * k, 101, "uniform")
    b = oracle.fill_f32(k * n, 102, "uniform")
    c = oracle.fill_f32(m * n, 103, "uniform")
    got = run_dev("N", "N", m, n, k, 1.5, a, m, b, k, 0.5, c, m, prec)
    want = H.oracle_gemm("N", "N", m, n, k, 1.5, a, m, b, k, 0.5, c, m)
    amat = a.reshape(k, m).T.astype(np.float64)
    bmat = b.reshape(n, k).T.astype(np.float64)
    bound = H.gemm_bound(prec, k, 1.5, np.abs(amat) @ np.abs(bmat), 0.5, c.reshape(n, m).T)
    err = np.abs(got.reshape(n, m).T - want)
    assert (err <= bound).all(), float((err / bound).max())
    # the host-pointer entry gives the same result bit for bit
    ch = c.copy()
    assert mb.gemm("N", "N", m, n, k, 1.5, a, m, b, k, 0.5, ch, m, precision=prec) == 0
    assert np.array_equal(ch, got)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_deterministic_and_beta_zero_does_not_read_c(prec):
    rng = np.random.default_rng(12)
    m, n, k = 257, 130, 300
    abuf, _ = H.col_major(rng, m, k, m)
    bbuf, _ = H.col_major(rng, k, n, k)
    c1 = run_dev("N", "N", m, n, k, 1.0, abuf, m, bbuf, k, 0.0, np.full(m * n, np.nan, np.float32), m, prec)
    c2 = run_dev("N", "N", m, n, k, 1.0, abuf, m, bbuf, k, 0.0, np.zeros(m * n, np.float32), m, prec)
    assert np.isfinite(c1).all()
    assert np.array_equal(c1, c2)


def test_alpha_zero_scales_only():
    m, n, k = 33, 17, 9
    a = np.full(m * k, np.nan, np.float32)  # not read when alpha == 0
    c = np.arange(m * n, dtype=np.float32)
    got = run_dev("N", "N", m, n, k, 0.0, a, m, a, k, 2.0, c, m, "fp32")
    assert np.array_equal(got, 2.0 * c)
    got = run_dev("N", "N", m, n, k, 0.0, a, m, a, k, 0.0, c, m, "fp32")
    assert (got == 0).all()


@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_dgemm_vs_oracle(ta, tb):
    rng = np.random.default_rng(13)
    for m, n, k in [(5, 3, 7), (64, 64, 64), (100, 77, 130)]:
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        abuf, amat = H.col_major(rng, ar, ac, ar + 1, dtype=np.float64)
        bbuf, bmat = H.col_major(rng, br, bc, br, dtype=np.float64)
        cbuf, cmat = H.col_major(rng, m, n, m, dtype=np.float64)
        got = run_dev(ta, tb, m, n, k, 1.25, abuf, ar + 1, bbuf, br, -0.5, cbuf, m, None)
        want = cbuf.copy()
        oracle.dgemm(ta, tb, m, n, k, 1.25, abuf, ar + 1, bbuf, br, -0.5, want, m)
        bound = 4 * k * 2.0 ** -53 * (1.25 * H.abs_product(ta, tb, amat, bmat) + 0.5 * np.abs(cmat)) + 1e-300
        assert (np.abs(got.reshape(n, m).T - want.reshape(n, m).T) <= bound).all()


@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
@pytest.mark.parametrize("k,pad", [(512, 0), (520, 0), (512, 1)])
def test_dgemm_large_sampled(ta, tb, k, pad):
    """2120 x 2088 fp64.  With k % 16 == 0 and 16-byte aligned operands the
    whole 128 x 128 tiles run the interior kernel and the 72-row / 40-column
    strips the generic one; k = 520 or an odd ld take the generic kernel
    throughout.  Checked on 4096 sampled entries plus every entry of the rows
    and columns next to the seams and the last ones, against fp64 dot
    products: |C - C_ref| <= 4k u64 (|alpha| |A||B| + |beta C|)."""
    m, n = 2048 + 72, 2048 + 40
    g = torch.Generator(device="cuda").manual_seed(31)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb = ar + pad, br + pad
    A = torch.randn(lda * ac, device="cuda", generator=g, dtype=torch.float64)
    B = torch.randn(ldb * bc, device="cuda", generator=g, dtype=torch.float64)
    C0 = torch.randn(m * n, device="cuda", generator=g, dtype=torch.float64)
    C = C0.clone()
    alpha, beta = 0.75, -1.5
    assert mb.gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, m) == 0
    torch.cuda.synchronize()
    a = A.cpu().numpy()[: lda * ac].reshape(ac, lda).T[:ar]
    b = B.cpu().numpy()[: ldb * bc].reshape(bc, ldb).T[:br]
    opa = a if ta == "N" else a.T
    opb = b if tb == "N" else b.T
    c0 = C0.cpu().numpy().reshape(n, m).T
    got = C.cpu().numpy().reshape(n, m).T
    rng = np.random.default_rng(32)
    seams = [0, 2047, 2048, m - 1]
    rows = np.concatenate([rng.integers(0, m, 4096), np.repeat(seams, n), np.tile(np.arange(m), 4)])
    cols = np.concatenate([rng.integers(0, n, 4096), np.tile(np.arange(n), 4), np.repeat([0, 2047, 2048, n - 1], m)])
    ra, cb = opa[rows], opb[:, cols].T
    want = alpha * np.einsum("ij,ij->i", ra, cb) + beta * c0[rows, cols]
    bound = 4 * k * 2.0 ** -53 * (abs(alpha) * np.einsum("ij,ij->i", np.abs(ra), np.abs(cb)) +
                                  abs(beta) * np.abs(c0[rows, cols])) + 1e-300
    err = np.abs(got[rows, cols] - want)
    assert (err <= bound).all(), float((err / bound).max())


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_large_sampled(ta, tb, prec):
    """2112 x 2080 x 2304 (ragged against the 128-tiles, 17 x 17 of them) on
    the vector path, priced on 4096 sampled entries by the oracle; plus
    linearity in alpha."""
    m, n, k = 2048 + 64, 2048 + 32, 2304
    g = torch.Generator(device="cuda").manual_seed(14)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A = torch.randn(ar * ac, device="cuda", generator=g)
    B = torch.randn(br * bc, device="cuda", generator=g)
    C = torch.empty(m * n, device="cuda")
    assert mb.gemm(ta, tb, m, n, k, 1.0, A, ar, B, br, 0.0, C, m, precision=prec) == 0
    C2 = torch.empty_like(C)
    assert mb.gemm(ta, tb, m, n, k, -2.0, A, ar, B, br, 0.0, C2, m, precision=prec) == 0
    torch.cuda.synchronize()
    assert torch.equal(C2, -2.0 * C)  # alpha applied exactly once, in the epilogue
    rng = np.random.default_rng(15)
    rows = rng.integers(0, m, 4096)
    cols = rng.integers(0, n, 4096)
    a, b, c = A.cpu().numpy(), B.cpu().numpy(), C.cpu().numpy()
    want = oracle.sgemm_sample(ta, tb, k, 1.0, a, ar, b, br, 0.0, None, rows, cols)
    absab = oracle.sgemm_sample_abs(ta, tb, k, a, ar, b, br, rows, cols)
    got = c[rows + cols * m].astype(np.float64)
    bound = H.gemm_bound(prec, k, 1.0, absab, 0.0, 0.0)
    assert (np.abs(got - want) <= bound).all(), float((np.abs(got - want) / bound).max())


def _tile_scaled_ints(rng, rows, cols, rows_tiles_first):
    """Integers in [-8, 8] times 2^e, e in [-2, 2] drawn per 128x128 tile of
    the (rows x cols) matrix: every value and every partial product sum is
    exact in e4m3 (after the per-tile power-of-two scale) and in fp32."""
    x = rng.integers(-8, 9, (rows, cols)).astype(np.float64)
    et = rng.integers(-2, 3, ((rows + 127) // 128, (cols + 127) // 128))
    e = np.repeat(np.repeat(et, 128, axis=0), 128, axis=1)[:rows, :cols]
    return x * np.exp2(e)


@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_fp8_exact_on_representable_inputs(ta, tb):
    """FP8 (e4m3, per-128x128-tile power-of-two scales) is bit-exact when every
    operand tile is representable: pins the per-tile scale lookup and layout."""
    rng = np.random.default_rng(16)
    m, n, k = 300, 260, 384
    opa = _tile_scaled_ints(rng, m, k, True)   # op(A), m x k: tiles over (row, k)
    opb = _tile_scaled_ints(rng, n, k, True).T  # op(B), k x n: tiles over (n, k)
    a = (opa if ta == "N" else opa.T).astype(np.float32)
    b = (opb if tb == "N" else opb.T).astype(np.float32)
    abuf = np.asfortranarray(a).ravel(order="F")
    bbuf = np.asfortranarray(b).ravel(order="F")
    got = run_dev(ta, tb, m, n, k, 1.0, abuf, a.shape[0], bbuf, b.shape[0], 0.0, np.zeros(m * n, np.float32), m,
                  "fp8")
    want = H.oracle_gemm(ta, tb, m, n, k, 1.0, abuf, a.shape[0], bbuf, b.shape[0], 0.0, np.zeros(m * n), m)
    assert np.array_equal(got.reshape(n, m).T.astype(np.float64), want)


def _fp8_bound(op_rows_by_k, k):
    """Per-element e4m3 error bound after per-128x128-tile scaling:
    |x - q s| <= 2^-4 |x| + 2^-10 s, s = 2^ceil(log2(amax_tile / 448))."""
    r = op_rows_by_k.shape[0]
    d = np.empty_like(op_rows_by_k)
    for i0 in range(0, r, 128):
        for k0 in range(0, k, 128):
            t = op_rows_by_k[i0:i0 + 128, k0:k0 + 128]
            amax = np.abs(t).max()
            s = 2.0 ** np.ceil(np.log2(amax / 448.0)) if amax > 0 else 1.0
            d[i0:i0 + 128, k0:k0 + 128] = 2.0 ** -4 * np.abs(t) + 2.0 ** -10 * s
    return d


@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
def test_fp8_within_rounding_bound(ta, tb):
    rng = np.random.default_rng(17)
    m, n, k = 257, 200, 300
    opa = rng.standard_normal((m, k))
    opb = rng.standard_normal((k, n))
    a = (opa if ta == "N" else opa.T).astype(np.float32)
    b = (opb if tb == "N" else opb.T).astype(np.float32)
    abuf = np.asfortranarray(a).ravel(order="F")
    bbuf = np.asfortranarray(b).ravel(order="F")
    got = run_dev(ta, tb, m, n, k, 1.0, abuf, a.shape[0], bbuf, b.shape[0], 0.0, np.zeros(m * n, np.float32), m,
                  "fp8")
    want = H.oracle_gemm(ta, tb, m, n, k, 1.0, abuf, a.shape[0], bbuf, b.shape[0], 0.0, np.zeros(m * n), m)
    oa = np.abs(opa.astype(np.float32).astype(np.float64))
    ob = np.abs(opb.astype(np.float32).astype(np.float64))
    da = _fp8_bound(oa, k)
    db = _fp8_bound(ob.T, k).T
    bound = da @ ob + oa @ db + da @ db + 2 * k * 2.0 ** -24 * (oa @ ob)
    err = np.abs(got.reshape(n, m).T - want)
    assert (err <= bound).all(), float((err / bound).max())
    # and it is a low-accuracy mode: clearly less accurate than bf16x3
    assert err.max() > 1e-3
