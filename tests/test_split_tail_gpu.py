"""Partial last round of the split-precision gemm: when the 256 x 256 tiles
leave a last round at most half full (17 x 17 = 289 tiles, or the 528-tile
block triangle of an n = 8192 syrk, on 256 CUs), the tiles of that round
run split along k with a deterministic reduction.  Checked against the
oracle on entries concentrated in the last tile-row and column (where those
tiles lie) and spread over the rest; results are bit-identical over two
runs, and alpha is applied exactly once.  The syrk / syr2k cases also pack
their shared operand once (op(A) == op(B)^T).

Bound: |C - C64| <= (3*2^-16 + (2k+2)u) |alpha| |A||B| + 4u|beta C|.
"""
import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
import mojoblas_amd as mb

pytestmark = pytest.mark.gpu


def _entries(rng, m, n, lo_m, lo_n, count=4096):
    """count random entries, plus count in rows >= lo_m and count in columns >= lo_n."""
    rows = np.concatenate([rng.integers(0, m, count), rng.integers(lo_m, m, count), rng.integers(0, m, count)])
    cols = np.concatenate([rng.integers(0, n, count), rng.integers(0, n, count), rng.integers(lo_n, n, count)])
    return rows.astype(np.int64), cols.astype(np.int64)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_gemm_with_partial_last_round(ta, tb):
    m = n = 17 * 256
    k = 2048 + 96
    g = torch.Generator(device="cuda").manual_seed(41)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A = torch.randn(ar * ac, device="cuda", generator=g)
    B = torch.randn(br * bc, device="cuda", generator=g)
    C0 = torch.randn(m * n, device="cuda", generator=g)
    C1, C2, C3 = C0.clone(), C0.clone(), torch.empty_like(C0)
    assert mb.gemm(ta, tb, m, n, k, 1.5, A, ar, B, br, 0.5, C1, m, precision="bf16x3") == 0
    assert mb.gemm(ta, tb, m, n, k, 1.5, A, ar, B, br, 0.5, C2, m, precision="bf16x3") == 0
    assert mb.gemm(ta, tb, m, n, k, -3.0, A, ar, B, br, 0.0, C3, m, precision="bf16x3") == 0
    C4 = torch.empty_like(C0)
    assert mb.gemm(ta, tb, m, n, k, 1.0, A, ar, B, br, 0.0, C4, m, precision="bf16x3") == 0
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)  # deterministic
    assert torch.equal(C3, -3.0 * C4)  # alpha applied once, after the reduction
    rng = np.random.default_rng(42)
    rows, cols = _entries(rng, m, n, 16 * 256, 15 * 256)
    a, b, c0 = A.cpu().numpy(), B.cpu().numpy(), C0.cpu().numpy()
    want = oracle.sgemm_sample(ta, tb, k, 1.5, a, ar, b, br, 0.5, c0[rows + cols * m], rows, cols)
    absab = oracle.sgemm_sample_abs(ta, tb, k, a, ar, b, br, rows, cols)
    got = C1.cpu().numpy()[rows + cols * m].astype(np.float64)
    bound = H.gemm_bound("bf16x3", k, 1.5, absab, 0.5, c0[rows + cols * m].astype(np.float64))
    err = np.abs(got - want)
    assert (err <= bound).all(), float((err / bound).max())


@pytest.mark.parametrize("routine", ["syrk", "syr2k"])
@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "T")])
def test_syrk_block_triangle_with_partial_last_round(routine, uplo, trans):
    n, k = 8192, 1024
    g = torch.Generator(device="cuda").manual_seed(43)
    ar, ac = (n, k) if trans == "N" else (k, n)
    A = torch.randn(ar * ac, device="cuda", generator=g)
    B = torch.randn(ar * ac, device="cuda", generator=g)
    C0 = torch.randn(n * n, device="cuda", generator=g)
    rng = np.random.default_rng(44)
    i, j = _entries(rng, n, n, n - 256, n - 4096)
    lower = uplo == "L"
    rows, cols = (np.maximum(i, j), np.minimum(i, j)) if lower else (np.minimum(i, j), np.maximum(i, j))
    if not lower:  # the upper triangle's last tiles: last tile-column, rows from 4096
        rows2 = rng.integers(n - 4096, n, 4096)
        cols2 = rng.integers(n - 256, n, 4096)
        rows, cols = np.concatenate([rows, np.minimum(rows2, cols2)]), np.concatenate([cols, np.maximum(rows2, cols2)])
    other = (rows < cols) if lower else (rows > cols)
    assert not other.any()
    outs = []
    for _ in range(2):
        C = C0.clone()
        if routine == "syrk":
            assert mb.syrk(uplo, trans, n, k, 0.75, A, ar, -0.5, C, n, precision="bf16x3") == 0
        else:
            assert mb.syr2k(uplo, trans, n, k, 0.75, A, ar, B, ar, -0.5, C, n, precision="bf16x3") == 0
        outs.append(C)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1])
    a, b, c0 = A.cpu().numpy(), B.cpu().numpy(), C0.cpu().numpy()
    c = outs[0].cpu().numpy()
    ta, tb = ("N", "T") if trans == "N" else ("T", "N")
    cs = c0[rows + cols * n].astype(np.float64)
    if routine == "syrk":
        want = oracle.sgemm_sample(ta, tb, k, 0.75, a, ar, a, ar, -0.5, c0[rows + cols * n], rows, cols)
        absab = oracle.sgemm_sample_abs(ta, tb, k, a, ar, a, ar, rows, cols)
    else:
        want = (oracle.sgemm_sample(ta, tb, k, 0.75, a, ar, b, ar, -0.5, c0[rows + cols * n], rows, cols) +
                oracle.sgemm_sample(ta, tb, k, 0.75, b, ar, a, ar, 0.0, None, rows, cols))
        absab = (oracle.sgemm_sample_abs(ta, tb, k, a, ar, b, ar, rows, cols) +
                 oracle.sgemm_sample_abs(ta, tb, k, b, ar, a, ar, rows, cols))
    got = c[rows + cols * n].astype(np.float64)
    bound = H.gemm_bound("bf16x3", k, 0.75, absab, -0.5, cs)
    if routine == "syr2k":  # two products, each rounded, the second added with beta = 1
        bound = 2 * bound + 4 * H.U32 * np.abs(got)
    err = np.abs(got - want)
    assert (err <= bound).all(), float((err / bound).max())
    # the other triangle is untouched
    orow, ocol = cols[rows != cols], rows[rows != cols]
    assert np.array_equal(c[orow + ocol * n], c0[orow + ocol * n])
