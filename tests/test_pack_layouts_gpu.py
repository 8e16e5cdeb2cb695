"""BF16x3 on the same op(B) stored either way: B k x n (not transposed, its
columns k-contiguous, packed by the direct path) or its transposed copy
(transposed, packed through the transposing path) must give bit-identical
C, since both pack passes write the same hi / lo planes; checked on 256 x 256
tiles with a partial last round, padded ld and clamped edge rows, and against
the oracle's BF16x3 bound on sampled entries."""
import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
import mojoblas_amd as mb

pytestmark = pytest.mark.gpu

CASES = [  # m, n, k, ldb pad, alpha, beta, transa
    (4096, 4096, 1024, 0, 1.0, 0.0, "N"),
    (4352, 4096 + 96, 2048, 4, 1.25, -0.5, "N"),   # partial last round, rows past n clamped
    (4096, 4096, 1536, 8, -0.75, 1.0, "T"),
]


@pytest.mark.parametrize("m,n,k,pad,alpha,beta,ta", CASES)
def test_both_layouts_of_b_agree_bitwise(m, n, k, pad, alpha, beta, ta):
    g = torch.Generator(device="cuda").manual_seed(51)
    ar, ac = (m, k) if ta == "N" else (k, m)
    ldb = k + pad
    A = torch.randn(ar * ac, device="cuda", generator=g)
    Bfull = torch.randn(ldb * n, device="cuda", generator=g)        # B: k x n, ld = k + pad
    Bmat = Bfull.view(n, ldb)[:, :k]  # row j = column j of B (the n x k matrix B^T, row-major)
    BT = Bmat.t().contiguous().view(-1)  # B^T stored column-major (n x k, ld = n)
    C0 = torch.randn(m * n, device="cuda", generator=g)
    C1, C2 = C0.clone(), C0.clone()
    assert mb.gemm(ta, "N", m, n, k, alpha, A, ar, Bfull, ldb, beta, C1, m, precision="bf16x3") == 0
    assert mb.gemm(ta, "T", m, n, k, alpha, A, ar, BT, n, beta, C2, m, precision="bf16x3") == 0
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)
    rng = np.random.default_rng(52)
    rows = np.concatenate([rng.integers(0, m, 2048), rng.integers(m - 256, m, 1024)])
    cols = np.concatenate([rng.integers(0, n, 2048), rng.integers(n - 300, n, 1024)])
    a, b, c0 = A.cpu().numpy(), Bfull.cpu().numpy(), C0.cpu().numpy()
    want = oracle.sgemm_sample(ta, "N", k, alpha, a, ar, b, ldb, beta, c0[rows + cols * m], rows, cols)
    absab = oracle.sgemm_sample_abs(ta, "N", k, a, ar, b, ldb, rows, cols)
    got = C1.cpu().numpy()[rows + cols * m].astype(np.float64)
    bound = H.gemm_bound("bf16x3", k, alpha, absab, beta, c0[rows + cols * m].astype(np.float64))
    assert (np.abs(got - want) <= bound).all()


def test_misaligned_or_ragged_k_layouts_agree():
    """The same with B's base off the 16-byte grid (scalar loads) or k not a
    multiple of 32 (zero-padded last k-tile)."""
    m, n = 4096, 4096
    g = torch.Generator(device="cuda").manual_seed(53)
    for k, off in ((1000, 0), (1024, 1)):
        A = torch.randn(m * k, device="cuda", generator=g)
        Bbuf = torch.randn(k * n + 4, device="cuda", generator=g)
        B = Bbuf[off: off + k * n]
        C1 = torch.empty(m * n, device="cuda")
        C2 = torch.empty(m * n, device="cuda")
        assert mb.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C1, m, precision="bf16x3") == 0
        BT = B.view(n, k).t().contiguous().view(-1)
        assert mb.gemm("N", "T", m, n, k, 1.0, A, m, BT, n, 0.0, C2, m, precision="bf16x3") == 0
        torch.cuda.synchronize()
        assert torch.equal(C1, C2)
