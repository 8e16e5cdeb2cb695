"""Adjoints of gemm and trsm behind torch.autograd.Function (forward and
backward on the HIP kernels), checked against
  * an analytic fp64 CPU reference gradient (numpy / scipy, i.e. OpenBLAS):
      gemm  dA = dC B^T, dB = A^T dC
      trsm  dB = A^-T dX, dA = -tril(dB X^T)  (and the other sides/triangles)
  * fp64 central finite differences of f = <W, y(A, B)> along 64 random
    directions, step h = 1e-6, f evaluated in fp64 (torch on the GPU).
Tolerances: the gradients are fp32 (or BF16x3) products, so agreement is
stated relative to the magnitude sums |dC||B^T| etc.: componentwise
2(k+2)u + e_rep for the analytic reference, 1e-5 normwise for the FD
directional derivatives.
"""
import numpy as np
import pytest
import scipy.linalg as sl

from tests import helpers as H

torch = pytest.importorskip("torch")
from mojoblas_amd import autograd as ag

pytestmark = pytest.mark.gpu


def _gemm_case(m, n, k, prec, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.randn(m, k, device="cuda", generator=g).requires_grad_(True)
    B = torch.randn(k, n, device="cuda", generator=g).requires_grad_(True)
    W = torch.randn(m, n, device="cuda", generator=g)  # upstream gradient dC
    C = ag.gemm(A, B, alpha=1.0, precision=prec)
    C.backward(W)
    return A, B, W, C


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("shape", [(1, 1, 1), (33, 17, 65), (130, 257, 70), (2048, 2048, 2048)])
def test_gemm_backward_vs_fp64_reference(shape, prec):
    m, n, k = shape
    A, B, W, C = _gemm_case(m, n, k, prec, 50)
    a, b, w = (t.detach().double().cpu().numpy() for t in (A, B, W))
    c_ref = a @ b
    da_ref, db_ref = w @ b.T, a.T @ w
    e = H.rep_error(prec)
    for got, want, bound_mag, kk in ((C, c_ref, np.abs(a) @ np.abs(b), k),
                                     (A.grad, da_ref, np.abs(w) @ np.abs(b.T), n),
                                     (B.grad, db_ref, np.abs(a.T) @ np.abs(w), m)):
        err = np.abs(got.detach().double().cpu().numpy() - want)
        bound = (e + (2 * kk + 2) * H.U32) * bound_mag + 1e-30
        assert (err <= bound).all(), (shape, prec, float((err / bound).max()))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_gemm_backward_vs_fp64_finite_differences(prec):
    """Config size m = n = k = 2048: 64 random directions (V, U), h = 1e-6,
    |fd - <dA, V> - <dB, U>| <= 1e-5 sum(|dA||V| + |dB||U|) for both
    precisions.  The tolerance is statistical, not a worst case: a
    gradient entry's rounding error (fp32 accumulation; BF16x3 also about
    2^-17 |a b| per product) is a sum of k random-signed terms, about
    2^-17 sqrt(k) ~ 3e-4 here against entries of size ~sqrt(k) ~ 45, and the
    directional sum over 4M entries adds them with random signs (~1 in all,
    against a tolerance ~1.5e3); a wrong gradient (a transposed or dropped
    term) moves <dA, V> by about its own size (~1e5)."""
    m = n = k = 2048
    A, B, W, _ = _gemm_case(m, n, k, prec, 51)
    A64, B64, W64 = (t.detach().double() for t in (A, B, W))
    f = lambda a, b: (W64 * (a @ b)).sum()
    g = torch.Generator(device="cuda").manual_seed(52)
    h = 1e-6
    dA, dB = A.grad.double(), B.grad.double()
    for _ in range(64):
        V = torch.randn(m, k, device="cuda", dtype=torch.float64, generator=g)
        U = torch.randn(k, n, device="cuda", dtype=torch.float64, generator=g)
        fd = (f(A64 + h * V, B64 + h * U) - f(A64 - h * V, B64 - h * U)) / (2 * h)
        ours = (dA * V).sum() + (dB * U).sum()
        scale = (dA.abs() * V.abs()).sum() + (dB.abs() * U.abs()).sum()
        assert abs(float(fd - ours)) <= 1e-5 * float(scale), (prec, float(fd), float(ours), float(scale))


def _tri(n, lower, seed, unit=False):
    g = torch.Generator(device="cuda").manual_seed(seed)
    a = (torch.rand(n, n, device="cuda", generator=g) * 2 - 1) / n
    a = a + torch.diag(1 + torch.rand(n, device="cuda", generator=g))
    a = torch.tril(a) if lower else torch.triu(a)
    if unit:
        # the unit diagonal is implied; store junk there (never read)
        a = a - torch.diag(torch.diag(a)) + 7.0 * torch.eye(n, device="cuda")
    return a


def _trsm_ref(a, b, side, uplo, trans, diag, alpha, w):
    """fp64 CPU reference: X and the gradients of <w, X> (scipy / OpenBLAS)."""
    lower = uplo == "L"
    t = np.tril(a) if lower else np.triu(a)
    if diag == "U":
        t = t - np.diag(np.diag(t)) + np.eye(a.shape[0])
    op = t if trans == "N" else t.T
    if side == "L":
        x = alpha * sl.solve(op, b)
        g = sl.solve(op.T, w)         # op(A)^-T dX
        dop = -g @ x.T
    else:
        x = alpha * sl.solve(op.T, b.T).T
        g = sl.solve(op, w.T).T       # dX op(A)^-T
        dop = -x.T @ g
    da = dop if trans == "N" else dop.T
    da = np.tril(da) if lower else np.triu(da)
    if diag == "U":
        da = da - np.diag(np.diag(da))
    return x, alpha * g, da


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsm_backward_vs_fp64_reference(side, uplo, trans, diag):
    rows, cols = (150, 70) if side == "L" else (70, 150)
    na = rows if side == "L" else cols
    A = _tri(na, uplo == "L", 60, unit=(diag == "U")).requires_grad_(True)
    g = torch.Generator(device="cuda").manual_seed(61)
    B = (torch.rand(rows, cols, device="cuda", generator=g) * 2 - 1).requires_grad_(True)
    W = torch.randn(rows, cols, device="cuda", generator=g)
    X = ag.trsm(A, B, side=side, uplo=uplo, trans=trans, diag=diag, alpha=1.5)
    X.backward(W)
    x, db, da = _trsm_ref(*(t.detach().double().cpu().numpy() for t in (A, B)), side, uplo, trans, diag, 1.5,
                          W.double().cpu().numpy())
    _check_trsm_grads(X, B.grad, A.grad, x, db, da, side, trans, 1.5, "fp32")


def _check_trsm_grads(X, dB, dA, x, db, da, side, trans, alpha, prec):
    """X and dB come from refined solves: normwise fp32 level.  dA is one
    product -G X^T (or -X^T G): componentwise gemm bound on |G||X|."""
    for got, want in ((X, x), (dB, db)):
        err = np.abs(got.detach().double().cpu().numpy() - want).max() / max(np.abs(want).max(), 1e-30)
        assert err <= 3e-6, (side, trans, prec, err)
    g = np.abs(db / alpha)
    ax = np.abs(x)
    mag = g @ ax.T if side == "L" else ax.T @ g
    kk = x.shape[1] if side == "L" else x.shape[0]
    if trans == "T":
        mag = mag.T
    # G itself carries the refined-solve error (normwise 3e-6) into the product
    bound = (H.rep_error(prec) + (2 * kk + 2) * H.U32 + 3e-6) * mag + 1e-30
    err = np.abs(dA.detach().double().cpu().numpy() - da)
    assert (err <= bound).all(), (side, trans, prec, float((err / bound).max()))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_trsm_backward_config5_2048(prec):
    """X = A^-1 B (L,L,N,N, m = n = 2048, the config-3 triangular
    distribution), dX iid N(0,1): analytic fp64 CPU gradient and 64 fp64 FD
    directions (A perturbed on its lower triangle only)."""
    n = 2048
    A = _tri(n, True, 62).requires_grad_(True)
    g = torch.Generator(device="cuda").manual_seed(63)
    B = (torch.rand(n, n, device="cuda", generator=g) * 2 - 1).requires_grad_(True)
    W = torch.randn(n, n, device="cuda", generator=g)
    X = ag.trsm(A, B, precision=prec)
    X.backward(W)
    a, b, w = (t.detach().double().cpu().numpy() for t in (A, B, W))
    x, db, da = _trsm_ref(a, b, "L", "L", "N", "N", 1.0, w)
    _check_trsm_grads(X, B.grad, A.grad, x, db, da, "L", "N", 1.0, prec)
    A64, B64, W64 = (t.detach().double() for t in (A, B, W))
    f = lambda aa, bb: (W64 * torch.linalg.solve_triangular(aa, bb, upper=False)).sum()
    dA, dB = A.grad.double(), B.grad.double()
    h = 1e-6
    for _ in range(64):
        V = torch.tril(torch.randn(n, n, device="cuda", dtype=torch.float64, generator=g))
        U = torch.randn(n, n, device="cuda", dtype=torch.float64, generator=g)
        fd = (f(A64 + h * V, B64 + h * U) - f(A64 - h * V, B64 - h * U)) / (2 * h)
        ours = (dA * V).sum() + (dB * U).sum()
        scale = (dA.abs() * V.abs()).sum() + (dB.abs() * U.abs()).sum()
        assert abs(float(fd - ours)) <= 1e-5 * float(scale), (float(fd), float(ours), float(scale))
