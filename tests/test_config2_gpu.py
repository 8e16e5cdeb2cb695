"""Config 2 at its own sizes: SGEMM m=n=k=8192 in all four transposes and
m=n=4096, k=16384, alpha=1, beta=0, A/B iid normal(0,1), in fp32, BF16x3
and FP8 (e4m3, per-128x128-tile power-of-two scales), through the C-ABI.

Checks per variant:
* the oracle (oracle.sgemm_sample / sgemm_sample_abs, FP64 dot products of
  the fp32 inputs) on 4096 random entries plus 1024 random entries along
  each row and column at the tile seams (0, 127, 128, 255, 256, last):
  fp32 / BF16x3 within |C - C64| <= (e_rep + (2k+2)u)|A||B| (u = 2^-24,
  e_rep = 0 / 3*2^-16); FP8 within the per-element rounding model
  sum_l (d_a|b| + |a|d_b + d_a d_b) + 2k u |A||B|, d_x = 2^-4|x| + 2^-10 s
  with s the operand tile's scale 2^ceil(log2(amax/448));
* every entry against an FP64 product of the same inputs on the GPU (the
  ROCm library's fp64 gemm through torch, itself checked against the oracle
  on the sampled entries to 1e-12 relative), with the same bounds (fp32,
  BF16x3), and the error record: max |C - C64|, max |C - C64| / (|A||B|)
  and relative Frobenius, printed as one JSON line per variant.
"""
import itertools
import json

import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
import mojoblas_amd as mb

pytestmark = pytest.mark.gpu

SHAPES = {"8192^3": (8192, 8192, 8192), "4096x4096x16384": (4096, 4096, 16384)}
CASES = ([("8192^3", ta + tb, p) for ta, tb in itertools.product("NT", "NT") for p in ("fp32", "bf16x3", "fp8")] +
         [("4096x4096x16384", "NN", p) for p in ("fp32", "bf16x3", "fp8")])


def _operands(shape, tt, seed):
    m, n, k = SHAPES[shape]
    ta, tb = tt
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.randn(ar * ac, device="cuda", generator=g)
    B = torch.randn(br * bc, device="cuda", generator=g)
    return m, n, k, ta, tb, ar, ac, br, bc, A, B


def _op(X, rows, cols, trans):
    """op(X) as a 2-D view of the column-major flat tensor."""
    M = X.view(cols, rows).t()
    return M if trans == "N" else M.t()


def _samples(m, n, rng):
    seams_m = np.array([0, 127, 128, 255, 256, m - 1])
    seams_n = np.array([0, 127, 128, 255, 256, n - 1])
    rows = np.concatenate([rng.integers(0, m, 4096), np.repeat(seams_m, 1024), rng.integers(0, m, 1024 * 6)])
    cols = np.concatenate([rng.integers(0, n, 4096), rng.integers(0, n, 1024 * 6), np.repeat(seams_n, 1024)])
    return rows.astype(np.int64), cols.astype(np.int64)


def _fp8_dev(opx):
    """Per-element e4m3 error bound of a (rows x k) operand, fp64 on the GPU:
    d = 2^-4 |x| + 2^-10 s, s = 2^ceil(log2(amax / 448)) per 128x128 tile."""
    r, k = opx.shape
    x = opx.abs().double()
    pr, pk = (-r) % 128, (-k) % 128
    xp = torch.nn.functional.pad(x, (0, pk, 0, pr))
    amax = xp.view((r + pr) // 128, 128, (k + pk) // 128, 128).amax(dim=(1, 3))
    s = torch.where(amax > 0, torch.exp2(torch.ceil(torch.log2(amax / 448.0))), torch.ones_like(amax))
    sfull = s.repeat_interleave(128, 0).repeat_interleave(128, 1)[:r, :k]
    return 2.0 ** -4 * x + 2.0 ** -10 * sfull


def _as_stored(opx, trans):
    """A (rows x k) op() operand back in the stored layout, flat fp32."""
    M = opx if trans == "N" else opx.t()
    return M.t().contiguous().view(-1).float()


@pytest.mark.parametrize("shape,tt,prec", CASES, ids=["%s-%s-%s" % c for c in CASES])
def test_config2_at_size(shape, tt, prec):
    seed = {"NN": 1, "NT": 2, "TN": 3, "TT": 4}[tt] + (10 if shape != "8192^3" else 0)
    m, n, k, ta, tb, ar, ac, br, bc, A, B = _operands(shape, tt, seed)
    C = torch.full((m * n,), float("nan"), device="cuda")  # beta = 0: C is never read
    assert mb.gemm(ta, tb, m, n, k, 1.0, A, ar, B, br, 0.0, C, m, precision=prec) == 0
    torch.cuda.synchronize()
    assert torch.isfinite(C).all()

    rng = np.random.default_rng(100 + seed)
    rows, cols = _samples(m, n, rng)
    a, b = A.cpu().numpy(), B.cpu().numpy()
    want = oracle.sgemm_sample(ta, tb, k, 1.0, a, ar, b, br, 0.0, None, rows, cols)
    absab = oracle.sgemm_sample_abs(ta, tb, k, a, ar, b, br, rows, cols)
    got = C.cpu().numpy()[rows + cols * m].astype(np.float64)
    err = np.abs(got - want)
    if prec == "fp8":
        opa = _op(A, ar, ac, ta)
        opbt = _op(B, br, bc, tb).t()  # n x k
        da = _as_stored(_fp8_dev(opa), ta).cpu().numpy()
        db = _as_stored(_fp8_dev(opbt).t(), tb).cpu().numpy()
        absa, absb = np.abs(a), np.abs(b)
        bound = (oracle.sgemm_sample_abs(ta, tb, k, da, ar, absb, br, rows, cols) +
                 oracle.sgemm_sample_abs(ta, tb, k, absa, ar, db, br, rows, cols) +
                 oracle.sgemm_sample_abs(ta, tb, k, da, ar, db, br, rows, cols)) * 1.001 + \
            2 * k * H.U32 * absab + 1e-30
    else:
        bound = H.gemm_bound(prec, k, 1.0, absab, 0.0, 0.0)
    assert (err <= bound).all(), (shape, tt, prec, float((err / bound).max()))

    # every entry against an FP64 product of the same inputs
    opa = _op(A, ar, ac, ta).double().contiguous()
    opb = _op(B, br, bc, tb).double().contiguous()
    c64 = opa @ opb
    abs64 = opa.abs() @ opb.abs()
    c64f = c64.t().contiguous().view(-1)  # column-major, like C
    ref_s = c64f[torch.from_numpy(rows + cols * m).cuda()].cpu().numpy()
    assert (np.abs(ref_s - want) <= 1e-12 * absab).all()  # the FP64 product agrees with the oracle
    d = (C.double() - c64f).abs()
    absf = abs64.t().contiguous().view(-1)
    rec = {"shape": shape, "trans": tt, "precision": prec,
           "max_abs": float(d.max()), "max_rel_componentwise": float((d / absf).max()),
           "rel_frobenius": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(c64f))}
    print("CONFIG2_ERROR " + json.dumps(rec))
    if prec != "fp8":
        e = H.rep_error(prec) + (2.0 * k + 2.0) * H.U32
        assert bool((d <= e * absf + 1e-30).all()), rec
