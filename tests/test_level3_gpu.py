"""GPU parity of trsm, trmm, syrk, syr2k and symm against the fp64 oracle,
through the C-ABI (device entries with torch tensors, host entries with numpy).

Bounds:
  trmm/syrk/syr2k/symm (products): componentwise
      |X - X_oracle| <= (e_rep + (2k+2) u)|alpha| P + 4u|beta C|
    with P the matching |A||B| product (helpers.gemm_bound);
  trsm: normwise on the solution, max|X - X_oracle| <= tol * max|X_oracle|,
    tol = 1.1 * 2^-24 for fp32 data (the refined entry solves in fp64 and
    rounds X to fp32 once), 1e-12 for fp64 data; the test
    matrices have diagonal U[1,2] and off-diagonal U[-1,1]/n (well conditioned).
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


@pytest.mark.parametrize("case", [t for t in KAT["tests"] if t["routine"] != "gemm"], ids=lambda t: t["name"])
@pytest.mark.parametrize("entry", ["dev", "host"])
def test_known_answers(case, entry):
    dt = np.float32 if case["dtype"] == "float32" else np.float64
    arrays = {k: np.array(v, dtype=dt) for k, v in case["arrays"].items()}
    if entry == "dev":
        arrays = {k: dev(v) for k, v in arrays.items()}
    args = [arrays[a["array"]] if isinstance(a, dict) else a for a in case["args"]]
    assert getattr(mb, case["routine"])(*args) == 0
    torch.cuda.synchronize()
    for e in case["expects"]:
        got = float(arrays[e["array"]][e["index"]])
        assert abs(got - e["value"]) <= e["atol"] + e["rtol"] * abs(e["value"]), (case["name"], e, got)


def tri_operand(rng, n, upper, dtype, unit_junk=False):
    """Stored triangular A: diag U[1,2], off-diag U[-1,1]/n in the `upper`
    triangle, NaN in the other triangle (must never be read)."""
    a = rng.uniform(-1, 1, (n, n)) / max(n, 1)
    a[np.diag_indices(n)] = rng.uniform(1, 2, n)
    mask = np.triu(np.ones((n, n), bool)) if upper else np.tril(np.ones((n, n), bool))
    a = np.where(mask, a, np.nan)
    if unit_junk:
        a[np.diag_indices(n)] = np.nan  # unit diagonal is implied, never read
    return np.asfortranarray(a).astype(dtype)


TRSM_SIZES = [(1, 1), (5, 7), (130, 33), (300, 129), (257, 400), (1100, 70), (70, 1100)]


@pytest.mark.parametrize("side,uplo,ta,diag", list(itertools.product("LR", "UL", "NT", "NU")))
@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "f64"])
def test_trsm_vs_oracle(side, uplo, ta, diag, prec):
    rng = np.random.default_rng(21)
    dt = np.float64 if prec == "f64" else np.float32
    for m, n in TRSM_SIZES:
        na = m if side == "L" else n
        a = tri_operand(rng, na, uplo == "U", dt, unit_junk=(diag == "U"))
        b = np.asfortranarray(rng.uniform(-1, 1, (m, n))).astype(dt)
        alpha = 1.5
        want = b.astype(np.float64).ravel(order="F").copy()
        assert oracle.dtrsm(side, uplo, ta, diag, m, n, alpha, a.astype(np.float64).ravel(order="F"), na, want, m) == 0
        A, B = dev(a.ravel(order="F")), dev(b.ravel(order="F"))
        kw = {} if prec == "f64" else {"precision": prec}
        assert mb.trsm(side, uplo, ta, diag, m, n, alpha, A, na, B, m, **kw) == 0
        got = B.cpu().numpy().astype(np.float64)
        tol = 1e-12 if prec == "f64" else 1.1 * 2.0 ** -24
        err = np.abs(got - want).max() / max(np.abs(want).max(), 1e-30)
        assert err <= tol, (m, n, side, uplo, ta, diag, prec, err)


def test_trsm_refinement_reaches_fp32_level():
    """Without refinement the BF16x3 updates leave ~1e-6-level error; the
    refined (fp64) solve must be at fp32 rounding level."""
    rng = np.random.default_rng(22)
    m, n = 1024, 256
    a = tri_operand(rng, m, False, np.float32)
    b = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    want = b.astype(np.float64).ravel(order="F").copy()
    oracle.dtrsm("L", "L", "N", "N", m, n, 1.0, a.astype(np.float64).ravel(order="F"), m, want, m)
    errs = {}
    for refine in (0, 1):
        A, B = dev(a.ravel(order="F")), dev(np.asfortranarray(b).ravel(order="F"))
        assert mb.trsm("L", "L", "N", "N", m, n, 1.0, A, m, B, m, precision="bf16x3", refine=refine) == 0
        errs[refine] = np.abs(B.cpu().numpy() - want).max() / np.abs(want).max()
    assert errs[1] <= 2e-7, errs
    assert errs[1] <= errs[0], errs


@pytest.mark.parametrize("side,uplo,ta", list(itertools.product("LR", "UL", "NT")))
@pytest.mark.parametrize("prec", ["f64", "fp32"])
def test_trsm_ill_conditioned_triangle(side, uplo, ta, prec):
    """Off-diagonal U[-1,1] without the 1/n scaling (diagonal U[1,2]): the
    condition number grows roughly tenfold per 16 rows (~1e4 at n = 96,
    ~1e8 at n = 260), so the explicit fp64 inverses of the diagonal blocks
    hold large entries and a wrong element of them shows.  Normwise bound
    against the fp64 oracle: fp64 data 64 kappa 2^-53; fp32 data (refined:
    solved in fp64, X rounded to fp32 once) 2^-24 + 64 kappa 2^-53."""
    rng = np.random.default_rng(31)
    dt = np.float64 if prec == "f64" else np.float32
    sizes = [96, 160, 260] if prec == "f64" else [48, 96]
    for na in sizes:
        a = rng.uniform(-1, 1, (na, na))
        a[np.diag_indices(na)] = rng.uniform(1, 2, na)
        a = np.triu(a) if uplo == "U" else np.tril(a)
        a = np.asfortranarray(a).astype(dt)
        kappa = np.linalg.cond(a.astype(np.float64))
        m, n = (na, 40) if side == "L" else (40, na)
        b = np.asfortranarray(rng.uniform(-1, 1, (m, n))).astype(dt)
        want = b.astype(np.float64).ravel(order="F").copy()
        assert oracle.dtrsm(side, uplo, ta, "N", m, n, 1.0, a.astype(np.float64).ravel(order="F"), na, want, m) == 0
        A, B = dev(a.ravel(order="F")), dev(b.ravel(order="F"))
        kw = {} if prec == "f64" else {"precision": prec}
        assert mb.trsm(side, uplo, ta, "N", m, n, 1.0, A, na, B, m, **kw) == 0
        got = B.cpu().numpy().astype(np.float64)
        tol = 64 * kappa * 2.0 ** -53 + (0.0 if prec == "f64" else 2.0 ** -24)
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= max(tol, 1e-12), (na, side, uplo, ta, prec, kappa, err, tol)


@pytest.mark.parametrize("side,uplo,ta,diag", list(itertools.product("LR", "UL", "NT", "NU")))
@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "f64"])
def test_trmm_vs_oracle(side, uplo, ta, diag, prec):
    rng = np.random.default_rng(23)
    dt = np.float64 if prec == "f64" else np.float32
    for m, n in [(1, 1), (7, 5), (140, 300), (300, 129)]:
        na = m if side == "L" else n
        a = tri_operand(rng, na, uplo == "U", dt, unit_junk=(diag == "U"))
        b = np.asfortranarray(rng.uniform(-1, 1, (m, n))).astype(dt)
        alpha = -0.75
        want = b.astype(np.float64).ravel(order="F").copy()
        a64 = a.astype(np.float64)
        assert oracle.dtrmm(side, uplo, ta, diag, m, n, alpha, a64.ravel(order="F"), na, want, m) == 0
        A, B = dev(a.ravel(order="F")), dev(b.ravel(order="F"))
        kw = {} if prec == "f64" else {"precision": prec}
        assert mb.trmm(side, uplo, ta, diag, m, n, alpha, A, na, B, m, **kw) == 0
        got = B.cpu().numpy().astype(np.float64)
        t = np.nan_to_num(a64)
        if diag == "U":
            t[np.diag_indices(na)] = 1.0
        t = t if ta == "N" else t.T
        bb = np.abs(b.astype(np.float64))
        p = (np.abs(t) @ bb) if side == "L" else (bb @ np.abs(t))
        if prec == "f64":
            bound = 4 * na * 2.0 ** -53 * abs(alpha) * p + 1e-300
        else:
            bound = H.gemm_bound(prec, na, alpha, p, 0.0, 0.0)
        err = np.abs(got.reshape(n, m).T - want.reshape(n, m).T)
        assert (err <= bound).all(), (m, n, side, uplo, ta, diag, prec, float((err / bound).max()))


@pytest.mark.parametrize("uplo,trans", list(itertools.product("UL", "NT")))
@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "f64"])
def test_syrk_syr2k_vs_oracle(uplo, trans, prec):
    rng = np.random.default_rng(24)
    dt = np.float64 if prec == "f64" else np.float32
    for n, k in [(1, 1), (9, 4), (200, 77), (300, 260)]:
        ar, ac = (n, k) if trans == "N" else (k, n)
        a = np.asfortranarray(rng.uniform(-1, 1, (ar, ac))).astype(dt)
        b = np.asfortranarray(rng.uniform(-1, 1, (ar, ac))).astype(dt)
        tri = np.triu(np.ones((n, n), bool)) if uplo == "U" else np.tril(np.ones((n, n), bool))
        c0 = rng.uniform(-1, 1, (n, n))
        c0[~tri] = np.nan  # the other triangle is neither read nor written
        c0 = np.asfortranarray(c0).astype(dt)
        for routine in ("syrk", "syr2k"):
            alpha, beta = 1.25, 0.5
            want = np.nan_to_num(c0.astype(np.float64)).ravel(order="F")
            af, bf = a.astype(np.float64).ravel(order="F"), b.astype(np.float64).ravel(order="F")
            if routine == "syrk":
                assert oracle.dsyrk(uplo, trans, n, k, alpha, af, ar, beta, want, n) == 0
            else:
                assert oracle.dsyr2k(uplo, trans, n, k, alpha, af, ar, bf, ar, beta, want, n) == 0
            C = dev(c0.ravel(order="F"))
            kw = {} if prec == "f64" else {"precision": prec}
            if routine == "syrk":
                assert mb.syrk(uplo, trans, n, k, alpha, dev(a.ravel(order="F")), ar, beta, C, n, **kw) == 0
            else:
                assert mb.syr2k(uplo, trans, n, k, alpha, dev(a.ravel(order="F")), ar, dev(b.ravel(order="F")), ar,
                                beta, C, n, **kw) == 0
            got = C.cpu().numpy().astype(np.float64).reshape(n, n).T
            w = want.reshape(n, n).T
            A64 = np.abs(a.astype(np.float64) if trans == "N" else a.astype(np.float64).T)
            B64 = np.abs(b.astype(np.float64) if trans == "N" else b.astype(np.float64).T)
            p = A64 @ A64.T if routine == "syrk" else A64 @ B64.T + B64 @ A64.T
            cabs = np.nan_to_num(np.abs(c0.astype(np.float64)))
            if prec == "f64":
                bound = 4 * max(k, 1) * 2.0 ** -53 * (alpha * p + beta * cabs) + 1e-300
            else:
                bound = H.gemm_bound(prec, 2 * k, alpha, p, beta, cabs)
            err = np.abs(got - w)
            assert (err[tri] <= bound[tri]).all(), (routine, n, k, uplo, trans, prec, float((err / bound)[tri].max()))
            assert np.isnan(got[~tri]).all(), (routine, "other triangle written")


@pytest.mark.parametrize("side,uplo", list(itertools.product("LR", "UL")))
@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "f64"])
def test_symm_vs_oracle(side, uplo, prec):
    rng = np.random.default_rng(25)
    dt = np.float64 if prec == "f64" else np.float32
    for m, n in [(1, 1), (6, 9), (200, 130), (129, 300)]:
        na = m if side == "L" else n
        s = rng.uniform(-1, 1, (na, na))
        s = s + s.T
        stored = s.copy()
        stored[(np.tril(np.ones((na, na)), -1) if uplo == "U" else np.triu(np.ones((na, na)), 1)) > 0] = np.nan
        a = np.asfortranarray(stored).astype(dt)
        b = np.asfortranarray(rng.uniform(-1, 1, (m, n))).astype(dt)
        b[0, 0] = 0.0  # exact zero in B (the reference's zero-skip must not matter)
        c = np.asfortranarray(rng.uniform(-1, 1, (m, n))).astype(dt)
        alpha, beta = -2.0, 0.5
        want = c.astype(np.float64).ravel(order="F").copy()
        assert oracle.dsymm(side, uplo, m, n, alpha, a.astype(np.float64).ravel(order="F"), na,
                            b.astype(np.float64).ravel(order="F"), m, beta, want, m) == 0
        C = dev(c.ravel(order="F"))
        kw = {} if prec == "f64" else {"precision": prec}
        assert mb.symm(side, uplo, m, n, alpha, dev(a.ravel(order="F")), na, dev(b.ravel(order="F")), m, beta, C, m,
                       **kw) == 0
        got = C.cpu().numpy().astype(np.float64).reshape(n, m).T
        sym = np.abs(s.astype(dt).astype(np.float64))
        bb = np.abs(b.astype(np.float64))
        p = sym @ bb if side == "L" else bb @ sym
        if prec == "f64":
            bound = 4 * na * 2.0 ** -53 * (2 * p + 0.5 * np.abs(c)) + 1e-300
        else:
            bound = H.gemm_bound(prec, na, alpha, p, beta, c.astype(np.float64))
        err = np.abs(got - want.reshape(n, m).T)
        assert (err <= bound).all(), (m, n, side, uplo, prec, float((err / bound).max()))


@pytest.mark.parametrize("routine,side,uplo,ta,diag",
                         [("symm", s, u, "N", "N") for s in "LR" for u in "UL"]
                         + [("trmm", s, u, t, d) for s in "LR" for u in "UL" for t in "NT" for d in "NU"])
def test_split_pack_views_equal_materialized_gemm(routine, side, uplo, ta, diag):
    """BF16x3 symm / trmm read A through the pack's operand views (the
    symmetric half mirrored per 64x128 block, per element on blocks that
    cross the diagonal; the triangle masked, the unit diagonal injected)
    instead of a copy.  The same product through gemm on the materialized
    operand (full symmetric matrix; triangle with zeros and the diagonal)
    packs the same values, so the results are equal bit for bit (a k-range
    only skips k-tiles of zeros).  Sizes with ragged blocks on both sides of
    the diagonal; the unused half of A and a unit diagonal hold NaN."""
    rng = np.random.default_rng(26)
    for m, n in [(1000, 777), (777, 1000), (129, 65)]:
        na = m if side == "L" else n
        full = rng.uniform(-1, 1, (na, na))
        keep = np.triu(np.ones((na, na), bool)) if uplo == "U" else np.tril(np.ones((na, na), bool))
        if routine == "symm":
            full = np.where(keep, full, full.T)  # symmetric
            op = full
        else:
            tri = np.where(keep, full, 0.0)
            if diag == "U":
                tri[np.diag_indices(na)] = 1.0
            op = tri if ta == "N" else tri.T
        stored = np.where(keep, full, np.nan)
        if routine == "trmm" and diag == "U":
            stored[np.diag_indices(na)] = np.nan
        a = dev(np.asfortranarray(stored).astype(np.float32).ravel(order="F"))
        opm = dev(np.asfortranarray(op).astype(np.float32).ravel(order="F"))
        b = dev(np.asfortranarray(rng.uniform(-1, 1, (m, n))).astype(np.float32).ravel(order="F"))
        c0 = dev(np.asfortranarray(rng.uniform(-1, 1, (m, n))).astype(np.float32).ravel(order="F"))
        alpha = 0.75
        beta = 0.5 if routine == "symm" else 0.0
        got, want = c0.clone(), c0.clone()
        if routine == "symm":
            assert mb.symm(side, uplo, m, n, alpha, a, na, b, m, beta, got, m, precision="bf16x3") == 0
        else:
            got = b.clone()
            assert mb.trmm(side, uplo, ta, diag, m, n, alpha, a, na, got, m, precision="bf16x3") == 0
        if side == "L":
            assert mb.gemm("N", "N", m, n, m, alpha, opm, na, b, m, beta, want, m, precision="bf16x3") == 0
        else:
            assert mb.gemm("N", "N", m, n, n, alpha, b, m, opm, na, beta, want, m, precision="bf16x3") == 0
        g, w = got.cpu().numpy(), want.cpu().numpy()
        assert np.isfinite(g).all(), (routine, m, n, side, uplo, ta, diag)
        assert np.array_equal(g, w), (routine, m, n, side, uplo, ta, diag, float(np.abs(g - w).max()))


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_config3_strsm_4096(prec):
    """strsm L,L,N,N m=n=4096 (config size): A lower, off-diagonal U[-1,1]/n,
    diagonal U[1,2], B U[-1,1].  Checked on the fp64 residual of the whole
    solution and on 16 columns against the oracle."""
    m = n = 4096
    g = torch.Generator(device="cuda").manual_seed(31)
    a = (torch.rand(m, m, device="cuda", generator=g, dtype=torch.float64) * 2 - 1) / m
    a = torch.tril(a, -1) + torch.diag(1 + torch.rand(m, device="cuda", generator=g, dtype=torch.float64))
    A = a.t().contiguous().float().reshape(-1)  # column-major buffer
    b = torch.rand(n, m, device="cuda", generator=g) * 2 - 1  # row j = column j of B
    B = b.reshape(-1).clone()
    assert mb.trsm("L", "L", "N", "N", m, n, 1.0, A, m, B, m, precision=prec) == 0
    torch.cuda.synchronize()
    X = B.view(n, m).t().double()
    L = A.view(m, m).t().double()
    res = (L @ X - b.t().double()).abs().max().item()
    assert res <= 1e-6, res
    cols = np.arange(0, n, n // 16)
    a_h = A.cpu().numpy().astype(np.float64)
    want = b[cols].cpu().numpy().astype(np.float64).ravel()  # 16 columns, ld = m
    oracle.dtrsm("L", "L", "N", "N", m, len(cols), 1.0, a_h, m, want, m)
    got = X[:, torch.from_numpy(cols).cuda()].t().cpu().numpy().ravel()
    assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_config3_ssyrk_8192x2048(prec):
    """ssyrk L,N n=8192 k=2048, A N(0,1), alpha=1, beta=0, priced on 2048
    sampled lower-triangle entries by the oracle; upper triangle untouched."""
    n, k = 8192, 2048
    g = torch.Generator(device="cuda").manual_seed(32)
    A = torch.randn(n * k, device="cuda", generator=g)
    C = torch.full((n * n,), float("nan"), device="cuda")
    assert mb.syrk("L", "N", n, k, 1.0, A, n, 0.0, C, n, precision=prec) == 0
    torch.cuda.synchronize()
    rng = np.random.default_rng(33)
    i = rng.integers(0, n, 2048)
    j = rng.integers(0, n, 2048)
    rows, cols = np.maximum(i, j), np.minimum(i, j)
    a = A.cpu().numpy()
    want = oracle.sgemm_sample("N", "T", k, 1.0, a, n, a, n, 0.0, None, rows, cols)
    absab = oracle.sgemm_sample_abs("N", "T", k, a, n, a, n, rows, cols)
    c = C.cpu().numpy()
    got = c[rows + cols * n].astype(np.float64)
    bound = H.gemm_bound(prec, k, 1.0, absab, 0.0, 0.0)
    assert (np.abs(got - want) <= bound).all()
    up_r, up_c = np.minimum(i, j), np.maximum(i, j)
    strict = up_r < up_c
    assert np.isnan(c[up_r[strict] + up_c[strict] * n]).all()


WHOLE_CASES = [("trmm", "L", "L", "N"), ("trmm", "L", "U", "N"), ("trmm", "R", "L", "T"), ("trmm", "R", "U", "N"),
               ("syrk", None, "L", "N"), ("syrk", None, "U", "T"), ("syr2k", None, "L", "N"), ("syr2k", None, "U", "T")]


@pytest.mark.parametrize("routine,side,uplo,trans,k", [c + (256,) for c in WHOLE_CASES]
                         + [(r, None, u, t, 512) for r in ("syrk", "syr2k") for u, t in (("L", "N"), ("U", "T"))])
@pytest.mark.parametrize("prec", ["fp32", "f64"])
def test_whole_tile_products_sampled(routine, side, uplo, trans, k, prec):
    """4096-sized strmm / ssyrk / ssyr2k (and the fp64 forms): C is whole
    128 x 128 tiles and k a multiple of the k-tile, so the triangular grids
    (lower and upper), the four k-ranges and the store masks run on the
    direct-to-LDS kernel.  With k = 512 the 528 triangle tiles leave a
    partial last round on 512 resident workgroups, whose 16 tiles are split
    along k into partial sums and reduced.  4096 sampled entries against
    fp64 dot products (bound (2k+2) u |alpha| |A||B| for fp32, 4k u64 for
    fp64); the untouched triangle of C stays NaN."""
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator(device="cuda").manual_seed(41)
    n = 4096
    rng = np.random.default_rng(42)
    kw = {} if prec == "f64" else {"precision": prec}
    alpha = 0.5
    if routine == "trmm":
        a = torch.randn(n * n, device="cuda", generator=g, dtype=dt)
        b0 = torch.randn(n * n, device="cuda", generator=g, dtype=dt)
        B = b0.clone()
        assert mb.trmm(side, uplo, trans, "N", n, n, alpha, a, n, B, n, **kw) == 0
        torch.cuda.synchronize()
        am = a.cpu().numpy().reshape(n, n).T
        am = np.tril(am) if uplo == "L" else np.triu(am)
        opa = am if trans == "N" else am.T
        bm = b0.cpu().numpy().reshape(n, n).T
        rows, cols = rng.integers(0, n, 4096), rng.integers(0, n, 4096)
        if side == "L":  # B := alpha op(A) B0
            ra, cb = opa[rows].astype(np.float64), bm[:, cols].T.astype(np.float64)
        else:  # B := alpha B0 op(A)
            ra, cb = bm[rows].astype(np.float64), opa[:, cols].T.astype(np.float64)
        got = B.cpu().numpy().reshape(n, n).T[rows, cols].astype(np.float64)
    else:
        a = torch.randn(n * k, device="cuda", generator=g, dtype=dt)
        b = torch.randn(n * k, device="cuda", generator=g, dtype=dt)
        C = torch.full((n * n,), float("nan"), device="cuda", dtype=dt)
        lda = n if trans == "N" else k
        if routine == "syrk":
            assert mb.syrk(uplo, trans, n, k, alpha, a, lda, 0.0, C, n, **kw) == 0
        else:
            assert mb.syr2k(uplo, trans, n, k, alpha, a, lda, b, lda, 0.0, C, n, **kw) == 0
        torch.cuda.synchronize()
        # op rows: trans N -> rows of the n x k A; T -> columns of the k x n A
        am = a.cpu().numpy().reshape(k, n).T if trans == "N" else a.cpu().numpy().reshape(n, k)
        bm = b.cpu().numpy().reshape(k, n).T if trans == "N" else b.cpu().numpy().reshape(n, k)
        am, bm = am.astype(np.float64), bm.astype(np.float64)
        i, j = rng.integers(0, n, 4096), rng.integers(0, n, 4096)
        lo, hi = np.minimum(i, j), np.maximum(i, j)
        rows, cols = (hi, lo) if uplo == "L" else (lo, hi)
        if routine == "syrk":
            ra, cb = am[rows], am[cols]
        else:
            ra = np.concatenate([am[rows], bm[rows]], axis=1)
            cb = np.concatenate([bm[cols], am[cols]], axis=1)
        c = C.cpu().numpy().reshape(n, n).T
        got = c[rows, cols].astype(np.float64)
        strict = lo < hi
        other_r, other_c = (lo, hi) if uplo == "L" else (hi, lo)
        assert np.isnan(c[other_r[strict], other_c[strict]]).all()
    kk = ra.shape[1]
    want = alpha * np.einsum("ij,ij->i", ra, cb)
    absab = abs(alpha) * np.einsum("ij,ij->i", np.abs(ra), np.abs(cb))
    u = 2.0 ** -53 if prec == "f64" else H.U32
    bound = (4 * kk * u if prec == "f64" else (2 * kk + 2) * u) * absab + 1e-300
    err = np.abs(got - want)
    assert (err <= bound).all(), float((err / bound).max())


@pytest.mark.parametrize("side,uplo,ta", [("L", "L", "N"), ("L", "U", "T"), ("R", "U", "N"), ("R", "L", "T")])
def test_refined_strsm_whole_tiles(side, uplo, ta):
    """strsm at 4096 (whole 128-tiles, so the refined solve's fp64 updates
    run on the direct-to-LDS fp64 kernel and its panels cover 4 diagonal
    blocks): against an fp64 solve of the same fp32 inputs, the refined
    solution must be within 4 u32 max|X| and well below the unrefined one's
    error."""
    m = n = 4096
    g = torch.Generator(device="cuda").manual_seed(43)
    a = (torch.rand(m, m, device="cuda", generator=g, dtype=torch.float64) * 2 - 1) / m
    a = (torch.tril(a, -1) if uplo == "L" else torch.triu(a, 1)) + torch.diag(
        1 + torch.rand(m, device="cuda", generator=g, dtype=torch.float64))
    A = a.t().contiguous().float().reshape(-1)  # column-major buffer
    b = (torch.rand(n, m, device="cuda", generator=g) * 2 - 1).reshape(-1)
    alpha = 0.75
    Xs = []
    for refine in (0, 1):
        B = b.clone()
        assert mb.trsm(side, uplo, ta, "N", m, n, alpha, A, m, B, m, precision="fp32", refine=refine) == 0
        Xs.append(B.view(n, m).t().double())
    torch.cuda.synchronize()
    a32 = A.view(m, m).t().double()
    opa = a32 if ta == "N" else a32.t().contiguous()
    upper = (uplo == "U") != (ta == "T")
    rhs = alpha * b.view(n, m).t().double()
    ref = torch.linalg.solve_triangular(opa, rhs, upper=upper, left=(side == "L"))
    e0, e1 = ((X - ref).abs().max().item() for X in Xs)
    assert e1 <= 4 * 2.0 ** -24 * ref.abs().max().item(), (e0, e1)
    assert e1 <= 0.6 * e0, (e0, e1)
