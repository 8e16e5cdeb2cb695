"""Batched gemm on bf16 operands (mbh_bgemm_vbatched_exec): A_p, B_p stored
as bf16 in HBM at any 2-byte alignment, C_p fp32.  Two checks per batch:
  * bit-identical C to the fp32-operand bf16 path (mbh_sgemm_vbatched_exec on
    the same plan) fed with fp32 copies of the same bf16 values: both stage
    the same bf16 tiles into LDS and run the same MFMAs in the same order;
  * every problem against the fp64 oracle on the bf16 values, within
    (2k+2)u|alpha||A||B| + 4u|beta C| (the operands are exact, so no
    representation term).
Operands are placed at element offsets of both parities, so columns start
in either half of a dword, and leading dimensions include odd paddings.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
import mojoblas_amd as mb
from mojoblas_amd import batched
from tests.test_batched_gpu import TABLES, _cases, _check_problems, table  # noqa: F401  (table: fixture)

pytestmark = pytest.mark.gpu


def _min8(m, n, k):
    """k raised where an operand of a problem with k > 0 would hold fewer
    than eight elements (the bf16-operand kernels read eight at a time)."""
    k = k.copy()
    for p in range(len(k)):
        while k[p] > 0 and m[p] > 0 and n[p] > 0 and (k[p] * m[p] < 8 or k[p] * n[p] < 8):
            k[p] += 1
    return k


def _pack16(m, n, k, g, fill="uniform"):
    """bf16 operands at element offsets of mixed parity (one flat buffer per
    operand, ld = rows), their exact fp32 copies at the same offsets, and an
    fp32 C buffer; returns (bf16 bufs, fp32 bufs, offsets, ptrs16, ptrs32)."""
    sa, sb, sc = m * k, k * n, m * n
    gap = np.arange(len(m)) % 3  # 0, 1, 2 extra elements: both parities
    aoff = 1 + np.concatenate([[0], np.cumsum(sa + gap)[:-1]]).astype(np.int64)
    boff = 2 + np.concatenate([[0], np.cumsum(sb + gap)[:-1]]).astype(np.int64)
    coff = np.concatenate([[0], np.cumsum((sc + 3) // 4 * 4)[:-1]]).astype(np.int64)
    tot = lambda off, s: int(off[-1] + s[-1] + 8) if len(off) else 8
    draw = (lambda t: torch.rand(t, device="cuda", generator=g) * 2 - 1) if fill == "uniform" else \
        (lambda t: torch.randn(t, device="cuda", generator=g))
    A16, B16 = draw(tot(aoff, sa)).to(torch.bfloat16), draw(tot(boff, sb)).to(torch.bfloat16)
    A32, B32 = A16.float(), B16.float()
    C = torch.zeros(max(int(((sc + 3) // 4 * 4).sum()), 1), device="cuda")
    ptr = lambda T, off, es: torch.from_numpy(T.data_ptr() + es * off).to("cuda")
    return ((A16, B16), (A32, B32, C), (aoff, boff, coff), (ptr(A16, aoff, 2), ptr(B16, boff, 2)),
            (ptr(A32, aoff, 4), ptr(B32, boff, 4), ptr(C, coff, 4)))


def _run_both(plan, alpha, beta, p16, p32, C, c0):
    """C from the bf16-operand exec and from the fp32-operand exec (same plan,
    same initial C); returns both."""
    C.copy_(torch.from_numpy(c0))
    plan.exec(alpha, p32[0], p32[1], beta, p32[2])
    torch.cuda.synchronize()
    ref = C.clone()
    C.copy_(torch.from_numpy(c0))
    plan.exec(alpha, p16[0], p16[1], beta, p32[2], operands="bf16")
    torch.cuda.synchronize()
    return C.clone(), ref


@pytest.mark.parametrize("table", [("default", "bf16"), ("all64", "bf16"), ("all128", "bf16")],
                         ids=["default", "all64", "all128"], indirect=True)
def test_ragged_bf16_operands(table):
    rng = np.random.default_rng(70)
    count = 400
    m = np.floor(np.exp(rng.uniform(0, np.log(300), count))).astype(np.int64)
    n = np.floor(np.exp(rng.uniform(0, np.log(300), count))).astype(np.int64)
    k = np.floor(np.exp(rng.uniform(0, np.log(300), count))).astype(np.int64)
    m[3], n[5], k[7] = 0, 0, 0
    k = _min8(m, n, k)
    g = torch.Generator(device="cuda").manual_seed(71)
    b16, (A32, B32, C), offs, p16, p32 = _pack16(m, n, k, g, fill="normal")
    c0 = (torch.rand(C.numel(), device="cuda", generator=g) * 2 - 1).cpu().numpy()
    plan = batched.VBatchPlan(m, n, k, precision="bf16")
    assert plan.info()["kernel"] == 1.0
    got, ref = _run_both(plan, 1.25, -0.5, p16, p32, C, c0)
    assert torch.equal(got, ref)
    _check_problems("fp32", m, n, k, (A32, B32, got), offs, 1.25, -0.5, c0, range(count))
    plan.close()


@pytest.mark.parametrize("table", [("all64", "bf16"), ("all128", "bf16")], ids=["all64", "all128"], indirect=True)
def test_tile_edges_bf16_operands(table):
    """m, n around the 32-row blocks and the tiles, k around the 8- and
    16-deep steps and the 32-deep k-tiles (the last k-tile's clamped and
    masked loads); beta != 0, and beta == 0 with a NaN C."""
    edges = [1, 2, 3, 7, 8, 9, 31, 32, 33, 63, 64, 65, 96, 127, 128, 129, 191, 192, 255, 257]
    ks = [0, 1, 2, 7, 8, 9, 15, 16, 17, 31, 32, 33, 40, 63, 64, 65, 100]
    rng = np.random.default_rng(72)
    cnt = 160
    m = np.array([edges[i % len(edges)] for i in range(cnt)], np.int64)
    n = np.array([edges[(7 * i + 3) % len(edges)] for i in range(cnt)], np.int64)
    k = np.array([ks[(5 * i + 1) % len(ks)] for i in range(cnt)], np.int64)
    rng.shuffle(k)
    k = _min8(m, n, k)
    g = torch.Generator(device="cuda").manual_seed(73)
    b16, (A32, B32, C), offs, p16, p32 = _pack16(m, n, k, g, fill="normal")
    c0 = (torch.rand(C.numel(), device="cuda", generator=g) * 2 - 1).cpu().numpy()
    plan = batched.VBatchPlan(m, n, k, precision="bf16")
    assert plan.info()["kernel"] == 1.0
    got, ref = _run_both(plan, 0.75, 1.5, p16, p32, C, c0)
    assert torch.equal(got, ref)
    _check_problems("fp32", m, n, k, (A32, B32, got), offs, 0.75, 1.5, c0, range(cnt))
    C.fill_(float("nan"))
    plan.exec(0.75, p16[0], p16[1], 0.0, p32[2], operands="bf16")
    torch.cuda.synchronize()
    _check_problems("fp32", m, n, k, (A32, B32, C), offs, 0.75, 0.0, np.zeros_like(c0), range(cnt))
    plan.close()


@pytest.mark.parametrize("table", [("default", "bf16"), ("mixed", "bf16")], ids=["default", "mixed"], indirect=True)
def test_config4_bf16_operands(table):
    """Config size: 4096 N,N problems, m, n, k log-uniform in [16, 512], the
    U[-1,1] operands rounded to bf16 (nearest even) once, alpha=1, beta=0,
    ld = rows.  Bit-identical to the fp32-operand bf16 path on the unrounded
    fp32 operands; every problem on 8 sampled entries, every 64th in full."""
    m, n, k = batched.config4_shapes()
    g = torch.Generator(device="cuda").manual_seed(42)
    bufs, offs, ptrs = batched.pack_problems(torch, m, n, k, gen=g)
    (A16, B16), (pa16, pb16) = batched.to_bf16(torch, bufs[:2], offs[:2])
    plan = batched.VBatchPlan(m, n, k, precision="bf16")
    bufs[2].fill_(float("nan"))
    plan.exec(1.0, ptrs[0], ptrs[1], 0.0, ptrs[2])
    torch.cuda.synchronize()
    ref = bufs[2].clone()
    bufs[2].fill_(float("nan"))
    plan.exec(1.0, pa16, pb16, 0.0, ptrs[2], operands="bf16")
    torch.cuda.synchronize()
    assert torch.equal(bufs[2].view(torch.int32), ref.view(torch.int32))  # bitwise (C's padding holds NaN)
    c0 = np.zeros(bufs[2].numel(), np.float32)
    exact = (A16.float(), B16.float(), bufs[2])
    _check_problems("fp32", m, n, k, exact, offs, 1.0, 0.0, c0, range(len(m)), sample=8)
    _check_problems("fp32", m, n, k, exact, offs, 1.0, 0.0, c0, range(0, len(m), 64))
    plan.close()


def test_padded_leading_dimensions_bf16_operands():
    """lda > m, ldb > k, ldc > m with odd paddings: columns start in either
    half of a dword; C's padding rows stay untouched."""
    rng = np.random.default_rng(74)
    count = 40
    m = rng.integers(1, 200, count)
    n = rng.integers(1, 200, count)
    k = _min8(m, n, rng.integers(1, 200, count))
    lda, ldb, ldc = m + rng.integers(1, 6, count), k + rng.integers(1, 6, count), m + rng.integers(1, 6, count)
    g = torch.Generator(device="cuda").manual_seed(75)
    As = [(torch.rand(int(lda[p] * k[p]), device="cuda", generator=g) * 2 - 1).to(torch.bfloat16)
          for p in range(count)]
    Bs = [(torch.rand(int(ldb[p] * n[p]), device="cuda", generator=g) * 2 - 1).to(torch.bfloat16)
          for p in range(count)]
    Cs = [torch.rand(int(ldc[p] * n[p]), device="cuda", generator=g) * 2 - 1 for p in range(count)]
    C0 = [c.clone() for c in Cs]
    ptr = lambda ts: torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device="cuda")
    plan = batched.VBatchPlan(m, n, k, lda=lda, ldb=ldb, ldc=ldc, precision="bf16")
    plan.exec(0.75, ptr(As), ptr(Bs), 1.5, ptr(Cs), operands="bf16")
    torch.cuda.synchronize()
    got16 = [c.clone() for c in Cs]
    A32, B32 = [a.float() for a in As], [b.float() for b in Bs]
    for c, c0 in zip(Cs, C0):
        c.copy_(c0)
    plan.exec(0.75, ptr(A32), ptr(B32), 1.5, ptr(Cs))
    torch.cuda.synchronize()
    for p in range(count):
        assert torch.equal(got16[p], Cs[p]), p
        ld = int(ldc[p])
        pad = torch.ones(ld, int(n[p]), dtype=torch.bool, device="cuda")
        pad[:int(m[p])] = False
        view = lambda t: t.view(int(n[p]), ld).t()
        assert torch.equal(view(got16[p])[pad], view(C0[p])[pad])
    # the fp32-operand path on these (exact) values is itself checked against
    # the oracle with padded leading dimensions in test_batched_gpu.py
    plan.close()


def test_graph_capture_replays_bf16_operands():
    m, n, k = batched.config4_shapes(count=512, seed=7)
    g = torch.Generator(device="cuda").manual_seed(76)
    bufs, offs, ptrs = batched.pack_problems(torch, m, n, k, gen=g)
    _, (pa16, pb16) = batched.to_bf16(torch, bufs[:2], offs[:2])
    plan = batched.VBatchPlan(m, n, k, precision="bf16")
    plan.exec(1.0, pa16, pb16, 0.0, ptrs[2], operands="bf16")
    torch.cuda.synchronize()
    eager = bufs[2].clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.exec(1.0, pa16, pb16, 0.0, ptrs[2], operands="bf16")
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.exec(1.0, pa16, pb16, 0.0, ptrs[2], operands="bf16")
    for rep in range(3):
        bufs[2].zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(bufs[2], eager), "replay %d differs from eager" % rep
        bufs[2].mul_(2.0)
        torch.cuda.synchronize()
    plan.close()


def test_bf16_operand_exec_rejections(monkeypatch):
    """Refused before any kernel: an fp32 plan, an operand under eight
    elements, a plan on the work-queue kernels."""
    p = torch.zeros(1, dtype=torch.int64, device="cuda")
    plan = batched.VBatchPlan([16], [16], [16], precision="fp32")
    with pytest.raises(mb.HipBlasError, match="bf16 operands need a BF16 plan"):
        plan.exec(1.0, p, p, 0.0, p, operands="bf16")
    plan.close()
    plan = batched.VBatchPlan([2], [16], [3], precision="bf16")  # A: 6 elements
    with pytest.raises(mb.HipBlasError, match="problem 0 has an operand of fewer than 8 elements"):
        plan.exec(1.0, p, p, 0.0, p, operands="bf16")
    plan.close()
    monkeypatch.setenv("MBH_VBATCH_KERNEL", "lean")
    plan = batched.VBatchPlan([16], [16], [16], precision="bf16")
    with pytest.raises(mb.HipBlasError, match="work-queue kernels"):
        plan.exec(1.0, p, p, 0.0, p, operands="bf16")
    plan.close()


def test_alpha_zero_and_empty_bf16_operands():
    """alpha == 0 scales C by beta without reading A or B (beta == 1 leaves C
    untouched); a batch whose problems are all empty or k == 0 runs."""
    m = np.array([16, 9, 0, 33], np.int64)
    n = np.array([16, 0, 5, 17], np.int64)
    k = np.array([16, 8, 8, 0], np.int64)
    g = torch.Generator(device="cuda").manual_seed(77)
    b16, (A32, B32, C), offs, p16, p32 = _pack16(m, n, k, g)
    c0 = (torch.rand(C.numel(), device="cuda", generator=g) * 2 - 1).cpu().numpy()
    plan = batched.VBatchPlan(m, n, k, precision="bf16")
    C.copy_(torch.from_numpy(c0))
    plan.exec(0.0, p16[0], p16[1], 1.0, p32[2], operands="bf16")
    torch.cuda.synchronize()
    assert np.array_equal(C.cpu().numpy(), c0)
    plan.exec(0.0, p16[0], p16[1], -2.0, p32[2], operands="bf16")
    torch.cuda.synchronize()
    _check_problems("fp32", m, n, k, (A32, B32, C), offs, 0.0, -2.0, c0, range(len(m)))
    # k == 0 problem with alpha != 0: C := beta C
    C.copy_(torch.from_numpy(c0))
    plan.exec(1.5, p16[0], p16[1], 0.5, p32[2], operands="bf16")
    torch.cuda.synchronize()
    _check_problems("fp32", m, n, k, (A32, B32, C), offs, 1.5, 0.5, c0, range(len(m)))
    plan.close()
