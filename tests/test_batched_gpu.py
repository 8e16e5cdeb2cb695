"""GPU parity of the variable-size batched gemm, problem by problem, against
the fp64 oracle (the CPU gemm), in fp32 and bf16-input modes; plus hipGraph
capture and replay of the exec.

Bounds per problem (componentwise): (e_rep + (2k+2) u)|alpha||A||B| + 4u|beta C|,
e_rep = 0 for fp32 and 2*2^-8 + 2^-16 for bf16 operands (|a - rn_bf16(a)| <=
2^-8 |a| per operand).
"""
import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
import mojoblas_amd as mb
from mojoblas_amd import batched

pytestmark = pytest.mark.gpu
E_REP = {"fp32": 0.0, "bf16": 2 * 2.0 ** -8 + 2.0 ** -16}


def _check_problems(prec, m, n, k, bufs, offs, alpha, beta, c0, which, sample=None):
    A, B, C = (t.cpu().numpy() for t in bufs)
    aoff, boff, coff = offs
    for p in which:
        mp, np_, kp = int(m[p]), int(n[p]), int(k[p])
        if mp == 0 or np_ == 0:
            continue
        a = A[aoff[p]:aoff[p] + mp * kp]
        b = B[boff[p]:boff[p] + kp * np_]
        got = C[coff[p]:coff[p] + mp * np_].astype(np.float64)
        cin = c0[coff[p]:coff[p] + mp * np_]
        if sample is None:
            want = cin.astype(np.float64).copy()
            oracle.dgemm("N", "N", mp, np_, kp, alpha, a.astype(np.float64), max(mp, 1), b.astype(np.float64),
                         max(kp, 1), beta, want, max(mp, 1))
            absab = (np.abs(a.reshape(kp, mp).T.astype(np.float64)) @ np.abs(b.reshape(np_, kp).T.astype(np.float64))
                     ).T.ravel() if kp else np.zeros(mp * np_)
            idx = slice(None)
        else:
            rng = np.random.default_rng(p)
            rows = rng.integers(0, mp, sample)
            cols = rng.integers(0, np_, sample)
            want = oracle.sgemm_sample("N", "N", kp, alpha, a, max(mp, 1), b, max(kp, 1), beta,
                                       cin[rows + cols * mp], rows, cols)
            absab = oracle.sgemm_sample_abs("N", "N", kp, a, max(mp, 1), b, max(kp, 1), rows, cols)
            idx = rows + cols * mp
            cin = cin[idx]
        e = E_REP[prec] + (2.0 * max(kp, 1) + 2.0) * H.U32
        bound = e * abs(alpha) * absab + 4 * H.U32 * np.abs(beta * cin.astype(np.float64)) + 1e-30
        err = np.abs(got[idx] - want)
        assert (err <= bound).all(), (prec, p, mp, np_, kp, float((err / bound).max()))


# Tile-class tables: "all64" puts every problem on 64x64 tiles, "all128" on
# 128x128 tiles, "mixed" 128x128 once both m and n exceed 64 (two queues);
# "default" is the compiled (tuned) table.
_MIXED = np.zeros((6, 6), np.int32)
_MIXED[3:, 3:] = 1
TABLES = {"default": None, "all64": np.zeros((6, 6), np.int32), "all128": np.ones((6, 6), np.int32),
          "mixed": _MIXED}


@pytest.fixture
def table(request):
    name, prec = request.param
    if TABLES[name] is not None:
        batched.set_tuning(TABLES[name], 0, prec)
    yield name
    batched.set_tuning(None, 0, prec)


def _cases(tables):
    return [pytest.param(p, (t, p), id="%s-%s" % (p, t)) for p in ("fp32", "bf16") for t in tables]


# Kernel path: "ws" = warp-specialized tiles on fixed per-workgroup lists
# (the default when every problem with k > 0 has at least four elements in
# each operand), "lean" = the work-queue tiles (MBH_VBATCH_KERNEL=lean, and
# the path for batches holding smaller operands).
@pytest.fixture(params=["ws", "lean"])
def path(request, monkeypatch):
    if request.param == "lean":
        monkeypatch.setenv("MBH_VBATCH_KERNEL", "lean")
    else:
        monkeypatch.delenv("MBH_VBATCH_KERNEL", raising=False)
    return request.param


def _ws_k(k, path):
    """k >= 4 where k > 0, so that a ws batch is eligible for the ws kernel."""
    return np.where((k > 0) & (k < 4), 4, k) if path == "ws" else k


def _assert_path(plan, path):
    assert plan.info()["kernel"] == (1.0 if path == "ws" else 0.0), plan.info()


@pytest.mark.parametrize("prec,table", _cases(TABLES), indirect=["table"])
def test_ragged_batch_vs_oracle(prec, table):
    rng = np.random.default_rng(40)
    count = 96
    m = np.floor(np.exp(rng.uniform(0, np.log(300), count))).astype(np.int64)
    n = np.floor(np.exp(rng.uniform(0, np.log(300), count))).astype(np.int64)
    k = np.floor(np.exp(rng.uniform(0, np.log(300), count))).astype(np.int64)
    m[3], n[5], k[7] = 0, 0, 0  # empty problems and a k == 0 problem
    g = torch.Generator(device="cuda").manual_seed(41)
    bufs, offs, ptrs = batched.pack_problems(torch, m, n, k, gen=g)
    bufs[2].uniform_(-1, 1, generator=g)
    c0 = bufs[2].cpu().numpy().copy()
    plan = batched.VBatchPlan(m, n, k, precision=prec)
    plan.exec(1.25, ptrs[0], ptrs[1], -0.5, ptrs[2])
    torch.cuda.synchronize()
    _check_problems(prec, m, n, k, bufs, offs, 1.25, -0.5, c0, range(count))
    info = plan.info()
    assert info["count"] == count and info["problems_64"] + info["problems_128"] == count - 2
    plan.close()


@pytest.mark.parametrize("prec,table", _cases(("default", "all64", "all128")), indirect=["table"])
def test_ragged_ws_lists_vs_oracle(prec, table):
    """The warp-specialized path on a ragged batch (m, n from 1, k from 4,
    empty and k == 0 problems), alpha, beta != 0, every problem in full;
    more problems than workgroups on some lists and empty lists on others."""
    rng = np.random.default_rng(46)
    count = 700
    m = np.floor(np.exp(rng.uniform(0, np.log(400), count))).astype(np.int64)
    n = np.floor(np.exp(rng.uniform(0, np.log(400), count))).astype(np.int64)
    k = np.floor(np.exp(rng.uniform(np.log(4), np.log(400), count))).astype(np.int64)
    m[3], n[5], k[7], k[11] = 0, 0, 0, 0
    g = torch.Generator(device="cuda").manual_seed(47)
    bufs, offs, ptrs = batched.pack_problems(torch, m, n, k, gen=g, fill="normal")
    bufs[2].uniform_(-1, 1, generator=g)
    c0 = bufs[2].cpu().numpy().copy()
    plan = batched.VBatchPlan(m, n, k, precision=prec)
    assert plan.info()["kernel"] == 1.0
    plan.exec(1.25, ptrs[0], ptrs[1], -0.5, ptrs[2])
    torch.cuda.synchronize()
    _check_problems(prec, m, n, k, bufs, offs, 1.25, -0.5, c0, range(count))
    # a second exec of the same plan (the records are rewritten each exec)
    bufs[2].copy_(torch.from_numpy(c0))
    plan.exec(1.25, ptrs[0], ptrs[1], -0.5, ptrs[2])
    torch.cuda.synchronize()
    _check_problems(prec, m, n, k, bufs, offs, 1.25, -0.5, c0, range(count))
    plan.close()


@pytest.mark.parametrize("prec,table", _cases(("all64", "all128")), indirect=["table"])
def test_tile_edges(prec, table, path):
    """Every m and n around the 32-row blocks and the 64 / 128 tiles (a load,
    MFMA block or store is skipped when it lies wholly past the edge), k
    around the 8- and 16-deep steps and the 32-deep k-tiles, operands of
    fewer than four elements; each problem checked in full."""
    edges = [1, 2, 3, 4, 5, 31, 32, 33, 63, 64, 65, 96, 97, 127, 128, 129, 160, 191, 192, 255, 257]
    ks = [0, 1, 2, 3, 7, 8, 9, 15, 16, 17, 31, 32, 33, 40, 63, 64, 65, 100]
    rng = np.random.default_rng(44)
    m = np.array([edges[i % len(edges)] for i in range(160)], np.int64)
    n = np.array([edges[(7 * i + 3) % len(edges)] for i in range(160)], np.int64)
    k = np.array([ks[(5 * i + 1) % len(ks)] for i in range(160)], np.int64)
    rng.shuffle(k)
    k = _ws_k(k, path)
    g = torch.Generator(device="cuda").manual_seed(45)
    bufs, offs, ptrs = batched.pack_problems(torch, m, n, k, gen=g, fill="normal")
    bufs[2].uniform_(-1, 1, generator=g)
    c0 = bufs[2].cpu().numpy().copy()
    plan = batched.VBatchPlan(m, n, k, precision=prec)
    _assert_path(plan, path)
    plan.exec(0.75, ptrs[0], ptrs[1], 1.5, ptrs[2])
    torch.cuda.synchronize()
    _check_problems(prec, m, n, k, bufs, offs, 0.75, 1.5, c0, range(len(m)))
    # beta == 0 never reads C (NaN stays out of the result)
    bufs[2].fill_(float("nan"))
    plan.exec(0.75, ptrs[0], ptrs[1], 0.0, ptrs[2])
    torch.cuda.synchronize()
    _check_problems(prec, m, n, k, bufs, offs, 0.75, 0.0, np.zeros_like(c0), range(len(m)))
    plan.close()


@pytest.mark.parametrize("prec,table", _cases(("default", "all64", "mixed")), indirect=["table"])
def test_config4_4096_problems(prec, table, path):
    """Config size: 4096 N,N problems, m, n, k log-uniform in [16, 512],
    U[-1,1], alpha=1, beta=0, ld = rows.  Every problem is priced on 8 sampled
    entries by the oracle, every 64th problem in full."""
    m, n, k = batched.config4_shapes()
    g = torch.Generator(device="cuda").manual_seed(42)
    bufs, offs, ptrs = batched.pack_problems(torch, m, n, k, gen=g)
    bufs[2].fill_(float("nan"))  # beta == 0: C is never read
    c0 = np.zeros(bufs[2].numel(), np.float32)
    plan = batched.VBatchPlan(m, n, k, precision=prec)
    _assert_path(plan, path)
    plan.exec(1.0, ptrs[0], ptrs[1], 0.0, ptrs[2])
    torch.cuda.synchronize()
    _check_problems(prec, m, n, k, bufs, offs, 1.0, 0.0, c0, range(len(m)), sample=8)
    _check_problems(prec, m, n, k, bufs, offs, 1.0, 0.0, c0, range(0, len(m), 64))
    plan.close()


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_graph_capture_replays_match_eager(prec, path):
    m, n, k = batched.config4_shapes(count=512, seed=7)
    g = torch.Generator(device="cuda").manual_seed(43)
    bufs, offs, ptrs = batched.pack_problems(torch, m, n, k, gen=g)
    plan = batched.VBatchPlan(m, n, k, precision=prec)
    _assert_path(plan, path)
    plan.exec(1.0, ptrs[0], ptrs[1], 0.0, ptrs[2])
    torch.cuda.synchronize()
    eager = bufs[2].clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        plan.exec(1.0, ptrs[0], ptrs[1], 0.0, ptrs[2])  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        plan.exec(1.0, ptrs[0], ptrs[1], 0.0, ptrs[2])
    for rep in range(3):
        bufs[2].zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(bufs[2], eager), "replay %d differs from eager" % rep
        # eager work between replays
        bufs[2].mul_(2.0)
        torch.cuda.synchronize()
    plan.close()


@pytest.mark.parametrize("prec,table", _cases(("all64", "all128", "mixed")), indirect=["table"])
def test_padded_leading_dimensions(prec, table, path):
    """lda > m, ldb > k, ldc > m with odd paddings (columns off 16-byte
    alignment); the padding rows of C must stay untouched."""
    rng = np.random.default_rng(44)
    count = 40
    m = rng.integers(1, 200, count)
    n = rng.integers(1, 200, count)
    k = _ws_k(rng.integers(1, 200, count), path)
    lda, ldb, ldc = m + rng.integers(1, 6, count), k + rng.integers(1, 6, count), m + rng.integers(1, 6, count)
    g = torch.Generator(device="cuda").manual_seed(45)
    As = [torch.rand(int(lda[p] * k[p]), device="cuda", generator=g) * 2 - 1 for p in range(count)]
    Bs = [torch.rand(int(ldb[p] * n[p]), device="cuda", generator=g) * 2 - 1 for p in range(count)]
    Cs = [torch.rand(int(ldc[p] * n[p]), device="cuda", generator=g) * 2 - 1 for p in range(count)]
    C0 = [c.cpu().numpy().copy() for c in Cs]
    ptr = lambda ts: torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device="cuda")
    plan = batched.VBatchPlan(m, n, k, lda=lda, ldb=ldb, ldc=ldc, precision=prec)
    _assert_path(plan, path)
    alpha, beta = 0.75, 1.5
    plan.exec(alpha, ptr(As), ptr(Bs), beta, ptr(Cs))
    torch.cuda.synchronize()
    for p in range(count):
        mp, np_, kp = int(m[p]), int(n[p]), int(k[p])
        a = As[p].cpu().numpy().astype(np.float64)
        b = Bs[p].cpu().numpy().astype(np.float64)
        want = C0[p].astype(np.float64).copy()
        oracle.dgemm("N", "N", mp, np_, kp, alpha, a, int(lda[p]), b, int(ldb[p]), beta, want, int(ldc[p]))
        got = Cs[p].cpu().numpy().astype(np.float64)
        am = a.reshape(kp, int(lda[p])).T[:mp]
        bm = b.reshape(np_, int(ldb[p])).T[:kp]
        absab = np.zeros((int(ldc[p]), np_))
        absab[:mp] = np.abs(am) @ np.abs(bm)
        absab = absab.T.ravel()
        e = E_REP[prec] + (2.0 * kp + 2.0) * H.U32
        bound = e * abs(alpha) * absab + 4 * H.U32 * np.abs(beta * C0[p].astype(np.float64)) + 1e-30
        err = np.abs(got - want)
        assert (err <= bound).all(), (prec, p, mp, np_, kp, float((err / bound).max()))
        pad = np.ones((int(ldc[p]), np_), bool)
        pad[:mp] = False
        assert np.array_equal(got.reshape(np_, int(ldc[p])).T[pad], C0[p].reshape(np_, int(ldc[p])).T[pad])
    plan.close()


def test_invalid_problem_is_rejected_before_gpu_work():
    with pytest.raises(mb.HipBlasError, match="problem 1 fails gemm argument 8"):
        batched.VBatchPlan([4, 4], [4, 4], [4, 4], lda=[4, 3])


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_operand_of_2_30_elements_takes_64bit_offsets(prec):
    """A problem whose A spans 2^30 elements (lda = 2^20, k = 1024: 4 GiB)
    does not fit the 32-bit offsets of the other kernels, so the plan runs
    the two-class kernels with 64-bit offsets (plan_info kernel = 2); both
    problems are checked in full against the fp64 oracle bound."""
    m, n, k = np.array([64, 100]), np.array([64, 70]), np.array([1024, 50])
    lda = np.array([1 << 20, 100])
    g = torch.Generator(device="cuda").manual_seed(48)
    A = [torch.rand(int((k[p] - 1) * lda[p] + m[p]), device="cuda", generator=g) * 2 - 1 for p in range(2)]
    B = [torch.rand(int(k[p] * n[p]), device="cuda", generator=g) * 2 - 1 for p in range(2)]
    C = [torch.rand(int(m[p] * n[p]), device="cuda", generator=g) * 2 - 1 for p in range(2)]
    C0 = [c.cpu().numpy().astype(np.float64) for c in C]
    ptr = lambda ts: torch.tensor([t.data_ptr() for t in ts], dtype=torch.int64, device="cuda")
    plan = batched.VBatchPlan(m, n, k, lda=lda, precision=prec)
    assert plan.info()["kernel"] == 2.0, plan.info()
    plan.exec(0.75, ptr(A), ptr(B), -1.25, ptr(C))
    torch.cuda.synchronize()
    for p in range(2):
        a = torch.as_strided(A[p], (int(m[p]), int(k[p])), (1, int(lda[p]))).cpu().numpy().astype(np.float64)
        b = B[p].cpu().numpy().astype(np.float64).reshape(int(n[p]), int(k[p])).T
        want = 0.75 * (a @ b) - 1.25 * C0[p].reshape(int(n[p]), int(m[p])).T
        got = C[p].cpu().numpy().astype(np.float64).reshape(int(n[p]), int(m[p])).T
        e = E_REP[prec] + (2.0 * int(k[p]) + 2.0) * H.U32
        bound = e * 0.75 * (np.abs(a) @ np.abs(b)) + 4 * H.U32 * 1.25 * np.abs(C0[p].reshape(int(n[p]), int(m[p])).T)
        assert (np.abs(got - want) <= bound + 1e-30).all(), (prec, p, float((np.abs(got - want) / bound).max()))
    plan.close()
    del A
    torch.cuda.empty_cache()
