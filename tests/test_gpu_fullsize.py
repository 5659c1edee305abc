"""Parity at the benchmark configurations' full sizes, against the FP64
oracle on sampled rows (the oracle stays cheap; every sampled entry is exact
to the metric).

  clustered 4096³ (64 prototypes, 2% flips), FP32 plain and row-clustered
  products at prototype densities 0.5 / 0.1 / 0.03125 / 0.01;
  the same mask at BF16 / FP16 / BF16X2 with B ~ N(0,1) and with
  B = ±10^U[-3,3] (log-uniform magnitudes);
  value-carrying 4096² mask (iid 10%, values N(0, 0.02²)) times 4096 x 2048:
  forward, dB = Aᵀ·dC and dA = (dC·Bᵀ)⊙mask at FP32 and BF16;
  grouping of the 4096² mask (tile 2048 rows, 2 groups), bit-exact;
  the 32-problem grouped BF16 forward + backward, eager and graph replays.

Bounds (|c - r| / Σ_k |a_ik||b_kj| against the FP64 oracle with the
unrounded operands): FP32 1e-5, BF16 1e-2, FP16 2e-3, BF16X2 1e-5; with the
rounded operands (kernel mechanics: exact products, FP32 accumulation) 1e-5.
"""
import numpy as np
import pytest
import torch

from _helpers import rel_err, sample_rows, sampled_err, to_dev

pytestmark = pytest.mark.gpu

M = N = K = 4096
ROWS = np.arange(5, M, 61)  # 67 sampled rows, spread over every row block


@pytest.fixture(scope="module")
def b_sets(bws, dev):
    out = {}
    for kind, args in (("normal", (0.0, 1.0)), ("logu", (-3.0, 3.0))):
        B = bws.gen_f32(1, (K, N), kind, *args)
        out[kind] = (B, to_dev(B, dev))
    return out


def clustered(bws, dev, p, seed=0):
    W = bws.gen_bitmap_clustered(seed, M, K, 64, p, 0.02)
    return W, bws.from_host_words(W, M, K, dev)


@pytest.mark.parametrize("p", [0.5, 0.1, 0.03125, 0.01])
def test_fp32_clustered_sweep(oracle, bws, dev, b_sets, p):
    """Both FP32 kernels on the clustered 4096³ mask at each sweep density."""
    B, Bd = b_sets["normal"]
    W, A = clustered(bws, dev, p)
    C = bws.spmm(A, Bd, "fp32")
    e = sampled_err(oracle, C, W, K, B, ROWS)
    assert e["max_rel_s"] <= 1e-5, e
    pl = bws.cluster_rows(A)
    Cc = bws.spmm_clustered(A, pl, Bd)
    ec = sampled_err(oracle, Cc, W, K, B, ROWS)
    assert ec["max_rel_s"] <= 1e-5, ec


@pytest.mark.parametrize("bkind", ["normal", "logu"])
@pytest.mark.parametrize("prec", ["bf16", "fp16", "bf16x2", "fp32"])
def test_precision_clustered_4096(oracle, bws, dev, b_sets, prec, bkind):
    """The precision configuration: the clustered 4096³ mask (5% dense) with B
    ~ N(0,1) and B = ±10^U[-3,3]; 16-bit B given as 16-bit and as FP32 (rounded
    once on the device), split BF16 from FP32."""
    B, Bd = b_sets[bkind]
    W, A = clustered(bws, dev, 0.03125)
    bound = {"bf16": 1e-2, "fp16": 2e-3, "bf16x2": 1e-5, "fp32": 1e-5}[prec]
    if prec in ("bf16", "fp16"):
        dt = torch.bfloat16 if prec == "bf16" else torch.float16
        rnd = oracle.round_bf16 if prec == "bf16" else oracle.round_f16
        C16 = bws.spmm(A, Bd.to(dt), prec)
        Cf = bws.spmm(A, Bd, prec)          # FP32 B rounded once on the device
        assert torch.equal(C16, Cf)
        assert sampled_err(oracle, C16, W, K, rnd(B), ROWS)["max_rel_s"] <= 1e-5
        e = sampled_err(oracle, C16, W, K, B, ROWS)
    elif prec == "bf16x2":
        C = bws.spmm(A, Bd, prec)
        hi, lo = oracle.split_bf16x2(B)
        assert sampled_err(oracle, C, W, K, (hi + lo).astype(np.float32), ROWS)["max_rel_s"] <= 1e-5
        e = sampled_err(oracle, C, W, K, B, ROWS)
    else:
        C = bws.spmm(A, Bd, prec)
        e = sampled_err(oracle, C, W, K, B, ROWS)
        pl = bws.cluster_rows(A)
        ec = sampled_err(oracle, bws.spmm_clustered(A, pl, Bd), W, K, B, ROWS)
        assert ec["max_rel_s"] <= bound, ec
    assert e["max_rel_s"] <= bound, e


@pytest.fixture(scope="module")
def value_mask(oracle, bws, dev):
    """iid 10% value mask 4096 x 4096, values N(0, 0.02²); B 4096 x 2048; dC."""
    Mv = Kv = 4096
    Nv = 2048
    W = bws.gen_bitmap_iid(3, Mv, Kv, 0.1)
    rp = oracle.row_ptr(W, Kv)
    vals = bws.gen_f32(4, int(rp[-1]), "normal", 0.0, 0.02)
    vals[vals == 0] = 1e-3
    B = bws.gen_f32(5, (Kv, Nv), "normal", 0.0, 1.0)
    dC = bws.gen_f32(6, (Mv, Nv), "normal", 0.0, 1.0)
    A = bws.Bitmap(to_dev(W, dev), Mv, Kv, to_dev(rp, dev), to_dev(vals, dev))
    return dict(W=W, rp=rp, vals=vals, B=B, dC=dC, A=A, Bd=to_dev(B, dev), dCd=to_dev(dC, dev),
                M=Mv, K=Kv, N=Nv)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_value_mask_forward_and_dB(oracle, bws, dev, value_mask, prec):
    v = value_mask
    W, rp, vals, K_ = v["W"], v["rp"], v["vals"], v["K"]
    dt = torch.bfloat16 if prec == "bf16" else torch.float32
    bound = 1e-2 if prec == "bf16" else 1e-5
    rows = np.arange(3, v["M"], 61)
    C = bws.spmm(v["A"], v["Bd"].to(dt), prec)
    e = sampled_err(oracle, C, W, K_, v["B"], rows, rp, vals)
    assert e["max_rel_s"] <= bound, e
    if prec == "bf16":  # mechanics: exact against the rounded values and B
        em = sampled_err(oracle, C, W, K_, oracle.round_bf16(v["B"]), rows, rp,
                         oracle.round_bf16(vals))
        assert em["max_rel_s"] <= 1e-5, em
    # dB = Aᵀ·dC through the transposed value bitmap
    AT = bws.transpose(v["A"])
    dB = bws.spmm_tA(AT, v["dCd"].to(dt), prec)
    WT, rpt, vt = oracle.transpose_values(W, v["M"], K_, rp, vals)
    rowsT = np.arange(7, K_, 61)
    e = sampled_err(oracle, dB, WT, v["M"], v["dC"], rowsT, rpt, vt)
    assert e["max_rel_s"] <= bound, e


def test_value_mask_dA_bf16(oracle, bws, dev, value_mask):
    """dA = (dC·Bᵀ)⊙mask on MFMA at BF16, sampled rows of the mask."""
    v = value_mask
    dA = bws.sddmm(v["A"], v["dCd"].to(torch.bfloat16), v["Bd"].to(torch.bfloat16), "bf16")
    rows = np.arange(11, v["M"], 97)
    Ws, rps, _ = sample_rows(v["W"], rows, v["rp"], v["vals"])
    dCs = np.ascontiguousarray(v["dC"][rows])
    R = oracle.sddmm_f64(Ws, v["K"], rps, dCs, v["B"])
    S = oracle.sddmm_f64(Ws, v["K"], rps, dCs, v["B"], absolute=True)
    got = np.concatenate([dA[int(v["rp"][r]):int(v["rp"][r + 1])].cpu().numpy() for r in rows])
    assert np.max(np.abs(got - R) / np.maximum(S, 1e-30)) <= 1e-2
    Rr = oracle.sddmm_f64(Ws, v["K"], rps, oracle.round_bf16(dCs), oracle.round_bf16(v["B"]))
    Sr = oracle.sddmm_f64(Ws, v["K"], rps, oracle.round_bf16(dCs), oracle.round_bf16(v["B"]),
                          absolute=True)
    assert np.max(np.abs(got - Rr) / np.maximum(Sr, 1e-30)) <= 1e-5


def test_grouping_clustered_4096(oracle, bws, dev):
    """Row grouping of the 4096² clustered mask at the reference's tile height
    (2048 rows) and group count (2): every output array bit-exact."""
    W, A = clustered(bws, dev, 0.03125)
    ref = oracle.group_rows(W, K, 2048, 2)
    g = bws.group_rows(A, 2048, 2)
    torch.cuda.synchronize()
    for k, b in ref.items():
        a = g[k].cpu().numpy()
        if k == "group_mask":
            a = a.view(np.uint32)
        assert np.array_equal(a, b), k


def test_grouped_32_problems_graph(oracle, bws, dev):
    """The grouped workload (32 problems, M, K in [256, 8192], N = 1024, BF16,
    forward + backward, autotuned tiles): the eager run against the FP64
    oracle on sampled rows of every problem, then one captured hipGraph
    replayed three times with eager work between, each replay equal to the
    eager result bit for bit."""
    import bench
    Nb = 1024
    probs = bench.grouped_problems(bws, dev, 32, Nb)
    fb, bb, (Cs, dBs) = bench.grouped_batches(bws, probs, Nb)
    assert any(L["tiles"] >= 384 for L in fb.launches)  # the whole-batch tile choice
    fb.run()
    bb.run()
    torch.cuda.synchronize()
    eager = [c.clone() for c in Cs] + [d.clone() for d in dBs]
    for i, p in enumerate(probs):
        Bh = p["B"].float().cpu().numpy()
        dCh = p["dC"].float().cpu().numpy()
        rows = np.unique(np.linspace(0, p["M"] - 1, 12).astype(np.int64))
        e = sampled_err(oracle, Cs[i], p["W"], p["K"], Bh, rows)
        assert e["max_rel_s"] <= 1e-5, (i, e)   # B is exactly BF16: mechanics bound
        WT = oracle.transpose_bitmap(p["W"], p["M"], p["K"])
        rowsT = np.unique(np.linspace(0, p["K"] - 1, 12).astype(np.int64))
        e = sampled_err(oracle, dBs[i], WT, p["M"], dCh, rowsT)
        assert e["max_rel_s"] <= 1e-5, (i, e)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            fb.run()
            bb.run()
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        for c in Cs + dBs:
            c.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(Cs + dBs, eager):
            assert torch.equal(a, b), rep
        torch.randn(1 << 20, device=dev).sum()  # eager work between replays
        torch.cuda.synchronize()
