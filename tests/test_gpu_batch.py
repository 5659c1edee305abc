"""GPU tests of the grouped launch (one kernel per tile configuration over
uneven problems), the autotuner, and hipGraph capture of a grouped
forward + backward."""
import numpy as np
import pytest
import torch

from _helpers import make_problem, rel_err, to_dev

pytestmark = pytest.mark.gpu


def _mixed(bws, oracle, dev, n, dt=torch.float32, seed=100, values_every=0, sizes=None):
    rng = np.random.default_rng(seed)
    probs, refs = [], []
    for i in range(n):
        if sizes:
            M, K = sizes[i]
        else:
            M, K = (int(x) for x in rng.integers(1, 40, 2) * 64)
        p = float(rng.uniform(0.01, 0.2))
        vals = bool(values_every) and i % values_every == 0
        A, W, rp, v, B, Bd = make_problem(bws, oracle, dev, M, K, 96, p, seed=seed + i,
                                          values=vals, clustered=(i % 2 == 1))
        Cd = torch.full((M, 96), float("nan"), device=dev)
        probs.append((A, Bd.to(dt), Cd))
        refs.append((W, K, B, rp if vals else None, v))
    return probs, refs


def test_batch_fp32_grouped(oracle, bws, dev):
    probs, refs = _mixed(bws, oracle, dev, 9, values_every=3)
    b = bws.spmm_batched(probs)
    kinds = {L["kind"] for L in b.launches}
    assert kinds == {1}, b.launches                   # all grouped vector-add
    assert sum(L["count"] for L in b.launches) == 9
    for (A, Bd, Cd), (W, K, B, rp, v) in zip(probs, refs):
        assert rel_err(oracle, Cd.cpu().numpy(), W, K, B, rp, v)["max_rel_s"] <= 1e-5


def test_batch_fp32_mixed_alignment(oracle, bws, dev):
    """An odd N and a strided B go to per-problem launches inside the batch."""
    probs, refs = _mixed(bws, oracle, dev, 4, seed=7)
    A, W, rp, v, B, Bd = make_problem(bws, oracle, dev, 200, 300, 37, 0.1, seed=9)
    Cd = torch.empty((200, 37), device=dev)
    probs.append((A, Bd, Cd))
    refs.append((W, 300, B, None, None))
    b = bws.spmm_batched(probs)
    assert sorted(L["kind"] for L in b.launches) == [1, 3]
    for (A, Bd, Cd), (W, K, B, rp, v) in zip(probs, refs):
        assert rel_err(oracle, Cd.cpu().numpy(), W, K, B, rp, v)["max_rel_s"] <= 1e-5


@pytest.mark.parametrize("prec,dt", [("bf16", torch.bfloat16), ("fp16", torch.float16),
                                     ("bf16x2", torch.float32)])
def test_batch_mfma_grouped(oracle, bws, dev, prec, dt):
    probs, refs = _mixed(bws, oracle, dev, 7, dt=dt, seed=200)
    b = bws.spmm_batched(probs, prec)
    assert {L["kind"] for L in b.launches} == {2}
    bound = {"bf16": 1e-2, "fp16": 2e-3, "bf16x2": 1e-5}[prec]
    for (A, Bd, Cd), (W, K, B, rp, v) in zip(probs, refs):
        assert rel_err(oracle, Cd.cpu().numpy(), W, K, B)["max_rel_s"] <= bound


def test_batch_plan_forces_tile(oracle, bws, dev):
    probs, refs = _mixed(bws, oracle, dev, 5, seed=300)
    plan = bws.make_plan(kernel=1, rows_per_wg=64, waves_per_wg=16, k_stage=32)
    b = bws.spmm_batched(probs, plan=plan)
    assert [(L["kind"], L["rows"], L["waves"], L["k_stage"]) for L in b.launches] == [(1, 64, 16, 32)]
    for (A, Bd, Cd), (W, K, B, rp, v) in zip(probs, refs):
        assert rel_err(oracle, Cd.cpu().numpy(), W, K, B)["max_rel_s"] <= 1e-5


def test_autotune_and_cache(oracle, bws, dev, tmp_path):
    bws.lib().bws_autotune_clear()
    A, W, rp, v, B, Bd = make_problem(bws, oracle, dev, 1024, 1024, 512, 0.05, seed=5)
    nnz = int(rp[-1])
    plan, ms = bws.autotune(A, Bd, "fp32", nnz=nnz)
    assert plan.kernel == 1 and ms > 0
    got = bws.autotune_lookup(1024, 1024, 512, nnz, "fp32")
    assert got is not None and got.rows_per_wg == plan.rows_per_wg
    # the tuned plan is valid and exact
    C = bws.spmm(A, Bd, plan=plan).cpu().numpy()
    assert rel_err(oracle, C, W, K=1024, B=B)["max_rel_s"] <= 1e-5
    planb, msb = bws.autotune(A, Bd.to(torch.bfloat16), "bf16", nnz=nnz)
    assert planb.kernel in (1, 2)
    path = str(tmp_path / "tune.txt").encode()
    assert bws.lib().bws_autotune_save(path) == 0
    bws.lib().bws_autotune_clear()
    assert bws.autotune_lookup(1024, 1024, 512, nnz, "fp32") is None
    assert bws.lib().bws_autotune_load(path) == 0
    assert bws.autotune_lookup(1024, 1024, 512, nnz, "fp32").rows_per_wg == plan.rows_per_wg
    # the batch picks the cached plan for problems with an nnz hint
    Cd = torch.empty((1024, 512), device=dev)
    b = bws.spmm_batched([(A, Bd, Cd)], nnz_hints=[nnz])
    assert b.launches[0]["rows"] == plan.rows_per_wg
    bws.lib().bws_autotune_clear()


def test_graph_capture_batch_fwd_bwd(oracle, bws, dev):
    """One captured hipGraph: grouped forward (BF16 MFMA) and the grouped
    backward dB = Aᵀ·dC over transposed bitmaps; replayed several times with
    eager work in between, every replay checked against the eager result."""
    probs, refs = _mixed(bws, oracle, dev, 6, dt=torch.bfloat16, seed=400)
    bwd = []
    for (A, Bd, Cd) in probs:
        AT = bws.transpose(A)
        dC = torch.randn(A.rows, Bd.shape[1], device=dev).to(torch.bfloat16)
        dB = torch.empty((A.cols, Bd.shape[1]), device=dev)
        bwd.append((AT, dC, dB))
    fwd_b = bws.Batch(probs, "bf16")
    bwd_b = bws.Batch(bwd, "bf16")
    fwd_b.run()
    bwd_b.run()
    torch.cuda.synchronize()
    eager_C = [c.clone() for (_, _, c) in probs]
    eager_dB = [d.clone() for (_, _, d) in bwd]
    for _, _, c in probs:
        c.fill_(float("nan"))
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            fwd_b.run()
            bwd_b.run()
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        for _, _, c in probs:
            c.fill_(float("nan"))
        for _, _, d in bwd:
            d.zero_()
        x = torch.randn(512, 512, device=dev)      # unrelated eager work between replays
        (x @ x).sum().item()
        g.replay()
        torch.cuda.synchronize()
        for c, e in zip([c for (_, _, c) in probs], eager_C):
            assert torch.equal(c, e), rep
        for d, e in zip([d for (_, _, d) in bwd], eager_dB):
            assert torch.equal(d, e), rep
    # and the eager values are right
    for (A, Bd, Cd), (W, K, B, rp, v) in zip(probs, refs):
        assert rel_err(oracle, Cd.cpu().numpy(), W, K, B)["max_rel_s"] <= 1e-2
