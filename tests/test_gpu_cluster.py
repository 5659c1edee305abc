"""GPU parity of the row-clustering plan and the clustered FP32 product.

The plan is integer work and must match the oracle's restatement bit for
bit; the product is FP32 and must stay within 1e-5 of the FP64 oracle in the
|A|·|B| metric (the FP32 path's bar), on clustered masks (most rows summed as
deltas), iid masks (few or none), ragged shapes, alpha/beta, and the
benchmark's 4096-row clustered mask."""
import numpy as np
import pytest
import torch

from _helpers import rel_err, to_dev, to_host_u32

pytestmark = pytest.mark.gpu

PLAN_CASES = [(1, 1, 0, False, 0.5), (15, 33, 0, False, 0.3), (17, 100, 3, True, 0.1),
              (256, 64, 0, True, 0.2), (300, 1000, 7, False, 0.05), (1000, 513, 0, True, 0.05),
              (2048, 2048, 0, True, 0.03125), (512, 768, 512, False, 0.05), (129, 40, 1, True, 0.3),
              # five 1024-row ranking blocks, more than 256 leaders (stride leaders)
              (5000, 300, 300, True, 0.05), (3000, 96, 257, False, 0.1),
              # the benchmark mask: 64 leaders searched over 256 candidates
              (4096, 4096, 0, True, 0.03125)]


def _bitmap(bws, dev, M, K, p, clustered, seed):
    W = (bws.gen_bitmap_clustered(seed, M, K, 16, p, 0.02) if clustered
         else bws.gen_bitmap_iid(seed, M, K, p))
    return W, bws.Bitmap(to_dev(W, dev), M, K)


@pytest.mark.parametrize("M,K,L,clustered,p", PLAN_CASES)
def test_plan_matches_oracle(oracle, bws, dev, M, K, L, clustered, p):
    W, A = _bitmap(bws, dev, M, K, p, clustered, seed=M + K)
    pl = bws.cluster_rows(A, max_clusters=L)
    slot, cm = oracle.cluster_rows(W, K, max_clusters=L)
    assert np.array_equal(pl.slot_row.cpu().numpy(), slot)
    assert np.array_equal(to_host_u32(pl.chunk_mask), cm)
    # same plan twice (deterministic)
    pl2 = bws.cluster_rows(A, max_clusters=L)
    assert torch.equal(pl.slot_row, pl2.slot_row) and torch.equal(pl.chunk_mask, pl2.chunk_mask)


def test_plan_graph_capture(oracle, bws, dev):
    """The whole plan captured into one hipGraph (no host synchronisation, no
    memset node): three replays, with eager work between them, each equal to
    the oracle's plan."""
    M, K = 3000, 1000
    W, A = _bitmap(bws, dev, M, K, 0.05, True, seed=11)
    slot, cm = oracle.cluster_rows(W, K)
    bws.cluster_rows(A)  # warm-up outside the capture
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            pl = bws.cluster_rows(A)
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        pl.slot_row.fill_(-1)
        pl.chunk_mask.fill_(0)
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(pl.slot_row.cpu().numpy(), slot), rep
        assert np.array_equal(to_host_u32(pl.chunk_mask), cm), rep
        torch.randn(1 << 20, device=dev).cumsum(0)  # eager work between replays
        torch.cuda.synchronize()


def test_plan_padded_ldw(oracle, bws, dev):
    M, K, ldw = 100, 70, 5
    W0 = bws.gen_bitmap_clustered(9, M, K, 4, 0.2, 0.02)
    W = np.full((M, ldw), 0xFFFFFFFF, np.uint32)  # junk beyond the row's words
    W[:, :W0.shape[1]] = W0
    W[:, W0.shape[1] - 1] |= 0x3F  # junk bits beyond column K in the last word
    A = bws.Bitmap(to_dev(W, dev), M, K)
    pl = bws.cluster_rows(A, 4)
    slot, cm = oracle.cluster_rows(W, K, 4)
    assert np.array_equal(pl.slot_row.cpu().numpy(), slot)
    assert np.array_equal(to_host_u32(pl.chunk_mask), cm)


SHAPES = [(1, 1, 4, 1.0, True), (16, 64, 256, 0.2, True), (100, 70, 36, 0.2, False),
          (257, 513, 300, 0.05, True), (1024, 1024, 1024, 0.1, True), (512, 2048, 96, 0.01, False),
          (300, 32, 1000, 0.3, True), (1000, 130, 4, 0.1, True), (4100, 700, 260, 0.05, True)]


@pytest.mark.parametrize("M,K,N,p,clustered", SHAPES)
def test_spmm_clustered(oracle, bws, dev, M, K, N, p, clustered):
    W, A = _bitmap(bws, dev, M, K, p, clustered, seed=M + N)
    B = bws.gen_f32(M + 1, (K, N), "normal", 0.0, 1.0)
    Bd = to_dev(B, dev)
    pl = bws.cluster_rows(A)
    C = bws.spmm_clustered(A, pl, Bd).cpu().numpy()
    e = rel_err(oracle, C, W, K, B)
    assert e["max_rel_s"] <= 1e-5, e
    C0 = bws.spmm(A, Bd, "fp32").cpu().numpy()
    S = oracle.spmm_f64(W, K, B, absolute=True)
    assert np.all(np.abs(C - C0) <= 2e-5 * S + 1e-30)


def test_spmm_clustered_alpha_beta_strides(oracle, bws, dev, rng):
    M, K, N = 300, 200, 148
    W, A = _bitmap(bws, dev, M, K, 0.1, True, seed=5)
    B = bws.gen_f32(6, (K, N + 4), "normal", 0.0, 1.0)
    Bd = to_dev(B, dev)[:, :N]  # ldb = N + 4
    C0 = rng.standard_normal((M, N + 8)).astype(np.float32)
    Cd = to_dev(C0, dev)
    out = Cd[:, 4:4 + N]  # ldc = N + 8, 16-byte aligned start
    pl = bws.cluster_rows(A, 8)
    bws.spmm_clustered(A, pl, Bd, alpha=0.5, beta=-2.0, out=out)
    R = oracle.spmm_f64(W, K, np.ascontiguousarray(B[:, :N]))
    S = oracle.spmm_f64(W, K, np.ascontiguousarray(B[:, :N]), absolute=True)
    expect = 0.5 * R - 2.0 * C0[:, 4:4 + N].astype(np.float64)
    got = Cd.cpu().numpy()
    assert np.all(np.abs(got[:, 4:4 + N] - expect) <= 1e-5 * (0.5 * S + 2 * np.abs(C0[:, 4:4 + N])) + 1e-6)
    assert np.array_equal(got[:, :4], C0[:, :4]) and np.array_equal(got[:, 4 + N:], C0[:, 4 + N:])


def test_spmm_clustered_empty_rows_and_k0(oracle, bws, dev):
    M, K, N = 200, 96, 64
    W = bws.gen_bitmap_clustered(2, M, K, 4, 0.3, 0.02)
    W[::3] = 0  # every third row empty
    A = bws.Bitmap(to_dev(W, dev), M, K)
    B = bws.gen_f32(3, (K, N), "normal", 0.0, 1.0)
    pl = bws.cluster_rows(A)
    C = bws.spmm_clustered(A, pl, to_dev(B, dev)).cpu().numpy()
    assert np.all(C[::3] == 0)
    assert rel_err(oracle, C, W, K, B)["max_rel_s"] <= 1e-5
    # K = 0: C = beta * C
    A0 = bws.Bitmap(torch.zeros((M, 1), dtype=torch.int32, device=dev), M, 0)
    pl0 = bws.cluster_rows(A0)
    out = torch.full((M, N), 3.0, device=dev)
    bws.spmm_clustered(A0, pl0, torch.empty((0, N), device=dev), beta=0.5, out=out)
    assert torch.all(out == 1.5)


def test_spmm_clustered_benchmark_mask(oracle, bws, dev):
    """The 4096 x 4096 clustered mask (64 prototypes, density 1/32, 2% flips):
    most rows are delta rows; rows are checked on a sample against FP64."""
    M = K = 4096
    N = 512
    W = bws.gen_bitmap_clustered(0, M, K, 64, 0.03125, 0.02)
    A = bws.Bitmap(to_dev(W, dev), M, K)
    pl = bws.cluster_rows(A)
    delta = ((pl.slot_row.cpu().numpy() & bws.SLOT_DELTA) != 0).mean()
    assert delta > 0.9, delta
    B = bws.gen_f32(1, (K, N), "normal", 0.0, 1.0)
    C = bws.spmm_clustered(A, pl, to_dev(B, dev)).cpu().numpy()
    rows = np.arange(0, M, 7)
    e = rel_err(oracle, C[rows], W[rows], K, B)
    assert e["max_rel_s"] <= 1e-5, e


def test_spmm_clustered_rejects(bws, dev):
    W = bws.gen_bitmap_iid(1, 64, 64, 0.1)
    A = bws.Bitmap(to_dev(W, dev), 64, 64)
    pl = bws.cluster_rows(A)
    with pytest.raises(bws.BwsError):
        bws.spmm_clustered(A, pl, torch.zeros((64, 6), device=dev))  # N % 4
    rp = torch.zeros(65, dtype=torch.int64, device=dev)
    with pytest.raises(bws.BwsError):
        bws.cluster_rows(A.with_values(torch.zeros(1, device=dev), rp))


def test_spmm_clustered_many_subtracts(oracle, bws, dev, rng):
    """Rows of one dense prototype (50%) with a fifth of its bits dropped and 1%
    added per row: nearly every slot subtracts in nearly every stage, so the
    subtract masks come from the stream's inline pair and from its full table
    (third and later flagged slots of a stage)."""
    M, K, N = 512, 1000, 132
    proto = rng.random(K) < 0.5
    drop = rng.random((M, K)) < 0.2
    add = rng.random((M, K)) < 0.01
    D = ((proto[None, :] & ~drop) | add).astype(np.float32)
    W = oracle.pack_bitmap(D)
    A = bws.Bitmap(to_dev(W, dev), M, K)
    B = bws.gen_f32(12, (K, N), "normal", 0.0, 1.0)
    pl = bws.cluster_rows(A)
    # the plan makes these delta rows, and some chunk-stage has three or more
    # subtracting slots
    slot = pl.slot_row.cpu().numpy().astype(np.int64)
    cm = to_host_u32(pl.chunk_mask)
    delta = (slot & bws.SLOT_DELTA) != 0
    assert delta.mean() > 0.9
    Wp = oracle.unpack_bitmap(W, K) != 0
    most = 0
    for c in range(M // 16):
        flagged = np.zeros((K + 63) // 64, np.int64)
        for j in range(16):
            sr = slot[16 * c + j]
            if sr < 0 or not (sr & bws.SLOT_DELTA):
                continue
            base = oracle.unpack_bitmap(cm[c, 1 if sr & bws.SLOT_BASE1 else 0][None, :], K)[0] != 0
            sub = base & ~Wp[sr & (bws.SLOT_BASE1 - 1)]
            flagged += np.add.reduceat(sub, np.arange(0, K, 64)) > 0
        most = max(most, int(flagged.max()))
    assert most >= 3
    C = bws.spmm_clustered(A, pl, to_dev(B, dev)).cpu().numpy()
    e = rel_err(oracle, C, W, K, B)
    assert e["max_rel_s"] <= 1e-5, e
