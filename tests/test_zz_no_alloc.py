"""bws_spmm_fp32_noalloc: the FP32 product that allocates nothing.

Its output must equal bws_spmm's with a caller workspace bit for bit (the
same kernels over the same mask stream) and stay within the FP32 bar (1e-5
of |A|·|B|) of the FP64 oracle on sampled rows, on two calls in a row (the
second smaller, so the library's device array still holds the first call's
stream), under graph capture, and it rejects a mask stream beyond its
maximum before launching anything."""
import numpy as np
import pytest
import torch

from _helpers import sampled_err, to_dev

pytestmark = pytest.mark.gpu


def _problem(bws, dev, M, K, N, p, seed):
    W = bws.gen_bitmap_clustered(seed, M, K, 16, p, 0.02)
    B = bws.gen_f32(seed + 1, (K, N), "normal", 0.0, 1.0)
    return W, bws.Bitmap(to_dev(W, dev), M, K), B, to_dev(B, dev)


# (M, K, N): both fill the chip with 256 x 256 tiles (the streamed kernel);
# the second is smaller in every dimension that sizes the stream and ragged
CALLS = [(4096, 4096, 3072, 0.05), (2000, 3001, 6144, 0.1)]


def test_noalloc_two_calls(oracle, bws, dev):
    assert bws.lib().bws_spmm_fp32_noalloc_max_bytes() == 16 << 20
    for i, (M, K, N, p) in enumerate(CALLS):
        W, A, B, Bd = _problem(bws, dev, M, K, N, p, seed=100 + i)
        s = A.struct()
        assert bws.lib().bws_spmm_workspace_size(bws.C.byref(s), N, bws.PRECISION["fp32"],
                                                  None) <= 16 << 20
        got = bws.spmm_fp32_noalloc(A, Bd)
        ref = bws.spmm(A, Bd, "fp32")
        torch.cuda.synchronize()
        assert torch.equal(got, ref), (i, M, K, N)
        rows = np.arange(3, M, 97)
        e = sampled_err(oracle, got, W, K, B, rows)
        assert e["max_rel_s"] <= 1e-5, (i, e)


def test_noalloc_alpha_beta_and_small_shapes(oracle, bws, dev, rng):
    # alpha/beta on the streamed kernel, then a shape below the streamed
    # kernel's size (the vector-add kernel runs; still no allocation)
    for M, K, N in [(4096, 1000, 3072), (300, 200, 36)]:
        W, A, B, Bd = _problem(bws, dev, M, K, N, 0.1, seed=M)
        C0 = rng.standard_normal((M, N)).astype(np.float32)
        out = to_dev(C0, dev)
        out2 = out.clone()
        bws.spmm_fp32_noalloc(A, Bd, alpha=0.5, beta=-2.0, out=out)
        bws.spmm(A, Bd, "fp32", alpha=0.5, beta=-2.0, out=out2)
        torch.cuda.synchronize()
        assert torch.equal(out, out2), (M, K, N)
        rows = np.arange(5, M, 61)
        Ws = np.ascontiguousarray(W[rows])
        R = oracle.spmm_f64(Ws, K, B)
        S = oracle.spmm_f64(Ws, K, B, absolute=True)
        expect = 0.5 * R - 2.0 * C0[rows].astype(np.float64)
        got = out.cpu().numpy()[rows]
        assert np.all(np.abs(got - expect) <= 1e-5 * (0.5 * S + 2 * np.abs(C0[rows])) + 1e-6)


def test_noalloc_graph_capture(bws, dev):
    """Captured into a hipGraph (the library allocates nothing during capture)
    and replayed three times with eager work between the replays."""
    M, K, N = 4096, 2048, 3072
    W, A, B, Bd = _problem(bws, dev, M, K, N, 0.05, seed=7)
    ref = bws.spmm(A, Bd, "fp32")
    out = torch.empty((M, N), device=dev)
    bws.spmm_fp32_noalloc(A, Bd, out=out)  # warm-up (scratch address lookup) outside
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            bws.spmm_fp32_noalloc(A, Bd, out=out)
    torch.cuda.current_stream().wait_stream(s)
    for rep in range(3):
        out.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), rep
        torch.randn(1 << 20, device=dev).cumsum(0)
        torch.cuda.synchronize()


def test_noalloc_rejects(bws, dev):
    # a mask stream of 32 MiB: over the maximum, rejected before any launch
    M, K = 16384, 16384
    A = bws.Bitmap(torch.zeros((M, K // 32), dtype=torch.int32, device=dev), M, K)
    B = torch.zeros((K, 4), device=dev)
    with pytest.raises(bws.BwsError):
        bws.spmm_fp32_noalloc(A, B)
    # value-carrying A: not this entry point's operand
    W = bws.gen_bitmap_iid(1, 64, 64, 0.1)
    A2 = bws.Bitmap(to_dev(W, dev), 64, 64)
    rp = torch.zeros(65, dtype=torch.int64, device=dev)
    with pytest.raises(bws.BwsError):
        bws.spmm_fp32_noalloc(A2.with_values(torch.zeros(1, device=dev), rp),
                              torch.zeros((64, 4), device=dev))
