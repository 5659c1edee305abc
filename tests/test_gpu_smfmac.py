"""The sparse-MFMA path (plan mfma_staging "sparse", v_smfmac_f32_16x16x64):
A 4:2-compressed per row and 4 K columns, nibbles with 3-4 bits through a
second sparse MFMA.  The products of 1.0 and 16-bit B values are exact in
FP32, so against the FP64 product of the 16-bit B the error is FP32
accumulation only: 1e-5 of |A|·|B| (far inside the 16-bit bounds), on ragged
shapes, K not a multiple of 64, empty stages, dense nibbles (every row-nibble
overflowing at p = 0.9), alpha/beta and the 4096³ clustered benchmark mask
(sampled rows)."""
import numpy as np
import pytest
import torch

from _helpers import make_problem, rel_err, sampled_err, to_dev

pytestmark = pytest.mark.gpu


def _plan(bws):
    return bws.make_plan(kernel=2, mfma_staging="sparse")


SHAPES = [(1000, 500, 304, 0.05), (257, 97, 136, 0.3), (300, 1000, 520, 0.1), (5, 3, 8, 0.9),
          (512, 256, 512, 0.0), (1000, 300, 264, 0.5), (33, 1000, 8, 0.2), (64, 4096, 512, 0.01),
          (700, 129, 256, 0.9), (256, 64, 256, 1.0)]


@pytest.mark.parametrize("M,K,N,p", SHAPES)
@pytest.mark.parametrize("prec,dt", [("bf16", torch.bfloat16), ("fp16", torch.float16)])
def test_smfmac_vs_fp64(oracle, bws, dev, M, K, N, p, prec, dt):
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M + 11 * N)
    Bh = Bd.to(dt)
    Bexact = Bh.float().cpu().numpy()  # the 16-bit B the product multiplies
    got = bws.spmm(A, Bh, prec, plan=_plan(bws)).cpu().numpy()
    e = rel_err(oracle, got, W, K, Bexact)
    assert e["max_rel_s"] <= 1e-5, e
    # the dense MFMA kernel on the same operands: equal up to FP32 summation order
    dense = bws.spmm(A, Bh, prec, plan=bws.make_plan(kernel=2)).cpu().numpy()
    S = oracle.spmm_f64(W, K, Bexact, absolute=True)
    assert np.all(np.abs(got - dense) <= 2e-5 * S + 1e-30)


def test_smfmac_alpha_beta_strides(oracle, bws, dev, rng):
    M, K, N = 300, 200, 136
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N + 8, 0.2, seed=4)
    Bh = Bd.to(torch.bfloat16)[:, :N]  # ldb = N + 8
    Bexact = np.ascontiguousarray(Bh.float().cpu().numpy())
    C0 = rng.standard_normal((M, N + 16)).astype(np.float32)
    Cd = to_dev(C0, dev)
    out = Cd[:, 8:8 + N]
    bws.spmm(A, Bh, "bf16", alpha=0.5, beta=-2.0, out=out, plan=_plan(bws))
    R = oracle.spmm_f64(W, K, Bexact)
    S = oracle.spmm_f64(W, K, Bexact, absolute=True)
    got = Cd.cpu().numpy()
    expect = 0.5 * R - 2.0 * C0[:, 8:8 + N].astype(np.float64)
    assert np.all(np.abs(got[:, 8:8 + N] - expect) <= 1e-5 * (0.5 * S + 2 * np.abs(C0[:, 8:8 + N])) + 1e-6)
    assert np.array_equal(got[:, :8], C0[:, :8]) and np.array_equal(got[:, 8 + N:], C0[:, 8 + N:])


def test_smfmac_benchmark_mask(oracle, bws, dev):
    """The 4096³ clustered mask at the BF16 bar of the benchmark, sampled rows."""
    M = N = K = 4096
    W = bws.gen_bitmap_clustered(0, M, K, 64, 0.03125, 0.02)
    A = bws.Bitmap(to_dev(W, dev), M, K)
    B = bws.gen_f32(1, (K, N), "normal", 0.0, 1.0)
    Bh = to_dev(B, dev).to(torch.bfloat16)
    got = bws.spmm(A, Bh, "bf16", plan=_plan(bws))
    rows = np.arange(7, M, 53)
    e = sampled_err(oracle, got, W, K, Bh.float().cpu().numpy(), rows)
    assert e["max_rel_s"] <= 1e-5, e
    e32 = sampled_err(oracle, got, W, K, B, rows)  # against the unrounded B
    assert e32["max_rel_s"] <= 1e-2, e32
