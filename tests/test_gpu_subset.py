"""The FP32 subset-sum product for dense 0/1 masks (plan kernel 1, k_stage
16): per 4 K rows an LDS table of the 16 subset sums of their B rows.  FP32
semantics: within 1e-5 of |A|·|B| of the FP64 oracle (the table pre-sums up
to four B rows, a different FP32 order than the bit walk), on dense and
sparse masks, ragged shapes (K not a multiple of 16, rows and columns past a
tile), alpha/beta and strides."""
import numpy as np
import pytest
import torch

from _helpers import make_problem, rel_err, sampled_err, to_dev

pytestmark = pytest.mark.gpu


def _plan(bws):
    return bws.make_plan(kernel=1, k_stage=16)


@pytest.mark.parametrize("M,K,N,p", [(300, 200, 36, 0.5), (1000, 3001, 260, 0.4),
                                     (257, 17, 1028, 0.9), (64, 16, 4, 1.0), (513, 100, 8, 0.0),
                                     (2000, 777, 300, 0.05), (4100, 1000, 512, 0.6)])
def test_subset_vs_fp64(oracle, bws, dev, M, K, N, p):
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M + N)
    got = bws.spmm(A, Bd, "fp32", plan=_plan(bws))
    rows = np.arange(0, M, max(1, M // 300))
    e = sampled_err(oracle, got, W, K, B, rows)
    assert e["max_rel_s"] <= 1e-5, e


def test_subset_alpha_beta_strides(oracle, bws, dev, rng):
    M, K, N = 600, 300, 132
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N + 4, 0.5, seed=3)
    Bv = Bd[:, :N]  # ldb = N + 4
    C0 = rng.standard_normal((M, N + 8)).astype(np.float32)
    Cd = to_dev(C0, dev)
    out = Cd[:, 4:4 + N]
    bws.spmm(A, Bv, "fp32", alpha=0.5, beta=-2.0, out=out, plan=_plan(bws))
    Bn = np.ascontiguousarray(B[:, :N])
    R = oracle.spmm_f64(W, K, Bn)
    S = oracle.spmm_f64(W, K, Bn, absolute=True)
    got = Cd.cpu().numpy()
    expect = 0.5 * R - 2.0 * C0[:, 4:4 + N].astype(np.float64)
    assert np.all(np.abs(got[:, 4:4 + N] - expect) <= 1e-5 * (0.5 * S + 2 * np.abs(C0[:, 4:4 + N])) + 1e-6)
    assert np.array_equal(got[:, :4], C0[:, :4]) and np.array_equal(got[:, 4 + N:], C0[:, 4 + N:])


def test_subset_dense_4096(oracle, bws, dev):
    """iid 50% at 4096³ (the density where the subset sums pay), sampled rows."""
    M = N = K = 4096
    W = bws.gen_bitmap_iid(0, M, K, 0.5)
    A = bws.Bitmap(to_dev(W, dev), M, K)
    B = bws.gen_f32(1, (K, N), "normal", 0.0, 1.0)
    got = bws.spmm(A, to_dev(B, dev), "fp32", plan=_plan(bws))
    e = sampled_err(oracle, got, W, K, B, np.arange(11, M, 97))
    assert e["max_rel_s"] <= 1e-5, e
