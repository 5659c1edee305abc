"""The MFMA density split (plan mfma_min_bits): at BF16 / FP16 on a 0/1 A,
128-row blocks with fewer than mfma_min_bits set bits per 64 K columns run on
the vector kernel and the others on the direct-to-LDS MFMA kernel.  Masks mix
dense (10%) and sparse (0.2%) row blocks; thresholds send every block to MFMA,
every block to the vector kernel, or split them.  The products of 1.0 and the
16-bit B are exact, so against the FP64 product of the 16-bit B the error is
FP32 accumulation only (1e-5 of |A|·|B|), and every split agrees with the
unsplit MFMA product to FP32 summation order."""
import numpy as np
import pytest
import torch

from _helpers import make_problem, to_dev

pytestmark = pytest.mark.gpu


def _mixed(bws, oracle, M, K, seed):
    rng = np.random.default_rng(seed)
    D = np.zeros((M, K), np.float32)
    for b in range(0, M, 128):
        p = 0.1 if (b // 128) % 2 == 0 else 0.002
        D[b:b + 128] = rng.random((min(128, M - b), K)) < p
    return oracle.pack_bitmap(D)


@pytest.mark.parametrize("M,K,N", [(1024, 1000, 264), (1000, 640, 128), (130, 257, 8)])
@pytest.mark.parametrize("prec,dt", [("bf16", torch.bfloat16), ("fp16", torch.float16)])
@pytest.mark.parametrize("fp32_b", [False, True])
def test_split_thresholds(oracle, bws, dev, M, K, N, prec, dt, fp32_b):
    W = _mixed(bws, oracle, M, K, seed=M + K)
    A = bws.Bitmap(to_dev(W, dev), M, K)
    B = bws.gen_f32(M + 3, (K, N), "normal", 0.0, 1.0)
    Bd = to_dev(B, dev)
    Bh = Bd.to(dt)
    Bexact = Bh.float().cpu().numpy()
    Bin = Bd if fp32_b else Bh
    R = oracle.spmm_f64(W, K, Bexact)
    S = oracle.spmm_f64(W, K, Bexact, absolute=True)
    ref = bws.spmm(A, Bin, prec, plan=bws.make_plan(kernel=2)).cpu().numpy()
    # 1: the dense blocks (about 6 bits per 64 columns and row) and every
    # sparse block with a bit go to MFMA; 40: dense blocks to MFMA, sparse to
    # the vector kernel; 10**6: every block to the vector kernel
    for thr in (1, 40, 10 ** 6):
        got = bws.spmm(A, Bin, prec, plan=bws.make_plan(mfma_min_bits=thr)).cpu().numpy()
        assert np.all(np.abs(got - R) <= 1e-5 * S + 1e-30), thr
        assert np.all(np.abs(got - ref) <= 2e-5 * S + 1e-30), thr


def test_split_alpha_beta_strides(oracle, bws, dev, rng):
    M, K, N = 512, 300, 136
    W = _mixed(bws, oracle, M, K, seed=3)
    A = bws.Bitmap(to_dev(W, dev), M, K)
    B = bws.gen_f32(4, (K, N + 8), "normal", 0.0, 1.0)
    Bh = to_dev(B, dev).to(torch.bfloat16)[:, :N]  # ldb = N + 8
    Bexact = np.ascontiguousarray(Bh.float().cpu().numpy())
    C0 = rng.standard_normal((M, N + 16)).astype(np.float32)
    Cd = to_dev(C0, dev)
    out = Cd[:, 8:8 + N]
    bws.spmm(A, Bh, "bf16", alpha=0.5, beta=-2.0, out=out, plan=bws.make_plan(mfma_min_bits=40))
    R = oracle.spmm_f64(W, K, Bexact)
    S = oracle.spmm_f64(W, K, Bexact, absolute=True)
    got = Cd.cpu().numpy()
    expect = 0.5 * R - 2.0 * C0[:, 8:8 + N].astype(np.float64)
    assert np.all(np.abs(got[:, 8:8 + N] - expect) <= 1e-5 * (0.5 * S + 2 * np.abs(C0[:, 8:8 + N])) + 1e-6)
    assert np.array_equal(got[:, :8], C0[:, :8]) and np.array_equal(got[:, 8 + N:], C0[:, 8 + N:])



@pytest.mark.parametrize("case", ["values", "n_not_mult_8"])
def test_split_not_applicable(oracle, bws, dev, case):
    """The split needs a 0/1 A and N % 8 == 0; otherwise the plan field is
    ignored and the product is the unsplit one (value-carrying A: its values
    rounded to BF16 too, hence the 1e-2 bound against FP64)."""
    M, K = 512, 384
    N = 136 if case == "values" else 131
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.05, seed=21,
                                         values=case == "values")
    Bh = Bd.to(torch.bfloat16)
    Bexact = Bh.float().cpu().numpy()
    got = bws.spmm(A, Bh, "bf16", plan=bws.make_plan(mfma_min_bits=40)).cpu().numpy()
    ref = bws.spmm(A, Bh, "bf16").cpu().numpy()
    R = oracle.spmm_f64(W, K, Bexact, rp, vals)
    S = oracle.spmm_f64(W, K, Bexact, rp, vals, absolute=True)
    bound = 1e-2 if case == "values" else 1e-5
    assert np.all(np.abs(got - R) <= bound * S + 1e-30)
    assert np.all(np.abs(got - ref) <= 2e-5 * S + 1e-30)
