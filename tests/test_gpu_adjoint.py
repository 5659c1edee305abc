"""GPU parity of the adjoints: SDDMM on the bitmap pattern (FP32 and MFMA
precisions), the transposed-bitmap SpMM, and the autograd harness checked
against FP64 central finite differences (on the CPU oracle) and a dense FP64
reference gradient.

Error scale for an SDDMM value: Σ_n |dC[r,n]|·|B[k,n]| (FP64)."""
import numpy as np
import pytest
import torch

from _helpers import make_problem, to_dev

pytestmark = pytest.mark.gpu


def sddmm_err(oracle, W, K, rp, dC, B, got):
    R = oracle.sddmm_f64(W, K, rp, dC, B)
    S = oracle.sddmm_f64(W, K, rp, dC, B, absolute=True)
    return float(np.max(np.abs(got - R) / np.maximum(S, 1e-300))) if len(R) else 0.0


@pytest.mark.parametrize("M,K,N,p", [(128, 128, 64, 0.1), (300, 500, 200, 0.05), (17, 33, 3, 0.5),
                                     (1024, 512, 1000, 0.02), (64, 64, 32, 1.0),
                                     (1000, 300, 256, 0.05), (129, 1000, 64, 0.3),
                                     (512, 512, 2048, 0.1), (4096, 2048, 192, 0.05),
                                     (4000, 2100, 64, 0.1)])
@pytest.mark.parametrize("prec,dt", [("bf16", torch.bfloat16), ("bf16", torch.float32),
                                     ("fp16", torch.float16), ("fp16", torch.float32),
                                     ("bf16x2", torch.float32)])
def test_sddmm_mfma(oracle, bws, dev, M, K, N, p, prec, dt):
    A, W, rp, _, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M + N)
    dC = bws.gen_f32(7, (M, N), "normal", 0.0, 1.0)
    A = bws.Bitmap(A.words, M, K, to_dev(rp, dev))
    dCd, Bx = to_dev(dC, dev).to(dt), Bd.to(dt)
    got = bws.sddmm(A, dCd, Bx, prec).cpu().numpy()
    if prec == "bf16x2":
        assert sddmm_err(oracle, W, K, rp, dC, B, got) <= 5e-5
        return
    rnd = oracle.round_bf16 if prec == "bf16" else oracle.round_f16
    # mechanics: against the operands the kernel actually multiplies
    assert sddmm_err(oracle, W, K, rp, rnd(dC), rnd(B), got) <= 2e-5
    bound = 1e-2 if prec == "bf16" else 2e-3
    assert sddmm_err(oracle, W, K, rp, dC, B, got) <= bound


@pytest.mark.parametrize("M,K,N", [(4096, 2048, 128), (4000, 2100, 192), (300, 500, 128)])
def test_sddmm_mfma_kernels_agree(bws, dev, M, K, N):
    """16-bit operands (direct-to-LDS kernels: 256 x 256 blocks for the large
    shapes, 128 x 128 for the small one) and FP32 operands rounded on the
    device (register-staged kernel) run the same MFMA sequence per output:
    the results must be identical."""
    W = bws.gen_bitmap_iid(M + K, M, K, 0.05)
    A0 = bws.Bitmap(to_dev(W, dev), M, K)
    A = A0.with_values(None, bws.row_ptr(A0))
    dC = torch.from_numpy(bws.gen_f32(5, (M, N), "normal", 0.0, 1.0)).to(dev)
    B = torch.from_numpy(bws.gen_f32(6, (K, N), "normal", 0.0, 1.0)).to(dev)
    for prec, dt in (("bf16", torch.bfloat16), ("fp16", torch.float16)):
        x = bws.sddmm(A, dC.to(dt), B.to(dt), prec)
        y = bws.sddmm(A, dC, B, prec)
        assert torch.equal(x, y), prec


@pytest.mark.parametrize("N,pad_c,pad_b", [(130, 3, 0),    # wave per row
                                            (128, 4, 8),    # 64 x 64 tiles, strided dC and B
                                            (192, 1, 0)])   # ldc not a multiple of 4: wave per row
def test_sddmm_fp32_strided(oracle, bws, dev, N, pad_c, pad_b):
    M, K = 200, 300
    A, W, rp, _, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.1, seed=3)
    dC = bws.gen_f32(8, (M, N + pad_c), "normal", 0.0, 1.0)
    Bp = np.zeros((K, N + pad_b), np.float32)
    Bp[:, :N] = B
    A = bws.Bitmap(A.words, M, K, to_dev(rp, dev))
    got = bws.sddmm(A, to_dev(dC, dev)[:, :N], to_dev(Bp, dev)[:, :N], "fp32").cpu().numpy()
    assert sddmm_err(oracle, W, K, rp, np.ascontiguousarray(dC[:, :N]), B, got) <= 1e-5


def _dense(oracle, W, K, rp, vals):
    M = W.shape[0]
    D = np.zeros((M, K), np.float64)
    mask = oracle.unpack_bitmap(W, K) != 0
    D[mask] = vals
    return D, mask


def test_sddmm_fp32_full_size(oracle, bws, dev):
    """The FP32 SDDMM at the value-mask benchmark's shape (4096 x 4096 at 10%,
    N = 2048; 64 x 64 tiles staged in LDS): every 16th row against the FP64
    oracle, and linearity in dC over the whole output."""
    M = K = 4096
    N = 2048
    W = bws.gen_bitmap_iid(21, M, K, 0.1)
    A0 = bws.Bitmap(to_dev(W, dev), M, K)
    rp_d = bws.row_ptr(A0)
    A = bws.Bitmap(A0.words, M, K, rp_d)
    rp = rp_d.cpu().numpy()
    dC = torch.randn((M, N), device=dev)
    B = torch.randn((K, N), device=dev)
    got = bws.sddmm(A, dC, B, "fp32")
    rows = np.arange(0, M, 16)
    Ws = W[rows]
    counts = rp[rows + 1] - rp[rows]
    rps = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    dCs, Bh = dC.cpu().numpy()[rows], B.cpu().numpy()
    R = oracle.sddmm_f64(Ws, K, rps, dCs, Bh)
    S = oracle.sddmm_f64(Ws, K, rps, dCs, Bh, absolute=True)
    g = got.cpu().numpy()
    gs = np.concatenate([g[rp[r]:rp[r + 1]] for r in rows])
    assert np.max(np.abs(gs - R) / np.maximum(S, 1e-30)) <= 1e-5
    dC2 = torch.randn_like(dC)
    lhs = bws.sddmm(A, dC + dC2, B, "fp32")
    rhs = got + bws.sddmm(A, dC2, B, "fp32")
    Sabs = bws.sddmm(A, dC.abs() + dC2.abs(), B.abs(), "fp32")
    assert torch.max((lhs - rhs).abs() / Sabs.clamp_min(1e-30)).item() <= 3e-5


def test_autograd_finite_differences(oracle, bws, dev):
    """C4's FD check: M = K = 256, N = 128, value-carrying mask, FP64 central
    differences of L = <G, C> computed on the CPU oracle."""
    from bitwise_spgemm_amd.autograd import bitmap_spmm
    M, K, N = 256, 256, 128
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.1, seed=11, values=True)
    G = bws.gen_f32(12, (M, N), "normal", 0.0, 1.0)
    v = A.values.clone().requires_grad_(True)
    Bt = Bd.clone().requires_grad_(True)
    C = bitmap_spmm(A, v, Bt, "fp32")
    (C * to_dev(G, dev)).sum().backward()
    gv = v.grad.cpu().numpy().astype(np.float64)
    gB = Bt.grad.cpu().numpy().astype(np.float64)

    def loss(values, Bm):
        return float((oracle.spmm_f64(W, K, Bm.astype(np.float32), rp, values.astype(np.float32)) *
                      G).sum())

    rng = np.random.default_rng(0)
    h = 1e-3
    vals64 = vals.astype(np.float64)
    for i in rng.choice(len(vals), 12, replace=False):
        vp, vm = vals64.copy(), vals64.copy()
        vp[i] += h
        vm[i] -= h
        fd = (loss(vp, B) - loss(vm, B)) / (2 * h)
        # L is linear in each value: the FD is exact up to the float32 rounding of v ± h
        assert abs(fd - gv[i]) <= 1e-3 * (abs(fd) + 1e-3), (i, fd, gv[i])
    for (r, c) in zip(rng.integers(0, K, 12), rng.integers(0, N, 12)):
        Bp, Bm_ = B.astype(np.float64), B.astype(np.float64)
        Bp = Bp.copy(); Bm_ = Bm_.copy()
        Bp[r, c] += h
        Bm_[r, c] -= h
        fd = (loss(vals64, Bp) - loss(vals64, Bm_)) / (2 * h)
        assert abs(fd - gB[r, c]) <= 1e-3 * (abs(fd) + 1e-3), (r, c, fd, gB[r, c])


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16", 1e-2), ("bf16x2", 5e-5)])
def test_autograd_dense_reference(oracle, bws, dev, prec, tol):
    """C4-shaped gradients (reduced to 1024 x 1024 x 512) against a dense FP64 reference."""
    from bitwise_spgemm_amd.autograd import bitmap_spmm
    M, K, N = 1024, 1024, 512
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.1, seed=21, values=True)
    G = bws.gen_f32(22, (M, N), "normal", 0.0, 1.0)
    v = A.values.clone().requires_grad_(True)
    Bt = Bd.clone().requires_grad_(True)
    C = bitmap_spmm(A, v, Bt, prec)
    (C * to_dev(G, dev)).sum().backward()
    D, mask = _dense(oracle, W, K, rp, vals)
    G64, B64 = G.astype(np.float64), B.astype(np.float64)
    # forward
    Cref = D @ B64
    Sref = np.abs(D) @ np.abs(B64)
    assert np.max(np.abs(C.detach().cpu().numpy() - Cref) / np.maximum(Sref, 1e-300)) <= tol
    # dB = Aᵀ G
    gB = Bt.grad.cpu().numpy()
    R = D.T @ G64
    S = np.abs(D).T @ np.abs(G64)
    assert np.max(np.abs(gB - R) / np.maximum(S, 1e-300)) <= tol
    # dvalues = (G Bᵀ) at the mask
    gv = v.grad.cpu().numpy()
    Rv = (G64 @ B64.T)[mask]
    Sv = (np.abs(G64) @ np.abs(B64).T)[mask]
    assert np.max(np.abs(gv - Rv) / np.maximum(Sv, 1e-300)) <= tol
