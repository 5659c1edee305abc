"""GPU parity of the grouped tile GEMM family against the FP64 oracle.

Tolerances (error metric |c - r| / Σ_k |a_ik||b_kj|, the componentwise
condition scale, against the FP64 oracle):
  FP32 path                          <= 1e-5   (also vs the FP32 restatement)
  BF16 / FP16 MFMA on rounded B      <= 1e-5   (kernel mechanics: exact 0/1 A, FP32 accumulate)
  BF16 vs the unrounded FP32 B       <= 1e-2   (BF16 rounding of B: 2^-9)
  FP16 vs the unrounded FP32 B       <= 2e-3   (2^-11)
  BF16X2 vs the unrounded FP32 B     <= 1e-5   (hi+lo split: 2^-17)
"""
import numpy as np
import pytest
import torch

from _helpers import make_problem, rel_err, to_dev

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1, 1, 1.0), (64, 64, 64, 0.1), (100, 70, 33, 0.2), (257, 513, 300, 0.05),
          (1024, 1024, 1024, 0.1), (512, 2048, 96, 0.01), (31, 1000, 1, 0.5), (300, 32, 1000, 0.3)]


@pytest.mark.parametrize("M,K,N,p", SHAPES)
@pytest.mark.parametrize("values", [False, True])
def test_spmm_fp32(oracle, bws, dev, M, K, N, p, values):
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M + K, values=values)
    C = bws.spmm(A, Bd, "fp32").cpu().numpy()
    e = rel_err(oracle, C, W, K, B, rp, vals)
    assert e["max_rel_s"] <= 1e-5, e
    C32 = oracle.spmm_f32_seq(W, K, B, rp, vals)
    S = oracle.spmm_f64(W, K, B, rp, vals, absolute=True)
    assert np.all(np.abs(C - C32) <= 2e-5 * S + 1e-30)


def test_spmm_fp32_c1_config(oracle, bws, dev):
    """1024^3, iid Bernoulli(0.1) 0/1 A, B ~ U[-1,1], seed 0."""
    M = N = K = 1024
    W = bws.gen_bitmap_iid(0, M, K, 0.1)
    B = bws.gen_f32(1, (K, N), "uniform", -1.0, 1.0)
    A = bws.Bitmap(to_dev(W, dev), M, K)
    C = bws.spmm(A, to_dev(B, dev)).cpu().numpy()
    e = rel_err(oracle, C, W, K, B)
    assert e["max_rel_s"] <= 1e-5, e
    # the reference's own CPU check loop on the dense 0/1 matrix, sampled rows
    Ad = oracle.unpack_bitmap(W, K)
    Cref = oracle.check_dense_f32(Ad, B, 0, 64, threads=4)
    S = oracle.spmm_f64(W[:64], K, B, absolute=True)
    assert np.all(np.abs(C[:64] - Cref[:64]) <= 2e-5 * S)


def test_spmm_alpha_beta_strides(oracle, bws, dev, rng):
    M, K, N = 200, 300, 150
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.1, seed=5)
    C0 = rng.standard_normal((M, N + 7)).astype(np.float32)
    Cd = to_dev(C0, dev)
    out = Cd[:, 3:3 + N]
    bws.spmm(A, Bd, "fp32", alpha=0.5, beta=-2.0, out=out)
    R = oracle.spmm_f64(W, K, B)
    expect = 0.5 * R - 2.0 * C0[:, 3:3 + N].astype(np.float64)
    got = Cd.cpu().numpy()
    np.testing.assert_allclose(got[:, 3:3 + N], expect, rtol=1e-5, atol=1e-4)
    assert np.array_equal(got[:, :3], C0[:, :3]) and np.array_equal(got[:, 3 + N:], C0[:, 3 + N:])
    # strided B (ldb > N)
    Bbig = to_dev(np.pad(B, ((0, 0), (0, 5))), dev)
    C2 = bws.spmm(A, Bbig[:, :N]).cpu().numpy()
    np.testing.assert_allclose(C2, R, rtol=1e-5, atol=1e-5)


def test_spmm_k_zero_and_empty(bws, dev):
    A = bws.Bitmap(torch.zeros((5, 1), dtype=torch.int32, device=dev), 5, 0)
    B = torch.zeros((0, 7), device=dev)
    out = torch.full((5, 7), 3.0, device=dev)
    bws.spmm(A, B, beta=0.5, out=out)
    assert torch.all(out == 1.5)
    bws.spmm(A, B, out=out)
    assert torch.all(out == 0)
    Z = bws.Bitmap(torch.zeros((64, 2), dtype=torch.int32, device=dev), 64, 64)
    C = bws.spmm(Z, torch.randn(64, 40, device=dev))
    assert torch.all(C == 0)


@pytest.mark.parametrize("dt", [torch.bfloat16, torch.float16])
def test_spmm_fp32_with_16bit_input(oracle, bws, dev, dt):
    M, K, N = 300, 200, 129
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.1, seed=9)
    B16 = Bd.to(dt)
    Bx = B16.float().cpu().numpy()
    C = bws.spmm(A, B16, "fp32").cpu().numpy()
    assert rel_err(oracle, C, W, K, Bx)["max_rel_s"] <= 1e-5


MFMA_SHAPES = [(128, 128, 128, 0.1), (64, 64, 64, 1.0), (1000, 500, 300, 0.05),
               (257, 97, 130, 0.3), (1024, 2048, 512, 0.02), (16, 32, 16, 0.5), (5, 3, 2, 0.9)]


@pytest.mark.parametrize("M,K,N,p", MFMA_SHAPES)
@pytest.mark.parametrize("prec", ["bf16", "fp16", "bf16x2"])
def test_spmm_mfma(oracle, bws, dev, M, K, N, p, prec):
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M * 3 + N)
    plan = bws.make_plan(kernel=2)
    if prec == "bf16":
        Bin, Br = Bd.to(torch.bfloat16), oracle.round_bf16(B)
    elif prec == "fp16":
        Bin, Br = Bd.to(torch.float16), oracle.round_f16(B)
    else:
        Bin, Br = Bd, None
    C = bws.spmm(A, Bin, prec, plan=plan).cpu().numpy()
    if Br is not None:
        # mechanics: exact against the rounded operand
        assert rel_err(oracle, C, W, K, Br)["max_rel_s"] <= 1e-5
        bound = 1e-2 if prec == "bf16" else 2e-3
    else:
        hi, lo = oracle.split_bf16x2(B)
        # hi + lo is exact in FP32 (8 + 8 significant bits); mechanics check first
        assert rel_err(oracle, C, W, K, (hi + lo).astype(np.float32))["max_rel_s"] <= 1e-5
        bound = 1e-5
    e = rel_err(oracle, C, W, K, B)
    assert e["max_rel_s"] <= bound, e
    # FP32-input 16-bit paths round on the device exactly like the oracle
    if prec in ("bf16", "fp16"):
        C2 = bws.spmm(A, Bd, prec, plan=plan).cpu().numpy()
        assert np.array_equal(C2, C)


@pytest.mark.parametrize("M,K,N,p,beta", [(1000, 500, 264, 0.05, 0.0), (257, 97, 136, 0.3, 0.5),
                                          (512, 1024, 1024, 0.1, 0.0), (33, 1000, 8, 0.2, -1.0)])
def test_spmm_bf16x2_two_forms(oracle, bws, dev, M, K, N, p, beta):
    """BF16X2 with the workspace (B split once into BF16 planes, two passes of
    the BF16 kernel) and without it (split in registers, one kernel): both
    within the split bound of the FP64 oracle, and within FP32 rounding of
    each other (the two-pass form rounds the partial sum once more)."""
    import ctypes as C
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M + N)
    C0 = torch.randn((M, N), device=dev)
    s = A.struct()
    lib = bws.lib()
    wsz = lib.bws_spmm_workspace_size(C.byref(s), N, bws.PRECISION["bf16x2"], None)
    assert wsz >= 2 * K * N * 2
    outs = []
    for ws in (None, bws._workspace(wsz, dev)):
        out = C0.clone()
        st = lib.bws_spmm(C.byref(s), bws._ptr(Bd), 0, Bd.stride(0), N, bws._ptr(out),
                          out.stride(0), C.c_float(0.75), C.c_float(beta), bws.PRECISION["bf16x2"],
                          None, bws._ptr(ws), 0 if ws is None else ws.numel(), bws._stream(dev))
        assert st == 0
        outs.append(out)
    torch.cuda.synchronize()
    ref = 0.75 * oracle.spmm_f64(W, K, B) + beta * C0.double().cpu().numpy()
    S = 0.75 * oracle.spmm_f64(W, K, np.abs(B)) + abs(beta) * np.abs(C0.double().cpu().numpy())
    for out in outs:
        e = np.abs(out.double().cpu().numpy() - ref) / np.maximum(S, 1e-30)
        assert e.max() <= 1e-5, e.max()
    d = (outs[0] - outs[1]).abs().double().cpu().numpy() / np.maximum(S, 1e-30)
    assert d.max() <= 2e-6, d.max()


@pytest.mark.parametrize("M,K,N,p", [(1000, 500, 264, 0.05), (257, 97, 136, 0.3), (300, 700, 520, 0.1)])
@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_spmm_fp32_b_rounded_once(oracle, bws, dev, M, K, N, p, prec):
    """FP32 B at BF16 / FP16: rounded once into the workspace (then the
    direct-to-LDS kernel) or in registers (no workspace) — the same result bit
    for bit, and the same as passing the rounded 16-bit B."""
    import ctypes as C
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=N + 5)
    s = A.struct()
    lib = bws.lib()
    P = bws.PRECISION[prec]
    wsz = lib.bws_spmm_workspace_size(C.byref(s), N, P, None)
    assert wsz >= K * N * 2
    outs = []
    for ws in (None, bws._workspace(wsz, dev)):
        out = torch.empty((M, N), device=dev)
        st = lib.bws_spmm(C.byref(s), bws._ptr(Bd), 0, N, N, bws._ptr(out), N, C.c_float(1.0),
                          C.c_float(0.0), P, None, bws._ptr(ws), 0 if ws is None else ws.numel(),
                          bws._stream(dev))
        assert st == 0
        outs.append(out)
    B16 = Bd.to(torch.bfloat16 if prec == "bf16" else torch.float16)
    ref = bws.spmm(A, B16, prec)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[1], ref)


@pytest.mark.parametrize("b_off", [4, 1])  # 16-byte aligned B rows (vector split) or not
def test_spmm_bf16x2_strided_alpha_beta(oracle, bws, dev, rng, b_off):
    """Split BF16 through the workspace with strided B and C, alpha and beta."""
    M, K, N = 300, 260, 136
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.1, seed=17)
    Bbig = np.zeros((K, N + 12), np.float32)
    Bbig[:, b_off:b_off + N] = B
    Bv = to_dev(Bbig, dev)[:, b_off:b_off + N]
    C0 = rng.standard_normal((M, N + 8)).astype(np.float32)
    Cbig = to_dev(C0, dev)
    out = Cbig[:, :N]
    bws.spmm(A, Bv, "bf16x2", alpha=0.5, beta=2.0, out=out)
    got = Cbig.cpu().numpy()
    ref = 0.5 * oracle.spmm_f64(W, K, B) + 2.0 * C0[:, :N].astype(np.float64)
    S = 0.5 * oracle.spmm_f64(W, K, np.abs(B)) + 2.0 * np.abs(C0[:, :N]).astype(np.float64)
    assert np.max(np.abs(got[:, :N] - ref) / np.maximum(S, 1e-30)) <= 1e-5
    assert np.array_equal(got[:, N:], C0[:, N:])  # the padding columns are untouched


TILES = {"bf16": [(128, 128), (128, 256), (256, 128), (256, 256)],
         "fp16": [(128, 128), (128, 256), (256, 128), (256, 256)],
         "bf16x2": [(128, 128), (256, 128)]}


@pytest.mark.parametrize("M,K,N,p", [(1000, 500, 300, 0.05), (257, 97, 130, 0.3),
                                     (300, 1000, 520, 0.1), (5, 3, 2, 0.9), (512, 256, 512, 0.0),
                                     (257, 97, 136, 0.3), (1000, 300, 264, 0.05), (33, 1000, 8, 0.2),
                                     (64, 4096, 512, 0.01)])
@pytest.mark.parametrize("prec", ["bf16", "fp16", "bf16x2"])
def test_spmm_mfma_tiles(oracle, bws, dev, M, K, N, p, prec):
    """Every MFMA workgroup tile and staging variant (register-staged B,
    direct-to-LDS B with 2 or 3 stages, and the table-lookup 64 x 64-per-wave
    kernel) gives bit-identical results: the K order of the accumulation is
    the same."""
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M + 7 * N)
    Bin = {"bf16": Bd.to(torch.bfloat16), "fp16": Bd.to(torch.float16), "bf16x2": Bd}[prec]
    ref = bws.spmm(A, Bin, prec, plan=bws.make_plan(kernel=2)).cpu().numpy()
    bound = {"bf16": 1e-2, "fp16": 2e-3, "bf16x2": 1e-5}[prec]
    assert rel_err(oracle, ref, W, K, B)["max_rel_s"] <= bound
    for bm, bn in TILES[prec]:
        for staging, ks in (("auto", 0), ("regs", 0), ("dma3", 0), ("dma2", 0), ("dma2", 128)):
            plan = bws.make_plan(kernel=2, rows_per_wg=bm, cols_per_wg=bn, mfma_staging=staging,
                                 k_stage=ks)
            got = bws.spmm(A, Bin, prec, plan=plan).cpu().numpy()
            assert np.array_equal(got, ref), (bm, bn, staging, ks)
    lut = bws.make_plan(kernel=2, mfma_staging="lut")
    assert np.array_equal(bws.spmm(A, Bin, prec, plan=lut).cpu().numpy(), ref)


@pytest.mark.parametrize("M,K,N,p", [(300, 500, 136, 0.1), (128, 64, 128, 0.5), (1000, 97, 264, 0.05),
                                     (17, 1000, 8, 0.2), (1024, 1024, 512, 0.1)])
@pytest.mark.parametrize("prec,dt", [("bf16", torch.bfloat16), ("fp16", torch.float16)])
def test_spmm_mfma_value_mask(oracle, bws, dev, M, K, N, p, prec, dt):
    """Value-carrying A on MFMA (dense 16-bit copy of A in the workspace):
    exact against the FP64 product of the rounded values and rounded B, and
    within the precision's bound of the unrounded product."""
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M + 3 * N, values=True)
    rnd = oracle.round_bf16 if prec == "bf16" else oracle.round_f16
    C = bws.spmm(A, Bd.to(dt), prec).cpu().numpy()
    e = rel_err(oracle, C, W, K, rnd(B), rp, rnd(vals))
    assert e["max_rel_s"] <= 1e-5, e
    bound = 1e-2 if prec == "bf16" else 2e-3
    assert rel_err(oracle, C, W, K, B, rp, vals)["max_rel_s"] <= bound
    # without a workspace the FP32 vector path runs (same A, same B)
    s = A.struct()
    out = torch.empty((M, N), device=dev)
    import ctypes
    assert bws.lib().bws_spmm(ctypes.byref(s), ctypes.c_void_p(Bd.to(dt).data_ptr()),
                              bws.DTYPE[dt], N, N, ctypes.c_void_p(out.data_ptr()), N, 1.0, 0.0,
                              bws.PRECISION[prec], None, None, 0, bws._stream(dev)) == 0
    torch.cuda.synchronize()
    Bx = (oracle.round_bf16(B) if prec == "bf16" else oracle.round_f16(B))
    assert rel_err(oracle, out.cpu().numpy(), W, K, Bx, rp, vals)["max_rel_s"] <= 1e-5


@pytest.mark.parametrize("prec,dt", [("bf16", torch.bfloat16), ("fp16", torch.float16)])
def test_spmm_mfma_value_mask_banded(oracle, bws, dev, prec, dt):
    """Value-carrying A whose set bits come in bands: 200 empty rows (whole
    waves of words without a value), 64 full rows (2048 values per wave of
    words), rows holding only their last column, and the matrix's last bit
    set, so the final value sits at the end of the value array."""
    M, K, N = 520, 1000, 72
    D = np.zeros((M, K), dtype=np.float32)
    D[200:264, :] = 1.0
    D[264:400, K - 1] = 1.0
    D[400:, 3::37] = 1.0
    D[M - 1, K - 1] = 1.0
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, K)
    nnz = int(rp[-1])
    vals = bws.gen_f32(5, nnz, "normal", 0.0, 0.02)
    vals[vals == 0] = 1e-3
    A = bws.Bitmap(to_dev(W, dev), M, K, to_dev(rp, dev), to_dev(vals, dev))
    B = bws.gen_f32(6, (K, N), "normal", 0.0, 1.0)
    rnd = oracle.round_bf16 if prec == "bf16" else oracle.round_f16
    C = bws.spmm(A, to_dev(B, dev).to(dt), prec).cpu().numpy()
    assert rel_err(oracle, C, W, K, rnd(B), rp, rnd(vals))["max_rel_s"] <= 1e-5
    assert not C[:200].any()


def test_spmm_mfma_tile_rejects_bad_tile(bws, dev):
    A = bws.pack_bitmap(torch.ones((64, 64), device=dev))
    B = torch.ones((64, 64), device=dev, dtype=torch.bfloat16)
    with pytest.raises(bws.BwsError):
        bws.spmm(A, B, "bf16", plan=bws.make_plan(kernel=2, rows_per_wg=512, cols_per_wg=256))


def test_spmm_groups_equals_spmm(oracle, bws, dev, rng):
    M, K, N, T, G = 700, 256, 200, 64, 2
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.05, seed=21, values=True)
    g = bws.group_rows(A, T, G)
    Cg = bws.spmm_groups(A, g, T, G, Bd).cpu().numpy()
    e = rel_err(oracle, Cg, W, K, B, rp, vals)
    assert e["max_rel_s"] <= 1e-5
    A01 = bws.Bitmap(A.words, M, K)
    g01 = bws.group_rows(A01, T, G)
    Cg01 = bws.spmm_groups(A01, g01, T, G, Bd).cpu().numpy()
    assert rel_err(oracle, Cg01, W, K, B)["max_rel_s"] <= 1e-5


@pytest.mark.parametrize("values", [False, True])
def test_spmm_tA(oracle, bws, dev, values):
    M, K, N = 600, 500, 130
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.1, seed=33, values=values)
    dC = bws.gen_f32(34, (M, N), "normal", 0.0, 1.0)
    AT = bws.transpose(A)
    dB = bws.spmm_tA(AT, to_dev(dC, dev)).cpu().numpy()
    Ad = np.zeros((M, K), np.float64)
    mask = oracle.unpack_bitmap(W, K) != 0
    Ad[mask] = vals if values else 1.0
    R = Ad.T @ dC.astype(np.float64)
    S = np.abs(Ad).T @ np.abs(dC).astype(np.float64)
    assert np.max(np.abs(dB - R) / np.maximum(S, 1e-30)) <= 1e-5


@pytest.mark.parametrize("M,K,N,p", [(100, 80, 64, 0.1), (300, 500, 1500, 0.05), (17, 33, 3, 0.5),
                                     # 64 x 64 tiles staged in LDS (N % 64 == 0): ragged edges,
                                     # several rounds of 512 bits, wave-per-bit sparse tiles
                                     (1000, 700, 256, 0.1), (300, 500, 1536, 0.05),
                                     (130, 200, 128, 0.6), (64, 64, 64, 1.0), (500, 400, 192, 0.003),
                                     (65, 129, 64, 0.3)])
def test_sddmm_fp32(oracle, bws, dev, M, K, N, p):
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=44)
    dC = bws.gen_f32(45, (M, N), "normal", 0.0, 1.0)
    A = bws.Bitmap(A.words, M, K, to_dev(rp, dev))
    dA = bws.sddmm(A, to_dev(dC, dev), Bd).cpu().numpy()
    R = oracle.sddmm_f64(W, K, rp, dC, B)
    S = oracle.sddmm_f64(W, K, rp, dC, B, absolute=True)
    assert np.max(np.abs(dA - R) / np.maximum(S, 1e-30)) <= 1e-5


@pytest.mark.parametrize("M,K,N,p", [(700, 900, 320, 0.08), (200, 300, 64, 0.7)])
def test_sddmm_fp32_two_forms(oracle, bws, dev, M, K, N, p):
    """FP32 SDDMM with the workspace (64 x 64 tiles with >= 64 set bits staged
    in LDS, the rest by a wave per row) and without it (every row by a wave):
    both within 1e-5 of the FP64 oracle relative to Σ|dC|·|B|, on a mask
    mixing dense, sparse (< 64 bits) and empty tiles."""
    import ctypes as C
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M + K)
    W = W.copy()
    W[M // 3: M // 2] = 0                      # a band of empty tiles
    lo, hi = M // 2, M // 2 + 70               # a band of sparse tiles
    W[lo:hi] = bws.gen_bitmap_iid(7, hi - lo, K, 0.004)[:, :W.shape[1]]
    A = bws.Bitmap(to_dev(W, dev), M, K)
    rp = bws.row_ptr(A)
    A = bws.Bitmap(A.words, M, K, rp)
    rph = rp.cpu().numpy()
    nnz = int(rph[-1])
    dC = bws.gen_f32(46, (M, N), "normal", 0.0, 1.0)
    dCd = to_dev(dC, dev)
    R = oracle.sddmm_f64(W, K, rph, dC, B)
    S = oracle.sddmm_f64(W, K, rph, dC, B, absolute=True)
    s = A.struct()
    lib = bws.lib()
    wsz = lib.bws_sddmm_workspace_size(C.byref(s))
    for ws in (None, bws._workspace(wsz, dev)):
        out = torch.full((nnz,), float("nan"), device=dev)
        st = lib.bws_sddmm(C.byref(s), bws._ptr(dCd), bws._ptr(Bd), 0, N, N, N, bws._ptr(out), 0,
                           bws._ptr(ws), 0 if ws is None else ws.numel(), bws._stream(dev))
        assert st == 0
        got = out.cpu().numpy()
        assert np.max(np.abs(got - R) / np.maximum(S, 1e-30)) <= 1e-5


def test_spmm_batched(oracle, bws, dev):
    probs, refs = [], []
    for i, (M, K, p) in enumerate([(64, 128, 0.1), (300, 64, 0.2), (1000, 256, 0.01)]):
        A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, 100, p, seed=100 + i)
        Cd = torch.empty((M, 100), device=dev)
        probs.append((A, Bd, Cd))
        refs.append((W, K, B))
    bws.spmm_batched(probs)
    for (A, Bd, Cd), (W, K, B) in zip(probs, refs):
        assert rel_err(oracle, Cd.cpu().numpy(), W, K, B)["max_rel_s"] <= 1e-5


def test_spmm_deterministic(bws, oracle, dev):
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, 1024, 1024, 512, 0.05, seed=77)
    a = bws.spmm(A, Bd)
    b = bws.spmm(A, Bd)
    assert torch.equal(a, b)
    p = bws.make_plan(kernel=2)
    a = bws.spmm(A, Bd.to(torch.bfloat16), "bf16", plan=p)
    b = bws.spmm(A, Bd.to(torch.bfloat16), "bf16", plan=p)
    assert torch.equal(a, b)


def test_spmm_4096_properties(oracle, bws, dev):
    """Full-size clustered problem (4096^3, ~5% dense): sampled rows against
    the FP64 oracle, and linearity A·(B1 + B2) = A·B1 + A·B2."""
    M = N = K = 4096
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.03125, seed=0, clustered=True)
    C = bws.spmm(A, Bd)
    rows = np.arange(0, M, 97)
    Cs = C[rows].cpu().numpy()
    e = rel_err(oracle, Cs, W[rows], K, B)
    assert e["max_rel_s"] <= 1e-5, e
    B2 = torch.randn_like(Bd)
    lhs = bws.spmm(A, Bd + B2)
    rhs = C + bws.spmm(A, B2)
    S = bws.spmm(A, Bd.abs() + B2.abs())
    assert torch.max((lhs - rhs).abs() / S.clamp_min(1e-30)).item() <= 3e-5
    # checksum of checksums: C·1 = A·(B·1)
    ones = torch.ones(N, 1, device=dev)
    lhs = C.double() @ ones.double()
    rhs = bws.spmm(A, (Bd.double() @ ones.double()).float().contiguous()).double()
    Sabs = bws.spmm(A, Bd.abs().sum(1, keepdim=True).contiguous()).double()
    assert torch.max((lhs - rhs).abs() / Sabs.clamp_min(1e-30)).item() <= 1e-5
