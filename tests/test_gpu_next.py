"""GPU parity of the sparse-operand phases: the symbolic boolean product with
output-row grouping (pre_spgemm), the numeric sparse x sparse product onto
that pattern and its grouped form, and the int8 bit-plane SpMM (eight passes
and one pass over B)."""
import numpy as np
import pytest
import torch

from _helpers import to_dev, to_host_u32

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,K,N,pa,pb", [(64, 64, 64, 0.05, 0.05), (300, 500, 257, 0.02, 0.01),
                                         (1024, 1024, 1024, 0.001, 0.01), (5, 33, 40, 0.5, 0.5)])
def test_symbolic_or_and_output_grouping(oracle, bws, dev, M, K, N, pa, pb):
    WA = bws.gen_bitmap_iid(1, M, K, pa)
    WB = bws.gen_bitmap_iid(2, K, N, pb)
    A = bws.Bitmap(to_dev(WA, dev), M, K)
    Bm = bws.Bitmap(to_dev(WB, dev), K, N)
    Cm = bws.symbolic_or(A, Bm)
    WC = oracle.symbolic_or(WA, K, WB, N)
    assert np.array_equal(to_host_u32(Cm.words), WC)
    # output-row grouping of C (pre_spgemm, transform.cuh:1065-1138)
    g = bws.group_rows(Cm, 512, 2)
    ref = oracle.group_rows(WC, N, 512, 2)
    for k in ("group_id", "spill_cnt", "spill_nnz", "spill_table", "spill_rev", "member_row"):
        assert np.array_equal(g[k].cpu().numpy(), ref[k]), k
    assert np.array_equal(g["group_mask"].cpu().numpy().view(np.uint32), ref["group_mask"])


def _i8_problem(bws, dev, M, K, N, density=0.1, seed=0):
    rng = np.random.default_rng(M + K + seed)
    D = rng.integers(-128, 128, size=(M, K)).astype(np.int8)
    D[rng.random((M, K)) >= density] = 0
    B = rng.standard_normal((K, N)).astype(np.float32)
    return D, B, bws.pack_bitplanes_i8(to_dev(D, dev))


@pytest.mark.parametrize("M,K,N", [(64, 64, 64), (300, 500, 130), (1024, 2048, 256)])
def test_bitplane_i8_spmm(oracle, bws, dev, M, K, N):
    """Eight weighted 0/1 products: within 1e-5 of the FP64 product relative
    to the planes' rounding scale Σ_b |w_b| (P_b |B|)."""
    D, B, P = _i8_problem(bws, dev, M, K, N)
    C = bws.spmm_bitplanes_i8(P, K, to_dev(B, dev), one_pass=False).cpu().numpy()
    R = D.astype(np.float64) @ B.astype(np.float64)
    planes = oracle.unpack_bitplanes_u8(D.view(np.uint8)).reshape(8, M, K).astype(np.float64)
    w = np.abs(oracle.bitplane_weights().astype(np.float64))
    S = sum(w[b] * planes[b] @ np.abs(B).astype(np.float64) for b in range(8))
    assert np.max(np.abs(C - R) / np.maximum(S, 1e-300)) <= 1e-5


@pytest.mark.parametrize("M,K,N,density", [(64, 64, 64, 0.1), (300, 500, 136, 0.1),
                                           (1024, 2048, 256, 0.1), (257, 97, 40, 0.6),
                                           (100, 33, 8, 1.0)])
def test_bitplane_i8_one_pass(oracle, bws, dev, M, K, N, density):
    """One pass over B.  FP32: the planes as a value-carrying bitmap of the
    exact int8 values, within 1e-5 of FP64 relative to |A|·|B|.  BF16 / FP16:
    the planes folded into one exact 16-bit A on MFMA — exact against the
    FP64 product of the rounded B (1e-5), within the precision's bound of the
    unrounded one."""
    D, B, P = _i8_problem(bws, dev, M, K, N, density)
    Bd = to_dev(B, dev)
    R = D.astype(np.float64) @ B.astype(np.float64)
    S = np.abs(D).astype(np.float64) @ np.abs(B).astype(np.float64)
    C = bws.spmm_bitplanes_i8(P, K, Bd).cpu().numpy()
    assert np.max(np.abs(C - R) / np.maximum(S, 1e-300)) <= 1e-5
    for prec, dt, rnd, bound in (("bf16", torch.bfloat16, oracle.round_bf16, 1e-2),
                                 ("fp16", torch.float16, oracle.round_f16, 2e-3)):
        if N % 8:
            continue
        C16 = bws.spmm_bitplanes_i8(P, K, Bd.to(dt), prec).cpu().numpy()
        Rr = D.astype(np.float64) @ rnd(B).astype(np.float64)
        Sr = np.abs(D).astype(np.float64) @ np.abs(rnd(B)).astype(np.float64)
        assert np.max(np.abs(C16 - Rr) / np.maximum(Sr, 1e-300)) <= 1e-5, prec
        assert np.max(np.abs(C16 - R) / np.maximum(S, 1e-300)) <= bound, prec


@pytest.mark.parametrize("M,K,N,pa,pb,a_vals,b_vals", [
    (64, 64, 64, 0.05, 0.05, True, True), (300, 500, 257, 0.02, 0.01, True, True),
    (200, 300, 5000, 0.02, 0.02, False, True),   # two accumulator slabs of 4096 columns
    (1024, 1024, 1024, 0.01, 0.01, True, False), (5, 33, 40, 0.5, 0.5, False, False),
    (130, 70, 90, 0.0, 0.1, True, True)])         # an empty A: an empty pattern
def test_spgemm_numeric_and_grouped(oracle, bws, dev, rng, M, K, N, pa, pb, a_vals, b_vals):
    """Numeric sparse x sparse product onto the symbolic pattern: structure
    (pattern, row pointers) bit-exact, values within 1e-5 of the FP64 product
    relative to |A|·|B|; the grouped form of the result (group_rows on the
    value-carrying C, the reference's dC_group_value) bit-exact against the
    oracle's grouping of the same values."""
    def vbitmap(W, rows, cols, vals, seed):
        rp = oracle.row_ptr(W, cols)
        v = bws.gen_f32(seed, int(rp[-1]), "normal", 0.0, 1.0) if vals else None
        A = bws.Bitmap(to_dev(W, dev), rows, cols, to_dev(rp, dev) if vals else None,
                       to_dev(v, dev) if vals else None)
        return A, rp, v
    WA = bws.gen_bitmap_iid(M, M, K, pa)
    WB = bws.gen_bitmap_iid(N, K, N, pb)
    A, rpA, vA = vbitmap(WA, M, K, a_vals, 3)
    Bm, rpB, vB = vbitmap(WB, K, N, b_vals, 4)
    Cb = bws.spgemm(A, Bm)
    WC = oracle.symbolic_or(WA, K, WB, N)
    rpC = oracle.row_ptr(WC, N)
    assert np.array_equal(to_host_u32(Cb.words), WC)
    assert np.array_equal(Cb.row_ptr.cpu().numpy(), rpC)
    R = oracle.spgemm_f64(WA, K, WB, N, WC, rpC, rpA if a_vals else None, vA,
                          rpB if b_vals else None, vB)
    S = oracle.spgemm_f64(WA, K, WB, N, WC, rpC, rpA if a_vals else None, vA,
                          rpB if b_vals else None, vB, absolute=True)
    got = Cb.values.cpu().numpy()
    assert got.shape == R.shape
    if R.size:
        assert np.max(np.abs(got - R) / np.maximum(S, 1e-300)) <= 1e-5
    g = bws.group_rows(Cb, 256, 2)
    ref = oracle.group_rows(WC, N, 256, 2, rpC, got)
    for k, b in ref.items():
        a = g[k].cpu().numpy()
        if k == "group_mask":
            a = a.view(np.uint32)
        assert np.array_equal(a, b), k
