"""GPU parity of the sparse-operand phases: the symbolic boolean product with
output-row grouping (pre_spgemm), and the int8 bit-plane SpMM."""
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


@pytest.mark.parametrize("M,K,N", [(64, 64, 64), (300, 500, 130), (1024, 2048, 256)])
def test_bitplane_i8_spmm(oracle, bws, dev, M, K, N):
    rng = np.random.default_rng(M + K)
    D = rng.integers(-128, 128, size=(M, K)).astype(np.int8)
    D[rng.random((M, K)) < 0.9] = 0
    B = rng.standard_normal((K, N)).astype(np.float32)
    P = bws.pack_bitplanes_i8(to_dev(D, dev))
    C = bws.spmm_bitplanes_i8(P, K, to_dev(B, dev)).cpu().numpy()
    R = D.astype(np.float64) @ B.astype(np.float64)
    # rounding scale of eight weighted 0/1 products
    planes = oracle.unpack_bitplanes_u8(D.view(np.uint8)).reshape(8, M, K).astype(np.float64)
    w = np.abs(oracle.bitplane_weights().astype(np.float64))
    S = sum(w[b] * planes[b] @ np.abs(B).astype(np.float64) for b in range(8))
    assert np.max(np.abs(C - R) / np.maximum(S, 1e-300)) <= 1e-5
