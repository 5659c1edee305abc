"""Pins the oracle's numerics (CPU only) against independent routines:
numpy/scipy dense and sparse products in float64, numpy's sequential float32
cumulative sums, torch's bf16/fp16 conversions, and closed forms of the error
metric.  The Matrix Market files shipped with the reference's comparator
(TileSpGEMM/data/*.mtx) are structural inputs only; their patterns are
regenerated here synthetically with the same shapes (the reference tree is
not available at test time)."""
import numpy as np
import pytest


def rand_mask_values(rng, M, K, p, values=True):
    D = rng.standard_normal((M, K)).astype(np.float32) if values else np.ones((M, K), np.float32)
    D[rng.random((M, K)) >= p] = 0
    return D


@pytest.mark.parametrize("M,K,N,p", [(1, 1, 1, 1.0), (17, 33, 5, 0.3), (64, 96, 70, 0.05),
                                     (40, 31, 9, 0.5), (8, 64, 0, 0.5)])
@pytest.mark.parametrize("values", [False, True])
def test_spmm_f64_vs_numpy(oracle, rng, M, K, N, p, values):
    D = rand_mask_values(rng, M, K, p, values)
    B = rng.standard_normal((K, N)).astype(np.float32)
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, K)
    vals = oracle.pack_values(D, W, rp) if values else None
    C = oracle.spmm_f64(W, K, B, rp, vals)
    ref = D.astype(np.float64) @ B.astype(np.float64)
    np.testing.assert_allclose(C, ref, rtol=1e-13, atol=1e-13)
    S = oracle.spmm_f64(W, K, B, rp, vals, absolute=True)
    np.testing.assert_allclose(S, np.abs(D).astype(np.float64) @ np.abs(B).astype(np.float64),
                               rtol=1e-13, atol=1e-13)


def test_spmm_f64_vs_scipy_sparse(oracle, rng):
    import scipy.sparse as sp
    M, K, N = 150, 150, 12   # the shape class of iris_dataset_30NN.mtx (150 x 150)
    A = sp.random(M, K, density=0.12, random_state=7, format="csr", dtype=np.float64)
    D = A.toarray().astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, K)
    vals = oracle.pack_values(D, W, rp)
    ref = sp.csr_matrix(D.astype(np.float64)) @ B.astype(np.float64)
    np.testing.assert_allclose(oracle.spmm_f64(W, K, B, rp, vals), ref, rtol=1e-13, atol=1e-13)


def test_spmm_f32_seq_is_sequential_fp32(oracle, rng):
    """Independent restatement: numpy float32 cumsum is a sequential left fold."""
    M, K, N = 9, 80, 13
    D = rand_mask_values(rng, M, K, 0.2)
    B = rng.standard_normal((K, N)).astype(np.float32)
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, K)
    vals = oracle.pack_values(D, W, rp)
    C = oracle.spmm_f32_seq(W, K, B, rp, vals)
    for r in range(M):
        ks = np.nonzero(D[r])[0]
        if len(ks) == 0:
            assert np.all(C[r] == 0)
            continue
        terms = (D[r, ks][:, None] * B[ks]).astype(np.float32)
        assert np.array_equal(C[r], np.cumsum(terms, axis=0, dtype=np.float32)[-1])


def test_check_dense_is_reference_loop(oracle, rng):
    """check() (backup/bit_wise_gemm.cu:13-32): sequential p loop in FP32."""
    m, n, k = 6, 7, 40
    A = rand_mask_values(rng, m, k, 0.4, values=False)
    B = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    C = oracle.check_dense_f32(A, B)
    for i in range(m):
        for j in range(n):
            terms = (A[i] * B[:, j]).astype(np.float32)
            expect = np.cumsum(terms, dtype=np.float32)[-1]
            # the C loop may contract into FMA; products of 0/1 and b are exact
            assert C[i, j] == expect
    C2 = oracle.check_dense_f32(A, B, threads=3)
    assert np.array_equal(C, C2)
    part = oracle.check_dense_f32(A, B, row_begin=2, row_end=4)
    assert np.array_equal(part[2:4], C[2:4]) and np.all(part[:2] == 0) and np.all(part[4:] == 0)


def test_sddmm_vs_numpy(oracle, rng):
    M, K, N = 30, 45, 11
    D = rand_mask_values(rng, M, K, 0.15)
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, K)
    dC = rng.standard_normal((M, N)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    dA = oracle.sddmm_f64(W, K, rp, dC, B)
    full = dC.astype(np.float64) @ B.astype(np.float64).T
    np.testing.assert_allclose(dA, full[D != 0], rtol=1e-13, atol=1e-13)
    dAs = oracle.sddmm_f64(W, K, rp, dC, B, absolute=True)
    np.testing.assert_allclose(dAs, (np.abs(dC).astype(np.float64) @ np.abs(B).astype(np.float64).T)[D != 0],
                               rtol=1e-13)


def test_symbolic_or_vs_boolean_product(oracle, rng):
    M, K, N = 21, 50, 70
    A = rng.random((M, K)) < 0.1
    Bm = rng.random((K, N)) < 0.05
    WA = oracle.pack_bitmap(A.astype(np.float32))
    WB = oracle.pack_bitmap(Bm.astype(np.float32))
    WC = oracle.symbolic_or(WA, K, WB, N)
    expect = (A.astype(np.int64) @ Bm.astype(np.int64)) > 0
    assert np.array_equal(WC, oracle.pack_bitmap(expect.astype(np.float32)))


@pytest.mark.parametrize("a_vals,b_vals", [(True, True), (False, True), (True, False),
                                           (False, False)])
def test_spgemm_f64_vs_scipy_sparse(oracle, rng, a_vals, b_vals):
    """The numeric sparse x sparse product onto the symbolic pattern equals
    scipy's CSR product read at the pattern's set bits (row-major order); a
    pattern superset reads explicit zeros there."""
    import scipy.sparse as sp
    M, K, N = 37, 70, 90
    DA = rand_mask_values(rng, M, K, 0.1, a_vals)
    DB = rand_mask_values(rng, K, N, 0.05, b_vals)
    WA, WB = oracle.pack_bitmap(DA), oracle.pack_bitmap(DB)
    rpA, rpB = oracle.row_ptr(WA, K), oracle.row_ptr(WB, N)
    vA = oracle.pack_values(DA, WA, rpA) if a_vals else None
    vB = oracle.pack_values(DB, WB, rpB) if b_vals else None
    ref = (sp.csr_matrix(DA.astype(np.float64)) @ sp.csr_matrix(DB.astype(np.float64))).toarray()
    for WC in (oracle.symbolic_or(WA, K, WB, N),
               oracle.pack_bitmap((rng.random((M, N)) < 0.5).astype(np.float32))):
        rpC = oracle.row_ptr(WC, N)
        got = oracle.spgemm_f64(WA, K, WB, N, WC, rpC, rpA, vA, rpB, vB)
        mask = oracle.unpack_bitmap(WC, N) != 0
        np.testing.assert_allclose(got, ref[mask], rtol=1e-13, atol=1e-13)
        S = oracle.spgemm_f64(WA, K, WB, N, WC, rpC, rpA, vA, rpB, vB, absolute=True)
        refS = (abs(sp.csr_matrix(DA.astype(np.float64))) @ abs(sp.csr_matrix(DB.astype(np.float64)))).toarray()
        np.testing.assert_allclose(S, refS[mask], rtol=1e-13, atol=1e-13)


def test_err_metrics_closed_form(oracle):
    R = np.array([[1.0, -2.0, 0.0, 1e-9]])
    S = np.array([[2.0, 4.0, 0.0, 1.0]])
    Cm = np.array([[1.5, -2.0, 0.0, 0.0]], np.float32)
    e = oracle.err_metrics(Cm, R, S)
    assert e["max_rel_s"] == pytest.approx(0.25)
    assert e["mean_rel_s"] == pytest.approx((0.25 + 0 + 0 + 1e-9) / 4)
    assert e["max_rel"] == pytest.approx(0.5)          # entry 0: 0.5 / 1
    assert e["n_excluded"] == 2                         # r == 0 and |r| < 1e-6 s
    assert e["max_abs"] == pytest.approx(0.5)
    bad = oracle.err_metrics(np.array([[1.0]], np.float32), np.array([[0.0]]), np.array([[0.0]]))
    assert bad["max_rel_s"] == float("inf")


def test_bf16_f16_rounding_vs_torch(oracle, rng):
    import torch
    x = np.concatenate([rng.standard_normal(5000).astype(np.float32) * 10.0 ** rng.integers(-40, 38, 5000),
                        np.array([0.0, -0.0, 65504, 65519, 65520, 1e-8, 6e-8, 3e-8, 2.9e-8,
                                  np.float32(2 ** -25), np.float32(3 * 2 ** -26), 1e38, -3e38,
                                  np.inf, -np.inf], np.float32)]).astype(np.float32)
    bt = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    ht = torch.from_numpy(x).to(torch.float16).to(torch.float32).numpy()
    assert np.array_equal(oracle.round_bf16(x), bt)
    assert np.array_equal(oracle.round_f16(x), ht)
    hi, lo = oracle.split_bf16x2(x[np.isfinite(x) & (np.abs(x) < 1e38)])
    y = x[np.isfinite(x) & (np.abs(x) < 1e38)]
    assert np.array_equal(hi, oracle.round_bf16(y))
    # hi + lo carries ~16 significant bits: |x - hi - lo| <= 2^-16 |x| (normal range)
    norm = np.abs(y) > 1e-30
    assert np.all(np.abs(y[norm].astype(np.float64) - hi[norm] - lo[norm]) <= 2.0 ** -16 * np.abs(y[norm]))
