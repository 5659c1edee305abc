"""GPU parity of the format converters (COO / CSR -> bitmap -> CSR),
bit-exact against the oracle's pack and scipy.sparse, including unsorted
entries, repeated and out-of-range entries, empty matrices, and the Matrix
Market fixtures driven through the SpMM against the FP64 oracle."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from _helpers import rel_err, to_dev, to_host_u32

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _random_coo(rng, M, K, nnz):
    flat = rng.choice(M * K, size=nnz, replace=False)
    r = (flat // K).astype(np.int32)
    c = (flat % K).astype(np.int32)
    v = rng.standard_normal(nnz).astype(np.float32)
    v[v == 0] = 1.0
    return r, c, v


def _check_bitmap(oracle, A, M, K, r, c, v):
    """Bits at the entries' positions (whatever their values), row pointers,
    and the values in (row, column) order."""
    P = np.zeros((M, K), np.float32)
    P[r, c] = 1.0
    W = oracle.pack_bitmap(P)
    assert np.array_equal(to_host_u32(A.words)[:, :W.shape[1]], W)
    rp = oracle.row_ptr(W, K)
    assert np.array_equal(A.row_ptr.cpu().numpy(), rp)
    if A.values is not None:
        order = np.lexsort((c, r))
        assert np.array_equal(A.values.cpu().numpy(), np.asarray(v)[order])


@pytest.mark.parametrize("M,K,nnz", [(100, 70, 500), (1, 1, 1), (257, 1000, 3000), (64, 33, 0),
                                     (3000, 64, 20000)])
def test_coo_to_bitmap_unsorted(oracle, bws, dev, rng, M, K, nnz):
    r, c, v = _random_coo(rng, M, K, nnz)          # random order
    A, bad = bws.coo_to_bitmap(to_dev(r, dev), to_dev(c, dev), to_dev(v, dev), (M, K))
    assert int(bad.item()) == 0
    _check_bitmap(oracle, A, M, K, r, c, v)


def test_coo_invalid_entries_counted(oracle, bws, dev, rng):
    M, K = 50, 40
    r, c, v = _random_coo(rng, M, K, 200)
    r2 = np.concatenate([r, r[:5], np.array([M, -1, 0], np.int32)])
    c2 = np.concatenate([c, c[:5], np.array([0, 0, K], np.int32)])
    v2 = np.concatenate([v, v[:5], np.ones(3, np.float32)])
    A, bad = bws.coo_to_bitmap(to_dev(r2, dev), to_dev(c2, dev), to_dev(v2, dev), (M, K))
    assert int(bad.item()) == 8                    # 5 repeats + 3 outside
    D = np.zeros((M, K), np.float32)
    D[r, c] = v
    assert np.array_equal(to_host_u32(A.words), oracle.pack_bitmap(D))


@pytest.mark.parametrize("M,K,density,shuffle", [(300, 500, 0.05, False), (300, 500, 0.05, True),
                                                 (1, 2000, 0.5, True), (500, 3, 0.9, False)])
def test_csr_round_trip(oracle, bws, dev, rng, M, K, density, shuffle):
    S = sp.random(M, K, density=density, format="csr", dtype=np.float32,
                  random_state=int(rng.integers(1 << 30)))
    S.data[S.data == 0] = 1.0
    indices = S.indices.copy()
    data = S.data.copy()
    if shuffle:                                     # columns out of order within rows
        for i in range(M):
            a, b = S.indptr[i], S.indptr[i + 1]
            p = rng.permutation(b - a)
            indices[a:b] = indices[a:b][p]
            data[a:b] = data[a:b][p]
    A, bad = bws.csr_to_bitmap(to_dev(S.indptr.astype(np.int64), dev),
                               to_dev(indices.astype(np.int32), dev), to_dev(data, dev), (M, K))
    assert int(bad.item()) == 0
    Sc = S.tocoo()
    _check_bitmap(oracle, A, M, K, Sc.row, Sc.col, Sc.data)
    rp, col, vals = bws.bitmap_to_csr(A)
    S.sort_indices()
    assert np.array_equal(rp.cpu().numpy(), S.indptr.astype(np.int64))
    assert np.array_equal(col.cpu().numpy(), S.indices)
    assert np.array_equal(vals.cpu().numpy(), S.data)


def test_bitmap_to_csr_mask(oracle, bws, dev):
    W = bws.gen_bitmap_iid(4, 200, 300, 0.1)
    A = bws.Bitmap(to_dev(W, dev), 200, 300)
    rp, col, vals = bws.bitmap_to_csr(A)
    D = oracle.unpack_bitmap(W, 300)
    S = sp.csr_matrix(D)
    assert np.array_equal(rp.cpu().numpy(), S.indptr)
    assert np.array_equal(col.cpu().numpy(), S.indices)
    assert np.all(vals.cpu().numpy() == 1.0)


@pytest.mark.parametrize("name", ["1138_bus.mtx", "iris_dataset_30NN.mtx"])
def test_matrix_market_spmm(oracle, bws, dev, name):
    A, bad = bws.load_matrix_market(os.path.join(GOLDEN, name), dev)
    assert int(bad.item()) == 0
    rows, cols, r, c, v = bws.read_matrix_market(os.path.join(GOLDEN, name))
    _check_bitmap(oracle, A, rows, cols, r, c, v)
    N = 96
    B = bws.gen_f32(9, (cols, N), "normal", 0.0, 1.0)
    W = to_host_u32(A.words)
    rp = A.row_ptr.cpu().numpy()
    vals = A.values.cpu().numpy()
    C = bws.spmm(A, to_dev(B, dev), "fp32").cpu().numpy()
    assert rel_err(oracle, C, W, cols, B, rp, vals)["max_rel_s"] <= 1e-5


@pytest.mark.parametrize("rp", [[0, 2, 1, 3, 5, 6],    # decreasing pointer (row 1)
                                [0, 2, 3, 3, 5, 9],    # past nnz (row 4)
                                [1, 2, 3, 3, 5, 6],    # nonzero first pointer (row 0)
                                [0, -1, 3, 3, 5, 6]])  # negative pointer (rows 0 and 1)
def test_csr_bad_row_pointers(bws, dev, rp):
    """Invalid CSR row pointers: each affected row counts one into n_invalid
    and reads only its range clamped to [0, nnz) and made non-decreasing;
    nothing outside col_idx is read."""
    M, K = 5, 40
    cols = np.array([1, 7, 3, 30, 12, 39], np.int32)  # distinct columns
    vals = np.arange(1, 7, dtype=np.float32)
    nnz = len(cols)
    A, bad = bws.csr_to_bitmap(to_dev(np.array(rp, np.int64), dev), to_dev(cols, dev),
                               to_dev(vals, dev), (M, K))
    expect = np.zeros((M, K), bool)
    n_bad = 0
    for r in range(M):
        a, b = rp[r], rp[r + 1]
        n_bad += not (a >= 0 and b >= a and b <= nnz and (r != 0 or a == 0))
        lo = min(max(a, 0), nnz)
        hi = min(max(b, lo), nnz)
        expect[r, cols[lo:hi]] = True
    assert int(bad.item()) == n_bad > 0
    W = to_host_u32(A.words)
    got = np.zeros((M, K), bool)
    for r in range(M):
        for c in range(K):
            got[r, c] = (W[r, c // 32] >> (31 - c % 32)) & 1
    assert np.array_equal(got, expect)
