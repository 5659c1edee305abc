"""GPU parity of the layout transforms against the oracle (bit-exact)."""
import numpy as np
import pytest
import torch

from _helpers import to_dev, to_host_u32

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,cols,ldw,p", [(1, 1, None, 1.0), (7, 33, None, 0.3),
                                             (64, 64, None, 0.05), (33, 1000, 40, 0.1),
                                             (300, 257, None, 0.5), (5, 0, 1, 0.0),
                                             (1024, 1024, None, 0.1)])
def test_pack_bitmap(oracle, bws, dev, rng, rows, cols, ldw, p):
    D = rng.standard_normal((rows, cols)).astype(np.float32)
    D[rng.random(D.shape) >= p] = 0
    if D.size:
        D.flat[rng.integers(0, D.size, 3)] = np.array([-0.0, np.nan, np.inf], np.float32)
    Wref = oracle.pack_bitmap(D, ldw)
    Dd = to_dev(D, dev)
    # garbage in the output buffer must be overwritten everywhere
    A = bws.pack_bitmap(Dd, ldw)
    torch.cuda.synchronize()
    assert np.array_equal(to_host_u32(A.words), Wref)
    if cols:
        assert np.array_equal(bws.unpack_bitmap(A).cpu().numpy(), oracle.unpack_bitmap(Wref, cols))


@pytest.mark.parametrize("ld,c0,cols", [(130, 3, 100),    # unaligned rows: scalar kernel
                                         (1036, 4, 1028),  # 16-byte rows: vector kernel, ragged span
                                         (1100, 1, 1024),  # cols % 4 == 0 but rows misaligned
                                         (2052, 0, 2048)])
def test_pack_strided_input(oracle, bws, dev, rng, ld, c0, cols):
    big = rng.standard_normal((50, ld)).astype(np.float32)
    big[rng.random(big.shape) < 0.7] = 0
    Dd = to_dev(big, dev)[:, c0:c0 + cols]
    A = bws.pack_bitmap(Dd)
    assert np.array_equal(to_host_u32(A.words), oracle.pack_bitmap(big[:, c0:c0 + cols]))


def test_bitplanes(oracle, bws, dev, rng):
    x = rng.standard_normal(1000).astype(np.float32)
    assert np.array_equal(bws.unpack_bitplanes_f32(to_dev(x, dev), 32).cpu().numpy(),
                          oracle.unpack_bitplanes_f32(x, 32))
    assert np.array_equal(bws.unpack_bitplanes_f32(to_dev(x, dev), 7).cpu().numpy(),
                          oracle.unpack_bitplanes_f32(x, 7))
    u = np.arange(256, dtype=np.uint8)
    assert np.array_equal(bws.unpack_bitplanes_u8(to_dev(u, dev)).cpu().numpy(),
                          oracle.unpack_bitplanes_u8(u))
    D = rng.integers(-128, 128, size=(37, 150)).astype(np.int8)
    P = bws.pack_bitplanes_i8(to_dev(D, dev))
    Pref = oracle.pack_bitplanes_i8(D)
    assert np.array_equal(to_host_u32(P), Pref)
    assert np.array_equal(bws.bitplane_row_counts(P, 150).cpu().numpy(),
                          oracle.bitplane_row_counts(Pref, 150))


@pytest.mark.parametrize("rows,cols,p", [(1, 5, 1.0), (300, 300, 0.05), (1000, 2000, 0.01),
                                         (4096, 100, 0.3), (0, 64, 0.1), (257, 4100, 0.5),
                                         (600, 256, 0.9)])
def test_row_ptr_values(oracle, bws, dev, rng, rows, cols, p):
    D = rng.standard_normal((rows, cols)).astype(np.float32)
    D[rng.random(D.shape) >= p] = 0
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, cols)
    A = bws.Bitmap(to_dev(W, dev), rows, cols)
    rpd = bws.row_ptr(A)
    assert np.array_equal(rpd.cpu().numpy(), rp)
    if rows:
        Av = bws.pack_values(to_dev(D, dev), A)
        assert np.array_equal(Av.values.cpu().numpy(), oracle.pack_values(D, W, rp))


@pytest.mark.parametrize("rows,cols,p", [(45, 70, 0.3), (64, 64, 0.1), (1, 1, 1.0),
                                         (513, 300, 0.02), (100, 33, 0.5), (700, 1000, 0.4),
                                         (130, 96, 1.0)])
def test_transpose(oracle, bws, dev, rng, rows, cols, p):
    D = rng.standard_normal((rows, cols)).astype(np.float32)
    D[rng.random(D.shape) >= p] = 0
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, cols)
    vals = oracle.pack_values(D, W, rp)
    WT, rpt, vt = oracle.transpose_values(W, rows, cols, rp, vals)
    A = bws.Bitmap(to_dev(W, dev), rows, cols)
    AT = bws.transpose(A)
    assert np.array_equal(to_host_u32(AT.words), WT)
    Av = bws.Bitmap(A.words, rows, cols, to_dev(rp, dev), to_dev(vals, dev))
    ATv = bws.transpose(Av)
    assert np.array_equal(to_host_u32(ATv.words), WT)
    assert np.array_equal(ATv.row_ptr.cpu().numpy(), rpt)
    assert np.array_equal(ATv.values.cpu().numpy(), vt)


@pytest.mark.parametrize("rows,cols,p", [(45, 70, 0.3), (513, 300, 0.02), (700, 1000, 0.4),
                                         (1, 1, 1.0)])
def test_transpose_perm_and_gather(oracle, bws, dev, rng, rows, cols, p):
    """perm_T maps AT's value order to A's: it equals the oracle's transpose of
    the values 0, 1, 2, ...; gathering A's values through it gives AT's values."""
    D = rng.standard_normal((rows, cols)).astype(np.float32)
    D[rng.random(D.shape) >= p] = 0
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, cols)
    nnz = int(rp[-1])
    vals = oracle.pack_values(D, W, rp)
    WT, rpt, vt = oracle.transpose_values(W, rows, cols, rp, vals)
    _, _, order = oracle.transpose_values(W, rows, cols, rp, np.arange(nnz, dtype=np.float32))
    A = bws.Bitmap(to_dev(W, dev), rows, cols, to_dev(rp, dev))
    AT, perm = bws.transpose_perm(A)
    assert np.array_equal(to_host_u32(AT.words), WT)
    assert np.array_equal(AT.row_ptr.cpu().numpy(), rpt)
    assert np.array_equal(perm.cpu().numpy(), order.astype(np.int64))
    vT = bws.gather(to_dev(vals, dev), perm)
    assert np.array_equal(vT.cpu().numpy(), vt)


@pytest.mark.parametrize("n", [1, 3, 4, 5, 4099, 1 << 20])
def test_gather_lengths(bws, dev, rng, n):
    """dst[i] = src[idx[i]] for short, odd and long lengths, with repeated
    indices."""
    src = rng.standard_normal(n + 7).astype(np.float32)
    idx = rng.integers(0, n + 7, size=n).astype(np.int64)
    got = bws.gather(to_dev(src, dev), to_dev(idx, dev)).cpu().numpy()
    assert np.array_equal(got, src[idx])

