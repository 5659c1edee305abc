"""Pins the oracle's layout transforms (CPU only).

Independent checks: numpy's packbits/unpackbits (a library routine with a
documented bit order), hand-built known answers, closed forms (int8 plane
weights over all 256 values), and round trips."""
import numpy as np
import pytest


def np_pack_msb(D, ldw=None):
    """Reference packing via numpy.packbits(bitorder='big') -> big-endian uint32:
    byte 0 bit 7 is column 0, so the uint32 has column 0 at bit 31."""
    rows, cols = D.shape
    ldw = ldw or (cols + 31) // 32
    nz = np.zeros((rows, ldw * 32), np.uint8)
    nz[:, :cols] = (D != 0)
    b = np.packbits(nz, axis=1, bitorder="big")
    return b.view(">u4").astype(np.uint32).reshape(rows, ldw)


def test_pack_known_answers(oracle):
    D = np.zeros((4, 64), np.float32)
    D[0, 0] = 1.0            # column 0 -> bit 31 of word 0
    D[1, 31] = -2.5          # column 31 -> bit 0 of word 0
    D[2, 32] = 3.0           # column 32 -> bit 31 of word 1
    D[3, :] = -0.0           # negative zero packs to 0
    D[3, 5] = np.nan         # NaN != 0 packs to 1
    D[3, 40] = np.inf
    D[3, 63] = np.float32(1e-45)  # subnormal is nonzero
    W = oracle.pack_bitmap(D)
    assert W.dtype == np.uint32
    assert W[0].tolist() == [0x80000000, 0]
    assert W[1].tolist() == [0x00000001, 0]
    assert W[2].tolist() == [0, 0x80000000]
    assert W[3].tolist() == [0x80000000 >> 5, (0x80000000 >> 8) | 1]


def test_pack_alternating_and_all_ones(oracle):
    D = np.ones((2, 32), np.float32)
    D[1, 1::2] = 0
    W = oracle.pack_bitmap(D)
    assert W[0, 0] == 0xFFFFFFFF
    assert W[1, 0] == 0xAAAAAAAA


@pytest.mark.parametrize("shape,ldw", [((7, 1), None), ((5, 33), None), ((9, 100), 5),
                                       ((3, 64), 4), ((1, 0), 1)])
def test_pack_vs_numpy_packbits(oracle, rng, shape, ldw):
    D = rng.standard_normal(shape).astype(np.float32)
    D[rng.random(shape) < 0.6] = 0
    W = oracle.pack_bitmap(D, ldw)
    assert np.array_equal(W, np_pack_msb(D, ldw))
    # padding bits and padding words are zero
    cols = shape[1]
    if cols % 32:
        assert np.all(W[:, cols // 32] & ((1 << (32 - cols % 32)) - 1) == 0)


def test_unpack_round_trip(oracle, rng):
    D = (rng.random((13, 77)) < 0.3).astype(np.float32)
    W = oracle.pack_bitmap(D)
    assert np.array_equal(oracle.unpack_bitmap(W, 77), D)


def test_bitplanes_f32_vs_numpy(oracle, rng):
    x = rng.standard_normal(257).astype(np.float32)
    x[:3] = [0.0, -0.0, np.inf]
    P = oracle.unpack_bitplanes_f32(x, 32)
    bits = np.unpackbits(x.view(np.uint8).reshape(-1, 4), axis=1, bitorder="little")
    # little-endian bytes, little bit order -> bit b of the uint32 at column b
    assert np.array_equal(P.T, bits.astype(np.int32))
    P8 = oracle.unpack_bitplanes_f32(x, 8)
    assert np.array_equal(P8, P[:8])


def test_bitplanes_u8_exhaustive(oracle):
    v = np.arange(256, dtype=np.uint8)
    P = oracle.unpack_bitplanes_u8(v)
    assert np.array_equal(P.T, np.unpackbits(v[:, None], axis=1, bitorder="little"))


def test_plane_weights_reconstruct_every_int8(oracle):
    """Closed form: Σ_b w_b·bit_b(v) = v for all 256 int8 values with the
    two's-complement weights [1,2,...,64,-128] (main.cu:889-912)."""
    w = oracle.bitplane_weights()
    assert w.tolist() == [1, 2, 4, 8, 16, 32, 64, -128]
    v = np.arange(-128, 128, dtype=np.int8)
    P = oracle.unpack_bitplanes_u8(v.view(np.uint8))
    assert np.array_equal((w[:, None] * P).sum(0).astype(np.int64), v.astype(np.int64))


def test_packed_planes_match_unpacked(oracle, rng):
    D = rng.integers(-128, 128, size=(6, 70)).astype(np.int8)
    P = oracle.pack_bitplanes_i8(D)
    U = oracle.unpack_bitplanes_u8(D.view(np.uint8))
    for b in range(8):
        plane = U[b].reshape(6, 70).astype(np.float32)
        assert np.array_equal(P[b], np_pack_msb(plane))
    counts = oracle.bitplane_row_counts(P, 70)
    assert np.array_equal(counts, U.reshape(8, 6, 70).sum(2))


def test_row_ptr_and_values(oracle, rng):
    D = rng.standard_normal((11, 90)).astype(np.float32)
    D[rng.random(D.shape) < 0.8] = 0
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, 90)
    nz = (D != 0)
    assert rp[0] == 0 and np.array_equal(np.diff(rp), nz.sum(1))
    vals = oracle.pack_values(D, W, rp)
    assert np.array_equal(vals, D[nz])   # numpy boolean indexing is row-major order


def test_transpose_bitmap_and_values(oracle, rng):
    D = rng.standard_normal((45, 70)).astype(np.float32)
    D[rng.random(D.shape) < 0.7] = 0
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, 70)
    vals = oracle.pack_values(D, W, rp)
    WT, rpt, vt = oracle.transpose_values(W, 45, 70, rp, vals)
    assert np.array_equal(WT, np_pack_msb(D.T.copy()))
    assert np.array_equal(oracle.transpose_bitmap(W, 45, 70), WT)
    assert np.array_equal(vt, D.T[D.T != 0])
    assert np.array_equal(np.diff(rpt), (D.T != 0).sum(1))
