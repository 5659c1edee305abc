"""Pins the oracle's row-clustering plan (CPU only).

The plan has no reference counterpart (it is this library's analysis step for
the FP32 product), so it is pinned by what fixes it:
1. hand-worked cases whose leaders, order, bases and flags follow from the
   rules;
2. exact clusters: rows that are copies of a few patterns must come out in
   chunks of identical rows, base 0 equal to the pattern, no base 1, and
   every row a zero-difference delta row;
3. outputs checked from first principles on random masks: slot_row is a
   permutation in nearest-leader order (leaders by farthest point over the
   stride-sampled candidate rows, or the candidates themselves past 256
   leaders), each base is the majority of its run, each flag follows the
   popcount rule; the 65535 clamp of the search distance decides a tie;
4. the algebra the clustered product relies on: T + (row & ~base)·B -
   (base & ~row)·B equals row·B, evaluated in FP64 against the FP64 oracle."""
import numpy as np
import pytest

DELTA = 1 << 30
BASE1 = 1 << 29
ROW = BASE1 - 1


def words(bits):
    """Bool matrix [rows, cols] -> MSB-first uint32 words."""
    rows, cols = bits.shape
    wpr = (cols + 31) // 32
    pad = np.zeros((rows, wpr * 32), bool)
    pad[:, :cols] = bits
    return np.packbits(pad.reshape(rows, wpr, 4, 8), axis=-1, bitorder="big").reshape(
        rows, wpr, 4).astype(np.uint32) @ np.array([1 << 24, 1 << 16, 1 << 8, 1], np.uint32)


def bits_of(W, cols):
    rows = W.shape[0]
    b = np.unpackbits(W.astype(">u4").view(np.uint8).reshape(rows, -1), axis=1)
    return b[:, :cols].astype(bool)


def test_words_helper_layout():
    b = np.zeros((1, 40), bool)
    b[0, 0] = b[0, 33] = True
    W = words(b)
    assert W[0, 0] == 0x80000000 and W[0, 1] == 0x40000000
    assert (bits_of(W, 40) == b).all()


def test_hand_worked_order(oracle):
    # A = 0xFFFF0000, A' = A | col 31, B = 0x0000FFFF, B' = B & ~col 31, rows
    # [A, B', A', B]: leader 0 = row 0; distances 0, 31, 1, 32 -> leader 1 =
    # row 3; row 1 is 1 from B (31 from A), so owners [0, 1, 0, 1] and the
    # order is rows 0, 2, 1, 3.  The chunk has runs {0, 2} and {1, 3} of two
    # rows: base 0 = bits in both of A, A' = A, base 1 = B' (B & B'); every
    # row is within one bit of its run's base.
    W = np.array([[0xFFFF0000], [0x0000FFFE], [0xFFFF0001], [0x0000FFFF]], np.uint32)
    slot, cm = oracle.cluster_rows(W, 32, max_clusters=2)
    assert slot.tolist() == [0 | DELTA, 2 | DELTA, 1 | DELTA | BASE1, 3 | DELTA | BASE1]
    assert cm.tolist() == [[[0xFFFF0000], [0x0000FFFE]]]
    # one leader: one run of four rows, no column set in three of them: no
    # base, no delta row, ascending order
    slot, cm = oracle.cluster_rows(W, 32, max_clusters=1)
    assert slot.tolist() == [0, 1, 2, 3]
    assert cm.tolist() == [[[0], [0]]]


def test_hand_worked_majority(oracle):
    # 3 copies of P and one empty row in one run (n = 4): bits of P are in 3
    # rows, 2*3 > 5, so base 0 is P; P rows are delta rows (0 < popcount P),
    # the empty row is not (popcount(P) > 0 and the empty base 1 ties).
    P = 0x0F0F00F1
    W = np.array([[P], [0], [P], [P]], np.uint32)
    slot, cm = oracle.cluster_rows(W, 32, max_clusters=1)
    assert cm.tolist() == [[[P], [0]]]
    assert slot.tolist() == [0 | DELTA, 1, 2 | DELTA, 3 | DELTA]


def test_hand_worked_runs(oracle):
    # 16 rows in one chunk from three patterns: X x7 (rows 0..6), Y x5, Z x4.
    # Leaders: row 0 (X), then the farthest pattern, then the last; runs in
    # leader order have lengths 7 (X), and 5 / 4: base 0 = X, base 1 = Y;
    # Z rows are direct (Z is far from X and Y).
    X, Y, Z = 0xFF000000, 0x00FF0000, 0x0000FFFF
    W = np.array([[X]] * 7 + [[Y]] * 5 + [[Z]] * 4, np.uint32)
    slot, cm = oracle.cluster_rows(W, 32, max_clusters=3)
    rows = slot & ROW
    assert sorted(rows.tolist()) == list(range(16))
    assert cm[0, 0, 0] == X and cm[0, 1, 0] == Y
    for s_, r in zip(slot, rows):
        want = {X: DELTA, Y: DELTA | BASE1, Z: 0}[int(W[r, 0])]
        assert s_ & (DELTA | BASE1) == want


@pytest.mark.parametrize("patterns,reps,cols", [(4, 32, 200), (7, 16, 96), (3, 48, 1000)])
def test_exact_clusters(oracle, rng, patterns, reps, cols):
    P = rng.random((patterns, cols)) < 0.3
    idx = rng.permutation(np.repeat(np.arange(patterns), reps))
    W = words(P[idx])
    slot, cm = oracle.cluster_rows(W, cols, max_clusters=patterns)
    rows = slot & ROW
    assert sorted(rows.tolist()) == list(range(len(idx)))
    assert ((slot & (DELTA | BASE1)) == DELTA).all()
    for c in range(len(idx) // 16):
        pats = idx[rows[16 * c:16 * c + 16]]
        assert len(set(pats.tolist())) == 1, "a chunk mixes clusters"
        assert (cm[c, 0] == words(P[pats[:1]])[0]).all()
        assert (cm[c, 1] == 0).all()


def rederive_owners(b, nl):
    """Leaders and nearest-leader owners restated from the rules: the last
    R = min(4, L - 1) leaders by exact farthest point over all rows (lowest
    row on ties); the first Lc = L - R by farthest point over the candidates
    j * rows // c (distance clamped at 65535, lowest candidate on ties) when
    Lc <= 256, else the candidates themselves."""
    import oracle
    rows = b.shape[0]
    R = min(4, nl - 1)
    assert R == oracle.cluster_exact_rounds(nl)
    Lc = nl - R
    c = Lc if Lc > 256 else min(rows, 256)
    assert c == oracle.cluster_candidates(rows, nl)
    cand = [j * rows // c for j in range(c)]
    d = lambda r, s: int((b[r] ^ b[s]).sum())
    if Lc > 256:
        leads = cand[:Lc]
    else:
        leads = [cand[0]]
        near = [min(d(x, cand[0]), 65535) for x in cand]
        for _ in range(1, Lc):
            j = int(np.argmax(near))  # first maximum: the lowest candidate
            leads.append(cand[j])
            near = [min(near[i], min(d(x, cand[j]), 65535)) for i, x in enumerate(cand)]
    dist = [min(d(r, l) for l in leads) for r in range(rows)]
    for _ in range(R):
        r = int(np.argmax(dist))  # first maximum: the lowest row
        leads.append(r)
        dist = [min(dist[x], d(x, r)) for x in range(rows)]
    owner = []
    for r in range(rows):
        ds = [d(r, l) for l in leads]
        owner.append(int(np.argmin(ds)))  # first minimum: the earliest leader
    return owner


@pytest.mark.parametrize("rows,cols,L,p", [(37, 45, 0, 0.2), (100, 64, 5, 0.1), (333, 129, 0, 0.05),
                                           (16, 300, 16, 0.3), (1, 10, 0, 0.5), (50, 33, 50, 0.4),
                                           (700, 40, 9, 0.2), (600, 70, 300, 0.1),
                                           (900, 50, 20, 0.15)])
def test_first_principles(oracle, rng, rows, cols, L, p):
    b = rng.random((rows, cols)) < p
    b[rng.random(rows) < 0.1] = False  # some empty rows
    W = words(b)
    slot, cm = oracle.cluster_rows(W, cols, max_clusters=L)
    order = slot & ROW
    assert sorted(order.tolist()) == list(range(rows))
    nl = oracle.cluster_leaders(rows, L)
    owner = rederive_owners(b, nl)
    expect = sorted(range(rows), key=lambda r: (owner[r], r))
    assert order.tolist() == expect
    valid = words(np.ones((1, cols), bool))[0]
    for c in range((rows + 15) // 16):
        rr = order[16 * c:16 * c + 16]
        runs = []  # (length, start) of runs of equal owner
        i = 0
        while i < len(rr):
            j = i
            while j < len(rr) and owner[rr[j]] == owner[rr[i]]:
                j += 1
            runs.append((j - i, i))
            i = j
        runs.sort(key=lambda t: (-t[0], t[1]))
        for j in range(2):
            if j < len(runs):
                n, st = runs[j]
                maj = 2 * b[rr[st:st + n]].sum(0) > n + 1
            else:
                maj = np.zeros(cols, bool)
            assert (bits_of(cm[c, j:j + 1], cols)[0] == maj).all()
            assert (cm[c, j] & ~valid == 0).all()
    for s in range(rows):
        r = order[s]
        base = [bits_of(cm[s // 16, j:j + 1], cols)[0] for j in range(2)]
        costs = [int(b[r].sum())] + [int((b[r] ^ m).sum()) for m in base]
        best = min(range(3), key=lambda i: (costs[i], i))
        assert slot[s] & (DELTA | BASE1) == [0, DELTA, DELTA | BASE1][best]


def test_delta_algebra_fp64(oracle, bws):
    M, K, N = 300, 200, 40
    W = bws.gen_bitmap_clustered(3, M, K, 8, 0.2, 0.03)
    B = bws.gen_f32(4, (K, N), "normal", 0.0, 1.0)
    slot, cm = oracle.cluster_rows(W, K, max_clusters=8)
    assert ((slot & DELTA) != 0).mean() > 0.8  # the clusters are found
    R = oracle.spmm_f64(W, K, B)
    Cp = np.zeros((M, N))
    T = oracle.spmm_f64(cm.reshape(-1, cm.shape[2]), K, B).reshape(-1, 2, N)  # base sums
    for s in range(M):
        r = slot[s] & ROW
        if slot[s] & DELTA:
            j = 1 if slot[s] & BASE1 else 0
            base = cm[s // 16, j]
            add = W[r] & ~base
            sub = base & ~W[r]
            Cp[r] = (T[s // 16, j] + oracle.spmm_f64(add[None], K, B)[0]
                     - oracle.spmm_f64(sub[None], K, B)[0])
        else:
            Cp[r] = oracle.spmm_f64(W[r:r + 1], K, B)[0]
    assert np.abs(Cp - R).max() <= 1e-12 * max(1.0, np.abs(R).max())


def test_search_distance_clamp(oracle):
    """Distances above 65535 tie in the candidate search.  Row 0 empty, row 1
    at distance 69990 and row 2 at 70000 from it, row 3 five bits, rows 4..11
    disjoint blocks of 1000 + i bits.  With 11 leaders (7 searched, 4 exact
    rounds) the search takes rows 0, 1 (the lower of the tied candidates) and
    11 .. 7; the exact rounds then rows 6, 5, 4 and last row 2.  Unclamped,
    row 2 would lead second and row 1 last, swapping their slots."""
    cols = 70000
    b = np.zeros((12, cols), bool)
    b[1, 10:] = True
    b[2, :] = True
    b[3, :5] = True
    for i in range(8):
        b[4 + i, 10000 + 1100 * i:11000 + 1101 * i] = True
    assert [int(b[4 + i].sum()) for i in range(8)] == [1000 + i for i in range(8)]
    W = words(b)
    slot, cm = oracle.cluster_rows(W, cols, max_clusters=11)
    assert (slot & ROW).tolist() == [0, 3, 1, 11, 10, 9, 8, 7, 6, 5, 4, 2]
    assert rederive_owners(b, 11) == [0, 1, 10, 0, 9, 8, 7, 6, 5, 4, 3, 2]


def test_candidate_count(oracle):
    assert oracle.cluster_candidates(0, 0) == 0
    assert oracle.cluster_candidates(4096, 64) == 256
    assert oracle.cluster_candidates(100, 1) == 100
    assert oracle.cluster_candidates(100000, 1000) == 996
    assert [oracle.cluster_exact_rounds(L) for L in (0, 1, 2, 9, 10, 64)] == [0, 0, 1, 4, 4, 4]


def test_leader_count(oracle):
    assert oracle.cluster_leaders(0, 0) == 0
    assert oracle.cluster_leaders(10, 0) == 1
    assert oracle.cluster_leaders(4096, 0) == 64
    assert oracle.cluster_leaders(100, 500) == 100
    assert oracle.cluster_leaders(10 ** 7, 10 ** 6) == 65536
