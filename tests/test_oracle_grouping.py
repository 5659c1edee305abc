"""Pins the oracle's row grouping (CPU only).

1. Hand-derived known answers (tests/golden/grouping_kats.json).
2. A characterisation that determines ascending first-fit uniquely, checked
   row by row from the outputs alone: a row in group g is disjoint from the
   members of g that precede it and conflicts with the earlier members of every
   group g' < g; a spilled row conflicts with the earlier members of every
   group; empty rows are never members.  Plus the invariants of
   source/src/main.cu:50-247 (group mask = OR of members, coverage,
   Σ popcount(group masks) + Σ spilled nnz = nnz).
3. Spilled-row CSR and un-grouping round trips."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "grouping_kats.json")


def kat_inputs(case):
    rows, cols = case["rows"], case["cols"]
    D = np.zeros((rows, cols), np.float32)
    for r, per_word in enumerate(case["row_cols"]):
        for w, cs in enumerate(per_word):
            for c in cs:
                D[r, 32 * w + c] = 10 * r + 32 * w + c + 1
    return D


def expand(per_tile, G, fill):
    out = np.full((len(per_tile), G, 32), fill, np.float64)
    for t, groups in enumerate(per_tile):
        for g, d in enumerate(groups):
            for k, v in d.items():
                out[t, g, int(k)] = v
    return out


@pytest.mark.parametrize("case", json.load(open(GOLDEN))["cases"], ids=lambda c: c["name"])
def test_grouping_known_answers(oracle, case):
    D = kat_inputs(case)
    cols, T, G = case["cols"], case["tile_rows"], case["max_groups"]
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, cols)
    vals = oracle.pack_values(D, W, rp)
    g = oracle.group_rows(W, cols, T, G, rp, vals)
    assert g["group_id"].tolist() == case["group_id"]
    assert g["group_mask"].tolist() == [[int(x, 16) for x in t] for t in case["group_mask"]]
    assert np.array_equal(g["member_row"], expand(case["member_cols"], G, -1))
    assert np.array_equal(g["group_vals"], expand(case["group_vals"], G, 0))
    assert g["spill_cnt"].tolist() == case["spill_cnt"]
    assert g["spill_nnz"].tolist() == case["spill_nnz"]
    assert g["spill_table"].tolist() == case["spill_table"]
    assert g["spill_rev"].tolist() == case["spill_rev"]
    assert int(g["grouped_nnz"][0]) == case["grouped_nnz"]
    s = oracle.group_spilled_csr(W, cols, T, g, rp, vals)
    for k, v in case["spilled_csr"].items():
        assert s[k].tolist() == v, k
    # 0/1 mask: group values are 1 where covered
    g01 = oracle.group_rows(W, cols, T, G)
    assert np.array_equal(g01["group_vals"], (g["member_row"] >= 0).astype(np.float32))


def check_first_fit(W, cols, T, G, g):
    """Row-by-row characterisation of ascending first fit (see module doc)."""
    rows, ldw = W.shape
    wpr = (cols + 31) // 32
    for rt in range((rows + T - 1) // T):
        for w in range(wpr):
            t = rt * wpr + w
            acc = [0] * G          # OR of members seen so far
            members = [[] for _ in range(G)]
            spilled = []
            nnz = 0
            for j in range(min(T, rows - rt * T)):
                r = rt * T + j
                m = int(W[r, w])
                gid = int(g["group_id"][r, w])
                nnz += bin(m).count("1")
                if m == 0:
                    assert gid == 0
                    continue
                if gid >= 0:
                    assert acc[gid] & m == 0
                    assert all(acc[h] & m for h in range(gid))
                    acc[gid] |= m
                    members[gid].append(j)
                else:
                    assert gid == -1 and all(acc[h] & m for h in range(G))
                    spilled.append(j)
            assert [int(x) for x in g["group_mask"][t]] == acc
            assert g["spill_cnt"][t] == len(spilled)
            assert g["spill_table"][t][:len(spilled)].tolist() == spilled
            assert np.all(g["spill_table"][t][len(spilled):] == 0)
            rev = np.full(T, -1)
            rev[spilled] = np.arange(len(spilled))
            assert np.array_equal(g["spill_rev"][t], rev)
            snnz = sum(bin(int(W[rt * T + j, w])).count("1") for j in spilled)
            assert g["spill_nnz"][t] == snnz
            assert sum(bin(a).count("1") for a in acc) + snnz == nnz
            for h in range(G):
                for i in range(32):
                    owner = [j for j in members[h] if int(W[rt * T + j, w]) & (0x80000000 >> i)]
                    assert len(owner) <= 1
                    assert g["member_row"][t, h, i] == (owner[0] if owner else -1)


@pytest.mark.parametrize("rows,cols,T,G,p", [(64, 32, 64, 2, 0.05), (100, 70, 16, 3, 0.1),
                                             (130, 64, 64, 1, 0.02), (50, 40, 7, 32, 0.3),
                                             (256, 96, 128, 4, 0.01)])
def test_first_fit_characterisation(oracle, rng, rows, cols, T, G, p):
    D = (rng.random((rows, cols)) < p).astype(np.float32)
    W = oracle.pack_bitmap(D)
    g = oracle.group_rows(W, cols, T, G)
    check_first_fit(W, cols, T, G, g)
    assert int(g["grouped_nnz"][0]) == sum(bin(int(x)).count("1") for x in g["group_mask"].ravel())


def test_spilled_csr_reconstructs_spilled_rows(oracle, rng):
    rows, cols, T, G = 90, 64, 32, 2
    D = rng.standard_normal((rows, cols)).astype(np.float32)
    D[rng.random(D.shape) < 0.85] = 0
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, cols)
    vals = oracle.pack_values(D, W, rp)
    g = oracle.group_rows(W, cols, T, G, rp, vals)
    s = oracle.group_spilled_csr(W, cols, T, g, rp, vals)
    wpr = 2
    R = np.zeros_like(D)
    for t in range(len(g["spill_cnt"])):
        rt, w = divmod(t, wpr)
        start = 0
        for slot in range(g["spill_cnt"][t]):
            r = rt * T + g["spill_table"][t, slot]
            end = s["row_end"][s["spill_row_off"][t] + slot]
            for e in range(start, end):
                R[r, 32 * w + s["col"][s["spill_nnz_off"][t] + e]] = s["val"][s["spill_nnz_off"][t] + e]
            start = end
    spilled_mask = np.repeat(g["group_id"] == -1, 32, axis=1)[:, :cols]
    assert np.array_equal(R, np.where(spilled_mask, D, 0))


def test_scatter_gather_round_trip(oracle, rng):
    rows, cols, T, G = 70, 50, 16, 2
    D = rng.standard_normal((rows, cols)).astype(np.float32)
    D[rng.random(D.shape) < 0.8] = 0
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, cols)
    vals = oracle.pack_values(D, W, rp)
    g = oracle.group_rows(W, cols, T, G, rp, vals)
    S = oracle.group_scatter(W, cols, g["group_id"], g["group_vals"], T, G, np.zeros_like(D))
    grouped = np.repeat(g["group_id"] >= 0, 32, axis=1)[:, :cols]
    assert np.array_equal(S, np.where(grouped, D, 0))
    gv = oracle.group_gather(D, W, cols, g["group_id"], T, G, np.zeros_like(g["group_vals"]))
    assert np.array_equal(gv, g["group_vals"])


def test_grouped_evaluation_equals_direct(oracle, rng):
    rows, cols, N, T, G = 60, 96, 17, 16, 2
    D = rng.standard_normal((rows, cols)).astype(np.float32)
    D[rng.random(D.shape) < 0.9] = 0
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, cols)
    vals = oracle.pack_values(D, W, rp)
    B = rng.standard_normal((cols, N)).astype(np.float32)
    g = oracle.group_rows(W, cols, T, G, rp, vals)
    Cg = oracle.spmm_groups_f64(W, cols, B, g["group_id"], g["group_vals"], T, G, rp, vals)
    Cd = oracle.spmm_f64(W, cols, B, rp, vals)
    np.testing.assert_allclose(Cg, Cd, rtol=0, atol=1e-12)
