"""GPU parity of row grouping, spilled CSR and un-grouping (bit-exact vs the
oracle), including the hand-derived known answers."""
import json
import os

import numpy as np
import pytest
import torch

from _helpers import to_dev, to_host_u32

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "grouping_kats.json")
KEYS = ("group_id", "group_mask", "group_vals", "member_row", "spill_cnt", "spill_nnz",
        "spill_table", "spill_rev", "grouped_nnz")


def compare_groups(g_dev, g_ref):
    for k in KEYS:
        a = g_dev[k].cpu().numpy()
        b = g_ref[k]
        if k == "group_mask":
            a = a.view(np.uint32)
        assert np.array_equal(a, b), k


def run_case(oracle, bws, dev, D, T, G, values):
    rows, cols = D.shape
    W = oracle.pack_bitmap(D)
    rp = oracle.row_ptr(W, cols)
    vals = oracle.pack_values(D, W, rp) if values else None
    ref = oracle.group_rows(W, cols, T, G, rp if values else None, vals)
    A = bws.Bitmap(to_dev(W, dev), rows, cols)
    if values:
        A = bws.Bitmap(A.words, rows, cols, to_dev(rp, dev), to_dev(vals, dev))
    g = bws.group_rows(A, T, G)
    torch.cuda.synchronize()
    compare_groups(g, ref)
    sref = oracle.group_spilled_csr(W, cols, T, ref, rp if values else None, vals)
    s = bws.group_spilled_csr(A, T, g)
    for k in sref:
        assert np.array_equal(s[k].cpu().numpy(), sref[k]), k
    return W, rp, vals, A, g, ref


@pytest.mark.parametrize("case", json.load(open(GOLDEN))["cases"], ids=lambda c: c["name"])
def test_grouping_known_answers_gpu(oracle, bws, dev, case):
    D = np.zeros((case["rows"], case["cols"]), np.float32)
    for r, per_word in enumerate(case["row_cols"]):
        for w, cs in enumerate(per_word):
            for c in cs:
                D[r, 32 * w + c] = 10 * r + 32 * w + c + 1
    _, _, _, _, g, _ = run_case(oracle, bws, dev, D, case["tile_rows"], case["max_groups"], True)
    assert g["group_id"].cpu().numpy().tolist() == case["group_id"]
    assert g["spill_table"].cpu().numpy().tolist() == case["spill_table"]


@pytest.mark.parametrize("rows,cols,T,G,p,values", [
    (2048, 64, 2048, 2, 0.05, False),     # the reference's tile height and group count
    (4096, 256, 2048, 2, 0.05, True),
    (1000, 100, 64, 4, 0.02, True),
    (777, 70, 100, 32, 0.3, False),
    (300, 32, 7, 1, 0.5, True),
    (5000, 33, 4096, 3, 0.01, False),
    (128, 96, 65536, 2, 0.1, False),
    # every compiled group count (1-4 with the masks in scalar registers, more
    # through the generic path), 0/1 and value-carrying, tiles of 1-3 load groups
    (1000, 100, 64, 4, 0.02, False),
    (3000, 64, 3000, 3, 0.03, True),
    (777, 70, 100, 8, 0.3, True),
    (2500, 40, 2500, 1, 0.02, False),
])
def test_group_rows_parity(oracle, bws, dev, rng, rows, cols, T, G, p, values):
    D = rng.standard_normal((rows, cols)).astype(np.float32)
    D[rng.random(D.shape) >= p] = 0
    run_case(oracle, bws, dev, D, T, G, values)


def test_group_scatter_gather(oracle, bws, dev, rng):
    rows, cols, T, G = 700, 100, 64, 2
    D = rng.standard_normal((rows, cols)).astype(np.float32)
    D[rng.random(D.shape) >= 0.05] = 0
    W, rp, vals, A, g, ref = run_case(oracle, bws, dev, D, T, G, True)
    base = rng.standard_normal((rows, cols)).astype(np.float32)
    S = bws.group_scatter(A, g, T, G, to_dev(base, dev))
    Sref = oracle.group_scatter(W, cols, ref["group_id"], ref["group_vals"], T, G, base)
    assert np.array_equal(S.cpu().numpy(), Sref)
    gv = bws.group_gather(to_dev(D, dev), A, g, T, G, torch.zeros_like(g["group_vals"]))
    gref = oracle.group_gather(D, W, cols, ref["group_id"], T, G, np.zeros_like(ref["group_vals"]))
    assert np.array_equal(gv.cpu().numpy(), gref)


def test_grouping_deterministic(oracle, bws, dev, rng):
    W = bws.gen_bitmap_clustered(3, 4096, 512, 64, 0.05, 0.02)
    A = bws.Bitmap(to_dev(W, dev), 4096, 512)
    a = bws.group_rows(A, 2048, 2)
    b = bws.group_rows(A, 2048, 2)
    for k in KEYS:
        assert torch.equal(a[k], b[k]), k
