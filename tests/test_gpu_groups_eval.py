"""bws_spmm_groups (grouped-form evaluation, the reference's grouped tile GEMM
reading dB_group_ele_val) against the FP64 oracle of the same tables, with the
group tables overwritten by independent random values so that every product
term shows whether it was read from the right table entry (tile t, group g,
bit i) or from the row's own values (spilled words).  Rows 16-byte aligned take
the LDS-staged tile kernel, others the row kernel; ragged K, tile heights that
do not divide M (1, 3, 100 rows), one to four groups, 0/1 and value-carrying
A, and the 4096² clustered mask at N = 1024 with the reference's tile shape
(2048 rows, 2 groups), sampled rows."""
import numpy as np
import pytest
import torch

from _helpers import make_problem, to_dev

pytestmark = pytest.mark.gpu


def _random_tables(g, seed):
    gv = g["group_vals"]
    r = np.random.default_rng(seed).standard_normal(tuple(gv.shape)).astype(np.float32)
    return torch.from_numpy(r).to(gv.device)


def _check(oracle, bws, W, K, A, g, gv, T, G, B, Bd, rp, vals, rows=None):
    gd = dict(g)
    gd["group_vals"] = gv
    C = bws.spmm_groups(A, gd, T, G, Bd)
    gid = g["group_id"].cpu().numpy()
    gvh = gv.cpu().numpy()
    if rows is None:
        rows = np.arange(W.shape[0])
    Ws = np.ascontiguousarray(W[rows])
    gids = np.ascontiguousarray(gid[rows])
    # the oracle's tile index uses the row number: evaluate on whole T-blocks
    # of rows and pick the sampled ones
    blocks = np.unique(rows // T)
    full = np.concatenate([np.arange(b * T, min((b + 1) * T, W.shape[0])) for b in blocks])
    Wf = np.ascontiguousarray(W[full])
    gidf = np.ascontiguousarray(gid[full])
    # re-base the tables: block b of the oracle's rows is block blocks[i]
    wpr = (K + 31) // 32
    gvf = np.concatenate([gvh[b * wpr:(b + 1) * wpr] for b in blocks])
    rpf = valsf = None
    if vals is not None:
        cnt = np.diff(rp)[full]
        rpf = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
        valsf = np.concatenate([vals[rp[r]:rp[r + 1]] for r in full]).astype(np.float32)
    R = oracle.spmm_groups_f64(Wf, K, B, gidf, gvf.reshape(-1, G, 32), T, G, rpf, valsf)
    S = oracle.spmm_groups_f64(Wf, K, np.abs(B), gidf, np.abs(gvf).reshape(-1, G, 32), T, G, rpf,
                               None if valsf is None else np.abs(valsf))
    pos = {r: i for i, r in enumerate(full)}
    sel = np.array([pos[r] for r in rows])
    got = C.cpu().numpy()[rows]
    err = np.abs(got - R[sel]) / np.maximum(S[sel], 1e-30)
    assert np.max(err) <= 1e-5, float(np.max(err))


CASES = [  # M, K, N, p, T, G
    (700, 256, 200, 0.05, 64, 2),
    (700, 256, 203, 0.05, 64, 2),   # N % 4 != 0: the row kernel
    (513, 1000, 256, 0.1, 100, 3),
    (300, 97, 64, 0.3, 1, 1),
    (257, 300, 128, 0.2, 3, 4),
    (1024, 2048, 512, 0.02, 512, 2),
]


@pytest.mark.parametrize("M,K,N,p,T,G", CASES)
@pytest.mark.parametrize("values", [False, True])
def test_groups_tables(oracle, bws, dev, M, K, N, p, T, G, values):
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, p, seed=M + K, values=values)
    g = bws.group_rows(A, T, G)
    gv = _random_tables(g, seed=M)
    _check(oracle, bws, W, K, A, g, gv, T, G, B, Bd, rp, vals)


def test_groups_tables_4096(oracle, bws, dev):
    """The reference's grouping shape on the 4096² clustered mask, N = 1024."""
    M = K = 4096
    N, T, G = 1024, 2048, 2
    A, W, rp, vals, B, Bd = make_problem(bws, oracle, dev, M, K, N, 0.03125, seed=0,
                                         values=True, clustered=True)
    g = bws.group_rows(A, T, G)
    gv = _random_tables(g, seed=5)
    rows = np.arange(3, 4096, 97)
    _check(oracle, bws, W, K, A, g, gv, T, G, B, Bd, rp, vals, rows=rows)
