"""Helpers shared by the GPU parity tests (host <-> device conversion, input
construction and the error metric against the FP64 oracle)."""
import numpy as np
import torch


def to_dev(x, dev):
    if x.dtype == np.uint32:
        x = x.view(np.int32)
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def to_host_u32(t):
    return t.cpu().numpy().view(np.uint32)


def make_problem(bws, oracle, dev, M, K, N, p, seed=0, values=False, clustered=False):
    """Bitmap A (M x K) from the product's generator, B (K x N) N(0,1).
    Returns (A_bitmap_on_device, W_host, rp_host, vals_host, B_host, B_dev)."""
    if clustered:
        W = bws.gen_bitmap_clustered(seed, M, K, 64, p, 0.02)
    else:
        W = bws.gen_bitmap_iid(seed, M, K, p)
    B = bws.gen_f32(seed + 1, (K, N), "normal", 0.0, 1.0)
    rp = oracle.row_ptr(W, K)
    vals = None
    A = bws.Bitmap(to_dev(W, dev), M, K)
    if values:
        nnz = int(rp[-1])
        vals = bws.gen_f32(seed + 2, nnz, "normal", 0.0, 0.02)
        vals[vals == 0] = 1e-3
        A = bws.Bitmap(A.words, M, K, to_dev(rp, dev), to_dev(vals, dev))
    return A, W, rp, vals, B, to_dev(B, dev)


def rel_err(oracle, C, W, K, B, rp=None, vals=None, B_exact=None):
    """Error metrics of C against the FP64 oracle; B_exact (if given) is the B
    the reference value uses (e.g. the unrounded FP32 B)."""
    Bref = B if B_exact is None else B_exact
    R = oracle.spmm_f64(W, K, Bref, rp, vals)
    S = oracle.spmm_f64(W, K, Bref, rp, vals, absolute=True)
    return oracle.err_metrics(C, R, S)


def sample_rows(W, rows, rp=None, vals=None):
    """Rows `rows` of a bitmap as their own bitmap: (W_sub, rp_sub, vals_sub)
    (rp_sub / vals_sub None for a 0/1 mask)."""
    Ws = np.ascontiguousarray(W[rows])
    if rp is None:
        return Ws, None, None
    cnt = rp[rows + 1] - rp[rows]
    rps = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    vs = (np.concatenate([vals[rp[r]:rp[r + 1]] for r in rows]).astype(np.float32)
          if vals is not None else None)
    return Ws, rps, vs


def sampled_err(oracle, C_dev, W, K, B, rows, rp=None, vals=None):
    """Error metrics of rows `rows` of the device result C_dev against the
    FP64 oracle (B: the exact host B the reference uses)."""
    Ws, rps, vs = sample_rows(W, rows, rp, vals)
    Cs = C_dev[torch.from_numpy(rows).to(C_dev.device)].cpu().numpy()
    return rel_err(oracle, Cs, Ws, K, B, rps, vs)
