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
