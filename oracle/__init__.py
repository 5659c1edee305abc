"""CPU oracle for the bit-wise SpGEMM path — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or the timed
CPU baseline).  The product package ``bitwise_spgemm_amd`` never imports it.

Thin numpy wrappers over ``liboracle.so`` (built from ``bws_oracle.c`` by
``oracle/Makefile``).  Each C function cites the reference lines it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32

_SIGS = {
    "oref_pack_bitmap_f32": [_P, _I64, _I64, _I64, _P, _I64],
    "oref_unpack_bitmap_f32": [_P, _I64, _I64, _I64, _P, _I64],
    "oref_unpack_bitplanes_f32": [_P, _I64, _I32, _P],
    "oref_unpack_bitplanes_u8": [_P, _I64, _P],
    "oref_pack_bitplanes_i8": [_P, _I64, _I64, _I64, _P, _I64],
    "oref_bitplane_weights": [_P],
    "oref_bitplane_row_counts": [_P, _I32, _I64, _I64, _I64, _P],
    "oref_row_ptr": [_P, _I64, _I64, _I64, _P],
    "oref_pack_values_f32": [_P, _I64, _I64, _I64, _P, _I64, _P, _P],
    "oref_transpose_bitmap": [_P, _I64, _I64, _I64, _P, _I64],
    "oref_transpose_values": [_P, _I64, _I64, _I64, _P, _P, _P, _I64, _P, _P],
    "oref_group_tiles": [_I64, _I64, _I32],
    "oref_group_rows": [_P, _I64, _I64, _I64, _P, _P, _I32, _I32] + [_P] * 9,
    "oref_group_spilled_csr": [_P, _I64, _I64, _I64, _P, _P, _I32] + [_P] * 8,
    "oref_group_scatter": [_P, _I64, _I64, _I64, _P, _P, _I32, _I32, _P, _I64],
    "oref_group_gather": [_P, _I64, _P, _I64, _I64, _I64, _P, _I32, _I32, _P],
    "oref_symbolic_or": [_P, _I64, _I64, _I64, _P, _I64, _I64, _P, _I64],
    "oref_spmm_f64": [_P, _I64, _I64, _I64, _P, _P, _P, _I64, _I64, _P, _I64, _I32],
    "oref_spgemm_f64": [_P, _I64, _I64, _I64, _P, _P, _P, _I64, _I64, _P, _P, _P, _I64, _P, _P,
                        _I32],
    "oref_spmm_f32_seq": [_P, _I64, _I64, _I64, _P, _P, _P, _I64, _I64, _P, _I64],
    "oref_spmm_groups_f64": [_P, _I64, _I64, _I64, _P, _P, _P, _P, _I32, _I32, _P, _I64,
                             _I64, _P, _I64],
    "oref_sddmm_f64": [_P, _I64, _I64, _I64, _P, _P, _I64, _P, _I64, _I64, _P, _I32],
    "oref_check_dense_f32": [_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I32],
    "oref_err_metrics": [_P, _I64, _P, _P, _I64, _I64, _P],
    "oref_cluster_leaders": [_I64, _I32],
    "oref_cluster_candidates": [_I64, _I32],
    "oref_cluster_exact_rounds": [_I32],
    "oref_cluster_rows": [_P, _I64, _I64, _I64, _I32, _P, _P],
    "oref_round_bf16_array": [_P, _I64, _P],
    "oref_round_f16_array": [_P, _I64, _P],
    "oref_split_bf16x2": [_P, _I64, _P, _P],
}
_RET = {"oref_group_tiles": _I64, "oref_cluster_leaders": _I32,
        "oref_cluster_candidates": _I64, "oref_cluster_exact_rounds": _I32}


def build() -> str:
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RET.get(name, None)
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def words_for(cols: int) -> int:
    return (cols + 31) // 32


# ----------------------------------------------------------------------------
# layout transforms
# ----------------------------------------------------------------------------
def pack_bitmap(D, ldw=None):
    D = _c(D, np.float32)
    rows, cols = D.shape
    ldw = ldw or words_for(cols)
    W = np.empty((rows, ldw), np.uint32)
    lib().oref_pack_bitmap_f32(_p(D), rows, cols, cols, _p(W), ldw)
    return W


def unpack_bitmap(W, cols):
    W = _c(W, np.uint32)
    rows, ldw = W.shape
    D = np.empty((rows, cols), np.float32)
    lib().oref_unpack_bitmap_f32(_p(W), rows, cols, ldw, _p(D), cols)
    return D


def unpack_bitplanes_f32(src, bit_width=32):
    src = _c(src, np.float32).ravel()
    dst = np.empty((bit_width, src.size), np.int32)
    lib().oref_unpack_bitplanes_f32(_p(src), src.size, bit_width, _p(dst))
    return dst


def unpack_bitplanes_u8(src):
    src = _c(src, np.uint8).ravel()
    dst = np.empty((8, src.size), np.int32)
    lib().oref_unpack_bitplanes_u8(_p(src), src.size, _p(dst))
    return dst


def pack_bitplanes_i8(D, ldw=None):
    D = _c(D, np.int8)
    rows, cols = D.shape
    ldw = ldw or words_for(cols)
    P = np.empty((8, rows, ldw), np.uint32)
    lib().oref_pack_bitplanes_i8(_p(D), rows, cols, cols, _p(P), ldw)
    return P


def bitplane_weights():
    w = np.empty(8, np.float32)
    lib().oref_bitplane_weights(_p(w))
    return w


def bitplane_row_counts(P, cols):
    P = _c(P, np.uint32)
    planes, rows, ldw = P.shape
    out = np.empty((planes, rows), np.int32)
    lib().oref_bitplane_row_counts(_p(P), planes, rows, cols, ldw, _p(out))
    return out


def row_ptr(W, cols):
    W = _c(W, np.uint32)
    rows, ldw = W.shape
    rp = np.empty(rows + 1, np.int64)
    lib().oref_row_ptr(_p(W), rows, cols, ldw, _p(rp))
    return rp


def pack_values(D, W, rp):
    D = _c(D, np.float32)
    W = _c(W, np.uint32)
    rows, cols = D.shape
    vals = np.empty(int(rp[-1]), np.float32)
    lib().oref_pack_values_f32(_p(D), rows, cols, cols, _p(W), W.shape[1], _p(_c(rp, np.int64)),
                               _p(vals))
    return vals


def transpose_bitmap(W, rows, cols, ldwt=None):
    W = _c(W, np.uint32)
    ldwt = ldwt or words_for(rows)
    WT = np.empty((cols, ldwt), np.uint32)
    lib().oref_transpose_bitmap(_p(W), rows, cols, W.shape[1], _p(WT), ldwt)
    return WT


def transpose_values(W, rows, cols, rp, vals):
    W = _c(W, np.uint32)
    ldwt = words_for(rows)
    WT = np.empty((cols, ldwt), np.uint32)
    rpt = np.empty(cols + 1, np.int64)
    nnz = int(rp[-1])
    vt = np.empty(nnz, np.float32)
    lib().oref_transpose_values(_p(W), rows, cols, W.shape[1], _p(_c(rp, np.int64)),
                                _p(None if vals is None else _c(vals, np.float32)),
                                _p(WT), ldwt, _p(rpt), _p(vt))
    return WT, rpt, vt


# ----------------------------------------------------------------------------
# grouping
# ----------------------------------------------------------------------------
def group_tiles(rows, cols, tile_rows):
    return int(lib().oref_group_tiles(rows, cols, tile_rows))


def group_rows(W, cols, tile_rows=2048, max_groups=2, rp=None, vals=None):
    """Canonical first-fit grouping; returns a dict of numpy arrays."""
    W = _c(W, np.uint32)
    rows, ldw = W.shape
    tiles = group_tiles(rows, cols, tile_rows)
    G, T = max_groups, tile_rows
    out = dict(
        group_id=np.full((rows, ldw), -7, np.int32),
        group_mask=np.empty((tiles, G), np.uint32),
        group_vals=np.empty((tiles, G, 32), np.float32),
        member_row=np.empty((tiles, G, 32), np.int32),
        spill_cnt=np.empty(tiles, np.int32),
        spill_nnz=np.empty(tiles, np.int32),
        spill_table=np.empty((tiles, T), np.int32),
        spill_rev=np.empty((tiles, T), np.int32),
        grouped_nnz=np.zeros(1, np.int64),
    )
    lib().oref_group_rows(
        _p(W), rows, cols, ldw, _p(None if rp is None else _c(rp, np.int64)),
        _p(None if vals is None else _c(vals, np.float32)), T, G,
        *[_p(out[k]) for k in ("group_id", "group_mask", "group_vals", "member_row",
                               "spill_cnt", "spill_nnz", "spill_table", "spill_rev",
                               "grouped_nnz")])
    return out


def cluster_leaders(rows, max_clusters=0):
    return int(lib().oref_cluster_leaders(rows, max_clusters))


def cluster_exact_rounds(L):
    return int(lib().oref_cluster_exact_rounds(L))


def cluster_candidates(rows, L):
    return int(lib().oref_cluster_candidates(rows, L))


def cluster_rows(W, cols, max_clusters=0):
    """Row-clustering plan: (slot_row int32 [rows], chunk_mask uint32
    [ceil(rows/16), 2, ldw] (two bases per chunk))."""
    W = _c(W, np.uint32)
    rows, ldw = W.shape
    slot_row = np.empty(rows, np.int32)
    chunk_mask = np.empty(((rows + 15) // 16, 2, ldw), np.uint32)
    if rows:
        lib().oref_cluster_rows(_p(W), rows, cols, ldw, max_clusters, _p(slot_row),
                                _p(chunk_mask))
    return slot_row, chunk_mask


def group_spilled_csr(W, cols, tile_rows, groups, rp=None, vals=None):
    W = _c(W, np.uint32)
    rows, ldw = W.shape
    tiles = group_tiles(rows, cols, tile_rows)
    nrow = int(groups["spill_cnt"].sum())
    nnz = int(groups["spill_nnz"].sum())
    out = dict(
        spill_row_off=np.empty(tiles + 1, np.int32),
        spill_nnz_off=np.empty(tiles + 1, np.int32),
        row_end=np.empty(max(nrow, 1), np.int32),
        col=np.empty(max(nnz, 1), np.int32),
        val=np.empty(max(nnz, 1), np.float32),
    )
    lib().oref_group_spilled_csr(
        _p(W), rows, cols, ldw, _p(None if rp is None else _c(rp, np.int64)),
        _p(None if vals is None else _c(vals, np.float32)), tile_rows,
        _p(groups["spill_cnt"]), _p(groups["spill_nnz"]), _p(groups["spill_table"]),
        *[_p(out[k]) for k in ("spill_row_off", "spill_nnz_off", "row_end", "col", "val")])
    out["row_end"] = out["row_end"][:nrow]
    out["col"] = out["col"][:nnz]
    out["val"] = out["val"][:nnz]
    return out


def group_scatter(W, cols, group_id, group_vals, tile_rows, max_groups, D):
    W = _c(W, np.uint32)
    D = np.array(D, dtype=np.float32, order="C", copy=True)
    rows, ldw = W.shape
    lib().oref_group_scatter(_p(W), rows, cols, ldw, _p(_c(group_id, np.int32)),
                             _p(_c(group_vals, np.float32)), tile_rows, max_groups, _p(D),
                             D.shape[1])
    return D


def group_gather(D, W, cols, group_id, tile_rows, max_groups, group_vals):
    W = _c(W, np.uint32)
    D = _c(D, np.float32)
    gv = np.array(group_vals, dtype=np.float32, order="C", copy=True)
    rows, ldw = W.shape
    lib().oref_group_gather(_p(D), D.shape[1], _p(W), rows, cols, ldw,
                            _p(_c(group_id, np.int32)), tile_rows, max_groups, _p(gv))
    return gv


def symbolic_or(WA, K, WB, N):
    WA = _c(WA, np.uint32)
    WB = _c(WB, np.uint32)
    M = WA.shape[0]
    ldwc = words_for(N)
    WC = np.empty((M, ldwc), np.uint32)
    lib().oref_symbolic_or(_p(WA), M, K, WA.shape[1], _p(WB), N, WB.shape[1], _p(WC), ldwc)
    return WC


# ----------------------------------------------------------------------------
# numerics
# ----------------------------------------------------------------------------
def spmm_f64(W, K, B, rp=None, vals=None, absolute=False):
    W = _c(W, np.uint32)
    B = _c(B, np.float32)
    M = W.shape[0]
    N = B.shape[1]
    Cm = np.empty((M, N), np.float64)
    lib().oref_spmm_f64(_p(W), M, K, W.shape[1], _p(None if rp is None else _c(rp, np.int64)),
                        _p(None if vals is None else _c(vals, np.float32)), _p(B), N, N,
                        _p(Cm), N, 1 if absolute else 0)
    return Cm


def spgemm_f64(WA, K, WB, N, WC, rpC, rpA=None, vA=None, rpB=None, vB=None, absolute=False):
    """A·B (both bitmaps, value-carrying when their values are given) at the
    set bits of WC, in WC's row-major bit order (FP64)."""
    WA, WB, WC = _c(WA, np.uint32), _c(WB, np.uint32), _c(WC, np.uint32)
    M = WA.shape[0]
    out = np.empty(int(rpC[-1]), np.float64)
    lib().oref_spgemm_f64(_p(WA), M, K, WA.shape[1],
                          _p(None if rpA is None else _c(rpA, np.int64)),
                          _p(None if vA is None else _c(vA, np.float32)), _p(WB), N, WB.shape[1],
                          _p(None if rpB is None else _c(rpB, np.int64)),
                          _p(None if vB is None else _c(vB, np.float32)), _p(WC), WC.shape[1],
                          _p(_c(rpC, np.int64)), _p(out), 1 if absolute else 0)
    return out


def spmm_f32_seq(W, K, B, rp=None, vals=None):
    W = _c(W, np.uint32)
    B = _c(B, np.float32)
    M = W.shape[0]
    N = B.shape[1]
    Cm = np.empty((M, N), np.float32)
    lib().oref_spmm_f32_seq(_p(W), M, K, W.shape[1],
                            _p(None if rp is None else _c(rp, np.int64)),
                            _p(None if vals is None else _c(vals, np.float32)), _p(B), N, N,
                            _p(Cm), N)
    return Cm


def spmm_groups_f64(W, K, B, group_id, group_vals, tile_rows, max_groups, rp=None, vals=None):
    W = _c(W, np.uint32)
    B = _c(B, np.float32)
    M = W.shape[0]
    N = B.shape[1]
    Cm = np.empty((M, N), np.float64)
    lib().oref_spmm_groups_f64(_p(W), M, K, W.shape[1],
                               _p(None if rp is None else _c(rp, np.int64)),
                               _p(None if vals is None else _c(vals, np.float32)),
                               _p(_c(group_id, np.int32)), _p(_c(group_vals, np.float32)),
                               tile_rows, max_groups, _p(B), N, N, _p(Cm), N)
    return Cm


def sddmm_f64(W, K, rp, dC, B, absolute=False):
    W = _c(W, np.uint32)
    dC = _c(dC, np.float32)
    B = _c(B, np.float32)
    M = W.shape[0]
    N = B.shape[1]
    dA = np.empty(int(rp[-1]), np.float64)
    lib().oref_sddmm_f64(_p(W), M, K, W.shape[1], _p(_c(rp, np.int64)), _p(dC), N, _p(B), N, N,
                         _p(dA), 1 if absolute else 0)
    return dA


def check_dense_f32(A, B, row_begin=0, row_end=None, threads=1, out=None):
    """The reference's CPU naive loop (backup/bit_wise_gemm.cu:13-32) over rows
    [row_begin, row_end); returns C (rows outside the range are untouched)."""
    A = _c(A, np.float32)
    B = _c(B, np.float32)
    m, k = A.shape
    n = B.shape[1]
    row_end = m if row_end is None else row_end
    Cm = np.zeros((m, n), np.float32) if out is None else out
    lib().oref_check_dense_f32(_p(A), _p(B), _p(Cm), m, n, k, row_begin, row_end, threads)
    return Cm


def err_metrics(Cm, R, S):
    """dict(max_rel_s, mean_rel_s, max_rel, mean_rel, n_excluded, max_abs, max_rel_ref)."""
    Cm = _c(Cm, np.float32)
    R = _c(R, np.float64)
    S = _c(S, np.float64)
    M, N = R.shape
    out = np.empty(7, np.float64)
    lib().oref_err_metrics(_p(Cm), Cm.shape[1], _p(R), _p(S), M, N, _p(out))
    keys = ("max_rel_s", "mean_rel_s", "max_rel", "mean_rel", "n_excluded", "max_abs",
            "max_rel_ref")
    return {k: float(v) for k, v in zip(keys, out)}


def round_bf16(x):
    x = _c(x, np.float32)
    y = np.empty_like(x)
    lib().oref_round_bf16_array(_p(x), x.size, _p(y))
    return y


def round_f16(x):
    x = _c(x, np.float32)
    y = np.empty_like(x)
    lib().oref_round_f16_array(_p(x), x.size, _p(y))
    return y


def split_bf16x2(x):
    x = _c(x, np.float32)
    hi = np.empty_like(x)
    lo = np.empty_like(x)
    lib().oref_split_bf16x2(_p(x), x.size, _p(hi), _p(lo))
    return hi, lo
