"""bitwise_spgemm_amd — MI355X (gfx950) bit-wise sparse GEMM.

Python host mirror of the C ABI in ``include/bws.h`` (``libbws.so``).  Device
memory and streams come from PyTorch-ROCm; every call enqueues HIP kernels of
``libbws.so`` on the current torch stream.  There is no fallback: if the
library is not built, or no GPU is present, calls raise.

Reference entry points and their counterparts here (ZSL98/bitwise_spgemm):

* ``dense2bitmask`` (source/src/main.cu:591-628)          -> :func:`pack_bitmap`
* ``layout_trans`` / ``uint8_layout_trans``
  (backup/layout_trans.cu:75-97)                         -> :func:`unpack_bitplanes_f32`,
                                                            :func:`unpack_bitplanes_u8`
* ``csr2tiledcsr`` prefix + values (source/src/main.cu:499-588)
                                                         -> :func:`row_ptr`, :func:`pack_values`
* ``generate_groups_large`` (source/src/main.cu:50-247)   -> :func:`group_rows`
* ``generate_spilled_csr`` (source/src/main.cu:445-496)   -> :func:`group_spilled_csr`
* ``group2dense`` / ``dense2group_from_idx``
  (source/src/main.cu:829-886)                           -> :func:`group_scatter`, :func:`group_gather`
* ``spgemm_compute_1dthread_tcore_v3``
  (source/src/transform.cuh:2062-2248), ``uint8_spmm_2d_1d``
  (backup/layout_trans.cu:158-171)                       -> :func:`spmm`, :func:`spmm_groups`

The reference-named aliases at the bottom of this module call the same code.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from dataclasses import dataclass
from typing import Optional

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbws.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "bws.h")

# status codes (include/bws.h)
BWS_OK = 0
BWS_EINVAL = -1
BWS_ECAPACITY = -2
BWS_EUNSUPPORTED = -3
BWS_EHIP = -4
BWS_EWORKSPACE = -5
BWS_ENODEVICE = -6
BWS_ENOTFOUND = -7

DTYPE = {torch.float32: 0, torch.bfloat16: 1, torch.float16: 2}
PRECISION = {"fp32": 0, "bf16": 1, "fp16": 2, "bf16x2": 3}


class BwsError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        msg = lib().bws_status_string(status).decode()
        if status == BWS_EHIP:
            msg += f" (hipError {lib().bws_last_hip_error()})"
        super().__init__(f"{where}: {msg}")


class _Bitmap(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("ldw", C.c_int64),
                ("words", C.c_void_p), ("row_ptr", C.c_void_p), ("values", C.c_void_p)]


class _Groups(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("group_id", "group_mask", "group_vals", "member_row",
                                          "spill_cnt", "spill_nnz", "spill_table", "spill_rev",
                                          "grouped_nnz")]


class _Plan(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("rows_per_wg", C.c_int32), ("waves_per_wg", C.c_int32),
                ("k_stage", C.c_int32), ("mfma_min_bits", C.c_int32), ("cols_per_wg", C.c_int32),
                ("mfma_staging", C.c_int32), ("reserved", C.c_int32)]


class _Problem(C.Structure):
    _fields_ = [("A", _Bitmap), ("B", C.c_void_p), ("ldb", C.c_int64), ("N", C.c_int64),
                ("C", C.c_void_p), ("ldc", C.c_int64), ("nnz_hint", C.c_int64)]


BATCH_MAX_LAUNCH = 16


class _BatchLaunch(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("kind", "rows_per_wg", "cols_per_wg", "waves_per_wg",
                                        "k_stage", "values", "vec", "first", "count")] + \
               [("total_tiles", C.c_int64)]


class _Batch(C.Structure):
    _fields_ = [("count", C.c_int32), ("b_dtype", C.c_int32), ("precision", C.c_int32),
                ("n_launch", C.c_int32), ("launch", _BatchLaunch * BATCH_MAX_LAUNCH),
                ("workspace", C.c_void_p), ("problems", C.c_void_p), ("plan", _Plan)]


_P, _I64, _I32, _U64, _F, _D, _SZ = (C.c_void_p, C.c_int64, C.c_int32, C.c_uint64, C.c_float,
                                     C.c_double, C.c_size_t)
_BP = C.POINTER(_Bitmap)

_SIGS = {
    "bws_status_string": ([_I32], C.c_char_p),
    "bws_last_hip_error": ([], _I32),
    "bws_version": ([], _I32),
    "bws_pack_bitmap_f32": ([_P, _I64, _I64, _I64, _P, _I64, _P], _I32),
    "bws_unpack_bitmap_f32": ([_P, _I64, _I64, _I64, _P, _I64, _P], _I32),
    "bws_unpack_bitplanes_f32": ([_P, _I64, _I32, _P, _P], _I32),
    "bws_unpack_bitplanes_u8": ([_P, _I64, _P, _P], _I32),
    "bws_pack_bitplanes_i8": ([_P, _I64, _I64, _I64, _P, _I64, _P], _I32),
    "bws_bitplane_weights": ([_P], _I32),
    "bws_bitplane_row_counts": ([_P, _I32, _I64, _I64, _I64, _P, _P], _I32),
    "bws_row_ptr_workspace_size": ([_I64], _SZ),
    "bws_row_ptr": ([_P, _I64, _I64, _I64, _P, _P, _SZ, _P], _I32),
    "bws_pack_values_f32": ([_P, _I64, _I64, _I64, _P, _I64, _P, _P, _P], _I32),
    "bws_transpose_bitmap": ([_P, _I64, _I64, _I64, _P, _I64, _P], _I32),
    "bws_transpose_workspace_size": ([_I64, _I64], _SZ),
    "bws_transpose_values": ([_BP, _P, _I64, _P, _P, _P, _SZ, _P], _I32),
    "bws_transpose_perm": ([_BP, _P, _I64, _P, _P, _P, _SZ, _P], _I32),
    "bws_gather_f32": ([_P, _P, _I64, _P, _P], _I32),
    "bws_convert_workspace_size": ([_I64, _I64], _SZ),
    "bws_coo_to_bitmap": ([_P, _P, _P, _I64, _I64, _I64, _P, _I64, _P, _P, _P, _P, _SZ, _P], _I32),
    "bws_csr_to_bitmap": ([_P, _P, _P, _I64, _I64, _I64, _P, _I64, _P, _P, _P, _P, _SZ, _P],
                          _I32),
    "bws_bitmap_to_csr": ([_BP, _P, _P, _P, _P, _SZ, _P], _I32),
    "bws_group_tiles": ([_I64, _I64, _I32], _I64),
    "bws_group_rows": ([_BP, _I32, _I32, C.POINTER(_Groups), _P], _I32),
    "bws_group_spilled_workspace_size": ([_I64, _I64, _I32], _SZ),
    "bws_group_spilled_csr": ([_BP, _I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _SZ, _P], _I32),
    "bws_group_scatter": ([_P, _I64, _I64, _I64, _P, _P, _I32, _I32, _P, _I64, _P], _I32),
    "bws_group_gather": ([_P, _I64, _P, _I64, _I64, _I64, _P, _I32, _I32, _P, _P], _I32),
    "bws_cluster_workspace_size": ([_I64, _I64, _I32], _SZ),
    "bws_cluster_rows": ([_BP, _I32, _P, _P, _P, _SZ, _P], _I32),
    "bws_cluster_stream_size": ([_I64, _I64], _SZ),
    "bws_cluster_stream": ([_BP, _P, _P, _P, _SZ, _P], _I32),
    "bws_spmm_clustered": ([_BP, _P, _P, _P, _I64, _I64, _P, _I64, _F, _F, _P], _I32),
    "bws_spmm_workspace_size": ([_BP, _I64, _I32, C.POINTER(_Plan)], _SZ),
    "bws_spmm": ([_BP, _P, _I32, _I64, _I64, _P, _I64, _F, _F, _I32, C.POINTER(_Plan), _P, _SZ,
                  _P], _I32),
    "bws_spmm_fp32_noalloc_max_bytes": ([], _SZ),
    "bws_spmm_fp32_noalloc": ([_BP, _P, _I64, _I64, _P, _I64, _F, _F, _P], _I32),
    "bws_spmm_groups": ([_BP, _P, _P, _I32, _I32, _P, _I64, _I64, _P, _I64, _P], _I32),
    "bws_spmm_tA": ([_BP, _P, _I32, _I64, _I64, _P, _I64, _F, _F, _I32, C.POINTER(_Plan), _P,
                     _SZ, _P], _I32),
    "bws_sddmm_workspace_size": ([_BP], _SZ),
    "bws_sddmm": ([_BP, _P, _P, _I32, _I64, _I64, _I64, _P, _I32, _P, _SZ, _P], _I32),
    "bws_spmm_batched_workspace_size": ([_I32], _SZ),
    "bws_spmm_batched": ([C.POINTER(_Problem), _I32, _I32, _I32, C.POINTER(_Plan), _P, _SZ, _P],
                         _I32),
    "bws_batch_workspace_size": ([_I32], _SZ),
    "bws_batch_prepare": ([C.POINTER(_Problem), _I32, _I32, _I32, C.POINTER(_Plan), _P, _SZ,
                           C.POINTER(_Batch), _P], _I32),
    "bws_batch_run": ([C.POINTER(_Batch), _P], _I32),
    "bws_autotune": ([_BP, _I64, _P, _I32, _I64, _I64, _P, _I64, _I32, _I32, C.POINTER(_Plan),
                      C.POINTER(C.c_float), _P], _I32),
    "bws_autotune_lookup": ([_I64, _I64, _I64, _I64, _I32, C.POINTER(_Plan)], _I32),
    "bws_autotune_clear": ([], _I32),
    "bws_autotune_save": ([C.c_char_p], _I32),
    "bws_autotune_load": ([C.c_char_p], _I32),
    "bws_symbolic_or": ([_P, _I64, _I64, _I64, _P, _I64, _I64, _P, _I64, _P], _I32),
    "bws_spgemm_workspace_size": ([_I64, _I64, _I64], _SZ),
    "bws_spgemm": ([_BP, _BP, _BP, _F, _P, _P, _SZ, _P], _I32),
    "bws_spmm_bitplanes_workspace_size": ([_I64, _I64, _I32], _SZ),
    "bws_spmm_bitplanes_i8": ([_P, _I64, _I64, _I64, _P, _I32, _I64, _I64, _P, _I64, _F, _F, _I32,
                               C.POINTER(_Plan), _P, _SZ, _P], _I32),
    "bws_gen_bitmap_iid": ([_U64, _I64, _I64, _D, _P, _I64], _I32),
    "bws_gen_bitmap_clustered": ([_U64, _I64, _I64, _I32, _D, _D, _P, _I64], _I32),
    "bws_gen_f32": ([_U64, _I64, _I32, _D, _D, _P], _I32),
}

_lib = None


def lib():
    """The loaded libbws.so (ctypes), with argument types set."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m bitwise_spgemm_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (args, ret) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ret
        _lib = L
    return _lib


def header_functions():
    """Names of every function declared in include/bws.h."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bws_[a-z0-9_]+)\s*\(", text)))


def _check(status: int, where: str):
    if status != BWS_OK:
        raise BwsError(status, where)


def _stream(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _require_gpu(*tensors):
    for t in tensors:
        if t is not None and not t.is_cuda:
            raise ValueError("bitwise_spgemm_amd operates on GPU tensors (device='cuda')")


def words_for(cols: int) -> int:
    return (cols + 31) // 32


# ----------------------------------------------------------------------------
# bitmap container
# ----------------------------------------------------------------------------
@dataclass
class Bitmap:
    """A packed MSB-first bitmap [rows x ldw] (int32 storage of uint32 words),
    optionally value-carrying (row_ptr [rows+1] int64 + values [nnz] f32)."""

    words: torch.Tensor
    rows: int
    cols: int
    row_ptr: Optional[torch.Tensor] = None
    values: Optional[torch.Tensor] = None

    @property
    def ldw(self) -> int:
        return self.words.shape[1] if self.words.dim() == 2 else words_for(self.cols)

    def struct(self) -> _Bitmap:
        return _Bitmap(self.rows, self.cols, self.ldw, self.words.data_ptr() if self.words.numel() else None,
                       None if self.row_ptr is None else self.row_ptr.data_ptr(),
                       None if self.values is None else self.values.data_ptr())

    @property
    def nnz(self) -> int:
        if self.row_ptr is not None:
            return int(self.row_ptr[-1].item())
        return int(popcount(self.words).sum().item())

    def with_values(self, values: Optional[torch.Tensor], row_ptr_t: Optional[torch.Tensor] = None):
        rp = row_ptr_t if row_ptr_t is not None else (self.row_ptr if self.row_ptr is not None
                                                      else row_ptr(self))
        return Bitmap(self.words, self.rows, self.cols, rp, values)


def popcount(words: torch.Tensor) -> torch.Tensor:
    """Per-word popcount of int32-stored uint32 words (torch ops, for host checks)."""
    x = words.to(torch.int64) & 0xFFFFFFFF
    n = torch.zeros_like(x)
    for _ in range(32):
        n += x & 1
        x >>= 1
    return n


def from_host_words(W, rows: int, cols: int, device="cuda") -> Bitmap:
    """Upload a numpy uint32 [rows x ldw] bitmap."""
    import numpy as np
    W = np.ascontiguousarray(W, dtype=np.uint32)
    return Bitmap(torch.from_numpy(W.view(np.int32)).to(device), rows, cols)


def to_host_words(bm_or_words):
    import numpy as np
    w = bm_or_words.words if isinstance(bm_or_words, Bitmap) else bm_or_words
    return w.cpu().numpy().view(np.uint32)


# ----------------------------------------------------------------------------
# (1) layout transforms
# ----------------------------------------------------------------------------
def pack_bitmap(D: torch.Tensor, ldw: Optional[int] = None) -> Bitmap:
    """dense2bitmask: bit (31-i) of word w set iff D[r, 32w+i] != 0."""
    _require_gpu(D)
    assert D.dtype == torch.float32 and D.dim() == 2 and (D.stride(1) == 1 or D.shape[1] <= 1)
    rows, cols = D.shape
    ldw = ldw or words_for(cols)
    W = torch.empty((rows, ldw), dtype=torch.int32, device=D.device)
    _check(lib().bws_pack_bitmap_f32(_ptr(D), rows, cols, max(D.stride(0), cols, 1), _ptr(W), ldw,
                                     _stream(D.device)), "bws_pack_bitmap_f32")
    return Bitmap(W, rows, cols)


def unpack_bitmap(A: Bitmap) -> torch.Tensor:
    D = torch.empty((A.rows, A.cols), dtype=torch.float32, device=A.words.device)
    _check(lib().bws_unpack_bitmap_f32(_ptr(A.words), A.rows, A.cols, A.ldw, _ptr(D), A.cols,
                                       _stream(D.device)), "bws_unpack_bitmap_f32")
    return D


def unpack_bitplanes_f32(src: torch.Tensor, bit_width: int = 32) -> torch.Tensor:
    """layout_trans: [bit_width, n] int32 planes of the raw FP32 bit pattern."""
    _require_gpu(src)
    src = src.contiguous().view(-1)
    dst = torch.empty((bit_width, src.numel()), dtype=torch.int32, device=src.device)
    _check(lib().bws_unpack_bitplanes_f32(_ptr(src), src.numel(), bit_width, _ptr(dst),
                                          _stream(src.device)), "bws_unpack_bitplanes_f32")
    return dst


def unpack_bitplanes_u8(src: torch.Tensor) -> torch.Tensor:
    """uint8_layout_trans: [8, n] int32 planes."""
    _require_gpu(src)
    src = src.contiguous().view(-1)
    dst = torch.empty((8, src.numel()), dtype=torch.int32, device=src.device)
    _check(lib().bws_unpack_bitplanes_u8(_ptr(src), src.numel(), _ptr(dst), _stream(src.device)),
           "bws_unpack_bitplanes_u8")
    return dst


def pack_bitplanes_i8(D: torch.Tensor, ldw: Optional[int] = None) -> torch.Tensor:
    """8 packed planes [8, rows, ldw] of an int8 matrix."""
    _require_gpu(D)
    assert D.dtype == torch.int8 and D.dim() == 2 and (D.stride(1) == 1 or D.shape[1] <= 1)
    rows, cols = D.shape
    ldw = ldw or words_for(cols)
    P = torch.empty((8, rows, ldw), dtype=torch.int32, device=D.device)
    _check(lib().bws_pack_bitplanes_i8(_ptr(D), rows, cols, D.stride(0), _ptr(P), ldw,
                                       _stream(D.device)), "bws_pack_bitplanes_i8")
    return P


def bitplane_weights() -> list:
    w = (C.c_float * 8)()
    _check(lib().bws_bitplane_weights(C.cast(w, C.c_void_p)), "bws_bitplane_weights")
    return list(w)


def bitplane_row_counts(P: torch.Tensor, cols: int) -> torch.Tensor:
    planes, rows, ldw = P.shape
    out = torch.empty((planes, rows), dtype=torch.int32, device=P.device)
    _check(lib().bws_bitplane_row_counts(_ptr(P), planes, rows, cols, ldw, _ptr(out),
                                         _stream(P.device)), "bws_bitplane_row_counts")
    return out


def _workspace(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def row_ptr(A: Bitmap) -> torch.Tensor:
    """Exclusive prefix of per-row set-bit counts, [rows+1] int64."""
    dev = A.words.device
    rp = torch.empty(A.rows + 1, dtype=torch.int64, device=dev)
    ws = _workspace(lib().bws_row_ptr_workspace_size(A.rows), dev)
    _check(lib().bws_row_ptr(_ptr(A.words), A.rows, A.cols, A.ldw, _ptr(rp), _ptr(ws), ws.numel(),
                             _stream(dev)), "bws_row_ptr")
    return rp


def pack_values(D: torch.Tensor, A: Bitmap) -> Bitmap:
    """Value-carrying bitmap of D's nonzeros (values in row, column order)."""
    _require_gpu(D)
    rp = A.row_ptr if A.row_ptr is not None else row_ptr(A)
    nnz = int(rp[-1].item())
    vals = torch.empty(max(nnz, 1), dtype=torch.float32, device=D.device)[:nnz]
    _check(lib().bws_pack_values_f32(_ptr(D), A.rows, A.cols, D.stride(0), _ptr(A.words), A.ldw,
                                     _ptr(rp), _ptr(vals) if nnz else _ptr(torch.empty(1, device=D.device)),
                                     _stream(D.device)), "bws_pack_values_f32")
    return Bitmap(A.words, A.rows, A.cols, rp, vals)


def transpose(A: Bitmap, with_values: bool = True) -> Bitmap:
    """Aᵀ as a bitmap (cols x rows); value-carrying if A is."""
    dev = A.words.device
    ldwt = words_for(A.rows)
    WT = torch.empty((A.cols, ldwt), dtype=torch.int32, device=dev)
    if A.row_ptr is None and (A.values is None or not with_values):
        _check(lib().bws_transpose_bitmap(_ptr(A.words), A.rows, A.cols, A.ldw, _ptr(WT), ldwt,
                                          _stream(dev)), "bws_transpose_bitmap")
        return Bitmap(WT, A.cols, A.rows)
    nnz = int(A.row_ptr[-1].item())
    rpt = torch.empty(A.cols + 1, dtype=torch.int64, device=dev)
    vt = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev) if A.values is not None else None
    ws = _workspace(lib().bws_transpose_workspace_size(A.rows, A.cols), dev)
    s = A.struct()
    _check(lib().bws_transpose_values(C.byref(s), _ptr(WT), ldwt, _ptr(rpt), _ptr(vt), _ptr(ws),
                                      ws.numel(), _stream(dev)), "bws_transpose_values")
    return Bitmap(WT, A.cols, A.rows, rpt, None if vt is None else vt[:nnz])


def transpose_perm(A: Bitmap):
    """Aᵀ's words and row pointers plus perm (int64 [nnz]): AT's j-th value is
    A's values[perm[j]].  For a fixed mask whose values change, AT's values are
    then gather(values, perm)."""
    dev = A.words.device
    if A.row_ptr is None:
        A = A.with_values(A.values, row_ptr(A))
    ldwt = words_for(A.rows)
    WT = torch.empty((A.cols, ldwt), dtype=torch.int32, device=dev)
    nnz = int(A.row_ptr[-1].item())
    rpt = torch.empty(A.cols + 1, dtype=torch.int64, device=dev)
    perm = torch.empty(max(nnz, 1), dtype=torch.int64, device=dev)
    ws = _workspace(lib().bws_transpose_workspace_size(A.rows, A.cols), dev)
    s = A.struct()
    _check(lib().bws_transpose_perm(C.byref(s), _ptr(WT), ldwt, _ptr(rpt), _ptr(perm), _ptr(ws),
                                    ws.numel(), _stream(dev)), "bws_transpose_perm")
    return Bitmap(WT, A.cols, A.rows, rpt), perm[:nnz]


def gather(src: torch.Tensor, idx: torch.Tensor, out: Optional[torch.Tensor] = None):
    """out[i] = src[idx[i]] (FP32 values, int64 indices) on the device."""
    _require_gpu(src, idx)
    assert src.dtype == torch.float32 and idx.dtype == torch.int64
    if out is None:
        out = torch.empty(idx.numel(), dtype=torch.float32, device=src.device)
    _check(lib().bws_gather_f32(_ptr(src), _ptr(idx), idx.numel(), _ptr(out),
                                _stream(src.device)), "bws_gather_f32")
    return out


# ----------------------------------------------------------------------------
# (2) row grouping
# ----------------------------------------------------------------------------
def coo_to_bitmap(row_idx: torch.Tensor, col_idx: torch.Tensor, values: Optional[torch.Tensor],
                  shape, with_values: bool = True):
    """COO entries (int32 device tensors, any order) -> (Bitmap, n_invalid).
    The Bitmap carries row_ptr, and values in bitmap order if given."""
    rows, cols = shape
    dev = row_idx.device
    nnz = row_idx.numel()
    ldw = max(words_for(cols), 1)
    W = torch.empty((max(rows, 1), ldw), dtype=torch.int32, device=dev)
    rp = torch.empty((rows + 1,), dtype=torch.int64, device=dev)
    vout = None
    if with_values and values is not None:
        vout = torch.empty((max(nnz, 1),), dtype=torch.float32, device=dev)
    bad = torch.empty((1,), dtype=torch.int64, device=dev)
    ws = _workspace(lib().bws_convert_workspace_size(rows, cols), dev)
    _check(lib().bws_coo_to_bitmap(_ptr(row_idx), _ptr(col_idx), _ptr(values), nnz, rows, cols,
                                   _ptr(W), ldw, _ptr(rp), _ptr(vout), _ptr(bad), _ptr(ws),
                                   ws.numel(), _stream(dev)), "bws_coo_to_bitmap")
    A = Bitmap(W[:rows], rows, cols, rp, None if vout is None else vout[:nnz])
    return A, bad


def csr_to_bitmap(row_ptr_t: torch.Tensor, col_idx: torch.Tensor, values: Optional[torch.Tensor],
                  shape, with_values: bool = True):
    """CSR (int64 row pointers, int32 columns, any column order) -> (Bitmap, n_invalid)."""
    rows, cols = shape
    dev = col_idx.device
    nnz = col_idx.numel()
    ldw = max(words_for(cols), 1)
    W = torch.empty((max(rows, 1), ldw), dtype=torch.int32, device=dev)
    rp = torch.empty((rows + 1,), dtype=torch.int64, device=dev)
    vout = None
    if with_values and values is not None:
        vout = torch.empty((max(nnz, 1),), dtype=torch.float32, device=dev)
    bad = torch.empty((1,), dtype=torch.int64, device=dev)
    ws = _workspace(lib().bws_convert_workspace_size(rows, cols), dev)
    _check(lib().bws_csr_to_bitmap(_ptr(row_ptr_t), _ptr(col_idx), _ptr(values), nnz, rows, cols,
                                   _ptr(W), ldw, _ptr(rp), _ptr(vout), _ptr(bad), _ptr(ws),
                                   ws.numel(), _stream(dev)), "bws_csr_to_bitmap")
    A = Bitmap(W[:rows], rows, cols, rp, None if vout is None else vout[:nnz])
    return A, bad


def bitmap_to_csr(A: Bitmap):
    """Bitmap -> (row_ptr int64 [rows+1], col_idx int32 [nnz], values f32 [nnz])."""
    dev = A.words.device
    rp = torch.empty((A.rows + 1,), dtype=torch.int64, device=dev)
    ws = _workspace(lib().bws_convert_workspace_size(A.rows, A.cols), dev)
    s = A.struct()
    _check(lib().bws_bitmap_to_csr(C.byref(s), _ptr(rp), None, None, _ptr(ws), ws.numel(),
                                   _stream(dev)), "bws_bitmap_to_csr")
    nnz = int(rp[-1].item())
    col = torch.empty((max(nnz, 1),), dtype=torch.int32, device=dev)
    vals = torch.empty((max(nnz, 1),), dtype=torch.float32, device=dev)
    _check(lib().bws_bitmap_to_csr(C.byref(s), _ptr(rp), _ptr(col), _ptr(vals), _ptr(ws),
                                   ws.numel(), _stream(dev)), "bws_bitmap_to_csr")
    return rp, col[:nnz], vals[:nnz]


def read_matrix_market(path: str):
    """Host reader for Matrix Market coordinate files (real / integer /
    pattern; general / symmetric / skew-symmetric).  Returns (rows, cols,
    row_idx int32, col_idx int32, values float32), 0-based, with symmetric
    off-diagonal entries mirrored."""
    import numpy as np
    with open(path, "r") as f:
        header = f.readline().split()
        if len(header) < 5 or header[0].lower() != "%%matrixmarket" or \
                header[1].lower() != "matrix" or header[2].lower() != "coordinate":
            raise ValueError(f"{path}: not a Matrix Market coordinate matrix")
        field, sym = header[3].lower(), header[4].lower()
        if field not in ("real", "integer", "pattern", "double"):
            raise ValueError(f"{path}: unsupported field {field}")
        if sym not in ("general", "symmetric", "skew-symmetric"):
            raise ValueError(f"{path}: unsupported symmetry {sym}")
        line = f.readline()
        while line.startswith("%") or not line.strip():
            line = f.readline()
        rows, cols, nnz = (int(x) for x in line.split())
        ncol = 2 if field == "pattern" else 3
        data = np.loadtxt(f, dtype=np.float64, ndmin=2, max_rows=nnz)
    if data.shape != (nnz, ncol) and not (nnz == 0):
        raise ValueError(f"{path}: expected {nnz} entries of {ncol} fields, got {data.shape}")
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int64) - 1
    v = data[:, 2].astype(np.float32) if ncol == 3 else np.ones(nnz, np.float32)
    if sym != "general":
        off = r != c
        sign = -1.0 if sym == "skew-symmetric" else 1.0
        r, c, v = (np.concatenate([r, c[off]]), np.concatenate([c, r[off]]),
                   np.concatenate([v, (sign * v[off]).astype(np.float32)]))
    return rows, cols, r.astype(np.int32), c.astype(np.int32), v


def load_matrix_market(path: str, device="cuda"):
    """Matrix Market file -> (value-carrying Bitmap on the device, n_invalid)."""
    rows, cols, r, c, v = read_matrix_market(path)
    dev = torch.device(device)
    return coo_to_bitmap(torch.from_numpy(r).to(dev), torch.from_numpy(c).to(dev),
                         torch.from_numpy(v).to(dev), (rows, cols))


def group_tiles(rows: int, cols: int, tile_rows: int) -> int:
    return int(lib().bws_group_tiles(rows, cols, tile_rows))


def group_rows(A: Bitmap, tile_rows: int = 2048, max_groups: int = 2, member_row: bool = True,
               spill_tables: bool = True) -> dict:
    """generate_groups_large (canonical ascending first fit); dict of tensors."""
    dev = A.words.device
    tiles = group_tiles(A.rows, A.cols, tile_rows)
    G, T = max_groups, tile_rows
    i32 = dict(dtype=torch.int32, device=dev)
    out = dict(
        group_id=torch.empty((A.rows, A.ldw), **i32),
        group_mask=torch.empty((tiles, G), **i32),
        group_vals=torch.empty((tiles, G, 32), dtype=torch.float32, device=dev),
        member_row=torch.empty((tiles, G, 32), **i32) if member_row else None,
        spill_cnt=torch.empty(tiles, **i32),
        spill_nnz=torch.empty(tiles, **i32),
        spill_table=torch.empty((tiles, T), **i32) if spill_tables else None,
        spill_rev=torch.empty((tiles, T), **i32) if spill_tables else None,
        grouped_nnz=torch.empty(1, dtype=torch.int64, device=dev),
    )
    g = _Groups(*[None if out[k] is None else out[k].data_ptr() for k in (
        "group_id", "group_mask", "group_vals", "member_row", "spill_cnt", "spill_nnz",
        "spill_table", "spill_rev", "grouped_nnz")])
    s = A.struct()
    _check(lib().bws_group_rows(C.byref(s), tile_rows, max_groups, C.byref(g), _stream(dev)),
           "bws_group_rows")
    return out


def group_spilled_csr(A: Bitmap, tile_rows: int, groups: dict) -> dict:
    dev = A.words.device
    tiles = group_tiles(A.rows, A.cols, tile_rows)
    nrow = int(groups["spill_cnt"].sum().item())
    nnz = int(groups["spill_nnz"].sum().item())
    i32 = dict(dtype=torch.int32, device=dev)
    out = dict(spill_row_off=torch.empty(tiles + 1, **i32),
               spill_nnz_off=torch.empty(tiles + 1, **i32),
               row_end=torch.empty(max(nrow, 1), **i32),
               col=torch.empty(max(nnz, 1), **i32),
               val=torch.empty(max(nnz, 1), dtype=torch.float32, device=dev))
    ws = _workspace(lib().bws_group_spilled_workspace_size(A.rows, A.cols, tile_rows), dev)
    s = A.struct()
    _check(lib().bws_group_spilled_csr(
        C.byref(s), tile_rows, _ptr(groups["spill_cnt"]), _ptr(groups["spill_nnz"]),
        _ptr(groups["spill_table"]), _ptr(out["spill_row_off"]), _ptr(out["spill_nnz_off"]),
        _ptr(out["row_end"]), _ptr(out["col"]), _ptr(out["val"]), _ptr(ws), ws.numel(),
        _stream(dev)), "bws_group_spilled_csr")
    out["row_end"] = out["row_end"][:nrow]
    out["col"] = out["col"][:nnz]
    out["val"] = out["val"][:nnz]
    return out


def group_scatter(A: Bitmap, groups: dict, tile_rows: int, max_groups: int,
                  D: Optional[torch.Tensor] = None) -> torch.Tensor:
    """group2dense: write group values into D at every grouped set bit."""
    dev = A.words.device
    if D is None:
        D = torch.zeros((A.rows, A.cols), dtype=torch.float32, device=dev)
    _check(lib().bws_group_scatter(_ptr(A.words), A.rows, A.cols, A.ldw, _ptr(groups["group_id"]),
                                   _ptr(groups["group_vals"]), tile_rows, max_groups, _ptr(D),
                                   D.stride(0), _stream(dev)), "bws_group_scatter")
    return D


def group_gather(D: torch.Tensor, A: Bitmap, groups: dict, tile_rows: int, max_groups: int,
                 group_vals: Optional[torch.Tensor] = None) -> torch.Tensor:
    """dense2group_from_idx: read group values back from D."""
    dev = A.words.device
    gv = groups["group_vals"].clone() if group_vals is None else group_vals
    _check(lib().bws_group_gather(_ptr(D), D.stride(0), _ptr(A.words), A.rows, A.cols, A.ldw,
                                  _ptr(groups["group_id"]), tile_rows, max_groups, _ptr(gv),
                                  _stream(dev)), "bws_group_gather")
    return gv


SLOT_DELTA = 1 << 30
SLOT_BASE1 = 1 << 29


class RowClusters:
    """Row-clustering plan of a 0/1 bitmap (bws_cluster_rows): slot_row int32
    [rows] (row | SLOT_DELTA for rows summed as a difference from a chunk
    base, | SLOT_BASE1 when that is the chunk's second base), chunk_mask
    int32 [ceil(rows/16), 2, ldw] (the chunks' two bases, uint32 words held as
    int32) and the operand stream of the clustered product (bws_cluster_stream,
    uint8 bytes)."""

    def __init__(self, slot_row: torch.Tensor, chunk_mask: torch.Tensor, stream: torch.Tensor,
                 rows: int, cols: int, max_clusters: int):
        self.slot_row = slot_row
        self.chunk_mask = chunk_mask
        self.stream = stream
        self.rows, self.cols, self.max_clusters = rows, cols, max_clusters


def cluster_rows(A: Bitmap, max_clusters: int = 0) -> RowClusters:
    """Row clustering for spmm_clustered (0/1 masks; max_clusters 0 = rows / 64),
    with the product's operand stream."""
    dev = A.words.device
    slot_row = torch.empty(max(A.rows, 1), dtype=torch.int32, device=dev)[:A.rows]
    chunk_mask = torch.empty(((A.rows + 15) // 16, 2, A.ldw), dtype=torch.int32, device=dev)
    ws = _workspace(lib().bws_cluster_workspace_size(A.rows, A.cols, max_clusters), dev)
    s = A.struct()
    _check(lib().bws_cluster_rows(C.byref(s), max_clusters, _ptr(slot_row), _ptr(chunk_mask),
                                  _ptr(ws), ws.numel(), _stream(dev)), "bws_cluster_rows")
    nbytes = lib().bws_cluster_stream_size(A.rows, A.cols)
    stream = _workspace(nbytes, dev)
    _check(lib().bws_cluster_stream(C.byref(s), _ptr(slot_row), _ptr(chunk_mask), _ptr(stream),
                                    stream.numel(), _stream(dev)), "bws_cluster_stream")
    return RowClusters(slot_row, chunk_mask, stream, A.rows, A.cols, max_clusters)


def spmm_clustered(A: Bitmap, clusters: RowClusters, B: torch.Tensor, alpha: float = 1.0,
                   beta: float = 0.0, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """C = alpha * A·B + beta * C through the row-clustering plan (FP32 B)."""
    _require_gpu(B)
    assert B.dim() == 2 and B.dtype == torch.float32 and B.shape[0] == A.cols
    assert (clusters.rows, clusters.cols) == (A.rows, A.cols)
    N = B.shape[1]
    dev = B.device
    if out is None:
        if beta != 0.0:
            raise ValueError("beta != 0 needs `out`")
        out = torch.empty((A.rows, N), dtype=torch.float32, device=dev)
    s = A.struct()
    _check(lib().bws_spmm_clustered(C.byref(s), _ptr(clusters.slot_row), _ptr(clusters.stream),
                                    _ptr(B), B.stride(0), N, _ptr(out), out.stride(0), alpha,
                                    beta, _stream(dev)),
           "bws_spmm_clustered")
    return out


# ----------------------------------------------------------------------------
# (3) grouped tile GEMM
# ----------------------------------------------------------------------------
MFMA_STAGING = {"auto": 0, "lut": 1, "regs": 2, "dma3": 3, "dma2": 4, "sparse": 5}


def make_plan(kernel: int = 0, rows_per_wg: int = 0, waves_per_wg: int = 0, k_stage: int = 0,
              mfma_min_bits: int = 0, cols_per_wg: int = 0, mfma_staging: str = "auto"):
    return _Plan(kernel, rows_per_wg, waves_per_wg, k_stage, mfma_min_bits, cols_per_wg,
                 MFMA_STAGING[mfma_staging], 0)


def spmm(A: Bitmap, B: torch.Tensor, precision: str = "fp32", alpha: float = 1.0,
         beta: float = 0.0, out: Optional[torch.Tensor] = None, plan=None) -> torch.Tensor:
    """C = alpha * A·B + beta * C (C FP32, M x N)."""
    _require_gpu(B)
    assert B.dim() == 2 and (B.stride(1) == 1 or B.shape[1] <= 1) and B.shape[0] == A.cols
    N = B.shape[1]
    dev = B.device
    if out is None:
        out = torch.empty((A.rows, N), dtype=torch.float32, device=dev)
        if beta != 0.0:
            raise ValueError("beta != 0 needs `out`")
    s = A.struct()
    pl = None if plan is None else C.byref(plan)
    wsz = lib().bws_spmm_workspace_size(C.byref(s), N, PRECISION[precision], pl)
    if (precision in ("bf16", "fp16") and A.values is None and B.dtype != torch.float32
            and (plan is None or plan.mfma_min_bits == 0)):
        wsz = 0  # the workspace would hold a rounded FP32 B: not needed for 16-bit B
    ws = _workspace(wsz, dev) if wsz else None
    _check(lib().bws_spmm(C.byref(s), _ptr(B), DTYPE[B.dtype], B.stride(0), N, _ptr(out),
                          out.stride(0), alpha, beta, PRECISION[precision], pl, _ptr(ws),
                          0 if ws is None else ws.numel(), _stream(dev)),
           "bws_spmm")
    return out


def spmm_fp32_noalloc(A: Bitmap, B: torch.Tensor, alpha: float = 1.0, beta: float = 0.0,
                      out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """C = alpha * A·B + beta * C at FP32 through bws_spmm_fp32_noalloc: the
    library allocates nothing (the mask stream lives in its own device array;
    calls overlapping on one device race on it)."""
    _require_gpu(B)
    assert B.dtype == torch.float32 and B.dim() == 2 and B.shape[0] == A.cols
    assert B.stride(1) == 1 or B.shape[1] <= 1
    N = B.shape[1]
    if out is None:
        out = torch.empty((A.rows, N), dtype=torch.float32, device=B.device)
        if beta != 0.0:
            raise ValueError("beta != 0 needs `out`")
    s = A.struct()
    _check(lib().bws_spmm_fp32_noalloc(C.byref(s), _ptr(B), B.stride(0), N, _ptr(out),
                                       out.stride(0), alpha, beta, _stream(B.device)),
           "bws_spmm_fp32_noalloc")
    return out


def spmm_tA(AT: Bitmap, dC: torch.Tensor, precision: str = "fp32", alpha: float = 1.0,
            beta: float = 0.0, out: Optional[torch.Tensor] = None, plan=None) -> torch.Tensor:
    """dB = alpha * Aᵀ·dC + beta * dB, given AT = transpose(A)."""
    _require_gpu(dC)
    N = dC.shape[1]
    dev = dC.device
    if out is None:
        out = torch.empty((AT.rows, N), dtype=torch.float32, device=dev)
    s = AT.struct()
    pl = None if plan is None else C.byref(plan)
    wsz = lib().bws_spmm_workspace_size(C.byref(s), N, PRECISION[precision], pl)
    if precision in ("bf16", "fp16") and AT.values is None and dC.dtype != torch.float32:
        wsz = 0  # the workspace would hold a rounded FP32 dC: not needed for 16-bit dC
    ws = _workspace(wsz, dev) if wsz else None
    _check(lib().bws_spmm_tA(C.byref(s), _ptr(dC), DTYPE[dC.dtype], dC.stride(0), N, _ptr(out),
                             out.stride(0), alpha, beta, PRECISION[precision], pl, _ptr(ws),
                             0 if ws is None else ws.numel(), _stream(dev)),
           "bws_spmm_tA")
    return out


def spmm_groups(A: Bitmap, groups: dict, tile_rows: int, max_groups: int,
                B: torch.Tensor) -> torch.Tensor:
    """C = A·B with values read from the group tables (grouped evaluation)."""
    N = B.shape[1]
    out = torch.empty((A.rows, N), dtype=torch.float32, device=B.device)
    s = A.struct()
    _check(lib().bws_spmm_groups(C.byref(s), _ptr(groups["group_id"]), _ptr(groups["group_vals"]),
                                 tile_rows, max_groups, _ptr(B), B.stride(0), N, _ptr(out),
                                 out.stride(0), _stream(B.device)), "bws_spmm_groups")
    return out


def sddmm(A: Bitmap, dC: torch.Tensor, B: torch.Tensor, precision: str = "fp32",
          out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """dA values = (dC·Bᵀ) at A's set bits, in A's value order."""
    _require_gpu(dC, B)
    assert dC.dtype == B.dtype
    rp = A.row_ptr if A.row_ptr is not None else row_ptr(A)
    nnz = int(rp[-1].item()) if out is None else out.numel()
    if out is None:
        out = torch.empty(max(nnz, 1), dtype=torch.float32, device=B.device)[:nnz]
    bm = Bitmap(A.words, A.rows, A.cols, rp, None)
    s = bm.struct()
    ws = _workspace(lib().bws_sddmm_workspace_size(C.byref(s)), B.device)
    _check(lib().bws_sddmm(C.byref(s), _ptr(dC), _ptr(B), DTYPE[B.dtype], dC.stride(0),
                           B.stride(0), B.shape[1], _ptr(out) if nnz else _ptr(ws),
                           PRECISION[precision], _ptr(ws), ws.numel(), _stream(B.device)),
           "bws_sddmm")
    return out


class Batch:
    """A prepared grouped launch (bws_batch_prepare); run() is capturable."""

    def __init__(self, problems, precision="fp32", plan=None, nnz_hints=None):
        n = len(problems)
        self.keep = problems                      # device tensors must outlive the batch
        self.arr = (_Problem * max(n, 1))()
        self.dtype = DTYPE[problems[0][1].dtype] if n else 0
        self.device = problems[0][1].device if n else torch.device("cuda")
        for i, (A, B, Cm) in enumerate(problems):
            hint = 0 if nnz_hints is None else int(nnz_hints[i])
            self.arr[i] = _Problem(A.struct(), B.data_ptr(), B.stride(0), B.shape[1],
                                   Cm.data_ptr(), Cm.stride(0), hint)
        self.ws = _workspace(lib().bws_batch_workspace_size(n), self.device)
        self.batch = _Batch()
        _check(lib().bws_batch_prepare(self.arr, n, self.dtype, PRECISION[precision],
                                       None if plan is None else C.byref(plan), _ptr(self.ws),
                                       self.ws.numel(), C.byref(self.batch),
                                       _stream(self.device)), "bws_batch_prepare")

    @property
    def launches(self):
        return [dict(kind=L.kind, rows=L.rows_per_wg, cols=L.cols_per_wg, waves=L.waves_per_wg,
                     k_stage=L.k_stage,
                     values=L.values, count=L.count, tiles=L.total_tiles)
                for L in self.batch.launch[:self.batch.n_launch]]

    def run(self):
        _check(lib().bws_batch_run(C.byref(self.batch), _stream(self.device)), "bws_batch_run")


def spmm_batched(problems, precision: str = "fp32", plan=None, nnz_hints=None):
    """problems: list of (A: Bitmap, B: Tensor, C: Tensor); C = A·B for each (one
    grouped launch per tile configuration)."""
    if not problems:
        return None
    b = Batch(problems, precision, plan, nnz_hints)
    b.run()
    return b


def autotune(A: Bitmap, B: torch.Tensor, precision: str = "fp32", nnz: Optional[int] = None,
             repeats: int = 3):
    """Time the candidate plans on this problem; returns (plan, ms) and caches it."""
    nnz = A.nnz if nnz is None else nnz
    N = B.shape[1]
    out = torch.empty((A.rows, N), dtype=torch.float32, device=B.device)
    plan = _Plan()
    ms = C.c_float()
    s = A.struct()
    _check(lib().bws_autotune(C.byref(s), nnz, _ptr(B), DTYPE[B.dtype], B.stride(0), N, _ptr(out),
                              out.stride(0), PRECISION[precision], repeats, C.byref(plan),
                              C.byref(ms), _stream(B.device)), "bws_autotune")
    return plan, ms.value


def autotune_lookup(M, K, N, nnz, precision="fp32"):
    plan = _Plan()
    st = lib().bws_autotune_lookup(M, K, N, nnz, PRECISION[precision], C.byref(plan))
    return plan if st == BWS_OK else None


def symbolic_or(A: Bitmap, Bm: Bitmap) -> Bitmap:
    """pre_spgemm's symbolic phase: the bitmap of A·B's structural nonzeros."""
    assert A.cols == Bm.rows
    dev = A.words.device
    ldwc = words_for(Bm.cols)
    WC = torch.empty((A.rows, max(ldwc, 1)), dtype=torch.int32, device=dev)
    _check(lib().bws_symbolic_or(_ptr(A.words), A.rows, A.cols, A.ldw, _ptr(Bm.words), Bm.cols,
                                 Bm.ldw, _ptr(WC), WC.shape[1], _stream(dev)), "bws_symbolic_or")
    return Bitmap(WC, A.rows, Bm.cols)


def spgemm(A: Bitmap, Bm: Bitmap, Cm: Optional[Bitmap] = None, alpha: float = 1.0) -> Bitmap:
    """Numeric sparse x sparse product: C = alpha * A·Bm on the pattern Cm (by
    default the symbolic product of A's and Bm's patterns), returned as a
    value-carrying bitmap.  bws.group_rows on it gives the grouped form."""
    assert A.cols == Bm.rows
    dev = A.words.device
    if Cm is None:
        Cm = symbolic_or(A, Bm)
    if Cm.row_ptr is None:
        Cm = Bitmap(Cm.words, Cm.rows, Cm.cols, row_ptr(Cm))
    nnz = int(Cm.row_ptr[-1].item()) if Cm.rows else 0
    store = torch.empty(max(nnz, 1), dtype=torch.float32, device=dev)  # never a null pointer
    ws = _workspace(lib().bws_spgemm_workspace_size(A.rows, A.cols, Bm.cols), dev)
    a, b, c = A.struct(), Bm.struct(), Cm.struct()
    _check(lib().bws_spgemm(C.byref(a), C.byref(b), C.byref(c), alpha, _ptr(store), _ptr(ws),
                            ws.numel(), _stream(dev)), "bws_spgemm")
    return Bitmap(Cm.words, Cm.rows, Cm.cols, Cm.row_ptr, store[:nnz])


def spmm_bitplanes_i8(P: torch.Tensor, cols: int, B: torch.Tensor, precision: str = "fp32",
                      alpha: float = 1.0, beta: float = 0.0, out=None, plan=None,
                      one_pass: bool = True) -> torch.Tensor:
    """C = Σ_b w_b P_b·B for int8 bit-planes P [8, M, ldw] (weights 1..64, -128).
    one_pass: give the library its workspace (B read once); False runs the
    eight weighted 0/1 products."""
    _, M, ldw = P.shape
    N = B.shape[1]
    if out is None:
        out = torch.empty((M, N), dtype=torch.float32, device=B.device)
    ws = None
    if one_pass:
        ws = _workspace(lib().bws_spmm_bitplanes_workspace_size(M, cols, PRECISION[precision]),
                        B.device)
    _check(lib().bws_spmm_bitplanes_i8(_ptr(P), M, cols, ldw, _ptr(B), DTYPE[B.dtype], B.stride(0),
                                       N, _ptr(out), out.stride(0), alpha, beta,
                                       PRECISION[precision],
                                       None if plan is None else C.byref(plan), _ptr(ws),
                                       0 if ws is None else ws.numel(), _stream(B.device)),
           "bws_spmm_bitplanes_i8")
    return out


# ----------------------------------------------------------------------------
# host generators (no GPU)
# ----------------------------------------------------------------------------
def gen_bitmap_iid(seed: int, rows: int, cols: int, p: float):
    import numpy as np
    W = np.empty((rows, words_for(cols)), np.uint32)
    _check(lib().bws_gen_bitmap_iid(seed, rows, cols, p, W.ctypes.data_as(C.c_void_p), W.shape[1]),
           "bws_gen_bitmap_iid")
    return W


def gen_bitmap_clustered(seed: int, rows: int, cols: int, n_proto: int, p_proto: float,
                         flip: float):
    import numpy as np
    W = np.empty((rows, words_for(cols)), np.uint32)
    _check(lib().bws_gen_bitmap_clustered(seed, rows, cols, n_proto, p_proto, flip,
                                          W.ctypes.data_as(C.c_void_p), W.shape[1]),
           "bws_gen_bitmap_clustered")
    return W


def gen_f32(seed: int, shape, kind: str = "normal", a: float = 0.0, b: float = 1.0):
    """kind: 'normal' N(a, b^2), 'uniform' U[a, b), 'logu' random sign * 10^U[a, b)."""
    import numpy as np
    out = np.empty(shape, np.float32)
    k = {"normal": 0, "uniform": 1, "logu": 2}[kind]
    _check(lib().bws_gen_f32(seed, out.size, k, a, b, out.ctypes.data_as(C.c_void_p)),
           "bws_gen_f32")
    return out


# reference-named aliases
dense2bitmask = pack_bitmap
layout_trans = unpack_bitplanes_f32
uint8_layout_trans = unpack_bitplanes_u8
generate_groups_large = group_rows
generate_spilled_csr = group_spilled_csr
group2dense = group_scatter
dense2group_from_idx = group_gather
spgemm_compute = spmm
