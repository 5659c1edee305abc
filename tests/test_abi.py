"""CPU checks of the C ABI: libbws.so loads, exports every function that
include/bws.h declares, and rejects invalid calls on the host (status codes,
no abort, no GPU needed).  Also the host-side generators."""
import ctypes as C
import subprocess

import numpy as np
import pytest


def test_exports_every_declared_symbol(bws):
    names = bws.header_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", bws.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(bws.lib(), n)


def test_status_strings_and_version(bws):
    L = bws.lib()
    for code in (0, -1, -2, -3, -4, -5, -6, -7):
        assert L.bws_status_string(code).startswith(b"BWS_")
    assert L.bws_status_string(123) == b"BWS_UNKNOWN"
    assert L.bws_version() >> 16 == 1
    assert L.bws_last_hip_error() == 0


def test_weights_host(bws):
    assert bws.bitplane_weights() == [1, 2, 4, 8, 16, 32, 64, -128]


def _bm(rows, cols, ldw, words=0x1000, row_ptr=None, values=None):
    import bitwise_spgemm_amd as b
    return b._Bitmap(rows, cols, ldw, words, row_ptr, values)


FAKE = 0x10000  # never dereferenced: validation fails first


def test_pack_validation(bws):
    L = bws.lib()
    assert L.bws_pack_bitmap_f32(FAKE, -1, 4, 4, FAKE, 1, None) == bws.BWS_EINVAL
    assert L.bws_pack_bitmap_f32(FAKE, 2, 40, 39, FAKE, 2, None) == bws.BWS_EINVAL  # ld < cols
    assert L.bws_pack_bitmap_f32(FAKE, 2, 40, 40, FAKE, 1, None) == bws.BWS_EINVAL  # ldw too small
    assert L.bws_pack_bitmap_f32(None, 2, 40, 40, FAKE, 2, None) == bws.BWS_EINVAL
    assert L.bws_pack_bitmap_f32(FAKE, 0, 40, 40, FAKE, 2, None) == bws.BWS_OK     # empty: no-op
    assert L.bws_unpack_bitplanes_f32(FAKE, 10, 33, FAKE, None) == bws.BWS_EINVAL
    assert L.bws_unpack_bitplanes_f32(FAKE, 10, 0, FAKE, None) == bws.BWS_EINVAL
    assert L.bws_row_ptr(FAKE, 10, 64, 2, FAKE, None, 0, None) == bws.BWS_EWORKSPACE
    assert L.bws_row_ptr(FAKE, 10, 64, 2, FAKE, FAKE, 1, None) == bws.BWS_EWORKSPACE
    ws = L.bws_row_ptr_workspace_size(10)
    assert L.bws_row_ptr(FAKE, 10, 64, 2, FAKE, FAKE + 8, ws, None) == bws.BWS_EWORKSPACE  # alignment


def test_transpose_validation(bws):
    L = bws.lib()
    assert L.bws_transpose_bitmap(FAKE, 40, 64, 2, FAKE, 1, None) == bws.BWS_EINVAL  # ldwt < 2
    a = _bm(40, 64, 2)
    assert L.bws_transpose_values(C.byref(a), FAKE, 2, FAKE, FAKE, FAKE, 1, None) == bws.BWS_EINVAL
    a = _bm(40, 64, 2, row_ptr=FAKE)
    assert L.bws_transpose_values(C.byref(a), FAKE, 2, FAKE, FAKE, FAKE, 1, None) == bws.BWS_EWORKSPACE


def test_group_validation(bws):
    L = bws.lib()
    a = _bm(100, 64, 2)
    g = bws._Groups(*([FAKE] * 9))
    assert L.bws_group_rows(C.byref(a), 0, 2, C.byref(g), None) == bws.BWS_EINVAL
    assert L.bws_group_rows(C.byref(a), 70000, 2, C.byref(g), None) == bws.BWS_EINVAL
    assert L.bws_group_rows(C.byref(a), 64, 0, C.byref(g), None) == bws.BWS_EINVAL
    assert L.bws_group_rows(C.byref(a), 64, 33, C.byref(g), None) == bws.BWS_EINVAL
    assert L.bws_group_rows(C.byref(a), 64, 2, None, None) == bws.BWS_EINVAL
    bad = _bm(100, 64, 1)
    assert L.bws_group_rows(C.byref(bad), 64, 2, C.byref(g), None) == bws.BWS_EINVAL
    vals_no_rp = _bm(100, 64, 2, values=FAKE)
    assert L.bws_group_rows(C.byref(vals_no_rp), 64, 2, C.byref(g), None) == bws.BWS_EINVAL
    g0 = bws._Groups(*([None] * 9))
    assert L.bws_group_rows(C.byref(a), 64, 2, C.byref(g0), None) == bws.BWS_EINVAL
    assert L.bws_group_tiles(100, 64, 64) == 4
    assert L.bws_group_tiles(100, 65, 64) == 6
    assert L.bws_group_tiles(-1, 65, 64) == -1
    assert L.bws_group_scatter(FAKE, 10, 64, 2, FAKE, FAKE, 64, 0, FAKE, 64, None) == bws.BWS_EINVAL
    assert L.bws_group_spilled_csr(C.byref(a), 64, FAKE, FAKE, FAKE, FAKE, FAKE, FAKE, FAKE,
                                   FAKE, None, 0, None) == bws.BWS_EWORKSPACE


def test_spmm_validation(bws):
    L = bws.lib()
    a = _bm(64, 64, 2)
    args = lambda **kw: dict(dict(B=FAKE, dt=0, ldb=64, N=64, Cp=FAKE, ldc=64, prec=0), **kw)

    def call(A=a, **kw):
        x = args(**kw)
        return L.bws_spmm(C.byref(A), x["B"], x["dt"], x["ldb"], x["N"], x["Cp"], x["ldc"], 1.0,
                          0.0, x["prec"], None, None, 0, None)
    assert call(prec=7) == bws.BWS_EINVAL
    assert call(dt=5) == bws.BWS_EINVAL
    assert call(N=-1) == bws.BWS_EINVAL
    assert call(ldb=63) == bws.BWS_EINVAL
    assert call(ldc=63) == bws.BWS_EINVAL
    assert call(Cp=None) == bws.BWS_EINVAL
    assert call(Cp=FAKE + 2) == bws.BWS_EINVAL          # misaligned FP32 output
    assert call(prec=3, dt=1) == bws.BWS_EUNSUPPORTED   # BF16X2 needs FP32 B
    assert call(prec=1, dt=2) == bws.BWS_EUNSUPPORTED   # BF16 with FP16 B
    assert call(prec=2, dt=1) == bws.BWS_EUNSUPPORTED
    assert call(N=0) == bws.BWS_OK                      # empty product: no-op
    assert call(A=_bm(0, 64, 2)) == bws.BWS_OK
    plan = bws.make_plan(kernel=5)
    assert L.bws_spmm(C.byref(a), FAKE, 0, 64, 64, FAKE, 64, 1.0, 0.0, 0, C.byref(plan), None, 0,
                      None) == bws.BWS_EINVAL
    plan = bws.make_plan(kernel=2)   # MFMA is not an FP32 path
    assert L.bws_spmm(C.byref(a), FAKE, 0, 64, 64, FAKE, 64, 1.0, 0.0, 0, C.byref(plan), None, 0,
                      None) == bws.BWS_EUNSUPPORTED
    plan = bws.make_plan(rows_per_wg=100)
    assert L.bws_spmm(C.byref(a), FAKE, 0, 64, 64, FAKE, 64, 1.0, 0.0, 0, C.byref(plan), None, 0,
                      None) == bws.BWS_EINVAL
    plan = bws.make_plan(mfma_min_bits=-1)   # 0 = off; no negative thresholds
    assert L.bws_spmm(C.byref(a), FAKE, 1, 64, 64, FAKE, 64, 1.0, 0.0, 1, C.byref(plan), None, 0,
                      None) == bws.BWS_EINVAL


def test_sddmm_and_groups_validation(bws):
    L = bws.lib()
    a = _bm(64, 64, 2)  # no row_ptr: SDDMM needs it
    assert L.bws_sddmm(C.byref(a), FAKE, FAKE, 0, 64, 64, 64, FAKE, 0, None, 0, None) == bws.BWS_EINVAL
    a = _bm(64, 64, 2, row_ptr=FAKE)
    assert L.bws_sddmm(C.byref(a), FAKE, FAKE, 0, 63, 64, 64, FAKE, 0, None, 0, None) == bws.BWS_EINVAL
    assert L.bws_sddmm(C.byref(a), FAKE, FAKE, 1, 64, 64, 64, FAKE, 0, None, 0, None) == bws.BWS_EUNSUPPORTED
    assert L.bws_sddmm(C.byref(a), FAKE, FAKE, 2, 64, 64, 64, FAKE, 1, None, 0, None) == bws.BWS_EUNSUPPORTED
    assert L.bws_sddmm(C.byref(a), FAKE, FAKE, 1, 64, 64, 64, FAKE, 1, None, 0, None) == bws.BWS_EWORKSPACE
    assert L.bws_sddmm_workspace_size(C.byref(a)) >= 64 * 2 * 4
    assert L.bws_spmm_groups(C.byref(a), FAKE, FAKE, 0, 2, FAKE, 64, 64, FAKE, 64, None) == bws.BWS_EINVAL
    assert L.bws_spmm_batched(None, 3, 0, 0, None, None, 0, None) == bws.BWS_EINVAL
    assert L.bws_spmm_batched(None, 0, 0, 0, None, FAKE, L.bws_batch_workspace_size(0), None) == bws.BWS_OK
    b = bws._Batch()
    assert L.bws_batch_prepare(None, 0, 0, 0, None, FAKE, 4096, C.byref(b), None) == bws.BWS_OK
    assert b.n_launch == 0 and L.bws_batch_run(C.byref(b), None) == bws.BWS_OK
    assert L.bws_batch_prepare(None, 0, 0, 0, None, None, 0, C.byref(b), None) == bws.BWS_EWORKSPACE
    assert L.bws_batch_prepare(None, 0, 1, 3, None, FAKE, 4096, C.byref(b), None) == bws.BWS_EUNSUPPORTED
    arr = (bws._Problem * 1)()
    arr[0] = bws._Problem(_bm(64, 64, 1), FAKE, 64, 64, FAKE, 64, 0)   # ldw too small
    ws = L.bws_batch_workspace_size(1)
    assert L.bws_batch_prepare(arr, 1, 0, 0, None, FAKE, ws, C.byref(b), None) == bws.BWS_EINVAL


def test_workspace_sizes(bws):
    """Workspace queries: on a 0/1 A with N % 8 == 0 and a plan that allows
    MFMA, BF16X2 asks for B's two BF16 planes and BF16 / FP16 for one rounded
    copy of B (used when B is FP32); FP32 with N % 4 == 0 asks for A's mask
    stream (8 bytes per row and 64-row stage, rows padded to 256); nothing
    otherwise; the SDDMM workspace holds the per-word offsets and one byte per
    64 x 64 tile."""
    L = bws.lib()
    P = bws.PRECISION
    a = _bm(300, 200, 7)
    assert L.bws_spmm_workspace_size(C.byref(a), 64, P["bf16x2"], None) >= 2 * 200 * 64 * 2
    assert L.bws_spmm_workspace_size(C.byref(a), 60, P["bf16x2"], None) == 0
    assert L.bws_spmm_workspace_size(C.byref(a), 64, P["bf16"], None) >= 200 * 64 * 2
    assert L.bws_spmm_workspace_size(C.byref(a), 64, P["fp16"], None) >= 200 * 64 * 2
    assert L.bws_spmm_workspace_size(C.byref(a), 60, P["bf16"], None) == 0
    assert L.bws_spmm_workspace_size(C.byref(a), 64, P["fp32"], None) == 512 * 4 * 8
    assert L.bws_spmm_workspace_size(C.byref(a), 62, P["fp32"], None) == 0
    assert L.bws_spmm_workspace_size(C.byref(a), 64, P["fp32"],
                                     C.byref(bws.make_plan(kernel=1, rows_per_wg=128))) == 0
    plan = bws.make_plan(kernel=1)
    assert L.bws_spmm_workspace_size(C.byref(a), 64, P["bf16x2"], C.byref(plan)) == 0
    av = _bm(300, 200, 7, row_ptr=FAKE, values=FAKE)
    assert L.bws_spmm_workspace_size(C.byref(av), 64, P["bf16x2"], None) == 0
    assert L.bws_spmm_workspace_size(C.byref(av), 64, P["bf16"], None) >= 300 * 256 * 2
    assert L.bws_sddmm_workspace_size(C.byref(a)) >= 300 * 7 * 4 + 5 * 4


def test_transpose_perm_gather_validation(bws):
    L = bws.lib()
    a = _bm(40, 64, 2, row_ptr=FAKE)
    ws = L.bws_transpose_workspace_size(40, 64)
    assert L.bws_transpose_perm(C.byref(a), FAKE, 2, FAKE, None, FAKE, ws, None) == bws.BWS_EINVAL
    assert L.bws_transpose_perm(C.byref(_bm(40, 64, 2)), FAKE, 2, FAKE, FAKE, FAKE, ws,
                                None) == bws.BWS_EINVAL          # needs row_ptr
    assert L.bws_transpose_perm(C.byref(a), FAKE, 2, FAKE, FAKE, FAKE, ws - 1,
                                None) == bws.BWS_EWORKSPACE
    assert L.bws_gather_f32(FAKE, FAKE, -1, FAKE, None) == bws.BWS_EINVAL
    assert L.bws_gather_f32(None, FAKE, 5, FAKE, None) == bws.BWS_EINVAL
    assert L.bws_gather_f32(None, None, 0, None, None) == bws.BWS_OK


def test_cluster_validation(bws):
    L = bws.lib()
    a = _bm(64, 64, 2)
    ws = L.bws_cluster_workspace_size(64, 64, 0)
    assert ws > 5 * 64 * 4 and L.bws_cluster_workspace_size(64, 64, -1) == 0
    assert L.bws_cluster_rows(C.byref(_bm(64, 64, 2, row_ptr=FAKE, values=FAKE)), 0, FAKE, FAKE,
                              FAKE, ws, None) == bws.BWS_EUNSUPPORTED
    assert L.bws_cluster_rows(C.byref(a), -1, FAKE, FAKE, FAKE, ws, None) == bws.BWS_EINVAL
    assert L.bws_cluster_rows(C.byref(a), 0, None, FAKE, FAKE, ws, None) == bws.BWS_EINVAL
    assert L.bws_cluster_rows(C.byref(a), 0, FAKE, FAKE, FAKE, ws - 1, None) == bws.BWS_EWORKSPACE
    assert L.bws_cluster_rows(C.byref(a), 0, FAKE, FAKE, FAKE + 16, ws, None) == bws.BWS_EWORKSPACE
    assert L.bws_cluster_rows(C.byref(_bm(0, 64, 2)), 0, None, None, None, 0, None) == bws.BWS_OK
    assert L.bws_cluster_rows(C.byref(_bm(1 << 30, 64, 2)), 0, FAKE, FAKE, FAKE, ws,
                              None) == bws.BWS_ECAPACITY

    def call(A=a, sr=FAKE, B=FAKE, ldb=64, N=64, Cp=FAKE, ldc=64):
        return L.bws_spmm_clustered(C.byref(A), sr, FAKE, B, ldb, N, Cp, ldc, 1.0, 0.0, None)
    assert call(A=_bm(64, 64, 2, row_ptr=FAKE, values=FAKE)) == bws.BWS_EUNSUPPORTED
    assert call(ldc=63) == bws.BWS_EINVAL
    assert call(ldb=63) == bws.BWS_EINVAL
    assert call(sr=None) == bws.BWS_EINVAL
    assert call(Cp=None) == bws.BWS_EINVAL
    assert call(N=6, ldb=6, ldc=6) == bws.BWS_EUNSUPPORTED   # N % 4
    assert call(B=FAKE + 4) == bws.BWS_EUNSUPPORTED           # B not 16-byte aligned
    assert call(N=0) == bws.BWS_OK
    assert call(A=_bm(0, 64, 2)) == bws.BWS_OK


def test_autotune_cache_host(bws, tmp_path):
    L = bws.lib()
    assert L.bws_autotune_clear() == bws.BWS_OK
    p = bws._Plan()
    assert L.bws_autotune_lookup(4096, 4096, 4096, 838861, 0, C.byref(p)) == bws.BWS_ENOTFOUND
    f = tmp_path / "cache.txt"
    f.write_text("# comment\n0 12 12 12 -4 1 256 16 64\n1 10 10 10 -3 2 0 0 0\n")
    assert L.bws_autotune_load(str(f).encode()) == bws.BWS_OK
    assert L.bws_autotune_lookup(4096, 4096, 4096, 838861, 0, C.byref(p)) == bws.BWS_OK
    assert (p.kernel, p.rows_per_wg, p.waves_per_wg, p.k_stage) == (1, 256, 16, 64)
    assert L.bws_autotune_lookup(1000, 1000, 1000, 100000, 1, C.byref(p)) == bws.BWS_OK
    assert p.kernel == 2
    g = tmp_path / "out.txt"
    assert L.bws_autotune_save(str(g).encode()) == bws.BWS_OK
    assert len([l for l in g.read_text().splitlines() if not l.startswith("#")]) == 2
    assert L.bws_autotune_load(b"/nonexistent/x") == bws.BWS_EINVAL
    L.bws_autotune_clear()


def test_generators_deterministic(bws):
    a = bws.gen_bitmap_iid(0, 64, 100, 0.1)
    b = bws.gen_bitmap_iid(0, 64, 100, 0.1)
    c = bws.gen_bitmap_iid(1, 64, 100, 0.1)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.all(a[:, -1] & ((1 << 28) - 1) == 0)      # columns >= 100 stay zero
    x = bws.gen_f32(3, (10, 10))
    assert np.array_equal(x, bws.gen_f32(3, (10, 10)))


def popcount_np(W):
    return np.unpackbits(W.view(np.uint8)).reshape(W.shape + (32,)).sum(-1)


@pytest.mark.parametrize("p", [0.1, 0.5, 0.01])
def test_iid_density(bws, p):
    n = 512 * 512
    W = bws.gen_bitmap_iid(7, 512, 512, p)
    k = int(popcount_np(W).sum())
    sd = (n * p * (1 - p)) ** 0.5
    assert abs(k - n * p) < 6 * sd


def test_clustered_density_and_structure(bws):
    rows, cols, P, p, f = 2048, 1024, 64, 0.05, 0.02
    W = bws.gen_bitmap_clustered(0, rows, cols, P, p, f)
    d = popcount_np(W).sum() / (rows * cols)
    expect = p * (1 - f) + (1 - p) * f
    assert abs(d - expect) < 0.003
    # rows come from 64 prototypes: nearest other row is much closer than iid
    bits = np.unpackbits(W.astype(">u4").view(np.uint8), axis=1)[:, :cols].astype(np.int32)
    sample = bits[:64]
    dist = (sample[:, None, :] != bits[None, :, :]).sum(-1)
    dist[np.arange(64), np.arange(64)] = cols
    assert np.median(dist.min(1)) < 0.7 * np.median(dist)


def test_normal_and_loguniform(bws):
    x = bws.gen_f32(11, 200000, "normal", 0.0, 1.0)
    assert abs(x.mean()) < 0.02 and abs(x.std() - 1) < 0.02
    y = bws.gen_f32(11, 100000, "logu", -3.0, 3.0)
    e = np.log10(np.abs(y))
    assert e.min() >= -3.0001 and e.max() <= 3.0001 and abs(np.mean(y > 0) - 0.5) < 0.01
    u = bws.gen_f32(2, 100000, "uniform", -1.0, 1.0)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01


def test_split_workspace_size(bws):
    """The MFMA density split asks for the block flags on top of the rounded B."""
    s = bws._Bitmap(300, 128, 4, None, None, None)
    L = bws.lib()
    plain = L.bws_spmm_workspace_size(C.byref(s), 64, bws.PRECISION["bf16"], None)
    split = L.bws_spmm_workspace_size(C.byref(s), 64, bws.PRECISION["bf16"],
                                      C.byref(bws.make_plan(mfma_min_bits=8)))
    assert split >= plain + 256  # the block flags on top of the rounded B
