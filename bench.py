#!/usr/bin/env python3
"""Bit-wise SpMM benchmark on one MI355X.

Workload (default): C = A·B, M = N = K = 4096, A a clustered 0/1 bitmap
(64 prototypes ~ Bernoulli(0.03125) with 2% bit flips, rows shuffled: ~5%
dense, i.e. "95% sparse"), B ~ N(0,1) FP32, FP32 path.  A step is one SpMM
launch over the resident bitmap and B (the reference's timed region is
likewise the GEMM on pre-packed, pre-grouped operands).  For FP32 the step is
the row-clustered product (bws_spmm_clustered) when the row-clustering plan,
built once per bitmap and timed separately, sums at least half of the rows
as differences from a chunk base and the clustered kernel is the faster of
the two on these operands; otherwise the plain bws_spmm.  Both
kernels' times are reported.  Prints ONE JSON line with effective TFLOP/s =
2·nnz·N / t, the HBM roofline of the kernel, and the reference CPU naive
loop timed on a bounded row sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFS = 2500.0  # dense BF16 / FP16
# FP32 vector adds per second: 157.3 TFLOP/s of FP32 vector FMA counts packed
# (two-wide) FMAs as 2 flops each; this toolchain builds with packed FP32 off,
# so the rate it can reach is half of the packed one.
ADD_RATE_PACKED = 157.3e12 / 2
ADD_RATE_UNPACKED = ADD_RATE_PACKED / 2
# HBM bytes per launch of the benchmark kernel, from the FETCH_SIZE / WRITE_SIZE
# counter runs summarised in profiles/ (one file per kernel)
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "spmm_{}_4096_clustered_traffic.json")


def sources_sha256(paths):
    """SHA-256 over the named source files (paths relative to the repository
    root, in order); the counter files record it for the kernel they measured,
    so that a reading taken before the kernel changed is not reported."""
    import hashlib
    h = hashlib.sha256()
    for rel in paths:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(rel.encode() + b"\0" + f.read())
    return h.hexdigest() if paths else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def popcount_u32(W):
    import numpy as np
    return int(np.unpackbits(W.view(np.uint8)).sum())


def spin_up(fn, seconds=0.3):
    """Run fn untimed for about `seconds` so the GPU leaves its idle clocks
    before anything is measured (a few milliseconds of work are not enough)."""
    import torch
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(20):
            fn()
        torch.cuda.synchronize()


def time_loop(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / steps, e0.elapsed_time(e1) / steps


def _check_rate(W, B, K, N, target_s, threads, rows_min):
    """Times the check() loop on the first rows of W (sized by one calibration
    pass to about target_s seconds); returns (rows, nnz, seconds)."""
    import numpy as np

    import oracle
    M = W.shape[0]

    def run(rows):
        Ad = oracle.unpack_bitmap(W[:rows], K)
        t0 = time.perf_counter()
        oracle.check_dense_f32(Ad, B, 0, rows, threads=threads)
        return time.perf_counter() - t0

    rows = min(M, rows_min)
    dt = run(rows)
    if dt < 0.5 * target_s and rows < M:
        rows = int(min(M, max(rows + 1, rows * target_s / max(dt, 1e-6))))
        rows = min(M, max(rows_min, rows - rows % rows_min))
        dt = run(rows)
    return rows, popcount_u32(np.ascontiguousarray(W[:rows])), dt


def cpu_baseline(bws, W, B, K, N, target_s, threads):
    """The reference's CPU naive loop (backup/bit_wise_gemm.cu:13-32: i, j, p
    order, FP32 accumulate, dense over all K, stride-N reads of B) on the dense
    0/1 form of a bounded row sample of the same A; reported in the metric's
    unit.  Main figure: OpenMP over rows on `threads` host cores.  Also: the
    same loop single-threaded (as the reference runs it), and the 1024³
    configuration (iid Bernoulli(0.1) A, B ~ U[-1, 1], seed 0) in full, single
    and all cores."""
    M = W.shape[0]
    rows, nnz, dt = _check_rate(W, B, K, N, target_s, threads, threads)
    out = {"value": round(2.0 * nnz * N / dt / 1e12, 6), "unit": "TFLOP/s", "cores": threads,
           "kind": "port",
           "sample": f"first {rows} of {M} rows of the same A (dense 0/1 form, nnz {nnz}) x the "
                     f"full {K}x{N} B: check() loop i,j,p in FP32, OpenMP over rows, "
                     f"{dt:.2f} s"}
    # for information: the oracle's sparse-aware FP64 product (one thread,
    # bitmap-driven, double accumulation) on the same row sample
    import numpy as np
    import oracle
    Ws = np.ascontiguousarray(W[:rows])
    t0 = time.perf_counter()
    oracle.spmm_f64(Ws, K, B)
    dtf = time.perf_counter() - t0
    out["fp64_oracle_sparse"] = {"value": round(2.0 * nnz * N / dtf / 1e12, 6), "cores": 1,
                                 "sample": f"the same {rows} rows, {dtf:.2f} s",
                                 "note": "information only: the checker's FP64 product"}
    rows1, nnz1, dt1 = _check_rate(W, B, K, N, target_s / 2, 1, 1)
    out["single_thread"] = {"value": round(2.0 * nnz1 * N / dt1 / 1e12, 6), "cores": 1,
                            "sample": f"first {rows1} rows, nnz {nnz1}, {dt1:.2f} s"}
    W1 = bws.gen_bitmap_iid(0, 1024, 1024, 0.1)
    B1 = bws.gen_f32(1, (1024, 1024), "uniform", -1.0, 1.0)
    c1 = {}
    for t in (1, threads):
        r, z, d = _check_rate(W1, B1, 1024, 1024, 1e9, t, 1024)   # all 1024 rows
        c1[f"threads_{t}"] = {"value": round(2.0 * z * 1024 / d / 1e12, 6), "seconds": round(d, 3),
                              "nnz": z}
    out["c1_1024_full"] = c1
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16", "bf16x2"])
    ap.add_argument("--m", type=int, default=4096)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--k", type=int, default=4096)
    ap.add_argument("--p", type=float, default=0.03125, help="prototype density")
    ap.add_argument("--flip", type=float, default=0.02,
                    help="bit-flip probability of the clustered mask (--p 0 --flip 0: all zero)")
    ap.add_argument("--iid", action="store_true", help="iid Bernoulli(p) mask instead of clustered")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--quick", action="store_true", help="extras: skip the C4/C5/comparator legs")
    ap.add_argument("--mfma-staging", default="auto", choices=["auto", "sparse"],
                    help="16-bit precisions: dense MFMA (auto) or the sparse MFMA")
    ap.add_argument("--fp32-path", default="auto", choices=["auto", "clustered", "plain"],
                    help="FP32 kernel: the row-clustered product, the plain one, or by the plan")
    ap.add_argument("--leg", default=None,
                    choices=["density_sweep", "value_mask", "int8_bitplanes", "mfma_split",
                             "grouped", "comparators"],
                    help="extras: run only this leg (for profiling one workload)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import bitwise_spgemm_amd as bws

    dev = torch.device("cuda:0")
    if args.copy_probe:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE readings give
        # the scale between the HBM counters and bytes moved
        n = args.copy_probe << 18
        src = torch.ones(n, device=dev)
        dst = torch.empty(n, device=dev)
        ms = time_loop(lambda: dst.copy_(src), 5, 1)[1]
        print(json.dumps({"copy_bytes_each_way": 4 * n, "ms": round(ms, 5),
                          "GBps_read_plus_write": round(8 * n / ms / 1e6, 1)}), flush=True)
        return
    M, N, K = args.m, args.n, args.k
    if args.iid:
        W = bws.gen_bitmap_iid(args.seed, M, K, args.p)
    else:
        W = bws.gen_bitmap_clustered(args.seed, M, K, 64, args.p, args.flip)
    B = bws.gen_f32(args.seed + 1, (K, N), "normal", 0.0, 1.0)
    nnz = popcount_u32(W)
    A = bws.from_host_words(W, M, K, dev)
    Bd = torch.from_numpy(B).to(dev)
    Bin = {"fp32": lambda: Bd, "bf16x2": lambda: Bd, "bf16": lambda: Bd.to(torch.bfloat16),
           "fp16": lambda: Bd.to(torch.float16)}[args.precision]()
    C = torch.empty((M, N), dtype=torch.float32, device=dev)
    log(f"workload M=N=K={M}x{N}x{K} nnz={nnz} density={nnz / (M * K):.4f} prec={args.precision}")

    mplan = (bws.make_plan(kernel=2, mfma_staging=args.mfma_staging)
             if args.precision in ("bf16", "fp16") and args.mfma_staging != "auto" else None)
    plain = lambda: bws.spmm(A, Bin, args.precision, out=C, plan=mplan)
    step, path, plan_info, clusters = plain, args.precision, None, None
    if args.precision == "fp32" and args.fp32_path != "plain":
        # the row-clustering plan: an analysis step per bitmap, timed apart
        clusters = bws.cluster_rows(A)
        plan_ms = time_loop(lambda: bws.cluster_rows(A), 3, 1)[1]
        delta = float(((clusters.slot_row & bws.SLOT_DELTA) != 0).float().mean().item())
        plan_info = {"plan_ms": round(plan_ms, 4), "delta_row_fraction": round(delta, 4)}
        clustered = lambda: bws.spmm_clustered(A, clusters, Bin, out=C)
        use = args.fp32_path == "clustered"
        if args.fp32_path == "auto" and delta >= 0.5:
            # tuned like any plan choice: the faster of the two kernels here
            spin_up(plain, 0.1)
            t_plain = time_loop(plain, 10, 2)[1]
            spin_up(clustered, 0.1)
            use = time_loop(clustered, 10, 2)[1] < t_plain
        if use:
            step = clustered
            path = "fp32_rowclusters"
    spin_up(step)
    ms_step, ms_kernel = time_loop(step, args.steps, args.warmup)
    flops = 2.0 * nnz * N
    value = flops / (ms_step * 1e-3) / 1e12
    if plan_info is not None:
        # the other FP32 kernel on the same operands, for comparison
        other = plain if path != "fp32" else (lambda: bws.spmm_clustered(A, clusters, Bin, out=C))
        spin_up(other, 0.1)
        o_ms = time_loop(other, max(args.steps // 2, 5), 2)[1]
        plan_info["plain_ms" if path != "fp32" else "clustered_ms"] = round(o_ms, 5)
        if path != "fp32" and o_ms > ms_kernel:
            plan_info["plan_break_even_calls"] = round(plan_info["plan_ms"] / (o_ms - ms_kernel), 1)
    s_b = {"fp32": 4, "bf16x2": 4, "bf16": 2, "fp16": 2}[args.precision]
    alg_bytes = 4 * M * ((K + 31) // 32) + s_b * K * N + 4 * M * N
    if path == "fp32_rowclusters":  # + slot rows and chunk bases
        alg_bytes += 4 * M + 4 * 2 * ((M + 15) // 16) * ((K + 31) // 32)
    achieved = alg_bytes / (ms_kernel * 1e-3) / 1e9
    kname = {"fp32_rowclusters": "k_spmm_stream<true>",
             "fp32": "k_spmm_stream<false>"}.get(path)
    traffic = None
    traffic_note = None
    tfile = PROFILE_TRAFFIC.format(path)
    if (os.path.exists(tfile) and not args.iid and M == N == K == 4096 and args.p == 0.03125
            and args.flip == 0.02):
        try:
            tj = json.load(open(tfile))
            # only a counter reading of the kernel that runs now, taken from
            # the kernel sources as they are now
            if kname is not None and tj.get("kernel") != kname:
                traffic_note = f"counters are for {tj.get('kernel')}"
            elif tj.get("sources_sha256") != sources_sha256(tj.get("sources", [])):
                traffic_note = "stale: the kernel sources changed since the counter run"
            else:
                traffic = tj.get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    t_hbm_us = alg_bytes / (HBM_PEAK_GBS * 1e9) * 1e6
    if args.precision == "fp32":
        # binding ceiling of the vector-add path: max(HBM time, nnz*N adds)
        t_add_us = nnz * N / ADD_RATE_PACKED * 1e6
        t_add_unpacked_us = nnz * N / ADD_RATE_UNPACKED * 1e6
        roofline = {"kernel": kname,
                    "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "traffic_note": traffic_note,
                    "kernel_ms": round(ms_kernel, 5), "algorithmic_bytes": alg_bytes,
                    "ceiling": {"hbm_us": round(t_hbm_us, 2), "fp32_add_us": round(t_add_us, 2),
                                "fp32_add_unpacked_us": round(t_add_unpacked_us, 2),
                                "binding": "fp32_add" if t_add_us > t_hbm_us else "hbm",
                                "frac_of_binding": round(max(t_add_us, t_hbm_us) / (ms_kernel * 1e3), 4),
                                "frac_of_unpacked": round(max(t_add_unpacked_us, t_hbm_us) / (ms_kernel * 1e3), 4)}}
    else:
        # the MFMA kernels run every 32-wide K step of a stage that has a set
        # bit: at these densities that is the dense product (2*M*N*K per pass,
        # two passes for the split BF16X2)
        passes = 2 if args.precision == "bf16x2" else 1
        mfma_flops = 2.0 * M * N * K * passes
        tf = mfma_flops / (ms_kernel * 1e-3) / 1e12
        t_mfma_us = mfma_flops / (MFMA_PEAK_TFS * 1e12) * 1e6
        roofline = {"bound": "mfma", "achieved": round(tf, 1), "peak": MFMA_PEAK_TFS,
                    "mfma_path": "sparse (v_smfmac, dense-equivalent flops)"
                    if mplan is not None else "dense",
                    "unit": "TFLOP/s", "frac": round(tf / MFMA_PEAK_TFS, 4), "traffic": traffic,
                    "traffic_note": traffic_note,
                    "kernel_ms": round(ms_kernel, 5), "mfma_flops": mfma_flops,
                    "hbm_achieved_GBps": round(achieved, 1), "algorithmic_bytes": alg_bytes,
                    "ceiling": {"hbm_us": round(t_hbm_us, 2), "mfma_us": round(t_mfma_us, 2),
                                "binding": "mfma" if t_mfma_us > t_hbm_us else "hbm",
                                "frac_of_binding": round(max(t_mfma_us, t_hbm_us) / (ms_kernel * 1e3), 4)}}

    extras = {}
    if not args.no_extras:
        extras = run_extras(bws, A, W, Bd, M, N, K, nnz, dev, args, clusters)

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(bws, W, B, K, N, args.cpu_seconds, threads)

    out = {
        "metric": "bitmap SpMM TFLOP/s (2\u00b7nnz\u00b7N/t) + %roofline; BF16 vs FP32 fwd+bwd at stated err",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16": "bf16", "fp16": "f16", "bf16x2": "bf16x2"}[args.precision],
        "data": "synthetic",
        "config": {"workload": ("iid" if args.iid else "clustered") +
                               f" 0/1 bitmap SpMM {M}x{K} @ {K}x{N}, nnz={nnz} "
                               f"(density {nnz / (M * K):.4f}), B N(0,1)",
                   "M": M, "N": N, "K": K, "nnz": nnz, "precision": args.precision,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if plan_info is not None:
        out["config"]["fp32_path"] = "row clusters" if path == "fp32_rowclusters" else "plain"
        out["fp32_paths"] = plan_info
    if extras:
        out["extras"] = extras
    print(json.dumps(out), flush=True)


def dense_rows(W, K, rows):
    """Dense 0/1 float64 form (host numpy) of the given rows of an MSB-first
    bitmap, built on the host from the words (independent of the library)."""
    import numpy as np
    Wr = np.ascontiguousarray(W[rows]).astype(">u4").view(np.uint8)
    bits = np.unpackbits(Wr, axis=1)[:, :K]
    return bits.astype(np.float64)


def sample_rows(n):
    """The rows every error figure samples: every 61st from row 5."""
    import numpy as np
    return np.arange(5, n, 61)


class Fp64Ref:
    """FP64 reference of C = A·B on a fixed sample of rows, computed with
    torch float64 on the device from the host bitmap (or a dense FP64 form of
    the sampled rows): r = A_s·B and s = |A_s|·|B| (the error metric's scale).
    err(C) gives max / mean of |c - r| / s over the sampled entries."""

    def __init__(self, W, K, rows, dev, dense=None):
        import torch
        self.idx = torch.from_numpy(rows).to(dev)
        Ad = dense_rows(W, K, rows) if dense is None else dense
        self.A = torch.from_numpy(Ad).to(dev)

    def set_b(self, Bx):
        B64 = Bx.double()
        self.R = self.A @ B64
        self.S = (self.A.abs() @ B64.abs()).clamp_min(1e-300)
        return self

    def err(self, C):
        e = (C[self.idx].double() - self.R).abs() / self.S
        return {"max_rel_s": float(e.max()), "mean_rel_s": float(e.mean())}


def run_extras(bws, A, W, Bd, M, N, K, nnz, dev, args, clusters=None):
    """Pipeline stages, the precision variants with their error, forward +
    backward FP32 vs BF16, the value-carrying mask workload, the grouped batch
    in one captured hipGraph, and library comparators (dense GEMM, CSR SpMM).

    Every error is |c - r| / s against an FP64 product on sampled rows
    (Fp64Ref: torch float64 on the host-built dense rows, s = |A|·|B|)."""
    import numpy as np
    import torch
    ex = {}
    leg_fns = {"density_sweep": lambda: _density_sweep(bws, dev, M, N, K, args),
            "value_mask": lambda: _value_mask_workload(bws, dev),
            "int8_bitplanes": lambda: _bitplanes_workload(bws, dev, Bd, M, N, K),
            "mfma_split": lambda: _split_workload(bws, dev, Bd, M, N, K),
            "grouped": lambda: _grouped_workload(bws, dev),
            "comparators": lambda: _comparators(bws, A, W, Bd, M, N, K, nnz, dev)}
    if args.leg:
        return {args.leg: leg_fns[args.leg]()}
    def tl(f, k=20, w=3):
        spin_up(f, 0.1)   # steady clocks before every comparison
        return round(time_loop(f, k, w)[1], 4)
    # the device and its sustained HBM rate, beside the published 8 TB/s peak:
    # a 1 GiB device copy (2 GiB moved) and a 1 GiB fill
    props = torch.cuda.get_device_properties(dev)
    big = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    big2 = torch.empty_like(big)
    copy_ms = tl(lambda: big2.copy_(big), 10, 2)
    fill_ms = tl(lambda: big.fill_(1.0), 10, 2)
    ex["platform"] = {"name": props.name, "arch": getattr(props, "gcnArchName", ""),
                      "cus": props.multi_processor_count,
                      "hbm_copy_TBps": round(2 * (1 << 30) / (copy_ms * 1e-3) / 1e12, 2),
                      "hbm_fill_TBps": round((1 << 30) / (fill_ms * 1e-3) / 1e12, 2)}
    del big, big2
    Ad = bws.unpack_bitmap(A)
    ex["pack_ms"] = tl(lambda: bws.pack_bitmap(Ad), 5, 2)
    del Ad
    ex["group_ms"] = tl(lambda: bws.group_rows(A, 2048, 2), 5, 2)
    ex["transpose_ms"] = tl(lambda: bws.transpose(A), 5, 2)

    # precision variants on the same A and B, errors against FP64 on sampled
    # rows, for B ~ N(0,1) and B = ±10^U[-3,3]
    ref = Fp64Ref(W, K, sample_rows(M), dev)
    C = torch.empty((M, N), device=dev)
    Blog = torch.from_numpy(bws.gen_f32(args.seed + 1, (K, N), "logu", -3.0, 3.0)).to(dev)
    prec = {}
    for bname, Bsrc in (("normal", Bd), ("logu", Blog)):
        ref.set_b(Bsrc)
        legs = {}
        for name, Bx in (("fp32", Bsrc), ("bf16", Bsrc.to(torch.bfloat16)),
                         ("fp16", Bsrc.to(torch.float16)), ("bf16x2", Bsrc)):
            ms = tl(lambda: bws.spmm(A, Bx, name, out=C))
            legs[name] = {"fwd_ms": ms, "tflops": round(2.0 * nnz * N / ms / 1e9, 2),
                          "err_vs_fp64": ref.err(C)}
        # BF16 on the sparse MFMA (4:2-compressed A): half the MFMA instructions
        Bh = Bsrc.to(torch.bfloat16)
        sp = bws.make_plan(kernel=2, mfma_staging="sparse")
        ms = tl(lambda: bws.spmm(A, Bh, "bf16", out=C, plan=sp))
        legs["bf16_sparse_mfma"] = {"fwd_ms": ms, "tflops": round(2.0 * nnz * N / ms / 1e9, 2),
                                    "err_vs_fp64": ref.err(C)}
        del Bh
        if clusters is not None:
            ms = tl(lambda: bws.spmm_clustered(A, clusters, Bsrc, out=C))
            legs["fp32_rowclusters"] = {"fwd_ms": ms,
                                        "tflops": round(2.0 * nnz * N / ms / 1e9, 2),
                                        "err_vs_fp64": ref.err(C)}
        prec["B_" + bname] = legs
    ex["precision"] = prec
    del Blog

    # forward + backward on the same 0/1 mask (trainable operand B): dB = Aᵀ·dC,
    # each output's error against FP64 on sampled rows (dB: sampled columns of A)
    AT = bws.transpose(A)
    WT = bws.to_host_words(AT)
    dC = torch.randn((M, N), device=dev)
    dB = torch.empty((K, N), device=dev)
    ref.set_b(Bd)
    refT = Fp64Ref(WT, M, sample_rows(K), dev).set_b(dC)

    def fb(p, Bx, dCx):
        bws.spmm(A, Bx, p, out=C)
        bws.spmm_tA(AT, dCx, p, out=dB)
    try:
        Bb, dCb = Bd.to(torch.bfloat16), dC.to(torch.bfloat16)
        f32 = tl(lambda: fb("fp32", Bd, dC), 10, 2)
        e32 = {"C": ref.err(C), "dB": refT.err(dB)}
        b16 = tl(lambda: fb("bf16", Bb, dCb), 10, 2)
        e16 = {"C": ref.err(C), "dB": refT.err(dB)}
        ex["fwd_bwd_01mask"] = {"fp32_ms": f32, "bf16_ms": b16, "speedup": round(f32 / b16, 3),
                                "fp32_err_vs_fp64": e32, "bf16_err_vs_fp64": e16}
        if clusters is not None:
            # FP32 with the row-clustered forward (the backward's Aᵀ is not
            # clustered by this plan)
            f32c = tl(lambda: (bws.spmm_clustered(A, clusters, Bd, out=C),
                               bws.spmm_tA(AT, dC, "fp32", out=dB)), 10, 2)
            ex["fwd_bwd_01mask"].update({"fp32_clustered_fwd_ms": f32c,
                                         "fp32_clustered_err_vs_fp64": {"C": ref.err(C),
                                                                        "dB": refT.err(dB)},
                                         "speedup_vs_clustered_fwd": round(f32c / b16, 3)})
    except Exception as e:  # reported, never hidden
        ex["fwd_bwd_01mask"] = {"error": str(e)}
    del AT, dC, dB, ref, refT

    if not args.quick:
        for name, leg in leg_fns.items():
            ex[name] = leg()
    return ex


def _density_sweep(bws, dev, M, N, K, args):
    """The clustered workload at nominal prototype densities 0.5 / 0.1 / 0.05 /
    0.01 (measured density p·0.98 + (1-p)·0.02): FP32 plain and row-clustered
    (with its plan's delta-row fraction), FP32 subset sums, and BF16 forward
    times on the dense and the sparse MFMA; then an iid 50% mask (no row
    structure) with FP32 plain, FP32 subset sums and BF16."""
    import torch
    out = []
    B = torch.from_numpy(bws.gen_f32(args.seed + 1, (K, N), "normal", 0.0, 1.0)).to(dev)
    Bb = B.to(torch.bfloat16)
    C = torch.empty((M, N), device=dev)
    def tl(f):
        spin_up(f, 0.05)
        return time_loop(f, 10, 2)[1]
    for p in (0.5, 0.1, 0.05, 0.01):
        W = bws.gen_bitmap_clustered(args.seed, M, K, 64, p, 0.02)
        nnz = popcount_u32(W)
        A = bws.from_host_words(W, M, K, dev)
        pl = bws.cluster_rows(A)
        delta = float(((pl.slot_row & bws.SLOT_DELTA) != 0).float().mean().item())
        f32 = tl(lambda: bws.spmm(A, B, "fp32", out=C))
        f32c = tl(lambda: bws.spmm_clustered(A, pl, B, out=C))
        b16 = tl(lambda: bws.spmm(A, Bb, "bf16", out=C))
        sp = bws.make_plan(kernel=2, mfma_staging="sparse")
        b16s = tl(lambda: bws.spmm(A, Bb, "bf16", out=C, plan=sp))
        ss = bws.make_plan(kernel=1, k_stage=16)
        f32s = tl(lambda: bws.spmm(A, B, "fp32", out=C, plan=ss))
        fl = 2.0 * nnz * N / 1e9
        out.append({"p_proto": p, "density": round(nnz / (M * K), 4), "nnz": nnz,
                    "fp32_ms": round(f32, 4), "fp32_tflops": round(fl / f32, 2),
                    "fp32_rowclusters_ms": round(f32c, 4), "delta_rows": round(delta, 3),
                    "bf16_ms": round(b16, 4), "bf16_tflops": round(fl / b16, 2),
                    "bf16_sparse_mfma_ms": round(b16s, 4), "fp32_subset_sums_ms": round(f32s, 4)})
    # iid 50%: no row structure for the clustering; the subset-sum kernel's case
    W = bws.gen_bitmap_iid(args.seed, M, K, 0.5)
    nnz = popcount_u32(W)
    A = bws.from_host_words(W, M, K, dev)
    ss = bws.make_plan(kernel=1, k_stage=16)
    f32 = tl(lambda: bws.spmm(A, B, "fp32", out=C))
    f32s = tl(lambda: bws.spmm(A, B, "fp32", out=C, plan=ss))
    b16 = tl(lambda: bws.spmm(A, Bb, "bf16", out=C))
    fl = 2.0 * nnz * N / 1e9
    out.append({"iid": 0.5, "density": round(nnz / (M * K), 4), "nnz": nnz,
                "fp32_ms": round(f32, 4), "fp32_tflops": round(fl / f32, 2),
                "fp32_subset_sums_ms": round(f32s, 4), "fp32_subset_sums_tflops": round(fl / f32s, 2),
                "bf16_ms": round(b16, 4)})
    return out


def _bitplanes_workload(bws, dev, Bd, M, N, K):
    """int8 A (M x K, 10% nonzero, values uniform over [-128, 127] without 0)
    held as 8 packed bit-planes, times B: B read once (the planes folded into
    one value-carrying or dense 16-bit A) against eight weighted 0/1 passes,
    FP32 (FP32 B) and BF16 (BF16 B); errors against FP64 on sampled rows."""
    import numpy as np
    import torch
    rng = np.random.default_rng(7)
    Ai = rng.integers(1, 128, size=(M, K), dtype=np.int16) * rng.choice([-1, 1], size=(M, K))
    Ai[rng.random((M, K)) >= 0.1] = 0
    Ai = np.clip(Ai, -128, 127).astype(np.int8)
    P = bws.pack_bitplanes_i8(torch.from_numpy(Ai).to(dev))
    rows = sample_rows(M)
    Bb = Bd.to(torch.bfloat16)
    C = torch.empty((M, N), device=dev)
    out = {"workload": f"int8 A {M}x{K} (10% nonzero) as 8 bit-planes @ {K}x{N}"}
    def tl(f):
        spin_up(f, 0.05)
        return round(time_loop(f, 10, 2)[1], 4)
    for prec, Bx in (("fp32", Bd), ("bf16", Bb)):
        ref = Fp64Ref(None, K, rows, dev, dense=Ai[rows].astype(np.float64)).set_b(Bx)
        one = tl(lambda: bws.spmm_bitplanes_i8(P, K, Bx, prec, out=C))
        e1 = ref.err(C)
        eight = tl(lambda: bws.spmm_bitplanes_i8(P, K, Bx, prec, out=C, one_pass=False))
        e8 = ref.err(C)
        out[prec] = {"one_pass_ms": one, "eight_passes_ms": eight,
                     "speedup": round(eight / one, 2),
                     "one_pass_err_vs_fp64": e1, "eight_passes_err_vs_fp64": e8}
    return out


def _split_workload(bws, dev, Bd, M, N, K):
    """BF16 on a 0/1 mask whose 128-row blocks alternate between 10% and
    0.05% density (every 128 x 64 block of the sparse ones still has bits, so
    the MFMA kernel skips none of its stages): every block on MFMA against the
    density split (plan mfma_min_bits 40: the sparse blocks on the vector
    kernel; mfma_min_bits 1 sends every block with a bit to the flagged-block
    MFMA kernel); errors against FP64 (of the BF16 B) on sampled rows."""
    import numpy as np
    import torch
    rng = np.random.default_rng(11)
    W = np.zeros((M, (K + 31) // 32), np.uint32)
    for b0 in range(0, M, 128):
        p = 0.1 if (b0 // 128) % 2 == 0 else 0.0005
        W[b0:b0 + 128] = bws.gen_bitmap_iid(int(rng.integers(1 << 30)), min(128, M - b0), K, p)
    A = bws.from_host_words(W, M, K, dev)
    nnz = int(np.unpackbits(W.view(np.uint8)).sum())
    Bb = Bd.to(torch.bfloat16)
    rows = sample_rows(M)
    ref = Fp64Ref(W, K, rows, dev).set_b(Bb)
    C = torch.empty((M, N), device=dev)
    out = {"workload": f"0/1 mask {M}x{K}, 128-row blocks alternating 10% / 0.05% "
                       f"(nnz {nnz}) @ {K}x{N}, BF16 B"}
    for name, plan in (("all_mfma", bws.make_plan(kernel=2)),
                       ("split_1", bws.make_plan(mfma_min_bits=1)),
                       ("split_40", bws.make_plan(mfma_min_bits=40))):
        f = lambda: bws.spmm(A, Bb, "bf16", out=C, plan=plan)
        spin_up(f, 0.05)
        ms = time_loop(f, 10, 2)[1]
        out[name + "_ms"] = round(ms, 4)
        out[name + "_tflops"] = round(2 * nnz * N / (ms * 1e-3) / 1e12, 2)
        out[name + "_err_vs_fp64"] = ref.err(C)
    out["speedup"] = round(out["all_mfma_ms"] / out["split_40_ms"], 3)
    return out


def _value_mask_workload(bws, dev):
    """Value-carrying mask (M = K = 4096, iid 10%, values N(0, 0.02²)) times
    B 4096x2048: forward, dB = Aᵀ·dC and dvalues = (dC·Bᵀ)⊙mask, FP32 and
    BF16.  The mask is fixed and its values are what training changes, so
    Aᵀ's pattern and the value permutation are built once and each step
    gathers Aᵀ's values through it."""
    import torch
    M = K = 4096
    N = 2048
    W = bws.gen_bitmap_iid(3, M, K, 0.1)
    A0 = bws.from_host_words(W, M, K, dev)
    rp = bws.row_ptr(A0)
    nnz = int(rp[-1].item())
    vals = torch.from_numpy(bws.gen_f32(4, (nnz,), "normal", 0.0, 0.02)).to(dev)
    A = A0.with_values(vals, rp)
    B = torch.from_numpy(bws.gen_f32(5, (K, N), "normal", 0.0, 1.0)).to(dev)
    dC = torch.randn((M, N), device=dev)
    C = torch.empty((M, N), device=dev)
    dB = torch.empty((K, N), device=dev)
    dv = torch.empty((nnz,), device=dev)
    AT = bws.transpose(A)
    ATs, perm = bws.transpose_perm(A)  # once per mask
    vT = torch.empty(nnz, device=dev)
    ATv = bws.Bitmap(ATs.words, ATs.rows, ATs.cols, ATs.row_ptr, vT)
    out = {"workload": f"iid 0.1 value mask {M}x{K} (nnz {nnz}) @ {K}x{N}"}

    def step(p, Bx, dCx):
        bws.spmm(A, Bx, p, out=C)
        bws.gather(vals, perm, out=vT)  # Aᵀ's values for this step's values
        bws.spmm_tA(ATv, dCx, p, out=dB)
        bws.sddmm(A, dCx, Bx, p, out=dv)
    def tl(f, k=10, w=2):
        spin_up(f, 0.1)
        return round(time_loop(f, k, w)[1], 4)

    # FP64 references on sampled rows: C rows, dB rows (columns of A), and the
    # dvalues of the sampled rows' set bits
    import numpy as np
    rows = sample_rows(M)
    rph = rp.cpu().numpy()
    vh = vals.double().cpu().numpy()
    Dr = np.zeros((len(rows), K))
    for i, r in enumerate(rows):
        Dr[i][dense_rows(W, K, [r])[0] != 0] = vh[rph[r]:rph[r + 1]]
    refC = Fp64Ref(None, K, rows, dev, dense=Dr).set_b(B)
    WT = bws.to_host_words(AT)
    rpT = AT.row_ptr.cpu().numpy()
    vT_h = AT.values.double().cpu().numpy()
    rowsT = sample_rows(K)
    DT = np.zeros((len(rowsT), M))
    for i, r in enumerate(rowsT):
        DT[i][dense_rows(WT, M, [r])[0] != 0] = vT_h[rpT[r]:rpT[r + 1]]
    refT = Fp64Ref(None, M, rowsT, dev, dense=DT).set_b(dC)
    idx = torch.from_numpy(rows).to(dev)
    G = dC[idx].double() @ B.double().T                       # (dC·Bᵀ) on the sampled rows
    Gs = dC[idx].double().abs() @ B.double().abs().T
    sel = torch.from_numpy(dense_rows(W, K, rows) != 0).to(dev)
    pos = torch.cat([torch.arange(int(rph[r]), int(rph[r + 1])) for r in rows]).to(dev)
    rA, sA = G[sel], Gs[sel].clamp_min(1e-300)                # row-major = value order

    def errs():
        e = (dv[pos].double() - rA).abs() / sA
        return {"C": refC.err(C), "dB": refT.err(dB),
                "dA": {"max_rel_s": float(e.max()), "mean_rel_s": float(e.mean())}}
    try:
        out["transpose_ms"] = tl(lambda: bws.transpose(A), 5, 1)
        out["values_gather_ms"] = tl(lambda: bws.gather(vals, perm, out=vT))
        out["fwd_ms"] = tl(lambda: bws.spmm(A, B, "fp32", out=C))
        out["dB_ms"] = tl(lambda: bws.spmm_tA(AT, dC, "fp32", out=dB))
        out["dA_fp32_ms"] = tl(lambda: bws.sddmm(A, dC, B, "fp32", out=dv), 5, 1)
        Bb, dCb = B.to(torch.bfloat16), dC.to(torch.bfloat16)
        out["fwd_bf16_ms"] = tl(lambda: bws.spmm(A, Bb, "bf16", out=C))
        out["dB_bf16_ms"] = tl(lambda: bws.spmm_tA(AT, dCb, "bf16", out=dB))
        out["dA_bf16_ms"] = tl(lambda: bws.sddmm(A, dCb, Bb, "bf16", out=dv), 5, 1)
        f32 = tl(lambda: step("fp32", B, dC), 5, 1)
        e32 = errs()
        b16 = tl(lambda: step("bf16", Bb, dCb), 5, 1)
        e16 = errs()
        out.update({"fwd_bwd_fp32_ms": f32, "fwd_bwd_bf16_ms": b16,
                    "speedup": round(f32 / b16, 3), "fp32_err_vs_fp64": e32,
                    "bf16_err_vs_fp64": e16})
    except Exception as e:
        out["error"] = str(e)
    return out


def grouped_problems(bws, dev, count=32, N=1024, seed=0):
    """The grouped workload: `count` problems, M, K uniform in [256, 8192]
    (multiples of 64), density uniform in [1%, 20%] (80-99% sparse), even
    problems iid and odd ones clustered masks, BF16 B (K x N) and dC (M x N).
    Returns one dict per problem with the host bitmap W, the device bitmap A,
    its transpose AT and the device B, dC."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    out = []
    for i in range(count):
        M, K = (int(x) for x in rng.integers(4, 129, 2) * 64)
        d = float(rng.uniform(0.01, 0.20))
        W = (bws.gen_bitmap_iid(100 + i, M, K, d) if i % 2 == 0 else
             bws.gen_bitmap_clustered(100 + i, M, K, 64, d, 0.02))
        A = bws.from_host_words(W, M, K, dev)
        B = torch.randn((K, N), device=dev, generator=gen).to(torch.bfloat16)
        dC = torch.randn((M, N), device=dev, generator=gen).to(torch.bfloat16)
        out.append({"W": W, "M": M, "K": K, "A": A, "AT": bws.transpose(A), "B": B, "dC": dC,
                    "nnz": popcount_u32(W)})
    return out


def grouped_batches(bws, probs, N=1024, tune=True):
    """Forward (C = A·B) and backward (dB = Aᵀ·dC) grouped launches over the
    problems, with tiles autotuned per shape class; returns (fwd, bwd, outputs)."""
    import torch
    fwd, bwd, hints = [], [], []
    tuned = set()
    for p in probs:
        dev = p["B"].device
        fwd.append((p["A"], p["B"], torch.empty((p["M"], N), device=dev)))
        bwd.append((p["AT"], p["dC"], torch.empty((p["K"], N), device=dev)))
        hints.append(p["nnz"])
        if tune:
            for (X, Y) in ((p["A"], p["B"]), (p["AT"], p["dC"])):
                key = (X.rows, X.cols)
                if key not in tuned:
                    bws.autotune(X, Y, "bf16", nnz=p["nnz"], repeats=2)
                    tuned.add(key)
    fb = bws.Batch(fwd, "bf16", nnz_hints=hints)
    bb = bws.Batch(bwd, "bf16", nnz_hints=hints)
    return fb, bb, ([c for (_, _, c) in fwd], [c for (_, _, c) in bwd])


def _grouped_workload(bws, dev):
    """32 problems (grouped_problems), forward + backward as grouped launches
    with autotuned tiles, all captured in one hipGraph; a step is one replay."""
    import torch
    N = 1024
    probs = grouped_problems(bws, dev, 32, N)
    flops = sum(2 * 2.0 * p["nnz"] * N for p in probs)
    fb, bb, _ = grouped_batches(bws, probs, N)
    spin_up(lambda: (fb.run(), bb.run()), 0.1)
    eager = time_loop(lambda: (fb.run(), bb.run()), 5, 1)[1]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            fb.run()
            bb.run()
    torch.cuda.current_stream().wait_stream(s)
    spin_up(g.replay, 0.1)
    ms = time_loop(g.replay, 10, 2)[1]
    return {"problems": 32, "N": N, "fwd_launches": len(fb.launches),
            "bwd_launches": len(bb.launches), "graph_ms": round(ms, 4),
            "eager_ms": round(eager, 4), "tflops": round(flops / ms / 1e9, 2)}


def _comparators(bws, A, W, Bd, M, N, K, nnz, dev):
    """Context numbers on the same operands: dense GEMM (hipBLASLt / rocBLAS
    through torch.matmul) on the dense 0/1 form of A, and CSR SpMM (rocSPARSE
    through torch.sparse) — effective TFLOP/s counted as 2·nnz·N like ours.
    Each comparator's C is first checked against the same FP64 reference on
    sampled rows as ours (|c - r| / s, bound 1e-5 for FP32, 1e-2 for BF16);
    a comparator that computes something else reports its error, not a time."""
    import torch
    out = {}
    ref = Fp64Ref(W, K, sample_rows(M), dev).set_b(Bd)
    Ad = bws.unpack_bitmap(A).to(torch.float32)

    def check(name, C, bound):
        e = ref.err(C)
        out[name + "_err_vs_fp64"] = e
        ok = e["max_rel_s"] <= bound
        if not ok:
            out[name + "_error"] = f"disagrees with the FP64 product: {e['max_rel_s']:.3g} > {bound}"
        return ok
    try:
        if check("dense_fp32", torch.matmul(Ad, Bd), 1e-5):
            spin_up(lambda: torch.matmul(Ad, Bd), 0.1)
            out["dense_fp32_ms"] = round(time_loop(lambda: torch.matmul(Ad, Bd), 5, 2)[1], 4)
        Ab, Bb = Ad.to(torch.bfloat16), Bd.to(torch.bfloat16)
        if check("dense_bf16", torch.matmul(Ab, Bb).float(), 1e-2):
            spin_up(lambda: torch.matmul(Ab, Bb), 0.1)
            out["dense_bf16_ms"] = round(time_loop(lambda: torch.matmul(Ab, Bb), 20, 2)[1], 4)
        # the same BF16 operands with C written in FP32, as ours is
        try:
            f = lambda: torch.mm(Ab, Bb, out_dtype=torch.float32)
            if check("dense_bf16_f32out", f(), 1e-2):
                spin_up(f, 0.1)
                out["dense_bf16_f32out_ms"] = round(time_loop(f, 20, 2)[1], 4)
        except Exception as e:
            out["dense_bf16_f32out_error"] = str(e)[:200]
        del Ab, Bb
    except Exception as e:
        out["dense_error"] = str(e)
    try:
        Acsr = Ad.to_sparse_csr()
        if check("csr_fp32", torch.sparse.mm(Acsr, Bd), 1e-5):
            spin_up(lambda: torch.sparse.mm(Acsr, Bd), 0.1)
            ms = time_loop(lambda: torch.sparse.mm(Acsr, Bd), 5, 2)[1]
            out["csr_fp32_ms"] = round(ms, 4)
            out["csr_fp32_tflops"] = round(2.0 * nnz * N / ms / 1e9, 3)
    except Exception as e:
        out["csr_error"] = str(e)
    return out


if __name__ == "__main__":
    main()
