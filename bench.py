#!/usr/bin/env python3
"""Bit-wise SpMM benchmark on one MI355X.

Workload (default): C = A·B, M = N = K = 4096, A a clustered 0/1 bitmap
(64 prototypes ~ Bernoulli(0.03125) with 2% bit flips, rows shuffled: ~5%
dense, i.e. "95% sparse"), B ~ N(0,1) FP32, FP32 path.  A step is one
bws_spmm launch over the resident bitmap and B (the reference's timed region
is likewise the GEMM on pre-packed operands).  Prints ONE JSON line with
effective TFLOP/s = 2·nnz·N / t, the HBM roofline of the kernel, and the
reference CPU naive loop timed on a bounded row sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "spmm_fp32_4096_clustered_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def popcount_u32(W):
    import numpy as np
    return int(np.unpackbits(W.view(np.uint8)).sum())


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


def cpu_baseline(W, B, K, N, target_s, threads):
    """The reference's CPU naive loop (backup/bit_wise_gemm.cu:13-32) on the
    dense 0/1 form of a bounded row sample; reported in the metric's unit."""
    import numpy as np

    import oracle
    rows = 8
    while True:
        Ad = oracle.unpack_bitmap(W[:rows], K)
        t0 = time.perf_counter()
        oracle.check_dense_f32(Ad, B, 0, rows, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= 0.2 * target_s or rows >= W.shape[0]:
            break
        rows = min(W.shape[0], max(rows * 2, int(rows * 0.25 * target_s / max(dt, 1e-3))))
    nnz = popcount_u32(np.ascontiguousarray(W[:rows]))
    return {"value": 2.0 * nnz * N / dt / 1e12, "unit": "TFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"{rows} of {W.shape[0]} rows of the same A (dense 0/1 form) x full B; "
                      f"check() loop i,j,p FP32, OpenMP over rows; {dt:.2f} s"}


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
    ap.add_argument("--iid", action="store_true", help="iid Bernoulli(p) mask instead of clustered")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import bitwise_spgemm_amd as bws

    dev = torch.device("cuda:0")
    M, N, K = args.m, args.n, args.k
    if args.iid:
        W = bws.gen_bitmap_iid(args.seed, M, K, args.p)
    else:
        W = bws.gen_bitmap_clustered(args.seed, M, K, 64, args.p, 0.02)
    B = bws.gen_f32(args.seed + 1, (K, N), "normal", 0.0, 1.0)
    nnz = popcount_u32(W)
    A = bws.from_host_words(W, M, K, dev)
    Bd = torch.from_numpy(B).to(dev)
    Bin = {"fp32": Bd, "bf16x2": Bd, "bf16": Bd.to(torch.bfloat16), "fp16": Bd.to(torch.float16)}[
        args.precision]
    C = torch.empty((M, N), dtype=torch.float32, device=dev)
    log(f"workload M=N=K={M}x{N}x{K} nnz={nnz} density={nnz / (M * K):.4f} prec={args.precision}")

    step = lambda: bws.spmm(A, Bin, args.precision, out=C)
    ms_step, ms_kernel = time_loop(step, args.steps, args.warmup)
    flops = 2.0 * nnz * N
    value = flops / (ms_step * 1e-3) / 1e12
    s_b = {"fp32": 4, "bf16x2": 4, "bf16": 2, "fp16": 2}[args.precision]
    alg_bytes = 4 * M * ((K + 31) // 32) + s_b * K * N + 4 * M * N
    achieved = alg_bytes / (ms_kernel * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC) and args.precision == "fp32" and not args.iid and M == N == K == 4096:
        try:
            traffic = json.load(open(PROFILE_TRAFFIC)).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel_ms": round(ms_kernel, 5), "algorithmic_bytes": alg_bytes}

    extras = {}
    if not args.no_extras:
        extras = run_extras(bws, A, W, Bd, M, N, K, nnz, dev)

    cpu = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(W, B, K, N, args.cpu_seconds, threads)

    out = {
        "metric": "bitmap SpMM TFLOP/s (2*nnz*N/t)",
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
    if extras:
        out["extras"] = extras
    print(json.dumps(out), flush=True)


def run_extras(bws, A, W, Bd, M, N, K, nnz, dev):
    """Pipeline stages and the BF16 vs FP32 forward+backward comparison."""
    import torch
    ex = {}
    Ad = bws.unpack_bitmap(A)
    ex["pack_ms"] = round(time_loop(lambda: bws.pack_bitmap(Ad), 5, 2)[1], 4)
    ex["group_ms"] = round(time_loop(lambda: bws.group_rows(A, 2048, 2), 5, 2)[1], 4)
    del Ad
    AT = bws.transpose(A)
    dC = torch.randn((M, N), device=dev)
    Bb = Bd.to(torch.bfloat16)
    dCb = dC.to(torch.bfloat16)
    C = torch.empty((M, N), device=dev)
    dB = torch.empty((K, N), device=dev)

    def fb(prec, Bx, dCx):
        # 0/1 mask: the trainable operand is B, so backward is dB = Aᵀ·dC
        bws.spmm(A, Bx, prec, out=C)
        bws.spmm_tA(AT, dCx, prec, out=dB)

    try:
        ex["fwd_bwd_fp32_ms"] = round(time_loop(lambda: fb("fp32", Bd, dC), 3, 1)[1], 4)
        ex["fwd_bwd_bf16_ms"] = round(time_loop(lambda: fb("bf16", Bb, dCb), 3, 1)[1], 4)
        ex["fwd_bf16_ms"] = round(time_loop(lambda: bws.spmm(A, Bb, "bf16", out=C), 10, 2)[1], 4)
    except Exception as e:  # reported, never hidden
        ex["fwd_bwd_error"] = str(e)
    return ex


if __name__ == "__main__":
    main()
