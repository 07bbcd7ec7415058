#!/usr/bin/env python3
"""Benchmark: SGEMM 8192 x 8192 x 8192, BF16x3 split precision vs plain FP32
MFMA, with normwise error against an FP64 reference (the four-product BF16x4
variant is timed and checked beside them).

One step = one C = A*B over the resident 8192^3 operands (A, B uniform[-1,1]
from MT19937 seed 1, beta = 0).  `value` is the BF16x3 GEMM's throughput
(2*M*N*K useful flops per step); the plain FP32-MFMA GEMM is timed the same
way in the same run and reported beside it, with both normwise errors
||C - C64||_F / (||A||_F ||B||_F) computed in FP64 on the GPU against an FP64
MFMA product of the same inputs (itself spot-checked against numpy float64
rows).  The cpu_baseline leg times the reference's CPU verification loop
(Matrix<T>::operator*, Kernel/cuda/src/Matrix.h:183-201, restated in the
oracle) on a bounded row sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SGEMM TFLOP/s + % peak: BF16x3-MFMA vs plain FP32, at normwise err vs FP64 oracle"
PEAK_BF16 = 2516.6  # TFLOP/s, MI355X dense bf16 MFMA
PEAK_BF16X3 = PEAK_BF16 / 3.0  # useful-flop peak of three bf16 products per FP32 product
PEAK_FP32 = 157.3  # TFLOP/s, MI355X FP32 MFMA
M = N = K = 8192


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_launches(fn, stream, steps: int, warmup: int, torch):
    """Warm up, then time `steps` launches with one HIP event pair per launch
    on the launch stream; return (wall ms per step, mean event ms per launch)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        fn()
        e.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    per = [s.elapsed_time(e) for s, e in ev]
    return wall, float(np.mean(per)), float(np.min(per))


def cpu_share() -> int:
    """Host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when set (the GPU boxes give one GPU a share of 16)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A: np.ndarray, B: np.ndarray, target_s: float = 12.0) -> dict:
    """Time the reference's CPU verification product (Matrix<T>::operator*,
    naive i-j-k with a float accumulator): on a row sample of this workload
    over the process's CPU share, and config 1 (1024^3) in full on one core
    and on the CPU share."""
    import oracle  # the CPU restatement; used here only as the timed baseline

    threads = cpu_share()
    rows = threads  # one row per thread per round; grow until the sample takes ~target_s
    while True:
        t0 = time.perf_counter()
        oracle.cpu_naive_sgemm(A, B, rows=rows, row0=0, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt >= target_s / 3 or rows >= 1024:
            break
        grow = max(2, int(target_s / max(dt, 1e-3)))
        rows = min(1024, rows * grow // threads * threads)
    flops = 2.0 * rows * N * K
    t0 = time.perf_counter()
    oracle.cpu_naive_sgemm(A, B, rows=1, row0=rows, nthreads=1)
    dt1 = time.perf_counter() - t0
    # Config 1 in full: 1024^3, the seed-0 stream of the plain-SGEMM workload.
    A1, B1, _ = oracle.workload_plain_1024()
    f1 = 2.0 * 1024**3
    t0 = time.perf_counter()
    oracle.cpu_naive_sgemm(A1, B1, nthreads=1)
    c1_one = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.cpu_naive_sgemm(A1, B1, nthreads=threads)
    c1_all = time.perf_counter() - t0
    return {
        "value": flops / dt / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"rows 0..{rows - 1} of the {M}x{N}x{K} product (naive i-j-k, float "
                   f"accumulator, OpenMP over rows, {threads} threads) in {dt:.1f} s; "
                   f"1 row on 1 core: {2.0 * N * K / dt1 / 1e9:.3f} GFLOP/s; extrapolated"),
        "single_core_value": 2.0 * N * K / dt1 / 1e12,
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "cpu_share_basis": "process affinity capped by OMP_NUM_THREADS (16 threads per GPU on the GPU boxes)",
        "config1_1024cubed_full": {
            "one_core_s": round(c1_one, 3), "one_core_gflops": round(f1 / c1_one / 1e9, 3),
            "all_cores_s": round(c1_all, 3), "all_cores_gflops": round(f1 / c1_all / 1e9, 3),
            "cores": threads},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 error reports")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy first (scale for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    import torch

    import cucosma_amd as cc

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log("generating inputs (numpy RandomState(1), uniform[-1,1])")
    rs = np.random.RandomState(1)
    A = rs.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rs.uniform(-1, 1, (K, N)).astype(np.float32)
    Ad = torch.from_numpy(A).to(dev)
    Bd = torch.from_numpy(B).to(dev)
    Cb = torch.empty((M, N), dtype=torch.float32, device=dev)
    Cp = torch.empty((M, N), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    lib = cc.lib
    a, b, cb, cp = Ad.data_ptr(), Bd.data_ptr(), Cb.data_ptr(), Cp.data_ptr()

    def run_bf16x3():
        st = lib.cosma_sgemm_bf16x3_8192x8192x8192(1.0, a, K, b, N, 0.0, cb, N)
        if st:
            raise cc.CosmaError(st, "bf16x3")

    def run_fp32():
        st = lib.cosma_sgemm_8192x8192x8192(1.0, a, K, b, N, 0.0, cp, N)
        if st:
            raise cc.CosmaError(st, "fp32")

    C4 = torch.empty((M, N), dtype=torch.float32, device=dev)

    def run_bf16x4():
        cc.gemm(Ad, Bd, C4, mode="bf16x4", stream=stream)

    if args.copy_calibration:
        # A plain device copy of known size (1 GiB read + 1 GiB write) in the
        # same process, so counter runs can scale FETCH_SIZE / WRITE_SIZE.
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # Both fixed-shape entry points launch on the default (null) stream, which
    # is torch's current stream here.
    flops = 2.0 * M * N * K
    log(f"timing bf16x3: warmup {args.warmup}, steps {args.steps}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wall_b, ev_b, min_b = time_launches(run_bf16x3, stream, args.steps, args.warmup, torch)
    torch.cuda.synchronize()
    log(f"bf16x3: {ev_b:.3f} ms/launch ({flops / ev_b / 1e9:.1f} TFLOP/s) in {time.perf_counter() - t0:.1f} s")
    wall_p, ev_p, min_p = time_launches(run_fp32, stream, args.steps, args.warmup, torch)
    log(f"fp32:   {ev_p:.3f} ms/launch ({flops / ev_p / 1e9:.1f} TFLOP/s)")
    wall_4, ev_4, min_4 = time_launches(run_bf16x4, stream, args.steps, args.warmup, torch)
    log(f"bf16x4: {ev_4:.3f} ms/launch ({flops / ev_4 / 1e9:.1f} TFLOP/s)")

    topo = cc.device_topology()
    peak_dev = cc.peak_tflops("bf16x3", topo)
    peak_basis_dev = (f"{topo['cus']} CUs x {topo['clock_khz'] / 1e3:.0f} MHz (device peak clock) x 4096 bf16 "
                      f"flops/clock/CU / 3 products = {peak_dev:.1f} TFLOP/s")
    cfg_b, split_b = cc.heuristic(cc.MODE_BF16X3, M, N, K)
    tile_b = cc.tile_configs(cc.MODE_BF16X3)[cfg_b]
    presplit = tile_b["pipe"] == 5
    shape_b = (f"{tile_b['BM']}x{tile_b['BN']}x{tile_b['BK']}, "
               f"{tile_b['threads'] // 64} waves of {tile_b['WM']}x{tile_b['WN']}")
    if presplit:
        # Each step splits A and B into bf16 planes (one HBM-bound launch) and
        # runs planar_kernel on them.  The roofline is the dominant kernel's:
        # split once, then time planar_kernel launches alone with HIP events.
        kname = f"planar_kernel<BF16x3, {shape_b}>"
        ap = cc.split_operand(Ad)
        bp = cc.split_operand(Bd, as_b=True)

        def run_kernel():
            cc.gemm(Ad, Bd, Cb, mode="bf16x3", a_planes=ap, b_planes=bp, stream=stream)

        _, ev_k, _ = time_launches(run_kernel, stream, args.steps, args.warmup, torch)
        log(f"planar_kernel alone: {ev_k:.3f} ms/launch ({flops / ev_k / 1e9:.1f} TFLOP/s); "
            f"split pass {ev_b - ev_k:.3f} ms per step")
        del ap, bp
    else:
        kname = f"gemm_kernel<BF16x3, {shape_b}, pipe {tile_b['pipe']}>"
        ev_k = ev_b
    tf_b = flops / wall_b / 1e9
    tf_p = flops / wall_p / 1e9
    tf_4 = flops / wall_4 / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_sgemm_8192.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tinfo = json.load(f)
        if tinfo.get("kernel") == kname:
            traffic = tinfo.get("bytes_per_launch")
    result = {
        "metric": METRIC,
        "value": round(tf_b, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_b, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3 (fp32 operands split to bf16 hi/lo, fp32 accumulate)",
        "data": "synthetic: A, B uniform[-1,1] from MT19937 seed 1 (numpy RandomState), beta=0",
        "config": {"workload": "sgemm 8192x8192x8192 row-major, alpha=1 beta=0, split-precision "
                               "bf16x3 vs plain fp32 mfma", "M": M, "N": N, "K": K},
        "bf16x3": {"tflops": round(tf_b, 3), "pct_useful_peak": round(100 * tf_b / PEAK_BF16X3, 2),
                   "pct_bf16_mfma_peak": round(100 * 3 * tf_b / PEAK_BF16, 2),
                   "ms_per_launch_events": round(ev_b, 4), "kernel": kname,
                   "kernel_ms_per_launch_events": round(ev_k, 4),
                   "split_pass_ms_per_step": round(ev_b - ev_k, 4) if presplit else 0.0},
        "plain_fp32": {"tflops": round(tf_p, 3), "pct_peak": round(100 * tf_p / PEAK_FP32, 2),
                       "ms_per_launch_events": round(ev_p, 4)},
        "bf16x4": {"tflops": round(tf_4, 3), "pct_bf16_mfma_peak": round(100 * 4 * tf_4 / PEAK_BF16, 2),
                   "ms_per_launch_events": round(ev_4, 4)},
        "speedup_bf16x3_vs_fp32": round(tf_b / tf_p, 3),
        "speedup_bf16x4_vs_fp32": round(tf_4 / tf_p, 3),
        "roofline": {"bound": "mfma", "achieved": round(flops / ev_k / 1e9, 3),
                     "peak": round(PEAK_BF16X3, 1), "unit": "TFLOP/s",
                     "frac": round(flops / ev_k / 1e9 / PEAK_BF16X3, 4), "traffic": traffic,
                     "kernel": kname,
                     "peak_basis": "MI355X dense bf16 MFMA 2516.6 TF / 3 products",
                     "peak_device": round(peak_dev, 1), "frac_device": round(flops / ev_k / 1e9 / peak_dev, 4),
                     "peak_device_basis": peak_basis_dev},
    }

    if not args.no_check:
        log("FP64 reference on the GPU (FP64 MFMA on exact copies)")
        R = cc.fp64_product(Ad, Bd)
        torch.cuda.synchronize()
        rows = [0, 4097, 8191]
        ref_rows = A[rows].astype(np.float64) @ B.astype(np.float64)
        dev_rows = R[rows].cpu().numpy()
        spot = float(np.max(np.abs(dev_rows - ref_rows)) / np.max(np.abs(ref_rows)))
        eb = cc.normwise_error(Ad, Bd, Cb, R)
        ep = cc.normwise_error(Ad, Bd, Cp, R)
        e4 = cc.normwise_error(Ad, Bd, C4, R)
        result["bf16x4"]["normwise_err"] = e4["normwise"]
        result["err_ratio_bf16x4_vs_fp32"] = e4["normwise"] / ep["normwise"] if ep["normwise"] else None
        result["bf16x3"]["normwise_err"] = eb["normwise"]
        result["bf16x3"]["max_rel_err"] = eb["max_rel"]
        result["plain_fp32"]["normwise_err"] = ep["normwise"]
        result["plain_fp32"]["max_rel_err"] = ep["max_rel"]
        result["max_rel_err_basis"] = ("max |C - R| / |R| over the entries with |R| >= 1e-3 max|R| "
                                       "(R: the GPU FP64 product)")
        result["err_ratio_bf16x3_vs_fp32"] = eb["normwise"] / ep["normwise"] if ep["normwise"] else None
        result["fp64_reference_spotcheck_rel"] = spot
        del R
        log(f"normwise: bf16x3 {eb['normwise']:.3e}  fp32 {ep['normwise']:.3e}  "
            f"bf16x4 {e4['normwise']:.3e}")

    if not args.no_cpu:
        log("cpu baseline (reference CPU verification loop on a row sample)")
        result["cpu_baseline"] = cpu_baseline(A, B)
    else:
        result["cpu_baseline"] = None

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
