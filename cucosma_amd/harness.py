"""Benchmark harness for every workload configuration of the hot path.

The reference's harness runs each implementation 10 times untimed and 100
times timed with a cache flush after every run (Kernel/cuda/src/benchmark.h:30-70,
Kernel/hip/src/benchmark.h:39-91, Cacheflusher.h:18-49) and writes
"<M>x<N>x<K><name>.csv" files with an "Implementation,ms" header.  This
module keeps that protocol (through the native ``cosma_benchmark`` entry
point) and that CSV format, and adds what the reference never reported:
TFLOP/s as a share of the MI355X peak for the arithmetic used, HBM GB/s for
the HBM-bound flat shape, and normwise error against an FP64 product.

Configurations:
  plain    1024^3, alpha 1, beta 0.5, A/B/C uniform[-1,1] from one MT19937(0) stream
  square   8192^3, A/B uniform[-1,1] seed 1, beta 0: fp32, bf16x3, bf16x4
  flat     256 x 256 x 1,048,576, seed 1, beta 0: fp32, bf16x3, bf16x4, the
           library's tile with split-K 8..512 (its default split marked)
  epilogue 8192^3 fp32 and bf16x3: plain, sigmoid fused into the store, and
           plain followed by a separate sigmoid pass (the reference's two
           variants, cuCOSMAV100.cuh:2466-2476 and sigmoid4.cuh:13-27)
  dgemm    32768x1024x1024 and 1024x65536x1024, normal(0,1) seed 2 with 1% of
           entries x1e6, C normal(0,1), alpha = beta = 1
  grouped  512 problems, M/N/K log-uniform in [32, 2048] rounded to 32, seed 3,
           fp32, bf16x3, bf16x4 and fp64 (FP64 storage for fp64):
           one hipGraph replay of the library's default grouped plan, of the
           grouped autotuner's plan, and one launch per problem
  grouped_tiles  uniform batches of 2048^2 outputs, K = 128 .. 2048, on the large
           tile: per-K-step and fixed per-tile cost of a grouped launch
  autograd X[16384x4096] W[4096x4096], forward + backward, bf16x3
  schedules the two on-the-fly 256x256 BF16 schedules (ping-pong cfg 6, spread
           loads cfg 12) on five shapes, orders alternated round by round
  heuristic the library's default tile and split against the autotuner's fastest
           candidate on five shape sets of 94, 58, 45, 48 and 45 shapes (BF16x3
           and BF16x4, or --modes)

Usage: python -m cucosma_amd.harness --config square [--runs 100] [--warmup 10]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

PEAK = {"fp32": 157.3, "fp64": 78.6, "bf16x3": 2516.6 / 3, "bf16x4": 2516.6 / 4}
HBM_PEAK = 8000.0  # GB/s
FLUSH_BYTES = 512 << 20


def _desc(cc, mode, A, B, Cm, alpha, beta, split_k=0, tile_cfg=-1, group_m=0, epilogue=None):
    from ._lib import MODES, GemmDesc
    from .gemm import _ld, _stream_ptr

    d = GemmDesc()
    cc.lib.cosma_gemm_desc_init(C.byref(d))
    d.mode = MODES[mode]
    d.m, d.k = A.shape
    d.n = B.shape[1]
    d.a, d.lda = A.data_ptr(), _ld(A, "A")
    d.b, d.ldb = B.data_ptr(), _ld(B, "B")
    d.c, d.ldc = Cm.data_ptr(), _ld(Cm, "C")
    d.alpha, d.beta = alpha, beta
    d.split_k, d.tile_cfg, d.group_m = split_k, tile_cfg, group_m
    d.epilogue = cc.EPILOGUES[epilogue]
    d.stream = _stream_ptr(None)
    return d


def bench_gemm(cc, mode, A, B, Cm, alpha=1.0, beta=0.0, runs=100, warmup=10, flush=True, **kw):
    """Reference protocol through cosma_benchmark; returns per-run ms."""
    d = _desc(cc, mode, A, B, Cm, alpha, beta, **kw)
    ms = (C.c_float * max(runs, 1))()
    cc.check(cc.lib.cosma_benchmark(C.byref(d), warmup, runs, FLUSH_BYTES if flush else 0, ms),
             "cosma_benchmark")
    return np.array(ms[:runs], dtype=np.float64)


def _summary(name, mode, M, N, K, ms, extra=None):
    med = float(np.median(ms))
    flops = 2.0 * M * N * K
    out = dict(config=name, mode=mode, M=M, N=N, K=K, runs=len(ms), ms_median=med,
               ms_mean=float(ms.mean()), ms_min=float(ms.min()),
               tflops=flops / med / 1e9, pct_peak=100 * flops / med / 1e9 / PEAK[mode])
    if extra:
        out.update(extra)
    return out


def _write_csv(csv_dir, M, N, K, name, ms):
    if not csv_dir:
        return
    os.makedirs(csv_dir, exist_ok=True)
    with open(os.path.join(csv_dir, f"{M}x{N}x{K}{name}.csv"), "w") as f:
        f.write("Implementation,ms\n")
        for v in ms:
            f.write(f"{name},{v:.6g}\n")


def _err(cc, A, B, Cm, R):
    e = cc.normwise_error(A, B, Cm, R)
    return dict(normwise_err=e["normwise"], max_rel_err=e["max_rel"], max_abs_err=e["max_abs"])


def cfg_plain(cc, torch, args, emit):
    M = N = K = 1024
    rs = np.random.RandomState(0)
    A = torch.from_numpy(rs.uniform(-1, 1, (M, K)).astype(np.float32)).cuda()
    B = torch.from_numpy(rs.uniform(-1, 1, (K, N)).astype(np.float32)).cuda()
    C0 = torch.from_numpy(rs.uniform(-1, 1, (M, N)).astype(np.float32)).cuda()
    R = cc.fp64_product(A, B, C0, 1.0, 0.5)
    for mode in ("fp32", "bf16x3"):
        Cm = C0.clone()
        cc.gemm(A, B, Cm, alpha=1.0, beta=0.5, mode=mode)
        err = _err(cc, A, B, Cm, R)
        ms = bench_gemm(cc, mode, A, B, C0.clone(), 1.0, 0.5, args.runs, args.warmup, not args.no_flush)
        _write_csv(args.csv_dir, M, N, K, "cosma_" + mode, ms)
        emit(_summary("plain", mode, M, N, K, ms, err))


def cfg_square(cc, torch, args, emit):
    M = N = K = 8192
    rs = np.random.RandomState(1)
    A = torch.from_numpy(rs.uniform(-1, 1, (M, K)).astype(np.float32)).cuda()
    B = torch.from_numpy(rs.uniform(-1, 1, (K, N)).astype(np.float32)).cuda()
    R = cc.fp64_product(A, B)
    for mode in ("fp32", "bf16x3", "bf16x4"):
        Cm = torch.empty((M, N), device="cuda")
        cc.gemm(A, B, Cm, mode=mode)
        err = _err(cc, A, B, Cm, R)
        ms = bench_gemm(cc, mode, A, B, Cm, 1.0, 0.0, args.runs, args.warmup, not args.no_flush)
        _write_csv(args.csv_dir, M, N, K, "cosma_" + mode, ms)
        emit(_summary("square", mode, M, N, K, ms, err))
    del R


def cfg_flat(cc, torch, args, emit):
    M, N, K = 256, 256, 1 << 20
    rs = np.random.RandomState(1)
    A = torch.from_numpy(rs.uniform(-1, 1, (M, K)).astype(np.float32)).cuda()
    B = torch.from_numpy(rs.uniform(-1, 1, (K, N)).astype(np.float32)).cuda()
    R = cc.fp64_product(A, B)
    alg_bytes = (M * K + K * N + M * N) * 4
    for mode in ("fp32", "bf16x3", "bf16x4"):
        cfg, default_split = cc.heuristic(cc._lib.MODES[mode], M, N, K)
        t = cc.tile_configs(cc._lib.MODES[mode])[cfg]
        tiles = (M // t["BM"]) * (N // t["BN"])
        for split in sorted({8, 16, 32, 64, 128, 256, 512, default_split}):
            if tiles * split > 4096:
                continue
            Cm = torch.empty((M, N), device="cuda")
            cc.gemm(A, B, Cm, mode=mode, split_k=split, tile_cfg=cfg)
            err = _err(cc, A, B, Cm, R)
            ms = bench_gemm(cc, mode, A, B, Cm, 1.0, 0.0, args.runs, args.warmup, not args.no_flush,
                            split_k=split, tile_cfg=cfg)
            med = float(np.median(ms))
            gbs = alg_bytes / (med * 1e-3) / 1e9
            _write_csv(args.csv_dir, M, N, K, f"cosma_{mode}_sk{split}", ms)
            emit(_summary("flat", mode, M, N, K, ms, dict(
                err, split_k=split, tile_cfg=cfg, tile=f"{t['BM']}x{t['BN']}x{t['BK']}",
                library_default=split == default_split, hbm_gbs=gbs, pct_hbm=100 * gbs / HBM_PEAK)))


def cfg_epilogue(cc, torch, args, emit):
    M = N = K = 8192
    rs = np.random.RandomState(1)
    A = torch.from_numpy(rs.uniform(-1, 1, (M, K)).astype(np.float32)).cuda()
    B = torch.from_numpy(rs.uniform(-1, 1, (K, N)).astype(np.float32)).cuda()
    Cm = torch.empty((M, N), device="cuda")
    flops = 2.0 * M * N * K
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for mode in ("fp32", "bf16x3"):
        plain = float(np.median(bench_gemm(cc, mode, A, B, Cm, 1.0, 0.0, args.runs, args.warmup,
                                           not args.no_flush)))
        fused = float(np.median(bench_gemm(cc, mode, A, B, Cm, 1.0, 0.0, args.runs, args.warmup,
                                           not args.no_flush, epilogue="sigmoid")))
        # separate pass: the GEMM, then an element-wise sigmoid over C
        for _ in range(3):
            cc.gemm(A, B, Cm, mode=mode)
            torch.sigmoid_(Cm)
        torch.cuda.synchronize()
        reps = max(3, args.runs // 5)
        ev[0].record()
        for _ in range(reps):
            cc.gemm(A, B, Cm, mode=mode)
            torch.sigmoid_(Cm)
        ev[1].record()
        torch.cuda.synchronize()
        separate = ev[0].elapsed_time(ev[1]) / reps
        emit(dict(config="epilogue", mode=mode, M=M, N=N, K=K, act="sigmoid", ms_plain=plain,
                  ms_fused=fused, ms_separate_pass=separate,
                  tflops_fused=flops / fused / 1e9, fused_overhead_pct=100 * (fused / plain - 1),
                  separate_overhead_pct=100 * (separate / plain - 1)))


def cfg_dgemm(cc, torch, args, emit):
    for (M, N, K) in ((32768, 1024, 1024), (1024, 65536, 1024)):
        rs = np.random.RandomState(2)
        A = rs.normal(0, 1, (M, K))
        A[rs.random_sample((M, K)) < 0.01] *= 1e6
        B = rs.normal(0, 1, (K, N))
        B[rs.random_sample((K, N)) < 0.01] *= 1e6
        C0 = rs.normal(0, 1, (M, N))
        Ad, Bd, Cd = (torch.from_numpy(x).cuda() for x in (A, B, C0))
        ms = bench_gemm(cc, "fp64", Ad, Bd, Cd, 1.0, 1.0, args.runs, args.warmup, not args.no_flush)
        _write_csv(args.csv_dir, M, N, K, "cosma_dgemm", ms)
        emit(_summary("dgemm", "fp64", M, N, K, ms))


def grouped_sizes(count=512, seed=3, lo=32, hi=2048):
    """M, N, K each log-uniform in [lo, hi], rounded to a multiple of 32."""
    rs = np.random.RandomState(seed)
    d = np.exp(rs.uniform(np.log(lo), np.log(hi), (count, 3)))
    d = np.clip(np.round(d / 32) * 32, lo, hi).astype(int)
    return [tuple(int(x) for x in row) for row in d]


def cfg_grouped(cc, torch, args, emit):
    sizes = grouped_sizes()
    flops = sum(2.0 * m * n * k for m, n, k in sizes)
    modes = args.grouped_modes.split(",") if getattr(args, "grouped_modes", None) else \
        ("fp32", "bf16x3", "bf16x4", "fp64")
    for mode in modes:
        dt = torch.float64 if mode == "fp64" else torch.float32
        g = torch.Generator(device="cuda").manual_seed(3)
        As = [torch.rand((m, k), device="cuda", generator=g, dtype=dt) * 2 - 1 for m, n, k in sizes]
        Bs = [torch.rand((k, n), device="cuda", generator=g, dtype=dt) * 2 - 1 for m, n, k in sizes]
        Cs = [torch.empty((m, n), device="cuda", dtype=dt) for m, n, k in sizes]
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

        def time_plan(plan, reps):
            # warm-up replays first, so that every plan (the first one timed
            # included) runs at the clocks of a busy GPU
            for _ in range(20):
                plan.run_graph()
            torch.cuda.synchronize()
            ev[0].record()
            for _ in range(reps):
                plan.run_graph()
            ev[1].record()
            torch.cuda.synchronize()
            return ev[0].elapsed_time(ev[1]) / reps

        def describe(plan):
            tiles = cc.tile_configs(cc._lib.MODES[mode])
            use = {}
            for c, s in plan.choices():
                key = f"{tiles[c]['BM']}x{tiles[c]['BN']}" + (f"/sk{s}" if s > 1 else "")
                use[key] = use.get(key, 0) + 1
            return use

        # The library's default plan: each problem on the first tile family's
        # variant with the least weighted padding, one persistent launch.
        plan = cc.GroupedGemm(As, Bs, Cs, mode=mode)
        ms_graph = time_plan(plan, args.runs)
        info, use = plan.info(), describe(plan)
        padded = sum(2.0 * -(-m // cc.tile_configs(cc._lib.MODES[mode])[c]["BM"])
                     * cc.tile_configs(cc._lib.MODES[mode])[c]["BM"]
                     * -(-n // cc.tile_configs(cc._lib.MODES[mode])[c]["BN"])
                     * cc.tile_configs(cc._lib.MODES[mode])[c]["BN"] * k
                     for (m, n, k), (c, s) in zip(sizes, plan.choices()))
        plan.close()
        # The grouped autotuner's per-size-bucket choices.
        tuned = None
        if not args.no_search:
            res = cc.grouped_autotune(As, Bs, Cs, mode=mode, warmup=3, reps=10)
            tp = cc.GroupedGemm(As, Bs, Cs, mode=mode, tile_cfgs=res["tile_cfgs"], splits=res["splits"])
            ms_tuned = time_plan(tp, args.runs)
            tuned = dict(ms=ms_tuned, tflops=flops / ms_tuned / 1e9, tiles=describe(tp),
                         candidates=len(res["tried"]))
            tp.close()
        # one launch per problem (library heuristic per shape)
        for A, B, Cm in zip(As, Bs, Cs):
            cc.gemm(A, B, Cm, mode=mode)
        torch.cuda.synchronize()
        reps = max(1, args.runs // 10)
        ev[0].record()
        for _ in range(reps):
            for A, B, Cm in zip(As, Bs, Cs):
                cc.gemm(A, B, Cm, mode=mode)
        ev[1].record()
        torch.cuda.synchronize()
        ms_loop = ev[0].elapsed_time(ev[1]) / reps
        emit(dict(config="grouped", mode=mode, problems=len(sizes), gflop=flops / 1e9,
                  padded_mfma_gflop=padded / 1e9, buckets=info["buckets"], tiles_used=use,
                  ms_graph=ms_graph, tflops_graph=flops / ms_graph / 1e9, tuned=tuned,
                  ms_per_problem_launches=ms_loop, tflops_per_problem_launches=flops / ms_loop / 1e9,
                  speedup=ms_loop / ms_graph,
                  pct_peak=100 * flops / ms_graph / 1e9 / PEAK[mode]))
        del As, Bs, Cs


def cfg_grouped_tiles(cc, torch, args, emit):
    """Per-tile cost of a grouped launch: uniform batches of 2048 x 2048
    outputs with K = 128 .. 2048 on the first family's large tile, timed as
    graph replays; the slope over K is the cost of a K-step, the intercept the
    fixed cost of a tile (prologue, epilogue stores)."""
    modes = args.grouped_modes.split(",") if getattr(args, "grouped_modes", None) else ("bf16x3", "fp32")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for mode in modes:
        dt = torch.float64 if mode == "fp64" else torch.float32
        big = cc.group_families(mode)[0][0]
        rows = []
        for (cnt, k) in ((64, 128), (64, 256), (32, 1024), (16, 2048)):
            m = n = 2048
            As = [torch.rand((m, k), device="cuda", dtype=dt) for _ in range(cnt)]
            Bs = [torch.rand((k, n), device="cuda", dtype=dt) for _ in range(cnt)]
            Cs = [torch.empty((m, n), device="cuda", dtype=dt) for _ in range(cnt)]
            plan = cc.GroupedGemm(As, Bs, Cs, mode=mode, tile_cfgs=[big] * cnt, splits=[1] * cnt)
            info = plan.info()
            for _ in range(10):
                plan.run_graph()
            torch.cuda.synchronize()
            ev[0].record()
            for _ in range(args.runs):
                plan.run_graph()
            ev[1].record()
            torch.cuda.synchronize()
            ms = ev[0].elapsed_time(ev[1]) / args.runs
            per_wg = info["total_tiles"] / info["buckets"][0]["blocks"]
            t = cc.tile_configs(cc._lib.MODES[mode])[big]
            rows.append((k // t["BK"], ms * 1e3 / per_wg))
            emit(dict(config="grouped_tiles", mode=mode, tile=f"{t['BM']}x{t['BN']}x{t['BK']}", K=k, problems=cnt,
                      ms=ms, us_per_tile=ms * 1e3 / per_wg, tflops=cnt * 2.0 * m * n * k / ms / 1e9))
            plan.close()
            del As, Bs, Cs
        x = np.array([r[0] for r in rows], float)
        y = np.array([r[1] for r in rows], float)
        slope, icpt = np.polyfit(x, y, 1)
        emit(dict(config="grouped_tiles_fit", mode=mode, us_per_kstep=float(slope), us_per_tile_fixed=float(icpt)))


def cfg_autograd(cc, torch, args, emit):
    g = torch.Generator(device="cpu").manual_seed(4)
    X = (torch.rand(16384, 4096, generator=g) * 2 - 1).cuda().requires_grad_()
    W = (torch.randn(4096, 4096, generator=g) / 64).cuda().requires_grad_()
    dY = torch.randn(16384, 4096, generator=g).cuda()
    flops = 2.0 * 16384 * 4096 * 4096
    for mode in ("bf16x3", "fp32"):
        for _ in range(3):
            Y = cc.matmul(X, W, mode)
            Y.backward(dY)
        torch.cuda.synchronize()
        reps = max(3, args.runs // 10)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        fwd = bwd = 0.0
        for _ in range(reps):
            X.grad = W.grad = None
            ev[0].record()
            Y = cc.matmul(X, W, mode)
            ev[1].record()
            Y.backward(dY)
            ev[2].record()
            torch.cuda.synchronize()
            fwd += ev[0].elapsed_time(ev[1])
            bwd += ev[1].elapsed_time(ev[2])
        fwd /= reps
        bwd /= reps
        emit(dict(config="autograd", mode=mode, M=16384, N=4096, K=4096, ms_forward=fwd,
                  ms_backward=bwd, tflops_forward=flops / fwd / 1e9,
                  tflops_backward=2 * flops / bwd / 1e9,
                  pct_peak_step=100 * 3 * flops / (fwd + bwd) / 1e9 / PEAK[mode]))


def cfg_schedules(cc, torch, args, emit):
    """The two on-the-fly 256x256 BF16 schedules side by side: ping-pong
    copies (cfg 6, the grouped launches' large tile) against next-chunk loads
    spread among the MFMAs (cfg 12, the single-launch large tile).  Same
    products in the same order (bitwise equal); the orders alternate from
    round to round and the first round is dropped, since a configuration timed
    later in a process tends to run faster."""
    shapes = [(256, 256, 1 << 20, 256), (8192, 8192, 8192, 1), (4096, 4096, 4096, 1),
              (1024, 1024, 32768, 16), (2048, 2048, 16384, 4)]
    rounds = 6
    for mode in ("bf16x3", "bf16x4"):
        for (M, N, K, split) in shapes:
            g = torch.Generator(device="cuda").manual_seed(1)
            A = torch.rand((M, K), device="cuda", generator=g) * 2 - 1
            B = torch.rand((K, N), device="cuda", generator=g) * 2 - 1
            Cm = torch.empty((M, N), device="cuda")
            res = {6: [], 12: []}
            for r in range(rounds):
                for cfg in ((6, 12) if r % 2 == 0 else (12, 6)):
                    ms = bench_gemm(cc, mode, A, B, Cm, runs=max(3, args.runs // 10), warmup=1,
                                    flush=not args.no_flush, split_k=split, tile_cfg=cfg)
                    if r > 0:
                        res[cfg] += list(ms)
            m6, m12 = float(np.median(res[6])), float(np.median(res[12]))
            flops = 2.0 * M * N * K
            emit(dict(config="schedules", mode=mode, M=M, N=N, K=K, split_k=split,
                      ms_pingpong_cfg6=m6, ms_spread_cfg12=m12, ratio_cfg12_over_cfg6=m12 / m6,
                      tflops_cfg6=flops / m6 / 1e9, tflops_cfg12=flops / m12 / 1e9))
            del A, B, Cm


def heuristic_shape_sets():
    """Shape sets whose 256x256 tiles mostly leave CUs idle: A, M = N from 384
    to 3072 (multiples of 256 and 128) with K from 512 to 32768; B, M = N from
    448 to 3328 (not multiples of 256) with K from 768 to 24576; C, M = N from
    320 to 3200 with K from 640 to 20480 (ten rectangular shapes each); S,
    skinny products, M from 16 to 192 against N from 1024 to 16384 and K from
    1024 to 16384, and three with N small; E, 256x256 tiles filling half the
    CUs or more, M = N from 2900 to 4000 (edge tiles but 3072 and 3328) with
    K from 512 to 8192, 4096^2 and four long rectangles."""
    a = [(m, m, k) for m in (384, 512, 640, 768, 1024, 1280, 1536, 1792, 2048, 2304, 2560, 3072)
         for k in (512, 1024, 2048, 4096, 8192, 16384, 32768)]
    a += [(1024, 2048, 4096), (2048, 1024, 8192), (512, 2048, 8192), (2048, 512, 8192), (1024, 4096, 2048),
          (4096, 1024, 2048), (768, 3072, 4096), (3072, 768, 4096), (1536, 3072, 2048), (512, 4096, 4096)]
    b = [(m, m, k) for m in (448, 896, 1152, 1408, 1664, 2176, 2816, 3328) for k in (768, 1536, 3072, 6144, 12288, 24576)]
    b += [(1000, 3000, 5000), (700, 5000, 3000), (4000, 600, 8000), (1500, 1500, 7000), (2900, 2900, 1100),
          (3000, 1000, 2000), (600, 600, 20000), (1200, 2400, 600), (3500, 3500, 3000), (2200, 2200, 9000)]
    c = [(m, m, k) for m in (320, 576, 960, 1216, 1856, 2432, 3200) for k in (640, 2560, 5120, 10240, 20480)]
    c += [(800, 2000, 6000), (2000, 800, 1500), (1300, 2600, 3500), (300, 3000, 12000), (3100, 1700, 900),
          (1750, 1750, 2500), (2600, 2600, 4500), (500, 1500, 30000), (2300, 1100, 16000), (3600, 3000, 1800)]
    s = [(m, n, k) for m in (16, 32, 64, 128, 192) for n in (1024, 4096, 16384) for k in (1024, 4096, 16384)]
    s += [(4096, 64, 4096), (16384, 128, 4096), (4096, 32, 16384)]
    e = [(m, m, k) for m in (2900, 3072, 3200, 3328, 3500, 3700, 3900) for k in (512, 1024, 2048, 4096, 8192)]
    e += [(m, n, k) for (m, n) in ((4000, 4000), (4096, 4096), (3900, 7900), (2000, 16000), (1500, 20000))
          for k in (512, 2048)]
    return {"A": a, "B": b, "C": c, "S": s, "E": e}


def cfg_heuristic(cc, torch, args, emit):
    """The library's default (tile, split) against the autotuner's fastest
    candidate (BF16x3 and BF16x4, or --modes): every tile configuration and
    group size at K splits 1, 2, 4, ..., floor(j CUs / tiles) for j = 1 to 4
    and the default's own, cache flushed.  When the default is not the
    fastest, the two are timed again against each other, four rounds of
    --runs/20 flushed calls each in alternating order, and the ratio is of
    their medians: the tuner's minimum over many single passes is biased low
    against one pass of the default.  One line per shape (default time /
    fastest, and the tuner's own ratio), then per set and mode the geometric
    mean, the largest ratio and the count above 1.10."""
    from . import autotune

    reps = max(3, args.runs // 20)
    g = torch.Generator(device="cuda").manual_seed(7)
    for mode in (args.modes.split(",") if args.modes else ("bf16x3", "bf16x4")):
        mi = cc._lib.MODES[mode]
        dt = torch.float64 if mode == "fp64" else torch.float32
        for name, shapes in heuristic_shape_sets().items():
            ratios = []
            for (M, N, K) in shapes:
                rec = autotune.tune(mode, M, N, K, warmup=1, reps=reps)
                cfg, split = cc.heuristic(mi, M, N, K)
                ok = [c for c in rec["candidates"] if c["status"] == 0]
                best = ok[0]
                mine = [c for c in ok if c["tile_cfg"] == cfg and c["split_k"] == split and c["group_m"] == 8]
                if not mine:
                    continue
                tuner_ratio = mine[0]["ms"] / best["ms"]
                ms_mine, ms_best = mine[0]["ms"], best["ms"]
                if mine[0] is not best:
                    A = torch.rand((M, K), device="cuda", generator=g, dtype=dt) * 2 - 1
                    B = torch.rand((K, N), device="cuda", generator=g, dtype=dt) * 2 - 1
                    Cm = torch.empty((M, N), device="cuda", dtype=dt)
                    t = {0: [], 1: []}
                    picks = ((cfg, split, 0), (best["tile_cfg"], best["split_k"], best["group_m"]))
                    for rnd in range(4):
                        for j in ((0, 1) if rnd % 2 == 0 else (1, 0)):
                            c, sk, gm = picks[j]
                            t[j].extend(bench_gemm(cc, mode, A, B, Cm, 1.0, 0.0, reps, 1, True, tile_cfg=c,
                                                   split_k=sk, group_m=gm).tolist())
                    ms_mine, ms_best = float(np.median(t[0])), float(np.median(t[1]))
                    del A, B, Cm
                ratio = ms_mine / ms_best
                ratios.append(ratio)
                emit(dict(config="heuristic", set=name, mode=mode, M=M, N=N, K=K, tile_cfg=cfg, split_k=split,
                          ms=ms_mine, best_tile_cfg=best["tile_cfg"], best_split_k=best["split_k"],
                          best_group_m=best["group_m"], best_ms=ms_best, ratio=ratio, tuner_ratio=tuner_ratio))
            if ratios:
                emit(dict(config="heuristic_summary", set=name, mode=mode, shapes=len(ratios),
                          geomean=float(np.exp(np.mean(np.log(ratios)))), max=float(max(ratios)),
                          above_1_10=int(sum(r > 1.10 for r in ratios))))


CONFIGS = dict(plain=cfg_plain, square=cfg_square, flat=cfg_flat, dgemm=cfg_dgemm,
               grouped=cfg_grouped, grouped_tiles=cfg_grouped_tiles, autograd=cfg_autograd, epilogue=cfg_epilogue,
               schedules=cfg_schedules, heuristic=cfg_heuristic)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="cosma benchmark harness")
    ap.add_argument("--config", default="all", choices=["all", *CONFIGS])
    ap.add_argument("--runs", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-flush", action="store_true", help="skip the cache flush between runs")
    ap.add_argument("--csv-dir", help="write <M>x<N>x<K><name>.csv files (reference format)")
    ap.add_argument("--tune-db", help="apply this autotuner DB before running")
    ap.add_argument("--out", help="append JSON lines here")
    ap.add_argument("--no-search", action="store_true",
                    help="grouped: time only the default plan (skip the grouped autotuner)")
    ap.add_argument("--grouped-modes", help="grouped: comma-separated modes (default all four)")
    ap.add_argument("--modes", help="heuristic: comma-separated modes (default bf16x3,bf16x4)")
    args = ap.parse_args(argv)

    import torch

    import cucosma_amd as cc

    if args.tune_db:
        from .autotune import TuneDB

        TuneDB(args.tune_db).apply()
    out = open(args.out, "a") if args.out else None

    def emit(rec):
        line = json.dumps(rec)
        print(line, flush=True)
        if out:
            out.write(line + "\n")
            out.flush()

    names = list(CONFIGS) if args.config == "all" else [args.config]
    for name in names:
        print(f"[harness] {name}", file=sys.stderr, flush=True)
        CONFIGS[name](cc, torch, args, emit)
        torch.cuda.empty_cache()
    if out:
        out.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
