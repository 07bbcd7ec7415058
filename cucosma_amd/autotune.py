"""Measured autotuning of tile configuration, split-K factor and XCD group size.

Replaces the reference's ScheduleGenerator (ScheduleGenerator/ScheduleGeneratorv4.h:13-391,
main.cpp:22-342), which ranks schedules by a communication-volume model.  The
reference's own thesis reports that volume is the wrong metric and that
occupancy and L2 behaviour decide (thesis p41-42, p44), so this tuner
measures instead:

* ``tune()`` runs the C++ tuner (``cosma_autotune_ex``): every built tile
  configuration x split-K factor, then XCD group sizes for the winner, each
  timed call after a 512 MiB cache flush (the benchmark protocol) with HIP
  events; each candidate carries its occupancy (workgroups per CU, waves per
  SIMD) and achieved fraction of the mode's dense MFMA peak.
* ``counters()`` re-runs a candidate under ``rocprofv3 --pmc`` and measures
  its MFMA busy share (SQ_VALU_MFMA_BUSY_CYCLES over busy SIMD cycles), L2
  hit rate (TCC_HIT / (TCC_HIT + TCC_MISS)) and HBM bytes (FETCH_SIZE,
  WRITE_SIZE); ``tune(with_counters=True)`` attaches them to the best
  candidate's record.
* ``cucosma_amd.grouped_autotune`` (``cosma_grouped_autotune``) picks tile
  configuration and split-K per size bucket of a grouped batch.
* ``TuneDB`` persists results as JSON, installs them into the library
  (``cosma_tune_set``) so every C-ABI call of that shape uses them, and
  imports / exports the reference's 22-column schedule CSV
  (ScheduleGenerator/main.cpp:227-249, read by Kernel/cuda/run.sh:189).

CLI: ``python -m cucosma_amd.autotune --mode bf16x3 --shape 8192 8192 8192``.
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile

from ._lib import MODES, GemmDesc, check, lib, tile_configs

MODE_NAMES = {v: k for k, v in MODES.items()}

CSV_COLUMNS = [
    "OCCUPANCY", "TYPE", "VECTORTYPE2", "VECTORTYPE4", "M", "N", "K", "THREADBLOCK_TILE_M",
    "THREADBLOCK_TILE_N", "THREADBLOCK_TILE_K", "LOAD_K", "WARP_TILE_M", "WARP_TILE_N",
    "THREAD_TILE_M", "THREAD_TILE_N", "SPLIT_K", "ALPHA", "BETA", "SWIZZLE", "A_OFFSET",
    "B_OFFSET", "ATOMIC_REDUCTION",
]


class TuneResult(C.Structure):
    """Mirror of struct cosma_tune_result."""

    _fields_ = [
        ("tile_cfg", C.c_int), ("split_k", C.c_int), ("group_m", C.c_int), ("status", C.c_int),
        ("ms", C.c_double), ("tflops", C.c_double), ("achieved_frac", C.c_double),
        ("blocks", C.c_int), ("occupancy", C.c_int), ("waves_per_simd", C.c_int),
        ("workspace_bytes", C.c_size_t),
    ]


lib.cosma_autotune.restype = C.c_int
lib.cosma_autotune.argtypes = [C.POINTER(GemmDesc), C.c_int, C.c_int, C.POINTER(TuneResult),
                               C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
lib.cosma_autotune_ex.restype = C.c_int
lib.cosma_autotune_ex.argtypes = [C.POINTER(GemmDesc), C.c_int, C.c_int, C.c_size_t, C.POINTER(TuneResult),
                                  C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
lib.cosma_tune_set.restype = C.c_int
lib.cosma_tune_set.argtypes = [C.c_int] * 9
lib.cosma_tune_lookup.restype = C.c_int
lib.cosma_tune_lookup.argtypes = [C.c_int] * 6 + [C.POINTER(C.c_int)] * 3
lib.cosma_tune_clear.restype = C.c_int
lib.cosma_tune_clear.argtypes = []


def _key(mode: str, m: int, n: int, k: int, ta: bool, tb: bool) -> str:
    return f"{mode}:{m}x{n}x{k}:{int(ta)}{int(tb)}"


def tune(mode: str, m: int, n: int, k: int, *, trans_a: bool = False, trans_b: bool = False,
         warmup: int = 2, reps: int = 5, alpha: float = 1.0, beta: float = 0.0,
         tile_cfg: int = -1, split_k: int = 0, group_m: int = 0, flush: bool = True,
         with_counters: bool = False) -> dict:
    """Time every candidate for one shape on random operands (caches flushed
    before each timed call unless flush is False); return the record, the
    fastest first.  with_counters adds the best candidate's measured MFMA busy
    share, L2 hit rate and HBM bytes (counters(), rocprofv3)."""
    import torch

    from .gemm import _ld, _stream_ptr

    dt = torch.float64 if mode == "fp64" else torch.float32
    dev = torch.device("cuda", torch.cuda.current_device())
    A = torch.rand((k, m) if trans_a else (m, k), dtype=dt, device=dev) * 2 - 1
    B = torch.rand((n, k) if trans_b else (k, n), dtype=dt, device=dev) * 2 - 1
    Cm = torch.zeros((m, n), dtype=dt, device=dev)
    d = GemmDesc()
    lib.cosma_gemm_desc_init(C.byref(d))
    d.mode, d.m, d.n, d.k = MODES[mode], m, n, k
    d.trans_a, d.trans_b = int(trans_a), int(trans_b)
    d.a, d.lda = A.data_ptr(), _ld(A, "A")
    d.b, d.ldb = B.data_ptr(), _ld(B, "B")
    d.c, d.ldc = Cm.data_ptr(), _ld(Cm, "C")
    d.alpha, d.beta = alpha, beta
    d.tile_cfg, d.split_k, d.group_m = tile_cfg, split_k, group_m
    d.stream = _stream_ptr(None)
    cap = 512
    res = (TuneResult * cap)()
    nres, best = C.c_int(), C.c_int()
    check(lib.cosma_autotune_ex(C.byref(d), warmup, reps, (512 << 20) if flush else 0, res, cap,
                                C.byref(nres), C.byref(best)), "cosma_autotune_ex")
    tiles = tile_configs(MODES[mode])
    cands = []
    for i in range(nres.value):
        r = res[i]
        t = tiles[r.tile_cfg]
        cands.append(dict(tile_cfg=r.tile_cfg, tile=f"{t['BM']}x{t['BN']}x{t['BK']}/{t['WM']}x{t['WN']}",
                          split_k=r.split_k, group_m=r.group_m, status=r.status,
                          ms=r.ms, tflops=r.tflops, achieved_frac=r.achieved_frac, blocks=r.blocks,
                          occupancy=r.occupancy, waves_per_simd=r.waves_per_simd,
                          workspace_bytes=int(r.workspace_bytes)))
    rec = dict(mode=mode, m=m, n=n, k=k, trans_a=bool(trans_a), trans_b=bool(trans_b), flushed=bool(flush),
               candidates=cands, best=cands[best.value] if best.value >= 0 else None)
    if with_counters and rec["best"]:
        b = rec["best"]
        cnt = counters(mode, m, n, k, b["tile_cfg"], b["split_k"], b["group_m"], trans_a, trans_b)
        b["counters"] = cnt
        b["mfma_busy"] = cnt.get("mfma_busy")
        b["l2_hit"] = cnt.get("l2_hit")
    return rec


# Counter passes of counters(): each within one pass's hardware limits (TCC:
# 4 counters, FETCH_SIZE takes 3 and WRITE_SIZE 2; SQ 8; GRBM 2).
COUNTER_PASSES = (["TCC_HIT_sum", "TCC_MISS_sum"], ["FETCH_SIZE"], ["WRITE_SIZE"],
                  ["SQ_VALU_MFMA_BUSY_CYCLES", "GRBM_GUI_ACTIVE"])


def counters(mode: str, m: int, n: int, k: int, tile_cfg: int, split_k: int, group_m: int = 8,
             trans_a: bool = False, trans_b: bool = False, reps: int = 2, timeout: float = 120.0) -> dict:
    """Hardware counters of one candidate's main kernel (gemm_kernel or
    planar_kernel), measured by rocprofv3 in a child process, one pass per
    counter group, averaged over the candidate's launches:

    * l2_hit: TCC_HIT / (TCC_HIT + TCC_MISS);
    * fetch_kib / write_kib: FETCH_SIZE / WRITE_SIZE per launch (KiB);
    * mfma_busy: SQ_VALU_MFMA_BUSY_CYCLES / (cycles the GPU was busy x
      SIMDs), with GRBM_GUI_ACTIVE summed over the XCDs (per-XCD cycles =
      GRBM_GUI_ACTIVE / XCDs) and 4 SIMDs per CU -- the share of SIMD cycles
      the matrix cores were busy, whatever the clock.
    Needs rocprofv3 and a GPU."""
    from ._lib import device_topology

    cmd_run = [sys.executable, "-m", "cucosma_amd.autotune", "--run-one", "--mode", mode,
               "--shape", str(m), str(n), str(k), "--tile-cfg", str(tile_cfg),
               "--split-k", str(split_k), "--group-m", str(group_m), "--reps", str(reps)]
    if trans_a:
        cmd_run.append("--trans-a")
    if trans_b:
        cmd_run.append("--trans-b")
    topo = device_topology()
    raw: dict[str, list[float]] = {}
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    cwd = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for pmc in COUNTER_PASSES:
        with tempfile.TemporaryDirectory() as td:
            cmd = ["rocprofv3", "--pmc", *pmc, "--kernel-trace", "--output-format", "csv", "-d", td,
                   "-o", "run", "--", *cmd_run]
            subprocess.run(cmd, check=True, capture_output=True, timeout=timeout, env=env, cwd=cwd)
            per: dict[tuple, dict] = {}
            for root, _, files in os.walk(td):
                for f in files:
                    if not f.endswith("counter_collection.csv"):
                        continue
                    for row in csv.DictReader(open(os.path.join(root, f))):
                        name = row["Kernel_Name"]
                        if "gemm_kernel" not in name and "planar_kernel" not in name:
                            continue
                        key = (row.get("Dispatch_Id", ""), name)
                        per.setdefault(key, {})[row["Counter_Name"]] = float(row["Counter_Value"])
            for vals in per.values():
                for c, v in vals.items():
                    raw.setdefault(c, []).append(v)
                if "SQ_VALU_MFMA_BUSY_CYCLES" in vals and vals.get("GRBM_GUI_ACTIVE", 0) > 0:
                    simd_cycles = vals["GRBM_GUI_ACTIVE"] / max(topo["xcds"], 1) * topo["cus"] * 4
                    raw.setdefault("mfma_busy", []).append(vals["SQ_VALU_MFMA_BUSY_CYCLES"] / simd_cycles)
    out = {c: sum(v) / len(v) for c, v in raw.items() if v}
    if "TCC_HIT_sum" in out and "TCC_MISS_sum" in out:
        tot = out["TCC_HIT_sum"] + out["TCC_MISS_sum"]
        out["l2_hit"] = out["TCC_HIT_sum"] / tot if tot else None
    if "FETCH_SIZE" in out:
        out["fetch_kib"] = out["FETCH_SIZE"]
    if "WRITE_SIZE" in out:
        out["write_kib"] = out["WRITE_SIZE"]
    return out


class TuneDB:
    """Persistent tuning results: {key: record with "best"}."""

    def __init__(self, path: str | None = None):
        self.path = path or os.environ.get(
            "COSMA_TUNE_DB", os.path.join(os.path.expanduser("~"), ".cache", "cucosma_amd", "tune.json"))
        self.entries: dict[str, dict] = {}
        if os.path.exists(self.path):
            with open(self.path) as f:
                self.entries = json.load(f)

    def add(self, rec: dict) -> None:
        self.entries[_key(rec["mode"], rec["m"], rec["n"], rec["k"], rec["trans_a"],
                          rec["trans_b"])] = rec

    def get(self, mode, m, n, k, trans_a=False, trans_b=False) -> dict | None:
        return self.entries.get(_key(mode, m, n, k, trans_a, trans_b))

    def save(self) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.entries, f, indent=1)
        os.replace(tmp, self.path)

    def apply(self) -> int:
        """Install every best choice into the library; return how many."""
        n = 0
        for rec in self.entries.values():
            b = rec.get("best")
            if not b:
                continue
            check(lib.cosma_tune_set(MODES[rec["mode"]], rec["m"], rec["n"], rec["k"],
                                     int(rec["trans_a"]), int(rec["trans_b"]), b["tile_cfg"],
                                     b["split_k"], b.get("group_m", 0)), "cosma_tune_set")
            n += 1
        return n

    def export_csv(self, path: str) -> None:
        """Write the best choices in the reference's schedule CSV schema."""
        with open(path, "w", newline="") as f:
            w = csv.writer(f)
            w.writerow(CSV_COLUMNS)
            for rec in self.entries.values():
                b = rec.get("best")
                if not b or rec["trans_a"] or rec["trans_b"]:
                    continue
                mode = rec["mode"]
                t = tile_configs(MODES[mode])[b["tile_cfg"]]
                elem = "double" if mode == "fp64" else "float"
                per = -(-rec["k"] // b["split_k"])
                kslice = -(-per // t["BK"]) * t["BK"]
                pad = 2 if mode == "fp64" else (8 if mode.startswith("bf16") else 4)
                w.writerow([b.get("occupancy", 1), elem, elem + "2", elem + "4", rec["m"], rec["n"],
                            rec["k"], t["BM"], t["BN"], kslice, t["BK"], t["WM"], t["WN"], 16, 16,
                            b["split_k"], rec.get("alpha", 1), rec.get("beta", 0), b["group_m"],
                            pad, pad, "false"])


def read_reference_csv(path: str) -> list[dict]:
    """Rows of a reference schedule CSV (Benchmarks/TESTS_TO_RUN/*.csv or
    ScheduleGenerator output) as workload dicts: shape, alpha, beta, the
    requested split-K, element type and the reference's tile sizes."""
    rows = []
    with open(path, newline="") as f:
        for r in csv.reader(f):
            if not r or r[0].strip() in ("", "OCCUPANCY", "ADDITIONAL_OCCUPANCY_SM"):
                continue
            vals = dict(zip(CSV_COLUMNS, [x.strip() for x in r]))
            if not vals.get("M", "").isdigit():
                continue
            rows.append(dict(type=vals["TYPE"], m=int(vals["M"]), n=int(vals["N"]), k=int(vals["K"]),
                             split_k=int(vals["SPLIT_K"]), alpha=float(vals["ALPHA"]),
                             beta=float(vals["BETA"]), swizzle=int(vals["SWIZZLE"]),
                             ref_tile=(int(vals["THREADBLOCK_TILE_M"]), int(vals["THREADBLOCK_TILE_N"])),
                             atomic=vals["ATOMIC_REDUCTION"].lower() == "true"))
    return rows


def _run_one(args) -> int:
    """Child of counters(): run one candidate a few times."""
    import torch

    from .gemm import gemm

    m, n, k = args.shape
    dt = torch.float64 if args.mode == "fp64" else torch.float32
    A = torch.rand((k, m) if args.trans_a else (m, k), dtype=dt, device="cuda")
    B = torch.rand((n, k) if args.trans_b else (k, n), dtype=dt, device="cuda")
    Cm = torch.empty((m, n), dtype=dt, device="cuda")
    for _ in range(args.reps):
        gemm(A, B, Cm, mode=args.mode, trans_a=args.trans_a, trans_b=args.trans_b,
             tile_cfg=args.tile_cfg, split_k=args.split_k, group_m=args.group_m)
    torch.cuda.synchronize()
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--mode", default="fp32", choices=sorted(MODES))
    ap.add_argument("--shape", type=int, nargs=3, action="append", metavar=("M", "N", "K"))
    ap.add_argument("--csv", help="tune every row of a reference schedule CSV")
    ap.add_argument("--trans-a", action="store_true")
    ap.add_argument("--trans-b", action="store_true")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--counters", action="store_true", help="add rocprofv3 L2/HBM counters for the best")
    ap.add_argument("--db", help="tuning DB path (JSON)")
    ap.add_argument("--export-csv", help="write the DB's choices in the reference CSV schema")
    ap.add_argument("--run-one", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--tile-cfg", type=int, default=-1)
    ap.add_argument("--split-k", type=int, default=0)
    ap.add_argument("--group-m", type=int, default=0)
    args = ap.parse_args(argv)
    if args.run_one:
        args.shape = args.shape[0]
        return _run_one(args)
    db = TuneDB(args.db)
    shapes = [tuple(s) for s in (args.shape or [])]
    if args.csv:
        shapes += [(r["m"], r["n"], r["k"]) for r in read_reference_csv(args.csv)]
    for (m, n, k) in shapes:
        rec = tune(args.mode, m, n, k, trans_a=args.trans_a, trans_b=args.trans_b,
                   warmup=args.warmup, reps=args.reps)
        if args.counters and rec["best"]:
            b = rec["best"]
            b["counters"] = counters(args.mode, m, n, k, b["tile_cfg"], b["split_k"], b["group_m"],
                                     args.trans_a, args.trans_b)
            b["mfma_busy"] = b["counters"].get("mfma_busy")
            b["l2_hit"] = b["counters"].get("l2_hit")
        db.add(rec)
        b = rec["best"]
        print(json.dumps(dict(mode=args.mode, m=m, n=n, k=k, best=b,
                              n_candidates=len(rec["candidates"]))), flush=True)
    if shapes:
        db.save()
    if args.export_csv:
        db.export_csv(args.export_csv)
    return 0


if __name__ == "__main__":
    sys.exit(main())
