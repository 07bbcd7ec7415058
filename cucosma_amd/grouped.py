"""Grouped GEMM: many independent row-major problems in one persistent
launch per tile-configuration bucket, optionally replayed as one hipGraph.

No reference counterpart: the reference builds one binary per problem shape
(Kernel/cuda/run.sh:185-313).  The C side is cosma_grouped_plan_* in
include/cosma/cosma.h.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import MODES, GroupProblem, GroupTuneResult, check, lib
from .gemm import _DTYPE, _check_operand, _ld, _stream_ptr


class GroupedGemm:
    """C_i = alpha * A_i @ B_i + beta * C_i for every i, as one plan.

    The plan keeps the operand tensors alive; their storage must not move
    while the plan exists.  tile_cfgs / splits: per-problem tile
    configuration and split-K factor (-1 / 0 = library choice), e.g. from
    grouped_autotune().
    """

    def __init__(self, As, Bs, Cs, *, mode: str = "fp32", alpha: float = 1.0, beta: float = 0.0,
                 tile_cfgs=None, splits=None, blocks_per_cu: int = 0):
        if not (len(As) == len(Bs) == len(Cs)):
            raise ValueError("As, Bs, Cs must have the same length")
        if mode not in MODES:
            raise ValueError(f"unknown mode {mode!r}")
        dt = _DTYPE[mode]
        self.mode = mode
        self._keep = (list(As), list(Bs), list(Cs))
        n = len(As)
        arr = (GroupProblem * max(n, 1))()
        dev = As[0].device if n else None
        for i, (A, B, Cm) in enumerate(zip(As, Bs, Cs)):
            for t, name in ((A, "A"), (B, "B"), (Cm, "C")):
                _check_operand(t, f"{name}[{i}]", dt, mode, dev)
            m, k = A.shape
            k2, nn = B.shape
            if k != k2 or tuple(Cm.shape) != (m, nn):
                raise ValueError(f"problem {i}: shapes {tuple(A.shape)} {tuple(B.shape)} {tuple(Cm.shape)}")
            arr[i] = GroupProblem(m, nn, k, A.data_ptr(), _ld(A, "A"), B.data_ptr(), _ld(B, "B"),
                                  Cm.data_ptr(), _ld(Cm, "C"), float(alpha), float(beta))
        self._arr, self._n = arr, n
        cfgs = None
        if tile_cfgs is not None:
            if len(tile_cfgs) != n:
                raise ValueError("tile_cfgs needs one entry per problem")
            cfgs = (C.c_int * max(n, 1))(*[int(c) for c in tile_cfgs])
        spl = None
        if splits is not None:
            if len(splits) != n:
                raise ValueError("splits needs one entry per problem")
            spl = (C.c_int * max(n, 1))(*[int(c) for c in splits])
        self._plan = C.c_void_p()
        check(lib.cosma_grouped_plan_create_ex(MODES[mode], n, arr, cfgs, spl, int(blocks_per_cu),
                                               C.byref(self._plan)), "cosma_grouped_plan_create_ex")

    def choices(self) -> list[tuple[int, int]]:
        """(tile_cfg, split_k) of every problem, in input order."""
        out = []
        c, s = C.c_int(), C.c_int()
        for i in range(self._n):
            check(lib.cosma_grouped_plan_choice(self._plan, i, C.byref(c), C.byref(s)), "choice")
            out.append((c.value, s.value))
        return out

    def run(self, stream=None) -> None:
        """One persistent launch per bucket on the stream."""
        check(lib.cosma_grouped_plan_run(self._plan, _stream_ptr(stream)), "grouped run")

    def run_graph(self, stream=None) -> None:
        """Replay the plan's hipGraph (captured on first use)."""
        check(lib.cosma_grouped_plan_run_graph(self._plan, _stream_ptr(stream)), "grouped graph")

    def info(self) -> dict:
        b, t, f = C.c_int(), C.c_int(), C.c_double()
        check(lib.cosma_grouped_plan_info(self._plan, C.byref(b), C.byref(t), C.byref(f)), "info")
        buckets = []
        for i in range(b.value):
            cfg, cnt, tiles, blocks = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            check(lib.cosma_grouped_plan_bucket(self._plan, i, C.byref(cfg), C.byref(cnt),
                                                C.byref(tiles), C.byref(blocks)), "bucket")
            buckets.append(dict(tile_cfg=cfg.value, problems=cnt.value, tiles=tiles.value,
                                blocks=blocks.value))
        return dict(buckets=buckets, total_tiles=t.value, flops=f.value)

    def close(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            lib.cosma_grouped_plan_destroy(self._plan)
        self._plan = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def group_families(mode: str) -> list[list[int]]:
    """Tile families of a mode: each list is run side by side by one
    persistent kernel (one configuration per problem)."""
    arr = (C.c_int * 4)()
    nf = lib.cosma_group_families(MODES[mode], -1, arr)
    out = []
    for f in range(nf):
        n = lib.cosma_group_families(MODES[mode], f, arr)
        out.append(list(arr[:n]))
    return out


def grouped_autotune(As, Bs, Cs, *, mode: str = "fp32", alpha: float = 1.0, beta: float = 0.0,
                     warmup: int = 3, reps: int = 10) -> dict:
    """Measured per-size-bucket choice of tile configuration and split-K for
    a grouped batch (cosma_grouped_autotune): returns the per-problem
    tile_cfgs / splits for GroupedGemm and every timed candidate.  Runs the
    batch on its buffers (C is overwritten)."""
    g = GroupedGemm(As, Bs, Cs, mode=mode, alpha=alpha, beta=beta)
    n = g._n
    cfgs = (C.c_int * max(n, 1))()
    spl = (C.c_int * max(n, 1))()
    cap = 1024
    res = (GroupTuneResult * cap)()
    nres = C.c_int()
    check(lib.cosma_grouped_autotune(MODES[mode], n, g._arr, warmup, reps, cfgs, spl, res, cap,
                                     C.byref(nres)), "cosma_grouped_autotune")
    g.close()
    tried = [dict(bucket=r.bucket, tile_cfg=r.tile_cfg, split_k=r.split_k, status=r.status, ms=r.ms)
             for r in res[:nres.value]]
    return dict(tile_cfgs=list(cfgs[:n]), splits=list(spl[:n]), tried=tried,
                ms_start=tried[0]["ms"] if tried else None,
                ms_best=min((t["ms"] for t in tried if t["status"] == 0), default=None))


def grouped_gemm(As, Bs, Cs, **kw) -> None:
    """One-shot grouped GEMM (plan, run, destroy)."""
    g = GroupedGemm(As, Bs, Cs, **kw)
    g.run()
    torch.cuda.current_stream().synchronize()
    g.close()
