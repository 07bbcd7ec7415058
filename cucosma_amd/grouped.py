"""Grouped GEMM: many independent row-major problems in one persistent
launch per tile-configuration bucket, optionally replayed as one hipGraph.

No reference counterpart: the reference builds one binary per problem shape
(Kernel/cuda/run.sh:185-313).  The C side is cosma_grouped_plan_* in
include/cosma/cosma.h.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import MODES, GroupProblem, check, lib
from .gemm import _ld, _stream_ptr


class GroupedGemm:
    """C_i = alpha * A_i @ B_i + beta * C_i for every i, as one plan.

    The plan keeps the operand tensors alive; their storage must not move
    while the plan exists.  tile_cfgs: per-problem tile configuration (-1 =
    library choice: one configuration for all such problems, picked from the
    batch's flop-weighted output sizes), e.g. from the autotuner.
    """

    def __init__(self, As, Bs, Cs, *, mode: str = "fp32", alpha: float = 1.0, beta: float = 0.0,
                 tile_cfgs=None, blocks_per_cu: int = 0):
        if not (len(As) == len(Bs) == len(Cs)):
            raise ValueError("As, Bs, Cs must have the same length")
        self.mode = mode
        self._keep = (list(As), list(Bs), list(Cs))
        n = len(As)
        arr = (GroupProblem * max(n, 1))()
        for i, (A, B, Cm) in enumerate(zip(As, Bs, Cs)):
            m, k = A.shape
            k2, nn = B.shape
            if k != k2 or tuple(Cm.shape) != (m, nn):
                raise ValueError(f"problem {i}: shapes {tuple(A.shape)} {tuple(B.shape)} {tuple(Cm.shape)}")
            arr[i] = GroupProblem(m, nn, k, A.data_ptr(), _ld(A, "A"), B.data_ptr(), _ld(B, "B"),
                                  Cm.data_ptr(), _ld(Cm, "C"), float(alpha), float(beta))
        cfgs = None
        if tile_cfgs is not None:
            cfgs = (C.c_int * max(n, 1))(*[int(c) for c in tile_cfgs])
        self._plan = C.c_void_p()
        check(lib.cosma_grouped_plan_create(MODES[mode], n, arr, cfgs, int(blocks_per_cu),
                                            C.byref(self._plan)), "cosma_grouped_plan_create")

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


def grouped_gemm(As, Bs, Cs, **kw) -> None:
    """One-shot grouped GEMM (plan, run, destroy)."""
    g = GroupedGemm(As, Bs, Cs, **kw)
    g.run()
    torch.cuda.current_stream().synchronize()
    g.close()
