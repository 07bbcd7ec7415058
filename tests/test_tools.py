"""Autotuner, harness and grouped-workload helpers: host-side logic on the
CPU, measurement paths on the GPU."""
import json
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_reference_csv_roundtrip(tmp_path):
    from cucosma_amd import autotune

    rows = autotune.read_reference_csv(
        os.path.join(ROOT, "tests", "golden", "schedule_sample.csv"))
    assert rows[0]["m"] == 128 and rows[0]["split_k"] == 1 and rows[0]["ref_tile"] == (128, 128)
    assert any(r["split_k"] > 1 for r in rows)
    db = autotune.TuneDB(str(tmp_path / "db.json"))
    db.add(dict(mode="fp32", m=1024, n=1024, k=1024, trans_a=False, trans_b=False,
                candidates=[], best=dict(tile_cfg=1, split_k=2, group_m=8, occupancy=2)))
    db.add(dict(mode="bf16x3", m=256, n=256, k=1 << 20, trans_a=False, trans_b=False,
                candidates=[], best=dict(tile_cfg=1, split_k=64, group_m=8, occupancy=2)))
    db.save()
    out = tmp_path / "sched.csv"
    autotune.TuneDB(str(tmp_path / "db.json")).export_csv(str(out))
    back = autotune.read_reference_csv(str(out))
    assert {(r["m"], r["n"], r["k"], r["split_k"]) for r in back} == {
        (1024, 1024, 1024, 2), (256, 256, 1 << 20, 64)}
    header = open(out).readline().strip().split(",")
    assert header == autotune.CSV_COLUMNS and len(header) == 22


def test_tune_table_roundtrip_on_host():
    import ctypes as C

    import cucosma_amd as cc
    from cucosma_amd import autotune

    lib = cc.lib
    assert lib.cosma_tune_set(0, 4096, 4096, 4096, 0, 0, 99, 1, 8) == -1
    assert lib.cosma_tune_set(0, 4096, 4096, 4096, 0, 0, 2, 3, 16) == 0
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    assert lib.cosma_tune_lookup(0, 4096, 4096, 4096, 0, 0, C.byref(a), C.byref(b), C.byref(c)) == 0
    assert (a.value, b.value, c.value) == (2, 3, 16)
    # the tuned split is what a workspace query now reports for that shape
    assert cc.workspace_bytes(4096, 4096, 4096) == 3 * 4096 * 4096 * 4
    assert lib.cosma_tune_clear() == 0
    assert cc.workspace_bytes(4096, 4096, 4096) == 0
    db = autotune.TuneDB("/nonexistent/never-written.json")
    assert db.entries == {}


def test_grouped_sizes_deterministic():
    from cucosma_amd.harness import grouped_sizes

    s1, s2 = grouped_sizes(), grouped_sizes()
    assert s1 == s2 and len(s1) == 512
    arr = np.array(s1)
    assert arr.min() >= 32 and arr.max() <= 2048 and np.all(arr % 32 == 0)


@pytest.mark.gpu
def test_autotune_picks_a_valid_fastest(dev):
    from cucosma_amd import autotune

    rec = autotune.tune("fp32", 512, 384, 2048, warmup=1, reps=2)
    ok = [c for c in rec["candidates"] if c["status"] == 0]
    assert ok and rec["best"] is not None
    assert rec["best"]["ms"] == min(c["ms"] for c in ok) == rec["candidates"][0]["ms"]
    assert rec["flushed"] and all(0 < c["achieved_frac"] <= 1 for c in ok)
    assert all(c["occupancy"] >= 1 for c in ok)
    assert any(c["split_k"] > 1 for c in ok)


@pytest.mark.gpu
def test_benchmark_protocol_and_tuned_dispatch(dev):
    import torch

    import cucosma_amd as cc
    from cucosma_amd.harness import bench_gemm

    A = torch.rand(256, 512, device=dev)
    B = torch.rand(512, 128, device=dev)
    Cm = torch.empty(256, 128, device=dev)
    ms = bench_gemm(cc, "fp32", A, B, Cm, runs=5, warmup=2, flush=True)
    assert ms.shape == (5,) and np.all(ms > 0)
    ref = (A.double() @ B.double())
    assert torch.allclose(Cm.double(), ref, rtol=1e-5, atol=1e-4)
    # a tuned entry changes what the library launches, not what it computes
    assert cc.lib.cosma_tune_set(0, 256, 128, 512, 0, 0, 1, 4, 2) == 0
    C2 = cc.gemm(A, B)
    assert cc.lib.cosma_tune_clear() == 0
    assert torch.allclose(C2.double(), ref, rtol=1e-5, atol=1e-4)


@pytest.mark.gpu
def test_autotune_record_carries_mfma_busy_and_l2_hit(dev):
    """The tuned record's best candidate carries counters measured by
    rocprofv3: MFMA busy share and L2 hit rate, each in (0, 1]."""
    import shutil

    from cucosma_amd import autotune

    if shutil.which("rocprofv3") is None:
        pytest.skip("rocprofv3 missing")
    rec = autotune.tune("bf16x3", 2048, 2048, 2048, warmup=1, reps=2, with_counters=True)
    b = rec["best"]
    assert 0.0 < b["mfma_busy"] <= 1.0, b["counters"]
    assert 0.0 < b["l2_hit"] <= 1.0, b["counters"]
    assert b["counters"]["fetch_kib"] > 0
