"""The C ABI without a GPU: the library loads, exports every function that
include/cosma/cosma.h declares, and rejects contract violations on the host
before any HIP call (so these run on a CPU-only machine)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cosma", "cosma.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cosma_[a-z0-9_x]+)\s*\(", text)) - {"cosma_sgemm_mnk_"})


def test_library_exports_every_declared_symbol():
    import cucosma_amd

    path = cucosma_amd._lib.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (cosma_\w+)", out))
    declared = set(_declared())
    # the cosmaSgemm drop-in macro expands to cosma_sgemm_mnk, declared above
    assert declared, "header parse found nothing"
    missing = declared - exported
    assert not missing, f"declared but not exported: {sorted(missing)}"
    for name in declared:
        getattr(cucosma_amd.lib, name)


def test_version_and_status_strings():
    import cucosma_amd

    assert cucosma_amd.__version__.count(".") == 2
    for st in (0, -1, -2, -3, -4):
        assert cucosma_amd.lib.cosma_status_string(st)


def test_tile_table_and_heuristic():
    import cucosma_amd

    for mode in range(3):
        cfgs = cucosma_amd.tile_configs(mode)
        assert cfgs
        for c in cfgs:
            assert c["BM"] % c["WM"] == 0 and c["BN"] % c["WN"] == 0
            assert c["threads"] == 64 * (c["BM"] // c["WM"]) * (c["BN"] // c["WN"])
            assert c["lds_bytes"] <= 160 * 1024
    for mode in range(4):
        cfg, split = cucosma_amd.heuristic(mode, 8192, 8192, 8192)
        t = cucosma_amd.tile_configs(mode)[cfg]
        assert split == 1 and t["BM"] * t["BN"] >= 128 * 128
        assert (8192 // t["BM"]) * (8192 // t["BN"]) >= 512
    # Deep K: the large tile, split until every CU holds one workgroup.
    for mode in (0, 2, 3):
        cfg, split = cucosma_amd.heuristic(mode, 256, 256, 1 << 20)
        t = cucosma_amd.tile_configs(mode)[cfg]
        assert (256 // t["BM"]) * (256 // t["BN"]) * split == 256
        assert (1 << 20) // split >= 512
    # An output smaller than the large tile keeps a 128-row tile.
    cfg, split = cucosma_amd.heuristic(0, 128, 128, 1 << 16)
    assert cucosma_amd.tile_configs(0)[cfg]["BM"] <= 128
    cfg, split = cucosma_amd.heuristic(0, 256, 256, 1 << 20)
    assert cucosma_amd.heuristic(1, 32768, 1024, 1024)[1] == 1
    assert cucosma_amd.workspace_bytes(256, 256, 1 << 20) == split * 256 * 256 * 4
    assert cucosma_amd.workspace_bytes(256, 256, 1 << 20, split_k=1) == 0


def test_heuristic_round_rules():
    """The split and tile rules that count rounds of workgroups (no GPU: the
    heuristic is host code; measurements behind each rule are in the
    library's comments)."""
    import cucosma_amd

    h = cucosma_amd.heuristic
    # FP64 medium outputs (fewer 128x128 tiles than four times the CUs): the
    # fitted cost model; each pick was the tuner's fastest or within 1.1%.
    assert h(1, 1280, 1280, 8192) == (7, 5)
    assert h(1, 1536, 1536, 16384) == (7, 7)
    assert h(1, 640, 640, 4096) == (8, 10)
    assert h(1, 384, 384, 512) == (10, 7)
    assert h(1, 3072, 3072, 2048) == (10, 1)
    assert h(1, 1024, 1024, 1024) == (8, 2)
    # FP64 large outputs: split in two when the last round of 512 is at most 1/4 full.
    assert h(1, 4096, 4096, 4096) == (7, 1)   # whole rounds
    assert h(1, 32768, 1024, 1024) == (7, 1)  # 2048 tiles: whole rounds
    # BF16: pre-split tile when ceil(tiles / 256) * 5504 >= M + N.
    for mode in (2, 3):
        assert h(mode, 4608, 4608, 4096) == (8, 1)
        assert h(mode, 4096, 12288, 4096) == (8, 1)
        assert h(mode, 4096, 4096, 4096) == (12, 1)
        # BF16x4's rounds constant is 6144: two rounds reach M + N = 11264.
        assert h(mode, 5632, 5632, 2048) == ((12, 1) if mode == 2 else (8, 1))
        assert h(mode, 4096, 8192, 4096) == ((12, 1) if mode == 2 else (8, 1))
        # One partly filled round (144 tiles): the pre-split tile from K = 512
        # (BF16x4: 1024) on, 128x128 tiles below (BF16x3, sides multiples of
        # 128: the 4-wave 128x128 tile).
        assert h(mode, 3072, 3072, 1024) == (8, 1)
        assert h(mode, 3072, 3072, 256) == ((0, 1) if mode == 2 else (11, 1))
        assert h(mode, 8192, 8192, 8192) == (8, 1)


def test_heuristic_fp32_cost_model():
    """FP32 outputs whose 256x128 tiles number fewer than twice the CUs (both
    sides >= 256, 512 <= K <= 32768) take the (tile, split) of least modelled
    time (host code, nominal 256 CUs); each pick below was the tuner's
    fastest configuration for the shape (cache flushed, one MI355X)."""
    import cucosma_amd

    h = cucosma_amd.heuristic
    assert h(0, 1536, 1536, 16384) == (7, 7)    # 72 tiles: two rounds at 98%, not split 4 (56%)
    assert h(0, 2304, 2304, 8192) == (7, 3)
    assert h(0, 3072, 3072, 16384) == (7, 8)    # 288 tiles: 9 rounds at 100%
    assert h(0, 1792, 1792, 2048) == (7, 5)
    assert h(0, 3072, 3072, 512) == (1, 1)      # short K: 64x64 tiles
    assert h(0, 1024, 1024, 1024) == (1, 3)     # the reference's first configuration (measured)
    assert h(0, 384, 384, 512)[0] == 1 and h(0, 384, 384, 512)[1] > 1   # K split even below K = 1024
    # Skinny products: 64x64 tiles (the rules before ran 256x128 split 2, 2.2-3x slower).
    assert h(0, 64, 16384, 16384) == (1, 4)
    assert h(0, 16, 16384, 4096) == (1, 4)
    assert h(1, 32, 16384, 16384) == (9, 4)
    # Outside the fitted range the rules stay: large outputs, the flat shape.
    assert h(0, 8192, 8192, 8192) == (7, 1)
    assert h(0, 256, 256, 1 << 20)[0] == 7


def test_heuristic_split_medium_rules():
    """BF16 outputs whose 256x256 tiles leave CUs idle (host code, nominal
    256 CUs): each case is a rule of split_medium, and each pick was the
    tuner's fastest configuration for the shape (cache flushed, one MI355X)."""
    import cucosma_amd

    h = cucosma_amd.heuristic
    # 256x256 split for one round of workgroups (64 tiles x 4; 49 x 5 with
    # 409-deep slices, 128x128 tiles filling 77% of their slots).
    assert h(2, 2048, 2048, 2048) == (12, 4)
    assert h(2, 1792, 1792, 2048) == (12, 5)
    assert h(2, 1024, 1024, 16384) == (12, 16)
    # ... or for two rounds when one leaves many CUs idle and K is deep.
    assert h(2, 2560, 2560, 8192) == (12, 2)
    assert h(2, 2560, 2560, 32768) == (12, 5)
    # ... over pre-split operands with edge tiles once M*N/(M+N) >= 768, and
    # in BF16x4 with two rounds once it reaches 1400; small edge outputs keep
    # the on-the-fly tile.
    assert h(2, 2432, 2432, 5120) == (8, 2)
    assert h(2, 1856, 1856, 10240) == (8, 4)
    assert h(3, 1856, 1856, 10240) == (8, 4)
    assert h(2, 2176, 2176, 6144) == (8, 3)
    assert h(3, 3072, 3072, 4096) == (8, 3)
    assert h(3, 3328, 3328, 8192) == (8, 3)
    assert h(2, 960, 960, 10240) == (12, 16)
    assert h(2, 1216, 1216, 10240) == (12, 10)
    # BF16x4 takes 128x128 tiles where they are as fast and fill better.
    assert h(3, 2048, 2048, 2048) == (11, 2)
    assert h(3, 2560, 2560, 2048) == (11, 1)
    assert h(3, 1024, 1024, 8192) == (11, 8)
    assert h(3, 3328, 3328, 1536) == (11, 1)
    # 256x256 tiles covering under 88% of the output: 128x128 tiles, K split
    # toward two workgroups per CU (BF16x3 with sides multiples of 128: the
    # 4-wave 128x128 tile at that split).
    assert h(2, 384, 384, 16384) == (0, 56)
    assert h(2, 1152, 1152, 3072) == (0, 6)
    assert h(2, 2560, 2560, 1024) == (0, 1)
    assert h(3, 640, 640, 16384) == (11, 20)
    assert h(3, 1408, 1408, 6144) == (11, 4)
    # One unsplit round filling half the CUs or more: the pre-split tile.
    assert h(2, 3500, 3500, 3000) == (8, 1)
    assert h(3, 2900, 2900, 1100) == (8, 1)
    # 128x128 tiles covering at most three quarters of the output: 64x64
    # tiles; a side below 128: K split toward four workgroups per CU, slices
    # at least 64 deep.
    assert h(2, 4096, 32, 16384) == (1, 16)
    assert h(2, 192, 16384, 1024) == (1, 1)
    assert h(3, 192, 16384, 1024) == (1, 1)
    assert h(2, 16, 16384, 16384) == (1, 4)
    assert h(3, 16, 16384, 16384) == (1, 4)
    assert h(2, 32, 1024, 1024) == (1, 16)
    # 64x64: K split toward three workgroups per CU, slices at least 128 deep.
    assert h(2, 384, 384, 512) == (1, 4)
    assert h(3, 1024, 1024, 1024) == (1, 3)


def test_heuristic_edge_tiles_presplit():
    """BF16 outputs with partly filled 256x256 tiles (a side not a multiple of
    256) take the pre-split tile once those tiles fill half the CUs, except
    BF16x4 outputs whose sides are multiples of 128 below K = 2048; aligned
    outputs keep the on-the-fly tiles.  Every pick here was the tuner's
    fastest configuration for the shape (cache flushed, one MI355X; host
    code, nominal 256 CUs)."""
    import cucosma_amd

    h = cucosma_amd.heuristic
    # One round of 144 tiles: the pre-split tile at the 256x256 split.
    assert h(2, 2900, 2900, 4096) == (8, 3)
    assert h(3, 2900, 2900, 4096) == (8, 3)
    assert h(3, 3700, 3700, 1536) == (8, 1)
    # Sides multiples of 128: BF16x4 keeps 128x128 tiles below K = 2048.
    assert h(3, 3200, 3200, 1536) == (11, 1)
    assert h(3, 3200, 3200, 4096) == (8, 3)
    # Aligned, BF16x3 from K = 1536: the spread-load tile.
    assert h(2, 3328, 3328, 1536) == (12, 1)
    assert h(2, 3072, 3072, 4096) == (12, 3)
    assert h(2, 3072, 3072, 1024) == (8, 1)
    # Full rounds: pre-split with edges from K = 512.
    assert h(2, 3900, 7900, 1024) == (8, 1)
    assert h(2, 3900, 3900, 1024) == (8, 1)
    assert h(2, 3900, 3900, 2048) == (8, 1)
    assert h(3, 4000, 4000, 512) == (8, 1)
    assert h(2, 4096, 4096, 2048) == (12, 1)


def test_transposed_a_takes_the_presplit_tile():
    """With A stored transposed (K >= 4096) the split modes take the pre-split
    tile, so the call needs workspace for the planes: (M + N) rows of 2 * K
    bf16.  Row-major A at one round keeps the on-the-fly tile (no workspace)."""
    import cucosma_amd

    for mode in (2, 3):
        for ta, k, want in ((1, 16384, (4096 + 4096) * 2 * 16384 * 2), (0, 16384, 0), (1, 1024, 0),
                            (1, 256, 0)):
            d = _desc(mode=mode, m=4096, n=4096, k=k, trans_a=ta, lda=4096 if ta else k, ldb=4096, ldc=4096)
            out = C.c_size_t()
            assert cucosma_amd.lib.cosma_workspace_query(C.byref(d), C.byref(out)) == 0
            assert out.value == want, (mode, ta, k, out.value)


def _desc(**kw):
    import cucosma_amd
    from cucosma_amd._lib import GemmDesc

    d = GemmDesc()
    cucosma_amd.lib.cosma_gemm_desc_init(C.byref(d))
    d.m = d.n = d.k = 64
    d.a, d.b, d.c = 0x1000, 0x2000, 0x3000
    d.lda = d.ldb = d.ldc = 64
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize(
    "kw,status",
    [
        (dict(m=-1), -1),
        (dict(k=-5), -1),
        (dict(mode=7), -1),
        (dict(c=None), -1),
        (dict(a=None), -1),
        (dict(lda=63), -1),
        (dict(ldb=10), -1),
        (dict(ldc=3), -1),
        (dict(trans_a=2), -1),
        (dict(split_k=5000), -1),
        (dict(tile_cfg=99), -1),
        (dict(epilogue=3), -1),
        (dict(epilogue=-1), -1),
        (dict(a_planes=0x5000), -1),  # planes with a non-split mode
        (dict(mode=2, b_planes=0x5008), -2),
        (dict(mode=2, a_planes=0x5000, tile_cfg=6), -1),  # planes need the pre-split tile
        (dict(a=0x1002), -2),
        (dict(b=0x2001), -2),
        (dict(mode=1, a=0x1004), -2),
        (dict(trans_a=1, trans_b=1), -3),
    ],
)
def test_contract_violations_rejected_on_host(kw, status):
    import cucosma_amd

    d = _desc(**kw)
    assert cucosma_amd.lib.cosma_gemm_ex(C.byref(d)) == status
    size = C.c_size_t()
    assert cucosma_amd.lib.cosma_workspace_query(C.byref(d), C.byref(size)) == status


def test_empty_problems_are_noops_on_host():
    import cucosma_amd

    for kw in (dict(m=0), dict(n=0), dict(m=0, c=None)):
        d = _desc(**kw)
        assert cucosma_amd.lib.cosma_gemm_ex(C.byref(d)) == 0
    # runtime-shape and fixed-shape entry points share the checks
    lib = cucosma_amd.lib
    assert lib.cosma_sgemm_mnk(-1, 4, 4, 1.0, 0x1000, 4, 0x1000, 4, 0.0, 0x1000, 4, None) == -1
    assert lib.cosma_sgemm_mnk(4, 4, 4, 1.0, 0x1000, 4, 0x1000, 4, 0.0, None, 4, None) == -1
    assert lib.cosma_sgemm_bf16x3_mnk(4, 4, 4, 1.0, 0x1001, 4, 0x1000, 4, 0.0, 0x1000, 4, None) == -2
    assert lib.cosma_dgemm_mnk(4, 4, 4, 1.0, 0x1000, 3, 0x1000, 4, 0.0, 0x1000, 4, None) == -1
    assert lib.cosma_sgemm_8192x8192x8192(1.0, 0x1000, 100, 0x1000, 8192, 0.0, 0x1000, 8192) == -1
    assert lib.cosma_dgemm_32768x1024x1024(1.0, 0x1004, 1024, 0x1000, 1024, 0.0, 0x1000, 1024) == -2
    m, n, k = C.c_int(), C.c_int(), C.c_int()
    lib.cosma_fixed_shape(C.byref(m), C.byref(n), C.byref(k))
    assert (m.value, n.value, k.value) == (1024, 1024, 1024)
    # split_operand: K rounded up to a 32-deep chunk must fit an int; empty is a no-op
    assert lib.cosma_split_operand(2, 1, 4, 0x7FFFFFF0, 0x1000, 0x7FFFFFF0, 0x2000, None) == -1
    assert lib.cosma_split_operand(2, 1, 0, 64, None, 64, None, None) == 0
    assert lib.cosma_split_operand(0, 1, 4, 64, 0x1000, 64, 0x2000, None) == -1  # not a split mode
    assert lib.cosma_error_stats_f32(-1, 2, None, 2, None, 2, (C.c_double * 5)(), None) == -1
    assert lib.cosma_convert_f32_f64(2, 2, None, 2, None, 2, None) == -1


def test_reference_macro_dropin_compiles():
    """A translation unit written against the reference call site,
    cosmaSgemm(A, lda, B, ldb, C, ldc) with config.h macros, compiles and
    links against the C ABI (host compiler only)."""
    import shutil
    import tempfile

    if shutil.which("g++") is None:
        pytest.skip("g++ missing")
    src = r'''
#define M 128
#define N 64
#define K 32
#define ALPHA 1
#define BETA 0.5
#define COSMA_REFERENCE_MACROS
#include "cosma/cosma.h"
int call(const float* A, const float* B, float* Cp) { return cosmaSgemm(A, K, B, N, Cp, N); }
'''
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "t.c")
        open(p, "w").write(src)
        subprocess.run(["gcc", "-c", "-I", os.path.join(ROOT, "include"), p, "-o",
                        os.path.join(td, "t.o")], check=True)


def test_heuristic_empty_outputs_return_at_once():
    """m == 0 or n == 0 returns a choice without looping or dividing by the
    empty tile count, for every mode and a deep K."""
    import cucosma_amd

    for mode in range(4):
        for m, n, k in ((0, 64, 64), (64, 0, 64), (0, 0, 0), (0, 64, 1 << 16), (64, 0, 1 << 20)):
            cfg, split = cucosma_amd.heuristic(mode, m, n, k)
            assert 0 <= cfg < cucosma_amd.lib.cosma_num_tile_configs(mode) and split == 1


def test_heuristic_follows_the_cu_count():
    """The heuristic takes the CU count as an input: on a 128-CU device (half
    an MI355X) the deep-K splits fill 128 CUs, large outputs keep the large
    tiles, and the round-counting rules count rounds of 128."""
    import cucosma_amd

    h = cucosma_amd.heuristic
    tc = cucosma_amd.tile_configs
    for cus in (256, 128, 64):
        for mode in (0, 2, 3):
            cfg, split = h(mode, 256, 256, 1 << 20, cus=cus)
            t = tc(mode)[cfg]
            assert (256 // t["BM"]) * (256 // t["BN"]) * split == cus, (mode, cus)
    # Nominal topology is the default without a device.
    assert h(2, 8192, 8192, 8192, cus=256) == h(2, 8192, 8192, 8192)
    assert h(2, 8192, 8192, 8192, cus=128) == (8, 1)
    # FP64 large outputs (at least four 128x128 tiles per CU): split in two
    # when the last round of 2 * cus is at most a quarter full.
    assert h(1, 3072, 3072, 3072, cus=128) == (7, 2)   # 576 tiles: last round 64 of 256
    assert h(1, 4608, 2304, 2048, cus=128) == (7, 1)   # 648 tiles: last round 136 of 256
    # The FP64 cost model counts rounds and per-CU loads of the given CUs: the
    # same shape gets another tile or split on fewer CUs.
    assert h(1, 2560, 2560, 2048, cus=256) == (10, 1)
    assert h(1, 2560, 2560, 2048, cus=128) == (3, 4)
    assert h(1, 1024, 1024, 1024, cus=256) == (8, 2)
    assert h(1, 1024, 1024, 1024, cus=64) == (3, 2)
    lib = cucosma_amd.lib
    out = (C.c_int(), C.c_int())
    assert lib.cosma_heuristic_topo(0, 64, 64, 64, 0, 0, C.byref(out[0]), C.byref(out[1])) == -1
    assert lib.cosma_heuristic_topo(0, 64, 64, 64, 2, 256, C.byref(out[0]), C.byref(out[1])) == -1


def test_device_topology_without_a_device_is_nominal():
    import cucosma_amd
    import torch

    t = cucosma_amd.device_topology()
    if torch.cuda.is_available():
        assert t["queried"] and t["cus"] > 0 and t["xcds"] >= 1
    else:
        assert not t["queried"] and (t["cus"], t["xcds"], t["clock_khz"]) == (256, 8, 2400000)
        assert abs(cucosma_amd.peak_tflops("bf16", t) - 2516.58) < 0.1
        assert abs(cucosma_amd.peak_tflops("fp32", t) - 157.29) < 0.01
        assert abs(cucosma_amd.peak_tflops("fp64", t) - 78.64) < 0.01
        assert abs(cucosma_amd.peak_tflops("bf16x3", t) - 838.86) < 0.01


def test_group_m_and_workspace_checks():
    import cucosma_amd

    lib = cucosma_amd.lib
    for kw, status in ((dict(group_m=70000), -1), (dict(workspace=0x5008, workspace_bytes=1 << 20), -2)):
        d = _desc(**kw)
        assert lib.cosma_gemm_ex(C.byref(d)) == status
    assert lib.cosma_workspace_set(0x5004, 4096) == -2
    assert lib.cosma_tune_set(0, 64, 64, 64, 0, 0, 0, 1, 1 << 20) == -1


def test_every_tile_configuration_in_one_group_family():
    """Each non-pre-split tile configuration belongs to exactly one grouped
    family; a family's configurations share one workgroup size; the first
    family of each mode holds its large tile and that tile's thin variants."""
    import cucosma_amd
    from cucosma_amd._lib import MODES

    for name, mode in MODES.items():
        fams = cucosma_amd.group_families(name)
        tiles = cucosma_amd.tile_configs(mode)
        seen = [c for f in fams for c in f]
        assert sorted(seen) == [t["cfg"] for t in tiles if t["pipe"] != 5], name
        for f in fams:
            assert len({tiles[c]["threads"] for c in f}) == 1, (name, f)
        big = tiles[fams[0][0]]
        assert len(fams[0]) >= 2
        for c in fams[0][1:]:
            assert tiles[c]["BM"] * tiles[c]["BN"] <= big["BM"] * big["BN"]
            assert (tiles[c]["BM"], tiles[c]["BN"]) != (big["BM"], big["BN"])


def test_grouped_plan_rejects_bad_problems_on_host():
    import cucosma_amd
    from cucosma_amd._lib import GroupProblem

    lib = cucosma_amd.lib
    plan = C.c_void_p()
    bad = (GroupProblem * 1)(GroupProblem(-1, 4, 4, 0x1000, 4, 0x1000, 4, 0x1000, 4, 1.0, 0.0))
    assert lib.cosma_grouped_plan_create(0, 1, bad, None, 0, C.byref(plan)) == -1
    bad = (GroupProblem * 1)(GroupProblem(4, 4, 4, 0x1000, 3, 0x1000, 4, 0x1000, 4, 1.0, 0.0))
    assert lib.cosma_grouped_plan_create(0, 1, bad, None, 0, C.byref(plan)) == -1
    ok = (GroupProblem * 1)(GroupProblem(4, 4, 4, 0x1000, 4, 0x1000, 4, 0x1000, 4, 1.0, 0.0))
    assert lib.cosma_grouped_plan_create(2, 1, ok, (C.c_int * 1)(8), 0, C.byref(plan)) == -3  # pre-split
    assert lib.cosma_grouped_plan_create_ex(0, 1, ok, None, (C.c_int * 1)(5000), 0, C.byref(plan)) == -1


def test_torch_wrappers_reject_host_tensors():
    import cucosma_amd
    import torch

    A = torch.zeros(8, 8)
    with pytest.raises(ValueError):
        cucosma_amd.gemm(A, A)
    with pytest.raises(ValueError):
        cucosma_amd.GroupedGemm([A], [A], [A])
    with pytest.raises(TypeError):
        cucosma_amd.GroupedGemm([A], [A], [A], mode="fp64")
