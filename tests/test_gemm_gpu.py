"""GPU parity of the HIP GEMM kernels against the CPU oracle.

Bars:
  * small-integer operands: every mode / layout / tile / split / edge case is
    bitwise equal to the exact integer product (sums stay below 2^24);
  * FP32 and BF16x3 on random data: componentwise |C - C*| <= c*K*u*(|A||B|)
    with u = 2^-24 (FP32 accumulation) plus, for BF16x3, 2^-15 |A||B| for the
    split; BF16x3 is also held to the exact three-product oracle;
  * config 1 (1024^3, alpha 1, beta 0.5): normwise error vs the FP64 oracle
    <= 1e-6, and agreement with the reference kernel's own arithmetic order
    (oracle.refsim_sgemm) within 2*K*u*|A||B|;
  * FP64: componentwise within 4*K*2^-53*(|A||B|) of the long-double oracle;
  * split-K: bitwise repeatable.
"""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def cc(dev):
    import cucosma_amd

    return cucosma_amd


def _t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def _int_operands(rng, M, N, K, dt, lo=-8, hi=9):
    A = rng.integers(lo, hi, (M, K)).astype(dt)
    B = rng.integers(lo, hi, (K, N)).astype(dt)
    C0 = rng.integers(lo, hi, (M, N)).astype(dt)
    return A, B, C0


MODES = [("fp32", np.float32), ("fp64", np.float64), ("bf16x3", np.float32), ("bf16x4", np.float32)]
LAYOUTS = [(False, False), (False, True), (True, False)]


@pytest.mark.parametrize("mode,dt", MODES)
@pytest.mark.parametrize("ta,tb", LAYOUTS)
def test_integer_exact_all_tiles(cc, dev, mode, dt, ta, tb):
    """Exact integer products: checks MFMA lane layouts, loaders (both
    layouts), edges, split-K slices and the epilogue for every tile config."""
    rng = np.random.default_rng(100)
    ncfg = len(cc.tile_configs(cc._lib.MODES[mode]))
    for (M, N, K) in ((256, 256, 64), (200, 136, 52)):
        A, B, C0 = _int_operands(rng, M, N, K, dt)
        want = (A.astype(np.int64) @ B.astype(np.int64)) * 2 - C0.astype(np.int64)
        Ad = _t(A.T if ta else A, dev)
        Bd = _t(B.T if tb else B, dev)
        for cfg in range(ncfg):
            for split in (1, 3):
                Cd = _t(C0, dev)
                cc.gemm(Ad, Bd, Cd, alpha=2.0, beta=-1.0, mode=mode, trans_a=ta, trans_b=tb,
                        tile_cfg=cfg, split_k=split)
                got = Cd.cpu().numpy()
                assert np.array_equal(got, want.astype(dt)), (M, N, K, cfg, split)


@pytest.mark.parametrize("mode,dt", MODES)
def test_padded_leading_dims_and_ragged_k(cc, dev, mode, dt):
    """K not a multiple of the vector width, rows padded (ld > cols)."""
    rng = np.random.default_rng(101)
    M, N, K = 70, 45, 50
    pad = 4 if dt == np.float32 else 2
    A, B, C0 = _int_operands(rng, M, N, K, dt)
    Abig = np.zeros((M, K + 2 + pad), dt)
    Abig[:, :K] = A
    Bbig = np.zeros((K, N + 3 + pad), dt)
    Bbig[:, :N] = B
    Ad = _t(Abig, dev)[:, :K]
    Bd = _t(Bbig, dev)[:, :N]
    assert Ad.stride(0) % pad == 0 and Bd.stride(0) % pad == 0
    Cd = _t(C0, dev)
    cc.gemm(Ad, Bd, Cd, alpha=1.0, beta=1.0, mode=mode)
    want = A.astype(np.int64) @ B.astype(np.int64) + C0.astype(np.int64)
    assert np.array_equal(Cd.cpu().numpy(), want.astype(dt))


def _bound_check(got, ref, A, B, slack, u, extra=0.0):
    absprod = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    K = A.shape[1]
    lim = (slack * K * u + extra) * absprod + 1e-30
    err = np.abs(got.astype(np.float64) - ref)
    assert np.all(err <= lim), float((err / lim).max())


@pytest.mark.parametrize("ta,tb", LAYOUTS)
@pytest.mark.parametrize("shape", [(64, 64, 64), (333, 129, 517), (1, 300, 2048), (257, 3, 4096)])
def test_fp32_random_vs_f64(cc, dev, ta, tb, shape):
    M, N, K = shape
    A, B = oracle.workload_square_flat(M, N, K, seed=1)
    ref = oracle.gemm_f64(A, B)
    Cd = cc.gemm(_t(A.T if ta else A, dev), _t(B.T if tb else B, dev), mode="fp32",
                 trans_a=ta, trans_b=tb)
    _bound_check(Cd.cpu().numpy(), ref, A, B, 2.0, 2.0**-24)


@pytest.mark.parametrize("mode", ["bf16x3", "bf16x4"])
def test_split_modes_vs_split_oracle_and_f64(cc, dev, mode):
    M, N, K = 300, 200, 1000
    A, B = oracle.workload_square_flat(M, N, K, seed=1)
    got = cc.gemm(_t(A, dev), _t(B, dev), mode=mode).cpu().numpy()
    exact = oracle.gemm_bf16x3(A, B) if mode == "bf16x3" else oracle.gemm_bf16x4(A, B)
    _bound_check(got, exact, A, B, 2.0, 2.0**-24)  # only FP32 accumulation error remains
    ref = oracle.gemm_f64(A, B)
    _bound_check(got, ref, A, B, 2.0, 2.0**-24, extra=2.0**-15)


def test_config1_plain_sgemm_parity(cc, dev):
    """1024^3, alpha 1, beta 0.5, A/B/C uniform[-1,1] from one MT19937(0)
    stream, through the compile-time specialisation and the fixed-shape
    entry point."""
    A, B, C0 = oracle.workload_plain_1024()
    ref = oracle.gemm_f64(A, B, C0, 1.0, 0.5)
    refsim = oracle.refsim_sgemm(A, B, C0, 1.0, 0.5)  # the reference kernel's arithmetic
    Ad, Bd = _t(A, dev), _t(B, dev)
    for fn in (cc.lib.cosma_sgemm_1024x1024x1024, cc.lib.cosma_sgemm):
        Cd = _t(C0, dev)
        torch.cuda.synchronize()
        st = fn(1.0, Ad.data_ptr(), 1024, Bd.data_ptr(), 1024, 0.5, Cd.data_ptr(), 1024)
        assert st == 0
        torch.cuda.synchronize()
        got = Cd.cpu().numpy()
        nw = oracle.normwise(got, ref, A, B)
        assert nw <= 1e-6, nw
        absprod = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
        diff = np.abs(got.astype(np.float64) - refsim)
        assert np.all(diff <= 2 * 1024 * 2.0**-24 * absprod + 1e-30)
        # and the reference's own harness check (max-abs vs a comparator, tol 0.1)
        assert diff.max() < 0.1


def test_splitk_deterministic_and_consistent(cc, dev):
    M, N, K = 96, 160, 8192
    A, B = oracle.workload_square_flat(M, N, K, seed=1)
    Ad, Bd = _t(A, dev), _t(B, dev)
    ref = oracle.gemm_f64(A, B)
    for mode in ("fp32", "bf16x3", "bf16x4"):
        for split in (2, 7, 16):
            first = cc.gemm(Ad, Bd, mode=mode, split_k=split, tile_cfg=1).cpu().numpy()
            for _ in range(3):
                again = cc.gemm(Ad, Bd, mode=mode, split_k=split, tile_cfg=1).cpu().numpy()
                assert np.array_equal(first.view(np.uint32), again.view(np.uint32))
            extra = 2.0**-15 if mode.startswith("bf16") else 0.0
            _bound_check(first, ref, A, B, 2.0, 2.0**-24, extra)


def test_splitk_reduce_is_ascending_slice_sum(cc, dev):
    """A split-K result equals, bitwise, the sum in ascending slice order of
    single launches over the K slices (same tile, the slice boundaries on BK
    multiples), i.e. the reduction adds the partials one slice after another
    from zero, like the reference's reduce loop."""
    M, N, K = 96, 160, 8192
    g = torch.Generator(device="cpu").manual_seed(11)
    # the reduction's load batches for this output size: 3 slices one load at a
    # time, 16 and 20 in 8-wide batches (20: a partial last batch), 64 and 67
    # in 32-wide (FP32) / 16-wide (FP64) batches (67: a partial last batch)
    cases = (("fp32", torch.float32, 16), ("fp64", torch.float64, 20), ("bf16x3", torch.float32, 16),
             ("fp32", torch.float32, 64), ("fp64", torch.float64, 64), ("fp32", torch.float32, 3))
    for mode, dt, S in cases:
        ks = K // S // 32 * 32
        Kc = ks * S
        A = (torch.rand(M, Kc, generator=g, dtype=dt) * 2 - 1).to(dev)
        B = (torch.rand(Kc, N, generator=g, dtype=dt) * 2 - 1).to(dev)
        got = cc.gemm(A, B, mode=mode, split_k=S, tile_cfg=1)
        acc = torch.zeros(M, N, dtype=dt, device=dev)
        for z in range(S):
            part = cc.gemm(A[:, z * ks:(z + 1) * ks], B[z * ks:(z + 1) * ks], mode=mode, split_k=1, tile_cfg=1)
            acc = acc + part
        assert torch.equal(got, acc), (mode, S)
    # 67 slices of 96
    S, ks = 67, 96
    A = (torch.rand(M, S * ks, generator=g) * 2 - 1).to(dev)
    B = (torch.rand(S * ks, N, generator=g) * 2 - 1).to(dev)
    got = cc.gemm(A, B, mode="fp32", split_k=S, tile_cfg=1)
    acc = torch.zeros(M, N, device=dev)
    for z in range(S):
        acc = acc + cc.gemm(A[:, z * ks:(z + 1) * ks], B[z * ks:(z + 1) * ks], mode="fp32", split_k=1, tile_cfg=1)
    assert torch.equal(got, acc)


def test_fp64_vs_long_double_with_outliers(cc, dev):
    A, B, C0 = oracle.workload_dgemm(200, 150, 300, seed=2)
    ref = oracle.gemm_ld(A, B, C0, 1.0, 1.0)
    for split in (1, 4):
        Cd = _t(C0, dev)
        cc.gemm(_t(A, dev), _t(B, dev), Cd, alpha=1.0, beta=1.0, mode="fp64", split_k=split)
        got = Cd.cpu().numpy()
        absprod = np.abs(A) @ np.abs(B) + np.abs(C0)
        assert np.all(np.abs(got - ref) <= 4 * 300 * 2.0**-53 * absprod)


def test_fp64_large_output_split_in_two(cc, dev):
    """FP64 4352x4096x2048: 1088 tiles of 128x128 leave the third round of 512
    workgroups 1/8 full, so the heuristic splits K in two; the result stays
    within 2*K*u*|A||B| (u = 2^-53) of torch's FP64 product."""
    M, N, K = 4352, 4096, 2048
    assert cc.heuristic(1, M, N, K) == (7, 2)
    g = torch.Generator(device="cpu").manual_seed(21)
    A = torch.randn(M, K, generator=g, dtype=torch.float64).to(dev)
    B = torch.randn(K, N, generator=g, dtype=torch.float64).to(dev)
    got = cc.gemm(A, B, mode="fp64")
    ref = A @ B
    bound = 2 * K * 2.0**-53 * (A.abs() @ B.abs())
    assert torch.all((got - ref).abs() <= bound)


@pytest.mark.parametrize("mode,dt", MODES)
def test_alpha_beta_semantics(cc, dev, mode, dt):
    rng = np.random.default_rng(102)
    A, B, C0 = _int_operands(rng, 40, 30, 20, dt)
    Ad, Bd = _t(A, dev), _t(B, dev)
    # beta == 0 never reads C: NaN in C must not leak
    Cd = torch.full((40, 30), float("nan"), dtype=Ad.dtype, device=dev)
    cc.gemm(Ad, Bd, Cd, alpha=1.0, beta=0.0, mode=mode)
    assert np.array_equal(Cd.cpu().numpy(), (A.astype(np.int64) @ B.astype(np.int64)).astype(dt))
    Cd = torch.full((40, 30), float("nan"), dtype=Ad.dtype, device=dev)
    cc.gemm(Ad, Bd, Cd, alpha=1.0, beta=0.0, mode=mode, split_k=3)
    assert np.isfinite(Cd.cpu().numpy()).all()
    # alpha == 0: C = beta*C, A/B not read
    Cd = _t(C0, dev)
    cc.gemm(Ad, Bd, Cd, alpha=0.0, beta=0.5, mode=mode)
    assert np.array_equal(Cd.cpu().numpy(), (C0 * 0.5).astype(dt))
    # K == 0: C = beta*C
    Cd = _t(C0, dev)
    cc.gemm(Ad[:, :0].contiguous(), Bd[:0, :].contiguous(), Cd, alpha=1.0, beta=2.0, mode=mode)
    assert np.array_equal(Cd.cpu().numpy(), (C0 * 2).astype(dt))


@pytest.mark.parametrize("mode,dt", MODES)
@pytest.mark.parametrize("act", ["relu", "sigmoid"])
def test_fused_epilogue(cc, dev, mode, dt, act):
    """C = act(alpha*AB + beta*C) with the activation fused into the store (or
    run after the split-K reduction) equals the activation of the unfused
    result: ReLU bitwise, sigmoid within 4 ulp of the oracle's double-precision
    sigmoid of the same values.  The alpha == 0 path applies it to beta*C."""
    rng = np.random.default_rng(300)
    ulp = 2.0**-23 if dt == np.float32 else 2.0**-52
    for (M, N, K, split) in ((256, 256, 128, 1), (200, 136, 300, 1), (128, 192, 1024, 3)):
        A = rng.uniform(-1, 1, (M, K)).astype(dt)
        B = rng.uniform(-1, 1, (K, N)).astype(dt)
        C0 = rng.uniform(-1, 1, (M, N)).astype(dt)
        Ad, Bd = _t(A, dev), _t(B, dev)
        plain, fused = _t(C0, dev), _t(C0, dev)
        cc.gemm(Ad, Bd, plain, alpha=1.5, beta=0.5, mode=mode, split_k=split)
        cc.gemm(Ad, Bd, fused, alpha=1.5, beta=0.5, mode=mode, split_k=split, epilogue=act)
        p, f = plain.cpu().numpy(), fused.cpu().numpy()
        want = oracle.activate(p, act)
        assert (p < 0).any() and (p > 0).any()
        if act == "relu":
            assert np.array_equal(f, want.astype(dt)), (M, N, K, split)
        else:
            assert np.all(np.abs(f - want) <= 4 * ulp * want), (M, N, K, split)
    Cd = _t(C0, dev)
    cc.gemm(Ad, Bd, Cd, alpha=0.0, beta=0.5, mode=mode, epilogue=act)
    want = oracle.activate(C0.astype(np.float64) * 0.5, act)
    got = Cd.cpu().numpy()
    if act == "relu":
        assert np.array_equal(got, want.astype(dt))
    else:
        assert np.all(np.abs(got - want) <= 4 * ulp * want)
    with pytest.raises(ValueError):
        cc.gemm(Ad, Bd, Cd, mode=mode, epilogue="tanh")


# pre-split configurations, and the on-the-fly configurations with the same
# 256x256 tile (the default one first: loads spread among the MFMAs; then the
# ping-pong one of grouped launches)
PRESPLIT = {"bf16x3": ((7, 8), (12, 6)), "bf16x4": ((7, 8), (12, 6))}


@pytest.mark.parametrize("mode", ["bf16x3", "bf16x4"])
@pytest.mark.parametrize("ta,tb", LAYOUTS)
def test_presplit_bitwise_equals_on_the_fly(cc, dev, mode, ta, tb):
    """The pre-split configuration (split pass into bf16 planes, direct-to-LDS
    loads) computes the same products in the same order as the on-the-fly
    kernel with the same 256x256 tile: bitwise equal results, for aligned and
    ragged shapes, with split-K, alpha and beta."""
    planars, fly = PRESPLIT[mode]
    same = fly[0]
    t = cc.tile_configs(cc._lib.MODES[mode])
    assert [c["cfg"] for c in t if c["pipe"] == 5] == list(planars)
    for cfg in planars + fly[1:]:
        assert (t[cfg]["BM"], t[cfg]["BN"]) == (t[same]["BM"], t[same]["BN"])
    rng = np.random.default_rng(400)
    for (M, N, K, split) in ((512, 512, 256, 1), (300, 200, 100, 1), (256, 512, 1000, 3), (33, 517, 77, 2)):
        A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
        B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
        C0 = rng.uniform(-1, 1, (M, N)).astype(np.float32)
        Ad, Bd = _t(A.T if ta else A, dev), _t(B.T if tb else B, dev)
        outs = []
        for cfg in fly + planars:
            Cd = _t(C0, dev)
            cc.gemm(Ad, Bd, Cd, alpha=1.25, beta=0.5, mode=mode, trans_a=ta, trans_b=tb, tile_cfg=cfg,
                    split_k=split)
            outs.append(Cd.cpu().numpy())
        for o in outs[1:]:
            assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32)), (M, N, K, split)
        # fused epilogue on both paths
        acts = []
        for cfg in (same, planars[-1]):
            Cd = _t(C0, dev)
            cc.gemm(Ad, Bd, Cd, alpha=1.25, beta=0.5, mode=mode, trans_a=ta, trans_b=tb, tile_cfg=cfg,
                    split_k=split, epilogue="sigmoid")
            acts.append(Cd.cpu().numpy())
        assert np.array_equal(acts[0].view(np.uint32), acts[1].view(np.uint32)), (M, N, K, split)
        ref = oracle.gemm_bf16x3(A, B, C0, 1.25, 0.5) if mode == "bf16x3" else oracle.gemm_bf16x4(A, B, C0, 1.25, 0.5)
        absprod = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64)) + 0.5 * np.abs(C0)
        assert np.all(np.abs(outs[0] - ref) <= 2 * (K + 2) * 2.0**-24 * 1.25 * absprod + 1e-30)


@pytest.mark.parametrize("mode", ["bf16x3", "bf16x4"])
def test_heuristic_presplit_for_underfilled_tiles(cc, dev, mode):
    """3072^2 x 1024 (144 256x256 tiles, one partly filled round) takes the
    pre-split configuration by default (library workspace for the planes),
    K = 256 128x128 tiles (BF16x3: the 4-wave one); either result is bitwise
    the on-the-fly 256x256 tile's unsplit (the same products in the same
    order)."""
    planars, fly = PRESPLIT[mode]
    same = fly[0]
    M, N, K = 3072, 3072, 1024
    assert cc.heuristic(cc._lib.MODES[mode], M, N, 256) == ((0, 1) if mode == "bf16x3" else (11, 1))
    cfg, split = cc.heuristic(cc._lib.MODES[mode], M, N, K)
    assert cfg in planars and split == 1
    assert cc.heuristic(cc._lib.MODES[mode], 4096, 4096, 4096)[0] == same
    rng = np.random.default_rng(403)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    for k in (K, 256):
        Ak, Bk = np.ascontiguousarray(A[:, :k]), np.ascontiguousarray(B[:k])
        for ta in (False, True):
            Ad, Bd = _t(Ak.T if ta else Ak, dev), _t(Bk, dev)
            got = cc.gemm(Ad, Bd, mode=mode, trans_a=ta).cpu().numpy()
            want = cc.gemm(Ad, Bd, mode=mode, trans_a=ta, tile_cfg=same, split_k=1).cpu().numpy()
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (k, ta)


@pytest.mark.parametrize("mode", ["bf16x3", "bf16x4"])
def test_heuristic_presplit_split_for_edge_tiles(cc, dev, mode):
    """2900^2 x 3072 (144 partly filled 256x256 tiles) takes the pre-split
    configuration with K split in three by default; bitwise the on-the-fly
    256x256 tile at the same split (both cut K into 32-aligned slices of the
    same length and sum the slices in the same order), and within the split
    modes' componentwise bound of the FP64 product."""
    planars, fly = PRESPLIT[mode]
    M, N, K = 2900, 2900, 3072
    cfg, split = cc.heuristic(cc._lib.MODES[mode], M, N, K)
    assert cfg in planars and split == 3
    rng = np.random.default_rng(405)
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    exact = A.astype(np.float64) @ B.astype(np.float64)
    for ta in (False, True):
        Ad, Bd = _t(A.T.copy() if ta else A, dev), _t(B, dev)
        got = cc.gemm(Ad, Bd, mode=mode, trans_a=ta).cpu().numpy()
        want = cc.gemm(Ad, Bd, mode=mode, trans_a=ta, tile_cfg=fly[0], split_k=3).cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), ta
    _bound_check(got, exact, A, B, 2.0, 2.0**-24, extra=2.0**-15)


@pytest.mark.parametrize("mode", ["bf16x3", "bf16x4"])
def test_spread_load_tile_equals_ping_pong(cc, dev, mode):
    """The spread-load 256x256 tile (every wave copies first and issues its
    next loads among its MFMAs) computes the ping-pong tile's products in the
    same order: bitwise equal results over random shapes (edges, K not a
    multiple of 32, unaligned leading dimensions), layouts, split-K factors,
    alpha/beta and fused epilogues."""
    planars, fly = PRESPLIT[mode]
    spread, pingpong = fly
    t = cc.tile_configs(cc._lib.MODES[mode])
    assert t[spread]["pipe"] == 7 and t[pingpong]["pipe"] == 4
    rng = np.random.default_rng(404)
    for case in range(12):
        M, N = (int(x) for x in rng.integers(1, 700, 2))
        K = int(rng.integers(1, 3000))
        ta, tb = LAYOUTS[case % 3]
        split = int(rng.integers(1, 6))
        pad = int(rng.integers(0, 3))  # extra columns: unaligned rows for pad 1, 2
        alpha, beta = float(rng.uniform(-2, 2)), (0.0 if case % 4 == 0 else float(rng.uniform(-1, 1)))
        act = (None, "relu", "sigmoid")[case % 3]
        A = rng.uniform(-1, 1, (K, M + pad) if ta else (M, K + pad)).astype(np.float32)
        B = rng.uniform(-1, 1, (N, K + pad) if tb else (K, N + pad)).astype(np.float32)
        C0 = rng.uniform(-1, 1, (M, N)).astype(np.float32)
        Ad = _t(A, dev)[:, :M] if ta else _t(A, dev)[:, :K]
        Bd = _t(B, dev)[:, :K] if tb else _t(B, dev)[:, :N]
        outs = []
        for cfg in (spread, pingpong):
            Cd = _t(C0, dev)
            cc.gemm(Ad, Bd, Cd, alpha=alpha, beta=beta, mode=mode, trans_a=ta, trans_b=tb, tile_cfg=cfg,
                    split_k=split, epilogue=act)
            outs.append(Cd.cpu().numpy())
        assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32)), (M, N, K, ta, tb, split, pad)


@pytest.mark.parametrize("ta,tb", LAYOUTS)
def test_presplit_caller_planes(cc, dev, ta, tb):
    """Operands split once by split_operand and passed as planes give the
    bitwise result of the in-call split; one or both operands may be given."""
    rng = np.random.default_rng(401)
    M, N, K = 300, 520, 200
    A = rng.uniform(-1, 1, (M, K)).astype(np.float32)
    B = rng.uniform(-1, 1, (K, N)).astype(np.float32)
    Ad, Bd = _t(A.T if ta else A, dev), _t(B.T if tb else B, dev)
    ap = cc.split_operand(Ad, trans=ta)
    bp = cc.split_operand(Bd, trans=tb, as_b=True)
    kp = (K + 31) // 32 * 32
    assert tuple(ap.shape) == (M, 2 * kp) and tuple(bp.shape) == (N, 2 * kp)
    assert ap.numel() * 2 == cc.lib.cosma_planes_bytes(M, K)
    want = cc.gemm(Ad, Bd, mode="bf16x3", trans_a=ta, trans_b=tb, tile_cfg=7).cpu().numpy()
    for kw in (dict(a_planes=ap, b_planes=bp), dict(a_planes=ap), dict(b_planes=bp)):
        got = cc.gemm(Ad, Bd, mode="bf16x3", trans_a=ta, trans_b=tb, **kw).cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), list(kw)
    # the planes hold the RNE hi/lo split of the operand, K zero-padded
    hi, lo = oracle.split_bf16(A)
    planes = ap.view(torch.int16).cpu().numpy().view(np.uint16).reshape(M, kp // 32, 64)
    hp, lp = planes[:, :, :32].reshape(M, kp)[:, :K], planes[:, :, 32:].reshape(M, kp)[:, :K]
    assert np.array_equal(hp, hi) and np.array_equal(lp, lo)
    assert not planes[:, :, :32].reshape(M, kp)[:, K:].any()


def test_presplit_workspace_and_grouped_rejection(cc, dev):
    from cucosma_amd._lib import GemmDesc

    M, N, K = 1000, 600, 333
    d = GemmDesc()
    cc.lib.cosma_gemm_desc_init(C.byref(d))
    d.mode, d.m, d.n, d.k = cc._lib.MODE_BF16X3, M, N, K
    d.a, d.b, d.c = 0x1000, 0x2000, 0x3000
    d.lda, d.ldb, d.ldc = K, N, N
    d.tile_cfg, d.split_k = 7, 1
    size = C.c_size_t()
    assert cc.lib.cosma_workspace_query(C.byref(d), C.byref(size)) == 0
    kp = (K + 31) // 32 * 32
    assert size.value == (M + N) * 2 * kp * 2
    d.split_k = 4
    assert cc.lib.cosma_workspace_query(C.byref(d), C.byref(size)) == 0
    part = 4 * M * N * 4
    assert size.value == (part + 255) // 256 * 256 + (M + N) * 2 * kp * 2
    A = torch.ones((64, 32), device=dev)
    B = torch.ones((32, 48), device=dev)
    Cm = torch.zeros((64, 48), device=dev)
    with pytest.raises(cc.CosmaError) as e:
        cc.GroupedGemm([A], [B], [Cm], mode="bf16x3", tile_cfgs=[7])
    assert e.value.status == -3


def test_residual_stats_on_gpu_match_host(cc, dev):
    rng = np.random.default_rng(103)
    R = rng.normal(0, 1, (513, 257))
    R[3, 4] = 0.0
    Cm = (R + rng.normal(0, 1e-5, R.shape)).astype(np.float32)
    st_gpu = cc.error_stats(_t(Cm, dev), _t(R, dev))
    st_cpu = oracle.err_stats(Cm, R)
    for key in ("sum_sq_err", "sum_sq_ref"):
        assert np.isclose(st_gpu[key], st_cpu[key], rtol=1e-12)
    for key in ("max_abs", "max_ref"):
        assert st_gpu[key] == st_cpu[key]
    assert st_gpu["max_rel_all"] == st_cpu["max_rel"]
    # the floored relative error: entries with |R| >= 1e-3 max|R| only
    d = np.abs(Cm.astype(np.float64) - R)
    keep = np.abs(R) >= 1e-3 * np.abs(R).max()
    assert st_gpu["max_rel"] == (d[keep] / np.abs(R)[keep]).max()
    assert cc.error_stats(_t(Cm, dev), _t(R, dev), rel_floor=0.0)["max_rel"] == st_cpu["max_rel"]
    # deterministic: bitwise repeatable
    again = cc.error_stats(_t(Cm, dev), _t(R, dev))
    assert again == st_gpu
    X = rng.normal(0, 1, (100, 77)).astype(np.float32)
    assert np.isclose(cc.sumsq(_t(X, dev)), oracle.sumsq(X), rtol=1e-12)
    Y = cc.to_f64(_t(X, dev)).cpu().numpy()
    assert np.array_equal(Y, X.astype(np.float64))


def test_caller_workspace(cc, dev):
    M, N, K = 64, 64, 4096
    A, B = oracle.workload_square_flat(M, N, K, seed=1)
    Ad, Bd = _t(A, dev), _t(B, dev)
    need = cc.workspace_bytes(M, N, K, split_k=8, tile_cfg=1)
    assert need == 8 * M * N * 4
    ws = torch.empty(need // 4, dtype=torch.float32, device=dev)
    c1 = cc.gemm(Ad, Bd, split_k=8, tile_cfg=1, workspace=ws).cpu().numpy()
    small = torch.empty(16, dtype=torch.float32, device=dev)
    with pytest.raises(cc.CosmaError):
        cc.gemm(Ad, Bd, split_k=8, tile_cfg=1, workspace=small)
    # installed workspace
    assert cc.lib.cosma_workspace_set(C.c_void_p(ws.data_ptr()), need) == 0
    c2 = cc.gemm(Ad, Bd, split_k=8, tile_cfg=1).cpu().numpy()
    assert cc.lib.cosma_workspace_set(None, 0) == 0
    assert np.array_equal(c1, c2)


def test_library_workspace_per_stream(cc, dev):
    """Calls on two streams without a caller workspace use separate library
    buffers: pre-split planes and split-K partials queued on both streams at
    once give the results of the same calls run one after another, bitwise.
    Releasing frees the buffers; later calls allocate again."""
    g = torch.Generator(device="cpu").manual_seed(9)
    M = N = 2048
    K = 4096
    ops = [((torch.rand(M, K, generator=g) - 0.5).to(dev), (torch.rand(K, N, generator=g) - 0.5).to(dev))
           for _ in range(2)]
    kws = [dict(mode="bf16x3", tile_cfg=8, split_k=1), dict(mode="fp32", tile_cfg=1, split_k=8)]
    need = max(cc.workspace_bytes(M, N, K, **kw) for kw in kws)
    assert need == 8 * M * N * 4
    refs = [[cc.gemm(A, B, **kw) for kw in kws] for A, B in ops]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [[torch.empty(M, N, device=dev) for _ in kws] for _ in ops]
    for s in streams:
        s.wait_stream(torch.cuda.current_stream())
    for _ in range(4):
        for j, kw in enumerate(kws):
            for (A, B), s, out in zip(ops, streams, outs):
                cc.gemm(A, B, out[j], stream=s, **kw)
    torch.cuda.synchronize()
    for i in range(2):
        for j in range(2):
            assert torch.equal(outs[i][j], refs[i][j]), (i, kws[j])
    # the current stream's buffer and one per side stream
    assert cc.release_workspace() >= 3 * need
    assert cc.release_workspace() == 0
    A, B = ops[0]
    assert torch.equal(cc.gemm(A, B, **kws[0]), refs[0][0])
    torch.cuda.synchronize()


def test_gemm_calls_in_a_graph(cc, dev):
    """A split-K call (partials in the library workspace) and a pre-split call
    (planes in the workspace), captured in one hipGraph on a stream whose
    workspace was sized by an eager call, replay bitwise equal to the eager
    results, with eager work between replays and after releasing the
    library's workspaces (the captured buffer is kept).  A call that needs a
    larger workspace while the stream is capturing fails before launching."""
    g = torch.Generator(device="cpu").manual_seed(12)
    A1 = (torch.rand(128, 4096, generator=g) * 2 - 1).to(dev)
    B1 = (torch.rand(4096, 192, generator=g) * 2 - 1).to(dev)
    A2 = (torch.rand(1024, 512, generator=g) * 2 - 1).to(dev)
    B2 = (torch.rand(512, 768, generator=g) * 2 - 1).to(dev)
    C1 = torch.empty(128, 192, device=dev)
    C2 = torch.empty(1024, 768, device=dev)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())

    def calls():
        cc.gemm(A1, B1, C1, mode="fp32", split_k=8, tile_cfg=1)
        cc.gemm(A2, B2, C2, mode="bf16x3", tile_cfg=8)

    with torch.cuda.stream(s):
        calls()
    torch.cuda.synchronize()
    ref1, ref2 = C1.clone(), C2.clone()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        calls()
    torch.cuda.synchronize()
    for rep in range(4):
        C1.zero_()
        C2.zero_()
        torch.cuda.synchronize()
        if rep == 2:
            cc.release_workspace()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(C1, ref1) and torch.equal(C2, ref2), rep
    A3 = (torch.rand(128, 16384, generator=g) * 2 - 1).to(dev)
    B3 = (torch.rand(16384, 192, generator=g) * 2 - 1).to(dev)
    C3 = torch.empty(128, 192, device=dev)
    assert cc.workspace_bytes(128, 192, 16384, split_k=64, tile_cfg=1) > cc.workspace_bytes(
        1024, 768, 512, mode="bf16x3", tile_cfg=8)
    graph2 = torch.cuda.CUDAGraph()
    with pytest.raises(cc.CosmaError) as err:
        with torch.cuda.graph(graph2, stream=s):
            cc.gemm(A3, B3, C3, mode="fp32", split_k=64, tile_cfg=1)
    assert err.value.status == -4
    torch.cuda.synchronize()
    del graph, graph2


def test_fp64_device_reference_matches_host_oracle(cc, dev):
    """The GPU FP64 product used for large-size error reports agrees with the
    host FP64 oracle."""
    A, B = oracle.workload_square_flat(256, 192, 3000, seed=1)
    R = cc.fp64_product(_t(A, dev), _t(B, dev)).cpu().numpy()
    ref = oracle.gemm_f64(A, B)
    absprod = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    assert np.all(np.abs(R - ref) <= 4 * 3000 * 2.0**-53 * absprod)


def test_small_caller_workspace_runs_on_the_fly_tile(cc, dev):
    """A default pre-split choice whose planes do not fit the caller's
    workspace runs the on-the-fly 256x256 tile at the same split: the call
    succeeds and is bitwise the library-workspace call.  A workspace too small
    for the split-K partials as well is refused before any launch."""
    from cucosma_amd._lib import MODES, GemmDesc

    def desc(A, B, Cm, ws=None, ws_bytes=0):
        d = GemmDesc()
        cc.lib.cosma_gemm_desc_init(C.byref(d))
        d.mode = MODES["bf16x3"]
        d.m, d.k = A.shape
        d.n = B.shape[1]
        d.a, d.lda = A.data_ptr(), A.shape[1]
        d.b, d.ldb = B.data_ptr(), B.shape[1]
        d.c, d.ldc = Cm.data_ptr(), Cm.shape[1]
        d.alpha, d.beta = 1.0, 0.0
        if ws is not None:
            d.workspace, d.workspace_bytes = ws.data_ptr(), ws_bytes
        d.stream = torch.cuda.current_stream().cuda_stream
        return d

    g = torch.Generator(device="cpu").manual_seed(406)
    for (M, N, K, split) in ((3072, 3072, 1024, 1), (2900, 2900, 3072, 3)):
        assert cc.heuristic(cc._lib.MODES["bf16x3"], M, N, K) == (8, split)
        A = (torch.rand(M, K, generator=g) * 2 - 1).to(dev)
        B = (torch.rand(K, N, generator=g) * 2 - 1).to(dev)
        ref = torch.empty(M, N, device=dev)
        assert cc.lib.cosma_gemm_ex(C.byref(desc(A, B, ref))) == 0
        partials = (split * M * N * 4) if split > 1 else 0
        ws = torch.empty(max(partials, 1 << 20), dtype=torch.uint8, device=dev)
        got = torch.empty(M, N, device=dev)
        assert cc.lib.cosma_gemm_ex(C.byref(desc(A, B, got, ws, ws.numel()))) == 0
        torch.cuda.synchronize()
        assert torch.equal(ref, got), (M, N, K)
        if split > 1:
            Cm = torch.full((M, N), 7.0, device=dev)
            assert cc.lib.cosma_gemm_ex(C.byref(desc(A, B, Cm, ws, partials - 256))) == -1
            torch.cuda.synchronize()
            assert bool((Cm == 7.0).all())
