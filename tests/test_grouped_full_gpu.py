"""Grouped GEMM at the benchmark workload and its new parts (GPU).

* The 512-problem batch the harness times (M, N, K log-uniform in
  [32, 2048], rounded to 32, seed 3, entries uniform[-1,1]), FP32 and BF16x3,
  run as the default plan's hipGraph: every element of every problem within
  the accumulation bound of the GPU FP64 product (itself checked against the
  host oracle on sampled rows), BF16x3 sampled rows against the exact
  three-product oracle, and every problem bitwise equal to a single launch
  with the same tile configuration and split.
* Split members (per-problem split-K, reduced in ascending slice order after
  the persistent launch): bitwise equal to single launches with the same
  split, and graph replays bitwise equal to the eager run.
* alpha == 0 with NULL operands and K == 0 give C = beta * C.
* The grouped autotuner's choices run and match single launches.
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


def _batch(sizes, seed, dev, dt=torch.float32):
    g = torch.Generator(device=dev).manual_seed(seed)
    As = [torch.rand((m, k), device=dev, generator=g, dtype=dt) * 2 - 1 for m, n, k in sizes]
    Bs = [torch.rand((k, n), device=dev, generator=g, dtype=dt) * 2 - 1 for m, n, k in sizes]
    return As, Bs


def _check_against_fp64(cc, As, Bs, Cs, mode, rows_per_problem=2):
    """Every element against the GPU FP64 product, with the accumulation
    bound of the mode; the FP64 product against the host oracle on rows."""
    u = 2.0**-24
    extra = 2.0**-15 if mode.startswith("bf16") else 0.0
    for i, (A, B, Cm) in enumerate(zip(As, Bs, Cs)):
        m, k = A.shape
        R = cc.fp64_product(A, B)
        absAB = cc.fp64_product(A.abs(), B.abs())
        bound = (2 * k * u + extra) * absAB
        diff = (Cm.double() - R).abs()
        assert bool((diff <= bound + 1e-30).all()), (i, tuple(A.shape), tuple(B.shape), float((diff - bound).max()))
        if i % 16 == 0:  # host oracle on sampled rows of the FP64 product
            An, Bn = A.cpu().numpy(), B.cpu().numpy()
            rows = sorted({0, m - 1})[:rows_per_problem]
            host = np.concatenate([oracle.gemm_f64(An, Bn, rows=1, row0=r) for r in rows])
            Rr = R[rows].cpu().numpy()
            ab = np.abs(An[rows].astype(np.float64)) @ np.abs(Bn.astype(np.float64))
            assert np.all(np.abs(Rr - host) <= 2 * k * 2.0**-53 * ab + 1e-300), i
            if mode == "bf16x3":
                exact3 = np.concatenate([oracle.gemm_bf16x3(An, Bn, rows=1, row0=r) for r in rows])
                got = Cm[rows].cpu().numpy().astype(np.float64)
                assert np.all(np.abs(got - exact3) <= 2 * k * u * ab + 1e-30), i


@pytest.mark.parametrize("mode", ["fp32", "bf16x3"])
def test_grouped_512_problem_batch(cc, dev, mode):
    from cucosma_amd.harness import grouped_sizes

    sizes = grouped_sizes()
    assert len(sizes) == 512
    As, Bs = _batch(sizes, 3, dev)
    Cs = [torch.empty((m, n), device=dev) for m, n, k in sizes]
    plan = cc.GroupedGemm(As, Bs, Cs, mode=mode)
    choices = plan.choices()
    fam0 = cc.group_families(mode)[0]
    assert all(c in fam0 for c, s in choices), "the default plan runs the first family"
    # thin variants are used where they cut padding
    assert len({c for c, s in choices}) >= 2
    plan.run_graph()
    torch.cuda.synchronize()
    _check_against_fp64(cc, As, Bs, Cs, mode)
    for i, ((cfg, split), A, B, Cg) in enumerate(zip(choices, As, Bs, Cs)):
        single = torch.empty_like(Cg)
        cc.gemm(A, B, single, mode=mode, tile_cfg=cfg, split_k=split)
        assert torch.equal(Cg, single), (i, sizes[i], cfg, split)
    # a second replay gives the same bits
    first = [c.clone() for c in Cs]
    for c in Cs:
        c.fill_(float("nan"))
    plan.run_graph()
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(first, Cs))
    plan.close()


@pytest.mark.parametrize("mode", ["fp32", "bf16x3", "fp64"])
def test_grouped_split_members(cc, dev, mode):
    """Deep-K members with few tiles are split by the default plan; their
    partials are reduced after the persistent launch in ascending slice order,
    so each result is bitwise the single launch with the same split."""
    dt = torch.float64 if mode == "fp64" else torch.float32
    sizes = [(256, 256, 32768)] * 3 + [(128, 96, 20000), (64, 64, 64), (300, 200, 96), (33, 17, 5)]
    As, Bs = _batch(sizes, 11, dev, dt)
    g = torch.Generator(device=dev).manual_seed(12)
    C0s = [torch.rand((m, n), device=dev, generator=g, dtype=dt) for m, n, k in sizes]
    Cs = [c.clone() for c in C0s]
    plan = cc.GroupedGemm(As, Bs, Cs, mode=mode, alpha=1.25, beta=0.5)
    choices = plan.choices()
    assert all(s > 1 for c, s in choices[:3]), choices
    assert choices[4][1] == 1 and choices[6][1] == 1
    plan.run()
    torch.cuda.synchronize()
    eager = [c.clone() for c in Cs]
    for i, ((cfg, split), A, B, C0) in enumerate(zip(choices, As, Bs, C0s)):
        single = C0.clone()
        cc.gemm(A, B, single, mode=mode, alpha=1.25, beta=0.5, tile_cfg=cfg, split_k=split)
        assert torch.equal(eager[i], single), (i, sizes[i], cfg, split)
        R = cc.gemm(A.double(), B.double(), C0.double().clone(), alpha=1.25, beta=0.5, mode="fp64")
        absAB = 1.25 * (A.double().abs() @ B.double().abs()) + 0.5 * C0.double().abs()
        u = 2.0**-53 if mode == "fp64" else 2.0**-24
        extra = 2.0**-15 if mode == "bf16x3" else 0.0
        k = A.shape[1]
        assert bool(((eager[i].double() - R).abs() <= (2 * (k + 2) * u + extra) * absAB + 1e-300).all()), i
    for rep in range(3):
        for Cd, C0 in zip(Cs, C0s):
            Cd.copy_(C0)
        torch.cuda.synchronize()
        plan.run_graph()
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip(Cs, eager)), rep
    plan.close()


def test_grouped_alpha_zero_and_empty_k(cc, dev):
    """alpha == 0 with NULL A and B, and K == 0, scale C by beta (the
    operands are never read); a regular problem in the same plan is
    unaffected."""
    from cucosma_amd._lib import GroupProblem

    C0 = torch.rand((100, 70), device=dev) * 2 - 1
    C1 = torch.rand((64, 64), device=dev) * 2 - 1
    A = torch.rand((40, 30), device=dev)
    B = torch.rand((30, 20), device=dev)
    Cz, Ck, Cr = C0.clone(), C1.clone(), torch.zeros((40, 20), device=dev)
    probs = (GroupProblem * 3)(
        GroupProblem(100, 70, 512, None, 512, None, 70, Cz.data_ptr(), 70, 0.0, 0.5),
        GroupProblem(64, 64, 0, None, 1, None, 64, Ck.data_ptr(), 64, 1.0, -2.0),
        GroupProblem(40, 20, 30, A.data_ptr(), 30, B.data_ptr(), 20, Cr.data_ptr(), 20, 1.0, 0.0),
    )
    lib = cc.lib
    for mode in (0, 2):
        Cz.copy_(C0)
        Ck.copy_(C1)
        plan = C.c_void_p()
        assert lib.cosma_grouped_plan_create(mode, 3, probs, None, 0, C.byref(plan)) == 0
        assert lib.cosma_grouped_plan_run(plan, None) == 0
        torch.cuda.synchronize()
        assert torch.equal(Cz, 0.5 * C0)
        assert torch.equal(Ck, -2.0 * C1)
        ref = A.double() @ B.double()
        assert float((Cr.double() - ref).abs().max()) < 30 * 2.0**-15 * 30
        lib.cosma_grouped_plan_destroy(plan)


def test_grouped_autotune_choices_run(cc, dev):
    from cucosma_amd.harness import grouped_sizes

    sizes = grouped_sizes(count=48, seed=7, hi=1024)
    As, Bs = _batch(sizes, 7, dev)
    Cs = [torch.empty((m, n), device=dev) for m, n, k in sizes]
    res = cc.grouped_autotune(As, Bs, Cs, mode="bf16x3", warmup=1, reps=2)
    assert res["tried"] and res["tried"][0]["bucket"] == -1
    assert res["ms_best"] <= res["ms_start"]
    fam0 = cc.group_families("bf16x3")[0]
    plan = cc.GroupedGemm(As, Bs, Cs, mode="bf16x3", tile_cfgs=res["tile_cfgs"], splits=res["splits"])
    assert [c for c, s in plan.choices()] == res["tile_cfgs"]
    assert all(c in fam0 or c in (0, 1) for c in res["tile_cfgs"])
    plan.run()
    torch.cuda.synchronize()
    for (cfg, split), A, B, Cg in zip(plan.choices(), As, Bs, Cs):
        single = torch.empty_like(Cg)
        cc.gemm(A, B, single, mode="bf16x3", tile_cfg=cfg, split_k=split)
        assert torch.equal(Cg, single)
    plan.close()


def test_grouped_planes_and_device_checks(cc, dev):
    A = torch.rand((64, 96), device=dev)
    B = torch.rand((96, 32), device=dev)
    small = torch.empty(16, dtype=torch.bfloat16, device=dev)
    with pytest.raises(ValueError):
        cc.gemm(A, B, mode="bf16x3", a_planes=small)
    with pytest.raises(ValueError):
        cc.gemm(A, B, torch.empty((64, 32)), mode="fp32")
    with pytest.raises(TypeError):
        cc.GroupedGemm([A], [B], [torch.empty((64, 32), device=dev)], mode="fp64")
