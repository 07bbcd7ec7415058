"""GPU tests of the grouped persistent GEMM (plan, launch, hipGraph replay)
and of the autograd wrapper (finite differences and torch FP64 gradients)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def cc(dev):
    import cucosma_amd

    return cucosma_amd


def _sizes(seed, count, lo=32, hi=512):
    """M, N, K log-uniform in [lo, hi], rounded to multiples of 32."""
    rs = np.random.RandomState(seed)
    out = []
    for _ in range(count):
        dims = np.exp(rs.uniform(np.log(lo), np.log(hi), 3))
        out.append(tuple(int(min(hi, max(lo, 32 * round(d / 32)))) for d in dims))
    return out


@pytest.mark.parametrize("mode", ["fp32", "bf16x3", "fp64"])
def test_grouped_matches_oracle_and_graph_replays(cc, dev, mode):
    dt = torch.float64 if mode == "fp64" else torch.float32
    sizes = _sizes(3, 24)
    sizes += [(33, 17, 5), (1, 1, 1), (7, 300, 64)]  # ragged, not multiples of 32
    rng = oracle.Rng(3)
    As, Bs, Cs, C0s = [], [], [], []
    for (m, n, k) in sizes:
        A = rng.uniform_f64((m, k))
        B = rng.uniform_f64((k, n))
        C0 = rng.uniform_f64((m, n))
        if mode != "fp64":
            A, B, C0 = A.astype(np.float32), B.astype(np.float32), C0.astype(np.float32)
        As.append(torch.from_numpy(A).to(dev))
        Bs.append(torch.from_numpy(B).to(dev))
        C0s.append(C0)
        Cs.append(torch.from_numpy(C0).to(dev))
    g = cc.GroupedGemm(As, Bs, Cs, mode=mode, alpha=1.0, beta=0.5)
    info = g.info()
    assert info["total_tiles"] > 0 and info["buckets"]
    assert abs(info["flops"] - sum(2.0 * m * n * k for m, n, k in sizes)) < 1
    g.run()
    torch.cuda.synchronize()
    eager = [c.clone() for c in Cs]
    u = 2.0**-53 if mode == "fp64" else 2.0**-24
    extra = 2.0**-15 if mode == "bf16x3" else 0.0
    for (m, n, k), A, B, C0, Cd in zip(sizes, As, Bs, C0s, Cs):
        An, Bn = A.cpu().numpy(), B.cpu().numpy()
        if mode == "fp64":
            ref = oracle.gemm_ld(An, Bn, C0, 1.0, 0.5)
        else:
            ref = oracle.gemm_f64(An, Bn, C0, 1.0, 0.5)
        absprod = np.abs(An.astype(np.float64)) @ np.abs(Bn.astype(np.float64)) + 0.5 * np.abs(C0)
        err = np.abs(Cd.cpu().numpy().astype(np.float64) - ref)
        assert np.all(err <= (2 * (k + 2) * u + extra) * absprod + 1e-300), (m, n, k)
    # hipGraph replay: reset C between replays with eager work and host
    # synchronisation, and check every replay bitwise against the eager run.
    for rep in range(3):
        for Cd, C0 in zip(Cs, C0s):
            Cd.copy_(torch.from_numpy(C0).to(dev))
        torch.cuda.synchronize()
        g.run_graph()
        torch.cuda.synchronize()
        for Cd, E in zip(Cs, eager):
            assert torch.equal(Cd, E), rep
    g.close()


def test_grouped_empty_and_explicit_tiles(cc, dev):
    g = cc.GroupedGemm([], [], [], mode="fp32")
    g.run()
    assert g.info()["total_tiles"] == 0
    g.close()
    A = torch.ones((64, 32), device=dev)
    B = torch.ones((32, 48), device=dev)
    Cm = torch.zeros((64, 48), device=dev)
    g = cc.GroupedGemm([A], [B], [Cm], tile_cfgs=[1])
    assert g.info()["buckets"][0]["tile_cfg"] == 1
    g.run()
    torch.cuda.synchronize()
    assert torch.all(Cm == 32)


@pytest.mark.parametrize("mode", ["fp32", "bf16x3", "bf16x4", "fp64"])
def test_grouped_default_plan_is_one_launch(cc, dev, mode):
    """Without tile_cfgs the whole batch is one persistent launch: the mode's
    first tile family (the large tile and its thin variants) when outputs
    with a side of at least 256 hold half the flops, else 128x128, else
    64x64."""
    from cucosma_amd.harness import grouped_sizes

    dt = torch.float64 if mode == "fp64" else torch.float32
    # the first tile family: FP32 256x128 (16 waves), FP64 128x128 (8 waves),
    # BF16 256x256 on the fly, each with its thin and small variants
    large = cc.group_families(mode)[0][0]
    for sizes, want in ((grouped_sizes(), large), ([(96, 160, 64)] * 3 + [(40, 32, 512)], 0),
                        ([(64, 64, 64), (32, 96, 128)], 1)):
        As = [torch.empty((m, k), device=dev, dtype=dt) for m, n, k in sizes]
        Bs = [torch.empty((k, n), device=dev, dtype=dt) for m, n, k in sizes]
        Cs = [torch.empty((m, n), device=dev, dtype=dt) for m, n, k in sizes]
        plan = cc.GroupedGemm(As, Bs, Cs, mode=mode)
        b = plan.info()["buckets"]
        assert len(b) == 1 and b[0]["tile_cfg"] == want and b[0]["problems"] == len(sizes), b
        assert 1 <= b[0]["blocks"] <= b[0]["tiles"]
        if want == large:
            assert {c for c, s in plan.choices()} <= set(cc.group_families(mode)[0])
        plan.close()
        del As, Bs, Cs


@pytest.mark.parametrize("mode", ["fp32", "bf16x3", "bf16x4", "fp64"])
def test_grouped_bitwise_equals_single_launches(cc, dev, mode):
    """Every tile configuration in one plan: each problem's result is bitwise
    the single-launch result of the same tile configuration without split-K
    (the grouped kernel runs the same per-tile body)."""
    dt = torch.float64 if mode == "fp64" else torch.float32
    # every tile configuration a grouped launch runs (not the pre-split ones, pipe 5)
    usable = [t["cfg"] for t in cc.tile_configs(cc._lib.MODES[mode]) if t["pipe"] != 5]
    ncfg = len(usable)
    sizes = _sizes(5, 2 * ncfg) + [(33, 17, 5), (7, 300, 64), (300, 7, 129)]
    cfgs = [usable[i % ncfg] for i in range(len(sizes))]
    g = torch.Generator(device="cpu").manual_seed(5)
    As = [(torch.rand(m, k, generator=g, dtype=dt) * 2 - 1).to(dev) for m, n, k in sizes]
    Bs = [(torch.rand(k, n, generator=g, dtype=dt) * 2 - 1).to(dev) for m, n, k in sizes]
    C0s = [(torch.rand(m, n, generator=g, dtype=dt) * 2 - 1).to(dev) for m, n, k in sizes]
    Cs = [c.clone() for c in C0s]
    plan = cc.GroupedGemm(As, Bs, Cs, mode=mode, alpha=1.5, beta=0.5, tile_cfgs=cfgs, splits=[1] * len(cfgs))
    assert [c for c, s in plan.choices()] == cfgs
    assert len(plan.info()["buckets"]) == len(cc.group_families(mode))
    plan.run()
    for A, B, C0, Cg, cfg in zip(As, Bs, C0s, Cs, cfgs):
        Cs1 = C0.clone()
        cc.gemm(A, B, Cs1, alpha=1.5, beta=0.5, mode=mode, tile_cfg=cfg, split_k=1)
        assert torch.equal(Cg, Cs1), (tuple(A.shape), tuple(B.shape), cfg)
    plan.close()


def test_autograd_fp64_gradcheck(cc, dev):
    X = torch.randn(12, 20, dtype=torch.float64, device=dev, requires_grad=True)
    W = torch.randn(20, 9, dtype=torch.float64, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x, w: cc.matmul(x, w), (X, W), eps=1e-6, atol=1e-8)


def test_autograd_central_differences_64(cc, dev):
    """64x64x64 slice: <dL/dX, V> and <dL/dW, U> against FP64 central
    differences of L = <X W, G> (fp64 mode), and bf16x3 gradients against
    torch's FP64 matmul gradient."""
    rs = np.random.RandomState(4)
    X = torch.from_numpy(rs.uniform(-1, 1, (64, 64))).to(dev).requires_grad_()
    W = torch.from_numpy(rs.normal(0, 1 / 8, (64, 64))).to(dev).requires_grad_()
    G = torch.from_numpy(rs.normal(0, 1, (64, 64))).to(dev)
    V = torch.from_numpy(rs.normal(0, 1, (64, 64))).to(dev)
    U = torch.from_numpy(rs.normal(0, 1, (64, 64))).to(dev)
    Y = cc.matmul(X, W)
    (Y * G).sum().backward()
    h = 1e-4

    def L(x, w):
        return float((cc.gemm(x.detach().contiguous(), w.detach().contiguous(), mode="fp64") * G).sum())

    fd_x = (L(X + h * V, W) - L(X - h * V, W)) / (2 * h)
    fd_w = (L(X, W + h * U) - L(X, W - h * U)) / (2 * h)
    assert abs(float((X.grad * V).sum()) - fd_x) <= 1e-8 * abs(fd_x) + 1e-9
    assert abs(float((W.grad * U).sum()) - fd_w) <= 1e-8 * abs(fd_w) + 1e-9

    # bf16x3 compute on FP32 storage vs torch FP64 gradient
    X32 = X.detach().float().requires_grad_()
    W32 = W.detach().float().requires_grad_()
    Y32 = cc.matmul(X32, W32)
    (Y32 * G.float()).sum().backward()
    Xr = X.detach().clone().requires_grad_()
    Wr = W.detach().clone().requires_grad_()
    (torch.matmul(Xr, Wr) * G).sum().backward()
    for got, ref, a, b in ((X32.grad, Xr.grad, G, W.detach().T), (W32.grad, Wr.grad, X.detach().T, G)):
        bound = (2.0**-15 + 2 * 64 * 2.0**-24) * (a.abs() @ b.abs()) + 1e-6
        assert torch.all((got.double() - ref).abs() <= bound)


def test_autograd_full_size_bf16x3(cc, dev):
    """X [16384 x 4096] uniform, W [4096 x 4096] normal(0, 1/64), dY normal:
    forward and both gradients against torch FP64 matmul, normwise."""
    g = torch.Generator(device="cpu").manual_seed(4)
    X = (torch.rand(16384, 4096, generator=g) * 2 - 1).to(dev)
    W = (torch.randn(4096, 4096, generator=g) / 64).to(dev)
    dY = torch.randn(16384, 4096, generator=g).to(dev)
    Xg = X.clone().requires_grad_()
    Wg = W.clone().requires_grad_()
    Y = cc.matmul(Xg, Wg)
    Y.backward(dY)
    torch.cuda.synchronize()
    X64, W64, dY64 = X.double(), W.double(), dY.double()
    checks = (
        (Y.detach(), X64 @ W64, X64, W64),
        (Xg.grad, dY64 @ W64.T.contiguous(), dY64, W64),
        (Wg.grad, X64.T.contiguous() @ dY64, X64, dY64),
    )
    for got, ref, a, b in checks:
        nw = float((got.double() - ref).norm() / (a.norm() * b.norm()))
        assert nw < 2e-6, nw
