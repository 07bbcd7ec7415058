"""cosma_gemm_ex_static -- the entry point for callers that cannot allocate
-- against cosma_gemm_ex, through the C ABI (GPU).

Its split-K partials and pre-split planes live in a static device array;
the kernels are those of cosma_gemm_ex, so results are bitwise equal.  Two
calls in a row, the second smaller, share the array; a call that needs more
than the array holds is refused; a BF16 product whose planes do not fit runs
the on-the-fly tile, bitwise the pre-split result.
"""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def cc(dev):
    import cucosma_amd

    return cucosma_amd


def _desc(cc, mode, A, B, Cm, alpha=1.0, beta=0.0, split_k=0, workspace=None):
    from cucosma_amd._lib import MODES, GemmDesc

    d = GemmDesc()
    cc.lib.cosma_gemm_desc_init(C.byref(d))
    d.mode = MODES[mode]
    d.m, d.k = A.shape
    d.n = B.shape[1]
    d.a, d.lda = A.data_ptr(), A.shape[1]
    d.b, d.ldb = B.data_ptr(), B.shape[1]
    d.c, d.ldc = Cm.data_ptr(), Cm.shape[1]
    d.alpha, d.beta = alpha, beta
    d.split_k = split_k
    if workspace is not None:
        d.workspace, d.workspace_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
    d.stream = torch.cuda.current_stream().cuda_stream
    return d


def _both(cc, mode, m, n, k, seed, beta=0.0):
    g = torch.Generator(device="cuda").manual_seed(seed)
    A = torch.rand((m, k), device="cuda", generator=g) * 2 - 1
    B = torch.rand((k, n), device="cuda", generator=g) * 2 - 1
    C0 = torch.rand((m, n), device="cuda", generator=g) * 2 - 1
    ref, got = C0.clone(), C0.clone()
    assert cc.lib.cosma_gemm_ex(C.byref(_desc(cc, mode, A, B, ref, 1.0, beta))) == 0
    assert cc.lib.cosma_gemm_ex_static(C.byref(_desc(cc, mode, A, B, got, 1.0, beta))) == 0
    torch.cuda.synchronize()
    return ref, got


def test_static_scratch_matches_gemm_ex(cc, dev):
    assert cc.lib.cosma_static_scratch_bytes() == 128 << 20
    # FP32 deep K: split-K partials in the array (32 slices of 512x512 = 32 MiB),
    # then a smaller split call reusing it.
    for (mode, m, n, k) in (("fp32", 512, 512, 65536), ("fp32", 128, 256, 16384)):
        cfg, split = cc.heuristic(cc._lib.MODES[mode], m, n, k)
        assert split > 1, (m, n, k)
        ref, got = _both(cc, mode, m, n, k, seed=m + k, beta=0.5)
        assert torch.equal(ref, got), (mode, m, n, k)
    # BF16x4 3072^2x4096: the pre-split 256x256 tile with K split for two
    # rounds of workgroups.  Planes and partials (96 + 108 MiB) exceed the
    # array, so the static entry runs the on-the-fly tile at the same split
    # (3 x 36 MiB of partials, more than half the array): bitwise the same.
    m, n, k = 3072, 3072, 4096
    assert cc.heuristic(cc.MODE_BF16X4, m, n, k) == (8, 3)
    ref, got = _both(cc, "bf16x4", m, n, k, seed=5, beta=0.5)
    assert torch.equal(ref, got)
    # BF16x3 on pre-split operands: planes in the array (61 MiB), then a
    # smaller pre-split call (24 MiB).
    for (m, n, k) in ((3900, 3900, 2048), (3072, 3072, 1024)):
        tiles = cc.tile_configs(cc.MODE_BF16X3)
        assert tiles[cc.heuristic(cc.MODE_BF16X3, m, n, k)[0]]["pipe"] == 5
        ref, got = _both(cc, "bf16x3", m, n, k, seed=m)
        assert torch.equal(ref, got), (m, n, k)


def test_static_scratch_limits(cc, dev):
    # Planes that do not fit (8192x8192x4096: 256 MiB): the on-the-fly tile,
    # bitwise the pre-split result of cosma_gemm_ex.
    ref, got = _both(cc, "bf16x3", 8192, 8192, 4096, seed=3)
    assert torch.equal(ref, got)
    # Partials that do not fit: refused before any launch, C untouched.
    A = torch.rand((512, 4096), device="cuda")
    B = torch.rand((4096, 512), device="cuda")
    Cm = torch.full((512, 512), 7.0, device="cuda")
    assert cc.lib.cosma_gemm_ex_static(C.byref(_desc(cc, "fp32", A, B, Cm, split_k=256))) == -5
    ws = torch.empty(1 << 20, dtype=torch.uint8, device="cuda")
    assert cc.lib.cosma_gemm_ex_static(C.byref(_desc(cc, "fp32", A, B, Cm, split_k=2, workspace=ws))) == -1
    torch.cuda.synchronize()
    assert bool((Cm == 7.0).all())
    # and a fitting call after the refusals still matches
    ref, got = _both(cc, "fp32", 256, 256, 8192, seed=9)
    assert torch.equal(ref, got)
