"""Parity at the full benchmark sizes.

The host oracle finishes only row samples at these sizes, so each shape is
held three ways:
  * sampled rows against the host oracle, componentwise, with the bound of
    the arithmetic used (FP32 accumulation: 2*K*2^-24*(|A||B|), plus
    2^-15*(|A||B|) for the bf16 split; FP64: 4*K*2^-53*(|A||B| + |C|));
  * whole-matrix normwise error ||C - R||_F / (||A||_F ||B||_F) against the
    GPU FP64 product R, which is itself checked against the host oracle on
    the same rows;
  * size-independent properties: bitwise repeatability of every launch, and
    agreement between two split-K factors within the FP32 bound;
  * BF16x3 against the exact sum of its three products (FP64 on the GPU):
    its own error, beyond what the split drops, within 1.5x of plain FP32's;
  * the compile-time specialisations of each shape, called through the C
    ABI, bitwise equal to the runtime-shape call.
Every product here goes through the library's own choice of tile and split
(the kernels the benchmarks time) unless a split is named.
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


def _absprod(A, B, rows):
    return np.abs(A[list(rows)].astype(np.float64)) @ np.abs(B.astype(np.float64))


def _check_rows(got, absprod, K, rows, ref_rows, u, extra):
    bound = (2 * K * u + extra) * absprod
    diff = np.abs(got[list(rows)].astype(np.float64) - ref_rows)
    assert np.all(diff <= bound + 1e-30), float((diff - bound).max())


def _fp64_reference_checked(cc, Ad, Bd, A, B, rows):
    """GPU FP64 product, checked against the host FP64 oracle on `rows`;
    returns it with the host rows and (|A||B|)[rows]."""
    R = cc.fp64_product(Ad, Bd)
    host = np.concatenate([oracle.gemm_f64(A, B, rows=1, row0=r) for r in rows])
    absprod = _absprod(A, B, rows)
    Rr = R[list(rows)].cpu().numpy()
    K = A.shape[1]
    assert np.all(np.abs(Rr - host) <= 2 * K * 2.0**-53 * absprod + 1e-300)
    return R, host, absprod


def _exact_split_reference(cc, Ad, Bd):
    """R3 = (Ah + Al)(Bh + Bl) - Al Bl = Ah Bh + Ah Bl + Al Bh in FP64 on the
    GPU: the exact sum of the three products BF16x3 computes, with hi =
    RNE(x) and lo = RNE(x - hi) (torch's float -> bfloat16 conversion rounds
    to nearest even; the oracle's split is pinned against it)."""
    def split(X):
        hi = X.to(torch.bfloat16)
        lo = (X - hi.float()).to(torch.bfloat16)
        return hi.double(), lo.double()

    Ah, Al = split(Ad)
    Bh, Bl = split(Bd)
    full = cc.gemm(Ah + Al, Bh + Bl, mode="fp64")
    del Ah, Bh
    full -= cc.gemm(Al, Bl, mode="fp64")
    return full


def _call_fixed(cc, name, alpha, A, lda, B, ldb, beta, Cm, ldc):
    st = getattr(cc.lib, name)(alpha, A.data_ptr(), lda, B.data_ptr(), ldb, beta, Cm.data_ptr(), ldc)
    assert st == 0, (name, st)
    torch.cuda.synchronize()


def test_square_8192(cc, dev):
    """8192^3, A/B uniform[-1,1] from MT19937 seed 1, beta 0: FP32, BF16x3, BF16x4."""
    M = N = K = 8192
    A, B = oracle.workload_square_flat(M, N, K, seed=1)
    Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    rows = (0, 1, 4097, 8191)
    R, host, absprod = _fp64_reference_checked(cc, Ad, Bd, A, B, rows)
    nw = {}
    for mode in ("fp32", "bf16x3", "bf16x4"):
        Cm = cc.gemm(Ad, Bd, mode=mode)
        again = cc.gemm(Ad, Bd, mode=mode)
        assert torch.equal(Cm, again), mode
        got = Cm.cpu().numpy()
        _check_rows(got, absprod, K, rows, host, 2.0**-24, 2.0**-15 if mode != "fp32" else 0.0)
        if mode == "bf16x3":  # the exact three-product sum, FP32 accumulation only
            exact3 = np.concatenate([oracle.gemm_bf16x3(A, B, rows=1, row0=r) for r in rows])
            _check_rows(got, absprod, K, rows, exact3, 2.0**-24, 0.0)
        nw[mode] = cc.normwise_error(Ad, Bd, Cm, R)["normwise"]
        # the fixed-shape specialisation computes the same bits
        fixed = {"fp32": "cosma_sgemm_8192x8192x8192", "bf16x3": "cosma_sgemm_bf16x3_8192x8192x8192"}
        if mode in fixed:
            Cf = torch.full_like(Cm, float("nan"))
            _call_fixed(cc, fixed[mode], 1.0, Ad, K, Bd, N, 0.0, Cf, N)
            assert torch.equal(Cf, Cm), mode
            del Cf
        if mode == "bf16x3":
            Cb = Cm
        del again
    assert nw["fp32"] <= 1e-6, nw
    # The four-product split is within 2x of plain FP32's error.
    assert nw["bf16x4"] <= 2.0 * nw["fp32"], nw
    # BF16x3's own error: its distance from the exact three-product sum (the
    # split's truncation removed) is FP32 accumulation error, within 1.5x of
    # plain FP32's normwise error; the split itself (lo*lo and the residual
    # terms dropped) is what puts its error vs FP64 at about 2.4x.
    del R
    R3 = _exact_split_reference(cc, Ad, Bd)
    own = cc.normwise_error(Ad, Bd, Cb, R3)["normwise"]
    assert own <= 1.5 * nw["fp32"], (own, nw)


def test_flat_256x256x1M(cc, dev):
    """256 x 256 x 1,048,576 (seed 1, beta 0): the deep split-K shape."""
    M, N, K = 256, 256, 1 << 20
    A, B = oracle.workload_square_flat(M, N, K, seed=1)
    Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    rows = (0, 255)
    R, host, absprod = _fp64_reference_checked(cc, Ad, Bd, A, B, rows)
    absAB = cc.fp64_product(Ad.abs(), Bd.abs()).cpu().numpy()  # |A||B| for every element
    nw = {}
    for mode in ("fp32", "bf16x3", "bf16x4"):
        cfg, split = cc.heuristic(cc._lib.MODES[mode], M, N, K)
        assert split > 1
        Cm = cc.gemm(Ad, Bd, mode=mode)
        assert torch.equal(Cm, cc.gemm(Ad, Bd, mode=mode)), mode
        got = Cm.cpu().numpy()
        extra = 2.0**-15 if mode != "fp32" else 0.0
        _check_rows(got, absprod, K, rows, host, 2.0**-24, extra)
        # another split factor: the same split products, other FP32 sums
        other = cc.gemm(Ad, Bd, mode=mode, tile_cfg=cfg, split_k=split // 4).cpu().numpy()
        bound = 4 * K * 2.0**-24 * absAB
        assert np.all(np.abs(other.astype(np.float64) - got) <= bound), mode
        nw[mode] = cc.normwise_error(Ad, Bd, Cm, R)["normwise"]
        fixed = {"fp32": "cosma_sgemm_256x256x1048576", "bf16x3": "cosma_sgemm_bf16x3_256x256x1048576"}
        if mode in fixed:
            Cf = torch.full_like(Cm, float("nan"))
            _call_fixed(cc, fixed[mode], 1.0, Ad, K, Bd, N, 0.0, Cf, N)
            assert torch.equal(Cf, Cm), mode
        if mode == "bf16x3":
            Cb = Cm
    assert nw["fp32"] <= 1e-6, nw
    assert nw["bf16x4"] <= 2.0 * nw["fp32"], nw
    R3 = _exact_split_reference(cc, Ad, Bd)
    own = cc.normwise_error(Ad, Bd, Cb, R3)["normwise"]
    assert own <= 1.5 * nw["fp32"], (own, nw)


@pytest.mark.parametrize("shape", [(32768, 1024, 1024), (1024, 65536, 1024)])
def test_dgemm_full(cc, dev, shape):
    """DGEMM at the benchmark shapes: normal(0,1) seed 2 with 1% of entries
    scaled by 1e6, C normal(0,1), alpha = beta = 1, against the long-double
    compensated oracle on sampled rows."""
    M, N, K = shape
    A, B, C0 = oracle.workload_dgemm(M, N, K)
    Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    C0d = torch.from_numpy(C0).to(dev)
    Cd = C0d.clone()
    cc.gemm(Ad, Bd, Cd, alpha=1.0, beta=1.0, mode="fp64")
    again = C0d.clone()
    cc.gemm(Ad, Bd, again, alpha=1.0, beta=1.0, mode="fp64")
    assert torch.equal(Cd, again)
    rows = (0, M // 2 + 1, M - 1)
    ref = np.concatenate([oracle.gemm_ld(A, B, C0, 1.0, 1.0, rows=1, row0=r) for r in rows])
    got = Cd[list(rows)].cpu().numpy()
    bound = 4 * K * 2.0**-53 * (np.abs(A[list(rows)]) @ np.abs(B) + np.abs(C0[list(rows)]))
    assert np.all(np.abs(got - ref) <= bound), float((np.abs(got - ref) - bound).max())
    # the compile-time specialisation of this shape computes the same bits
    Cf = C0d.clone()
    _call_fixed(cc, f"cosma_dgemm_{M}x{N}x{K}", 1.0, Ad, K, Bd, N, 1.0, Cf, N)
    assert torch.equal(Cf, Cd)


def test_fixed_shape_dgemm_1024(cc, dev):
    """cosma_dgemm: the build-time shape (1024^3 by default), DGEMM workload
    inputs, against the long-double oracle on sampled rows."""
    m, n, k = C.c_int(), C.c_int(), C.c_int()
    cc.lib.cosma_fixed_shape(C.byref(m), C.byref(n), C.byref(k))
    M, N, K = m.value, n.value, k.value
    A, B, C0 = oracle.workload_dgemm(M, N, K)
    Ad, Bd = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    Cd = torch.from_numpy(C0).to(dev)
    _call_fixed(cc, "cosma_dgemm", 1.0, Ad, K, Bd, N, 1.0, Cd, N)
    rows = (0, M // 3, M - 1)
    ref = np.concatenate([oracle.gemm_ld(A, B, C0, 1.0, 1.0, rows=1, row0=r) for r in rows])
    got = Cd[list(rows)].cpu().numpy()
    bound = 4 * K * 2.0**-53 * (np.abs(A[list(rows)]) @ np.abs(B) + np.abs(C0[list(rows)]))
    assert np.all(np.abs(got - ref) <= bound)
    again = torch.from_numpy(C0).to(dev)
    cc.gemm(Ad, Bd, again, alpha=1.0, beta=1.0, mode="fp64")
    assert torch.equal(again, Cd)
