"""Pins for the CPU oracle (oracle/oracle.c).

Each check is chosen so that a plausible mistake in the oracle fails it:
library routines (numpy's MT19937 streams, libstdc++'s std::mt19937 and
uniform_real_distribution, numpy matmul, torch's bfloat16 rounding), closed
forms, and planted cases whose float result depends on the summation order
the reference kernel uses.
"""
import os
import shutil
import struct
import subprocess
import tempfile

import numpy as np
import pytest

import oracle


# ---------------------------------------------------------------- generators

def test_mt19937_known_answer():
    # C++ standard [rand.predef]: the 10000th invocation of a default
    # constructed mt19937 (seed 5489) produces 4123659995.
    assert oracle.Rng(5489).u32(10000)[-1] == 4123659995


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 4, 12345, 2**32 - 1])
def test_streams_match_numpy_randomstate(seed):
    rs = np.random.RandomState(seed)
    r = oracle.Rng(seed)
    assert np.array_equal(rs.randint(0, 2**32, size=777, dtype=np.uint32), r.u32(777))
    assert np.array_equal(rs.random_sample(1001), r.random_sample(1001))
    assert np.array_equal(rs.uniform(-1, 1, 999).astype(np.float32), r.uniform_f32(999))
    assert np.array_equal(rs.normal(0, 1, 1001), r.normal_f64(1001))  # odd: leaves a cached deviate
    assert np.array_equal(rs.random_sample(5), r.random_sample(5))
    assert np.array_equal(rs.normal(0, 1, 6), r.normal_f64(6))


def test_workload_generators_match_numpy():
    A, B, C0 = oracle.workload_plain_1024(64, 48, 32, seed=0)
    rs = np.random.RandomState(0)
    assert np.array_equal(A, rs.uniform(-1, 1, (64, 32)).astype(np.float32))
    assert np.array_equal(B, rs.uniform(-1, 1, (32, 48)).astype(np.float32))
    assert np.array_equal(C0, rs.uniform(-1, 1, (64, 48)).astype(np.float32))
    A, B, C0 = oracle.workload_dgemm(40, 30, 20, seed=2)
    rs = np.random.RandomState(2)
    eA = rs.normal(0, 1, (40, 20))
    eA[rs.random_sample((40, 20)) < 0.01] *= 1e6
    eB = rs.normal(0, 1, (20, 30))
    eB[rs.random_sample((20, 30)) < 0.01] *= 1e6
    eC = rs.normal(0, 1, (40, 30))
    assert np.array_equal(A, eA) and np.array_equal(B, eB) and np.array_equal(C0, eC)


CPP_FILL = r"""
#include <cstdio>
#include <cstring>
#include <random>
int main(int argc, char** argv) {
  unsigned seed = (unsigned)atoi(argv[1]);
  int n = atoi(argv[2]);
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> uniform(0, 1);
  for (int i = 0; i < n; ++i) {
    float f = (float)uniform(rng);
    unsigned u; std::memcpy(&u, &f, 4);
    std::printf("%08x\n", u);
  }
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_ref_fill_random_matches_libstdcxx():
    """Matrix<float>::fillRandom (Kernel/cuda/src/Matrix.h:75-87) uses
    std::mt19937 + std::uniform_real_distribution<double>(0,1); compare with
    libstdc++ itself."""
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "fill.cpp")
        exe = os.path.join(td, "fill")
        with open(src, "w") as f:
            f.write(CPP_FILL)
        subprocess.run(["g++", "-O2", "-std=c++17", src, "-o", exe], check=True)
        for seed in (0, 1, 42):
            out = subprocess.run([exe, str(seed), "3000"], check=True, capture_output=True,
                                 text=True).stdout.split()
            want = np.array([int(h, 16) for h in out], np.uint32).view(np.float32)
            got = oracle.ref_fill_random(seed, 3000)
            assert np.array_equal(got, want)
            assert got.min() >= 0.0 and got.max() < 1.0 + 1e-7


# ---------------------------------------------------------------- FP64 / long double

def _int_mats(rng, M, N, K, lo=-16, hi=17):
    return (rng.integers(lo, hi, (M, K)).astype(np.float32),
            rng.integers(lo, hi, (K, N)).astype(np.float32),
            rng.integers(lo, hi, (M, N)).astype(np.float32))


@pytest.mark.parametrize("ta,tb", [(False, False), (False, True), (True, False), (True, True)])
def test_gemm_f64_exact_on_integers(ta, tb):
    rng = np.random.default_rng(5)
    A, B, C0 = _int_mats(rng, 37, 29, 53)
    want = 2 * (A.astype(np.int64) @ B.astype(np.int64)) - 3 * C0.astype(np.int64)
    Aa = np.ascontiguousarray(A.T) if ta else A
    Bb = np.ascontiguousarray(B.T) if tb else B
    got = oracle.gemm_f64(Aa, Bb, C0, alpha=2.0, beta=-3.0, ta=ta, tb=tb)
    assert np.array_equal(got, want.astype(np.float64))
    # row windows
    part = oracle.gemm_f64(Aa, Bb, C0, alpha=2.0, beta=-3.0, ta=ta, tb=tb, rows=5, row0=11)
    assert np.array_equal(part, want[11:16].astype(np.float64))


def test_gemm_f64_closed_forms():
    rng = np.random.default_rng(6)
    B = rng.uniform(-1, 1, (20, 30)).astype(np.float32)
    C0 = rng.uniform(-1, 1, (20, 30)).astype(np.float32)
    eye = np.eye(20, dtype=np.float32)
    # A = I  =>  C = alpha*B + beta*C0
    got = oracle.gemm_f64(eye, B, C0, alpha=1.5, beta=0.25)
    assert np.array_equal(got, 1.5 * B.astype(np.float64) + 0.25 * C0.astype(np.float64))
    # alpha = 0 => beta*C0 ; beta = 0 never reads C0 (NaN C0 ignored)
    got = oracle.gemm_f64(eye, B, C0, alpha=0.0, beta=0.5)
    assert np.array_equal(got, 0.5 * C0.astype(np.float64))
    nanC = np.full_like(C0, np.nan)
    assert np.isfinite(oracle.gemm_f64(eye, B, nanC, alpha=1.0, beta=0.0)).all()
    # K = 1 outer product
    a = rng.uniform(-1, 1, (7, 1)).astype(np.float32)
    b = rng.uniform(-1, 1, (1, 9)).astype(np.float32)
    assert np.array_equal(oracle.gemm_f64(a, b), a.astype(np.float64) @ b.astype(np.float64))


def test_gemm_f64_vs_numpy_float64():
    rng = np.random.default_rng(7)
    A = rng.uniform(-1, 1, (64, 300)).astype(np.float32)
    B = rng.uniform(-1, 1, (300, 50)).astype(np.float32)
    got = oracle.gemm_f64(A, B)
    want = A.astype(np.float64) @ B.astype(np.float64)
    assert np.max(np.abs(got - want)) < 1e-12


def test_gemm_ld_exact_and_compensated():
    rng = np.random.default_rng(8)
    A = rng.integers(-1000, 1000, (17, 23)).astype(np.float64)
    B = rng.integers(-1000, 1000, (23, 19)).astype(np.float64)
    C0 = rng.integers(-1000, 1000, (17, 19)).astype(np.float64)
    want = (A.astype(np.int64) @ B.astype(np.int64) + C0.astype(np.int64)).astype(np.float64)
    assert np.array_equal(oracle.gemm_ld(A, B, C0, 1.0, 1.0), want)
    # Planted cancellation: 1e16 + 1 - 1e16 = 1 exactly; a plain double sum gives 0.
    a = np.array([[1e16, 1.0, -1e16]])
    b = np.ones((3, 1))
    assert oracle.gemm_ld(a, b)[0, 0] == 1.0
    assert oracle.gemm_f64_d(a, b)[0, 0] == 0.0
    # Outliers at 1e6 scale and transposes against numpy float64
    A = rng.normal(0, 1, (30, 40))
    A[rng.random((30, 40)) < 0.05] *= 1e6
    B = rng.normal(0, 1, (40, 20))
    want = A @ B
    got = oracle.gemm_ld(np.ascontiguousarray(A.T), B, ta=True)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-3)


# ---------------------------------------------------------------- refsim

def test_refsim_exact_on_integers():
    rng = np.random.default_rng(9)
    A, B, C0 = _int_mats(rng, 33, 21, 70)
    want = A.astype(np.int64) @ B.astype(np.int64)
    for split, lk in ((1, 8), (3, 8), (5, 4), (1, 1)):
        got = oracle.refsim_sgemm(A, B, C0, 1.0, 0.0, split_k=split, load_k=lk)
        assert np.array_equal(got, want.astype(np.float32)), (split, lk)
    # beta = 0.5 with even C0 is exact
    C0e = 2 * C0
    got = oracle.refsim_sgemm(A, B, C0e, 1.0, 0.5)
    assert np.array_equal(got, (want + C0.astype(np.int64)).astype(np.float32))


def test_refsim_descending_chunk_order():
    """K = 16, LOAD_K = 8: the kernel walks chunk k = 8..15 first, then 0..7
    (Kernel/cuda/src/cuCOSMAV100.cuh:2623-2746).  With p_8 = 2^24 and
    p_0..p_7 = 1, that order gives 2^24 (each +1 rounds away); ascending order
    would give 2^24 + 8."""
    A = np.zeros((1, 16), np.float32)
    B = np.zeros((16, 1), np.float32)
    A[0, :8] = 1.0
    B[:8, 0] = 1.0
    A[0, 8] = 4096.0
    B[8, 0] = 4096.0
    got = oracle.refsim_sgemm(A, B, None, 1.0, 0.0, split_k=1, load_k=8)
    assert got[0, 0] == np.float32(2**24)
    # With LOAD_K = 16 the single chunk is ascending: 8 + 2^24.
    got = oracle.refsim_sgemm(A, B, None, 1.0, 0.0, split_k=1, load_k=16)
    assert got[0, 0] == np.float32(2**24 + 8)


def test_refsim_splitk_z_order():
    """Partials summed z = 0, 1, 2 (Kernel/cuda/src/cosmaSplitKReduce.cuh:33-36):
    P = (2^24, 1, 1) gives 2^24; the reverse order would give 2^24 + 2."""
    A = np.zeros((1, 12), np.float32)
    B = np.zeros((12, 1), np.float32)
    A[0, 0], B[0, 0] = 4096.0, 4096.0   # slice 0 (k 0..3)
    A[0, 4], B[4, 0] = 1.0, 1.0         # slice 1 (k 4..7)
    A[0, 8], B[8, 0] = 1.0, 1.0         # slice 2 (k 8..11)
    got = oracle.refsim_sgemm(A, B, None, 1.0, 0.0, split_k=3, load_k=4, tbk=4)
    assert got[0, 0] == np.float32(2**24)
    rev = oracle.refsim_sgemm(A[:, ::-1].copy(), B[::-1].copy(), None, 1.0, 0.0, split_k=3,
                              load_k=4, tbk=4)
    assert rev[0, 0] == np.float32(2**24 + 2)


def test_refsim_alpha_zero_and_beta_fusion():
    A = np.ones((2, 4), np.float32)
    B = np.ones((4, 3), np.float32)
    C0 = np.full((2, 3), 3.0, np.float32)
    assert np.array_equal(oracle.refsim_sgemm(A, B, C0, 0.0, 0.5), np.full((2, 3), 1.5, np.float32))
    # acc = 1 (one product), beta = 0.1f, c = 3: fused gives RN(1 + 0.1f*3) once.
    A = np.ones((1, 1), np.float32)
    B = np.ones((1, 1), np.float32)
    C0 = np.full((1, 1), 3.0, np.float32)
    b = np.float32(0.1)
    fused = np.float32(np.float64(b) * 3.0 + 1.0)  # product exact in double, sum exact too
    unfused = np.float32(np.float32(b * np.float32(3.0)) + np.float32(1.0))
    assert oracle.refsim_sgemm(A, B, C0, 1.0, float(b), fused_beta=True)[0, 0] == fused
    assert oracle.refsim_sgemm(A, B, C0, 1.0, float(b), fused_beta=False)[0, 0] == unfused


def test_refsim_within_error_bound_of_f64():
    A, B, C0 = oracle.workload_plain_1024(64, 64, 512, seed=0)
    rs = oracle.refsim_sgemm(A, B, C0, 1.0, 0.5)
    ref = oracle.gemm_f64(A, B, C0, 1.0, 0.5)
    absprod = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64)) + 0.5 * np.abs(C0)
    u = 2.0**-24
    assert np.all(np.abs(rs - ref) <= 513 * u * absprod)


# ---------------------------------------------------------------- cpu_naive

def test_cpu_naive_bitwise_vs_numpy_float32_loop():
    rng = np.random.default_rng(10)
    A = rng.uniform(0, 1, (3, 50)).astype(np.float32)
    B = rng.uniform(0, 1, (50, 4)).astype(np.float32)
    want = np.zeros((3, 4), np.float32)
    for i in range(3):
        for j in range(4):
            acc = np.float32(0)
            for k in range(50):
                acc = np.float32(acc + np.float32(A[i, k] * B[k, j]))
            want[i, j] = acc
    assert np.array_equal(oracle.cpu_naive_sgemm(A, B), want)


def test_cpu_naive_threads_bit_identical():
    A, B = oracle.workload_square_flat(40, 64, 300, seed=1)
    one = oracle.cpu_naive_sgemm(A, B, nthreads=1)
    many = oracle.cpu_naive_sgemm(A, B, nthreads=4)
    part = oracle.cpu_naive_sgemm(A, B, rows=7, row0=13, nthreads=3)
    assert np.array_equal(one, many) and np.array_equal(one[13:20], part)


# ---------------------------------------------------------------- bf16 split

def test_bf16_rne_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    x = np.concatenate([
        rng.normal(0, 1, 20000).astype(np.float32),
        rng.uniform(-1e30, 1e30, 2000).astype(np.float32),
        # exact ties: low 16 bits == 0x8000, both parities of bit 16
        (np.arange(0x3f800000, 0x3f800000 + 512 * 0x10000, 0x10000, dtype=np.uint32) | 0x8000)
        .view(np.float32),
        np.array([0.0, -0.0, 1e-40, -1e-40, 3.4e38, -3.4e38], np.float32),
    ])
    hi, _ = oracle.split_bf16(x)
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(hi, want)


def test_bf16_split_closed_forms():
    rng = np.random.default_rng(12)
    # <= 8 significant bits: hi == x, lo == 0
    x8 = (rng.integers(-255, 256, 1000) * 2.0 ** rng.integers(-20, 20, 1000)).astype(np.float32)
    hi, lo = oracle.split_bf16(x8)
    assert np.array_equal(oracle.bf16_bits_to_f32(hi), x8) and np.all(oracle.bf16_bits_to_f32(lo) == 0)
    # <= 16 significant bits: hi + lo == x exactly
    x16 = (rng.integers(-65535, 65536, 1000) * 2.0 ** rng.integers(-20, 20, 1000)).astype(np.float32)
    hi, lo = oracle.split_bf16(x16)
    s = oracle.bf16_bits_to_f32(hi).astype(np.float64) + oracle.bf16_bits_to_f32(lo)
    assert np.array_equal(s, x16.astype(np.float64))
    # general floats: |x - hi - lo| <= 2^-17 |x|
    x = rng.uniform(-1, 1, 100000).astype(np.float32)
    hi, lo = oracle.split_bf16(x)
    r = x.astype(np.float64) - oracle.bf16_bits_to_f32(hi) - oracle.bf16_bits_to_f32(lo)
    assert np.all(np.abs(r) <= 2.0**-17 * np.abs(x))


def test_gemm_bf16x3_oracle():
    rng = np.random.default_rng(13)
    # 8-bit entries: the split is exact, so the three-product sum is the product.
    A = rng.integers(-128, 128, (20, 40)).astype(np.float32)
    B = rng.integers(-128, 128, (40, 30)).astype(np.float32)
    assert np.array_equal(oracle.gemm_bf16x3(A, B), (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64))
    # Random: within the dropped lo*lo term (<= 2^-16 |a||b|) plus the split residue.
    A, B = oracle.workload_square_flat(32, 32, 256, seed=1)
    got = oracle.gemm_bf16x3(A, B)
    ref = oracle.gemm_f64(A, B)
    absprod = np.abs(A.astype(np.float64)) @ np.abs(B.astype(np.float64))
    assert np.all(np.abs(got - ref) <= 2.0**-15 * absprod)
    # and it is materially better than hi*hi alone (guards a dropped cross term)
    hiA = oracle.bf16_bits_to_f32(oracle.split_bf16(A)[0])
    hiB = oracle.bf16_bits_to_f32(oracle.split_bf16(B)[0])
    hh = hiA.astype(np.float64) @ hiB.astype(np.float64)
    assert np.abs(hh - ref).max() > 50 * np.abs(got - ref).max()
    # four products: exactly (ah+al)(bh+bl); strictly closer to FP64 in norm,
    # and differs from the three-product sum by exactly al*bl
    got4 = oracle.gemm_bf16x4(A, B)
    loA = oracle.bf16_bits_to_f32(oracle.split_bf16(A)[1]).astype(np.float64)
    loB = oracle.bf16_bits_to_f32(oracle.split_bf16(B)[1]).astype(np.float64)
    # (differences of two double sums of 256 terms: absolute tolerance)
    assert np.allclose(got4 - got, loA @ loB, rtol=0, atol=1e-10)
    assert np.linalg.norm(got4 - ref) < np.linalg.norm(got - ref)
    full = (hiA.astype(np.float64) + loA) @ (hiB.astype(np.float64) + loB)
    assert np.allclose(got4, full, rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- metrics

def test_err_stats_vs_numpy():
    rng = np.random.default_rng(14)
    R = rng.normal(0, 1, (30, 40))
    R[0, 0] = 0.0
    Cm = (R + rng.normal(0, 1e-6, R.shape)).astype(np.float32)
    st = oracle.err_stats(Cm, R)
    d = Cm.astype(np.float64) - R
    assert np.isclose(st["sum_sq_err"], np.linalg.norm(d) ** 2, rtol=1e-12)
    assert np.isclose(st["sum_sq_ref"], np.linalg.norm(R) ** 2, rtol=1e-12)
    assert st["max_abs"] == np.abs(d).max()
    nz = R != 0
    assert st["max_rel"] == (np.abs(d)[nz] / np.abs(R)[nz]).max()
    A = rng.normal(0, 1, (5, 6)).astype(np.float32)
    assert np.isclose(oracle.sumsq(A), np.linalg.norm(A.astype(np.float64)) ** 2, rtol=1e-12)


# ---------------------------------------------------------------- epilogue

def test_activation_relu_closed_forms():
    x = np.array([-2.5, -0.0, 0.0, 3.0, np.nan, -np.inf, np.inf, 1e-300, -1e-300])
    got = oracle.activate(x, "relu")
    want = np.array([0.0, 0.0, 0.0, 3.0, 0.0, 0.0, np.inf, 1e-300, 0.0])
    assert np.array_equal(got, want)
    assert np.array_equal(oracle.activate(x[:4], None), x[:4])


def test_activation_sigmoid_vs_library_and_identities():
    special = pytest.importorskip("scipy.special")
    rng = np.random.default_rng(15)
    x = np.concatenate([rng.normal(0, 8, 4000), [0.0, 1.0, -1.0, 30.0, -30.0, 700.0, -700.0]])
    got = oracle.activate(x, "sigmoid")
    ref = special.expit(x)
    assert np.allclose(got, ref, rtol=4e-16, atol=0)
    assert oracle.activate([0.0], "sigmoid")[0] == 0.5
    # s(x) + s(-x) = 1 and s'(0) = 1/4 (a sign or a dropped exp would fail one of these)
    assert np.allclose(got + oracle.activate(-x, "sigmoid"), 1.0, rtol=0, atol=5e-16)
    h = 1e-6
    d = (oracle.activate([h], "sigmoid")[0] - oracle.activate([-h], "sigmoid")[0]) / (2 * h)
    assert abs(d - 0.25) < 1e-9
