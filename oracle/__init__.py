"""Python access to the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package
(cucosma_amd) never imports it.

Every function here is a thin numpy wrapper over liboracle.so; the C file's
header lists the reference file:line each routine restates and how each one
is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    res = subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB_PATH


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, i, i64, d, f = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_float
    u32 = C.c_uint32
    sig = {
        "orc_rng_new": (vp, [u32]),
        "orc_rng_free": (None, [vp]),
        "orc_rng_u32": (None, [vp, i64, vp]),
        "orc_rng_random_sample": (None, [vp, i64, vp]),
        "orc_rng_uniform_f32": (None, [vp, i64, d, d, vp]),
        "orc_rng_uniform_f64": (None, [vp, i64, d, d, vp]),
        "orc_rng_normal_f64": (None, [vp, i64, d, d, vp]),
        "orc_rng_normal_f32": (None, [vp, i64, d, d, vp]),
        "orc_rng_scale_outliers_f64": (None, [vp, i64, d, d, vp]),
        "orc_ref_fill_random": (None, [u32, i64, vp]),
        "orc_gemm_f64": (None, [i, i, i, d, vp, i, i, vp, i, i, d, vp, i, vp, i, i, i]),
        "orc_gemm_f64_d": (None, [i, i, i, d, vp, i, i, vp, i, i, d, vp, i, vp, i, i, i]),
        "orc_gemm_ld": (None, [i, i, i, d, vp, i, i, vp, i, i, d, vp, i, vp, i, i, i]),
        "orc_gemm_bf16x3_f64": (None, [i, i, i, d, vp, i, i, vp, i, i, d, vp, i, vp, i, i, i]),
        "orc_gemm_bf16x4_f64": (None, [i, i, i, d, vp, i, i, vp, i, i, d, vp, i, vp, i, i, i]),
        "orc_refsim_sgemm": (None, [i, i, i, f, f, vp, i, vp, i, vp, i, i, i, i, i, i]),
        "orc_cpu_naive_sgemm": (None, [i, i, i, vp, vp, vp, i, i, i]),
        "orc_omp_max_threads": (i, []),
        "orc_bf16_rne": (C.c_uint16, [f]),
        "orc_bf16_to_f32": (f, [C.c_uint16]),
        "orc_split_bf16": (None, [i64, vp, vp, vp]),
        "orc_err_f32": (None, [i, i, vp, i, vp, i, vp]),
        "orc_err_f64": (None, [i, i, vp, i, vp, i, vp]),
        "orc_sumsq_f32": (d, [i64, vp]),
        "orc_sumsq_f64": (d, [i64, vp]),
        "orc_activate_f64": (None, [i64, vp, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] or a.ndim == 0
    return a.ctypes.data_as(C.c_void_p)


class Rng:
    """One MT19937 stream (numpy.random.RandomState(seed) compatible)."""

    def __init__(self, seed: int):
        self._h = lib.orc_rng_new(seed & 0xFFFFFFFF)

    def __del__(self):
        if getattr(self, "_h", None):
            lib.orc_rng_free(self._h)
            self._h = None

    def u32(self, n):
        out = np.empty(n, np.uint32)
        lib.orc_rng_u32(self._h, n, _p(out))
        return out

    def random_sample(self, n):
        out = np.empty(n, np.float64)
        lib.orc_rng_random_sample(self._h, n, _p(out))
        return out

    def uniform_f32(self, shape, lo=-1.0, hi=1.0):
        out = np.empty(shape, np.float32)
        lib.orc_rng_uniform_f32(self._h, out.size, lo, hi, _p(out))
        return out

    def uniform_f64(self, shape, lo=-1.0, hi=1.0):
        out = np.empty(shape, np.float64)
        lib.orc_rng_uniform_f64(self._h, out.size, lo, hi, _p(out))
        return out

    def normal_f64(self, shape, mean=0.0, std=1.0):
        out = np.empty(shape, np.float64)
        lib.orc_rng_normal_f64(self._h, out.size, mean, std, _p(out))
        return out

    def normal_f32(self, shape, mean=0.0, std=1.0):
        out = np.empty(shape, np.float32)
        lib.orc_rng_normal_f32(self._h, out.size, mean, std, _p(out))
        return out

    def scale_outliers_f64(self, x, p=0.01, scale=1e6):
        assert x.dtype == np.float64 and x.flags["C_CONTIGUOUS"]
        lib.orc_rng_scale_outliers_f64(self._h, x.size, p, scale, _p(x))
        return x


def ref_fill_random(seed: int, shape) -> np.ndarray:
    """Matrix<float>::fillRandom(seed) (Kernel/cuda/src/Matrix.h:75-87)."""
    out = np.empty(shape, np.float32)
    lib.orc_ref_fill_random(seed, out.size, _p(out))
    return out


def _gemm(fn, dt, A, B, C0, alpha, beta, ta, tb, rows, row0):
    A = np.ascontiguousarray(A, dt)
    B = np.ascontiguousarray(B, dt)
    M, K = (A.shape[1], A.shape[0]) if ta else A.shape
    N = B.shape[0] if tb else B.shape[1]
    rows = M - row0 if rows is None else rows
    out = np.empty((rows, N), np.float64)
    if beta != 0.0:
        C0 = np.ascontiguousarray(C0, dt)
        cp, ldc = _p(C0), C0.shape[1]
    else:
        cp, ldc = None, N
    fn(M, N, K, alpha, _p(A), A.shape[1], int(ta), _p(B), B.shape[1], int(tb), beta, cp, ldc,
       _p(out), N, row0, rows)
    return out


def gemm_f64(A, B, C0=None, alpha=1.0, beta=0.0, ta=False, tb=False, rows=None, row0=0):
    """FP64 oracle of float32 operands: alpha*op(A)op(B) + beta*C0, rows [row0, row0+rows)."""
    return _gemm(lib.orc_gemm_f64, np.float32, A, B, C0, alpha, beta, ta, tb, rows, row0)


def gemm_f64_d(A, B, C0=None, alpha=1.0, beta=0.0, ta=False, tb=False, rows=None, row0=0):
    """FP64 oracle of float64 operands."""
    return _gemm(lib.orc_gemm_f64_d, np.float64, A, B, C0, alpha, beta, ta, tb, rows, row0)


def gemm_ld(A, B, C0=None, alpha=1.0, beta=0.0, ta=False, tb=False, rows=None, row0=0):
    """Long-double compensated oracle of float64 operands."""
    return _gemm(lib.orc_gemm_ld, np.float64, A, B, C0, alpha, beta, ta, tb, rows, row0)


def gemm_bf16x3(A, B, C0=None, alpha=1.0, beta=0.0, ta=False, tb=False, rows=None, row0=0):
    """What the three-product bf16 split computes, with exact (double) accumulation."""
    return _gemm(lib.orc_gemm_bf16x3_f64, np.float32, A, B, C0, alpha, beta, ta, tb, rows, row0)


def gemm_bf16x4(A, B, C0=None, alpha=1.0, beta=0.0, ta=False, tb=False, rows=None, row0=0):
    """The four-product bf16 split (adds lo*lo), with exact (double) accumulation."""
    return _gemm(lib.orc_gemm_bf16x4_f64, np.float32, A, B, C0, alpha, beta, ta, tb, rows, row0)


def refsim_sgemm(A, B, C0, alpha, beta, split_k=1, load_k=8, tbk=None, fused_beta=True,
                 beta_is_double=False):
    """The reference kernel's per-element arithmetic order (row-major only)."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    M, K = A.shape
    N = B.shape[1]
    if tbk is None:
        if split_k == 1:
            tbk = K
        else:  # ceil(K / split) rounded up to a multiple of 4, as ScheduleGenerator emits
            tbk = -(-K // split_k)
            tbk = -(-tbk // 4) * 4
    Cm = np.array(C0, np.float32, copy=True) if C0 is not None else np.zeros((M, N), np.float32)
    lib.orc_refsim_sgemm(M, N, K, alpha, beta, _p(A), K, _p(B), N, _p(Cm), N, split_k, load_k,
                         tbk, int(fused_beta), int(beta_is_double))
    return Cm


def cpu_naive_sgemm(A, B, rows=None, row0=0, nthreads=1):
    """Matrix<float>::operator* (Kernel/cuda/src/Matrix.h:183-201) on rows [row0, row0+rows)."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    M, K = A.shape
    N = B.shape[1]
    rows = M - row0 if rows is None else rows
    out = np.empty((rows, N), np.float32)
    lib.orc_cpu_naive_sgemm(M, N, K, _p(A), _p(B), _p(out), row0, rows, nthreads)
    return out


def omp_max_threads() -> int:
    return lib.orc_omp_max_threads()


def bf16_rne(x: float) -> int:
    return lib.orc_bf16_rne(x)


def split_bf16(x: np.ndarray):
    """(hi, lo) as uint16 bf16 bit patterns."""
    x = np.ascontiguousarray(x, np.float32)
    hi = np.empty(x.shape, np.uint16)
    lo = np.empty(x.shape, np.uint16)
    lib.orc_split_bf16(x.size, _p(x), _p(hi), _p(lo))
    return hi, lo


def bf16_bits_to_f32(h: np.ndarray) -> np.ndarray:
    return (np.asarray(h, np.uint32) << 16).view(np.float32)


def err_stats(Cm: np.ndarray, R: np.ndarray) -> dict:
    """[sum (C-R)^2, sum R^2, max|C-R|, max rel, max|R|] on the host."""
    R = np.ascontiguousarray(R, np.float64)
    out = np.zeros(5, np.float64)
    if Cm.dtype == np.float32:
        Cm = np.ascontiguousarray(Cm)
        lib.orc_err_f32(R.shape[0], R.shape[1], _p(Cm), Cm.shape[1], _p(R), R.shape[1], _p(out))
    else:
        Cm = np.ascontiguousarray(Cm, np.float64)
        lib.orc_err_f64(R.shape[0], R.shape[1], _p(Cm), Cm.shape[1], _p(R), R.shape[1], _p(out))
    return dict(sum_sq_err=out[0], sum_sq_ref=out[1], max_abs=out[2], max_rel=out[3],
                max_ref=out[4])


ACTIVATIONS = {None: 0, "none": 0, "relu": 1, "sigmoid": 2}


def activate(x, act) -> np.ndarray:
    """ReLU / sigmoid of the fused epilogue, computed in double (copy of x)."""
    y = np.array(x, dtype=np.float64, copy=True, order="C")
    lib.orc_activate_f64(y.size, _p(y), ACTIVATIONS[act])
    return y


def sumsq(x: np.ndarray) -> float:
    x = np.ascontiguousarray(x)
    if x.dtype == np.float32:
        return lib.orc_sumsq_f32(x.size, _p(x))
    return lib.orc_sumsq_f64(x.size, _p(np.ascontiguousarray(x, np.float64)))


def normwise(Cm, R, A, B) -> float:
    """||C - R||_F / (||A||_F ||B||_F)."""
    st = err_stats(Cm, R)
    return float(np.sqrt(st["sum_sq_err"]) / np.sqrt(sumsq(A) * sumsq(B)))


# --------------------------------------------------------------------------
# Workload generators (one MT19937 stream per workload, as the benchmark
# configurations define them).
# --------------------------------------------------------------------------

def workload_plain_1024(M=1024, N=1024, K=1024, seed=0):
    """A, B, C0 uniform[-1,1] from one stream, in that order."""
    r = Rng(seed)
    A = r.uniform_f32((M, K))
    B = r.uniform_f32((K, N))
    C0 = r.uniform_f32((M, N))
    return A, B, C0


def workload_square_flat(M, N, K, seed=1):
    """A, B uniform[-1,1] from one stream (beta = 0, no C)."""
    r = Rng(seed)
    return r.uniform_f32((M, K)), r.uniform_f32((K, N))


def workload_dgemm(M, N, K, seed=2, p=0.01, scale=1e6):
    """A, B normal(0,1) each followed by its 1% outlier mask, then C normal(0,1)."""
    r = Rng(seed)
    A = r.normal_f64((M, K))
    r.scale_outliers_f64(A, p, scale)
    B = r.normal_f64((K, N))
    r.scale_outliers_f64(B, p, scale)
    C0 = r.normal_f64((M, N))
    return A, B, C0
