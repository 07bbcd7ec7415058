/*
 * oracle.c -- CPU oracle for the COSMA-tiled GEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  It is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker (or, for cpu_naive, as
 * the timed CPU baseline), never as the thing that is measured or shipped.
 *
 * What it restates (reference paths relative to the cuCOSMA source tree):
 *
 *   orc_refsim_sgemm   per-element arithmetic order of cosmaSgemm_kernel:
 *                      K slice of blockIdx.z            Kernel/cuda/src/cuCOSMAV100.cuh:59,125,298
 *                      LOAD_K chunks walked downwards  Kernel/cuda/src/cuCOSMAV100.cuh:2623-2746
 *                      fma accumulation from 0          Kernel/cuda/src/cuCOSMAV100.cuh:657-675
 *                      acc *= ALPHA, acc += BETA*C      Kernel/cuda/src/cuCOSMAV100.cuh:1575-1595, 876-950
 *                      split-K reduce, z ascending      Kernel/cuda/src/cosmaSplitKReduce.cuh:20-42
 *                      alpha == 0 -> C = BETA*C only    Kernel/cuda/src/cuCOSMAV100.cuh:2782-2797
 *   orc_cpu_naive_sgemm the CPU verification product      Kernel/cuda/src/Matrix.h:183-201
 *   orc_ref_fill_random Matrix<T>::fillRandom              Kernel/cuda/src/Matrix.h:75-87
 *   orc_activate_f64    fused ReLU / sigmoid epilogue       Kernel/cuda/src/cuCOSMAV100.cuh:2466-2476,
 *                                                           Kernel/cuda/src/sigmoid4.cuh:13-27
 *   orc_err_*           max-abs comparison                  Kernel/cuda/src/Matrix.h:129-173
 *                       (plus the normwise and max-relative metrics this
 *                       project reports)
 *
 * and adds the FP64 / long-double oracles, seeded generators and the BF16
 * split arithmetic that the reference never had.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   mt19937 / uniform / normal  bit-exact against numpy.random.RandomState and
 *                               the C++ standard's known answer (10000th draw
 *                               of mt19937(5489) is 4123659995)
 *   orc_ref_fill_random         bit-exact against std::mt19937 +
 *                               std::uniform_real_distribution<double>, built
 *                               with g++ at test time
 *   orc_gemm_f64 / orc_gemm_ld  exact on small-integer matrices vs numpy
 *                               int64 matmul; closed forms (A = I, B = I,
 *                               alpha = 0, K = 1); numpy float64 matmul
 *   orc_refsim_sgemm            exact on small-integer inputs; a planted case
 *                               whose float result depends on the descending
 *                               LOAD_K chunk order; split-K z-order case; bound
 *                               vs orc_gemm_f64
 *   orc_cpu_naive_sgemm         bit-exact vs a numpy float32 loop without fma;
 *                               threaded == serial
 *   orc_bf16_*                  RNE vs torch.bfloat16 conversion; closed forms
 *   orc_gemm_bf16x{3,4}_f64     exact on 8-bit integers; bounds vs orc_gemm_f64;
 *                               dropped-term planted checks
 *   orc_activate_f64            ReLU closed forms (negatives and NaN -> 0);
 *                               sigmoid vs scipy.special.expit, sigmoid(0) =
 *                               1/2, sigmoid(x) + sigmoid(-x) = 1
 *   orc_err_*                   vs numpy.linalg.norm
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off -shared).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* MT19937 (Matsumoto & Nishimura 1998), the generator of std::mt19937 and of */
/* numpy.random.RandomState.                                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint32_t mt[624];
    int idx;
    int has_gauss;
    double gauss;
} orc_rng;

static void mt_seed(orc_rng *r, uint32_t seed) {
    r->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
    r->idx = 624;
    r->has_gauss = 0;
    r->gauss = 0.0;
}

static uint32_t mt_next(orc_rng *r) {
    if (r->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (r->mt[i] & 0x80000000u) | (r->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = r->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            r->mt[i] = v;
        }
        r->idx = 0;
    }
    uint32_t y = r->mt[r->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* 53-bit double in [0,1) from two draws: (a>>5)*2^26 + (b>>6), over 2^53.
 * This is numpy's RandomState.random_sample construction, so every stream
 * below can be regenerated bit-for-bit with numpy. */
static double mt_double(orc_rng *r) {
    uint32_t a = mt_next(r) >> 5, b = mt_next(r) >> 6;
    return ((double)a * 67108864.0 + (double)b) / 9007199254740992.0;
}

/* Marsaglia polar method with one cached deviate: numpy's legacy gauss. */
static double mt_gauss(orc_rng *r) {
    if (r->has_gauss) {
        r->has_gauss = 0;
        return r->gauss;
    }
    double x1, x2, r2;
    do {
        x1 = 2.0 * mt_double(r) - 1.0;
        x2 = 2.0 * mt_double(r) - 1.0;
        r2 = x1 * x1 + x2 * x2;
    } while (r2 >= 1.0 || r2 == 0.0);
    double f = sqrt(-2.0 * log(r2) / r2);
    r->gauss = f * x1;
    r->has_gauss = 1;
    return f * x2;
}

orc_rng *orc_rng_new(uint32_t seed) {
    orc_rng *r = (orc_rng *)malloc(sizeof(orc_rng));
    if (r) mt_seed(r, seed);
    return r;
}
void orc_rng_free(orc_rng *r) { free(r); }

void orc_rng_u32(orc_rng *r, int64_t n, uint32_t *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = mt_next(r);
}
void orc_rng_random_sample(orc_rng *r, int64_t n, double *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = mt_double(r);
}
/* numpy: RandomState.uniform(lo, hi, n) == lo + (hi - lo) * random_sample(). */
void orc_rng_uniform_f32(orc_rng *r, int64_t n, double lo, double hi, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = (float)(lo + (hi - lo) * mt_double(r));
}
void orc_rng_uniform_f64(orc_rng *r, int64_t n, double lo, double hi, double *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * mt_double(r);
}
/* numpy: RandomState.normal(mean, std, n) == mean + std * gauss(). */
void orc_rng_normal_f64(orc_rng *r, int64_t n, double mean, double std, double *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = mean + std * mt_gauss(r);
}
void orc_rng_normal_f32(orc_rng *r, int64_t n, double mean, double std, float *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = (float)(mean + std * mt_gauss(r));
}
/* Outlier mask for the DGEMM workload: x[i] *= scale where random_sample() < p,
 * one draw per element in index order. */
void orc_rng_scale_outliers_f64(orc_rng *r, int64_t n, double p, double scale, double *x) {
    for (int64_t i = 0; i < n; ++i)
        if (mt_double(r) < p) x[i] *= scale;
}

/* Matrix<T>::fillRandom (Kernel/cuda/src/Matrix.h:75-87): a fresh
 * std::mt19937(seed) and std::uniform_real_distribution<double>(0, 1), cast
 * to float.  libstdc++ builds the double with generate_canonical<double,53>
 * from two 32-bit draws: (x0 + x1 * 2^32) / 2^64, clamped below 1. */
void orc_ref_fill_random(uint32_t seed, int64_t n, float *out) {
    orc_rng r;
    mt_seed(&r, seed);
    for (int64_t i = 0; i < n; ++i) {
        double x0 = (double)mt_next(&r);
        double x1 = (double)mt_next(&r);
        double u = (x0 + x1 * 4294967296.0) / 18446744073709551616.0;
        if (u >= 1.0) u = nextafter(1.0, 0.0);
        out[i] = (float)u;
    }
}

/* ------------------------------------------------------------------------ */
/* Operand access.  A is logically M x K, B is K x N.                        */
/*   ta == 0: A(i,k) = A[i*lda + k]      ta == 1: A(i,k) = A[k*lda + i]      */
/*   tb == 0: B(k,j) = B[k*ldb + j]      tb == 1: B(k,j) = B[j*ldb + k]      */
/* ------------------------------------------------------------------------ */
#define AIX(i, k) ((ta) ? ((int64_t)(k) * lda + (i)) : ((int64_t)(i) * lda + (k)))
#define BIX(k, j) ((tb) ? ((int64_t)(j) * ldb + (k)) : ((int64_t)(k) * ldb + (j)))

/* FP64 oracle for FP32 inputs: out = alpha * sum_k A(i,k) B(k,j) + beta * C0.
 * Products of floats are exact in double; the sum is an i-k-j blocked double
 * accumulation.  beta == 0 never reads C0.  Rows [row0, row0+rows). */
void orc_gemm_f64(int M, int N, int K, double alpha, const float *A, int lda, int ta,
                  const float *B, int ldb, int tb, double beta, const float *C0, int ldc,
                  double *out, int ldo, int row0, int rows) {
    (void)M;
    const int KB = 256;
#pragma omp parallel for schedule(dynamic, 1)
    for (int ii = 0; ii < rows; ++ii) {
        int i = row0 + ii;
        double *o = out + (int64_t)ii * ldo;
        for (int j = 0; j < N; ++j) o[j] = 0.0;
        for (int k0 = 0; k0 < K; k0 += KB) {
            int k1 = k0 + KB < K ? k0 + KB : K;
            for (int k = k0; k < k1; ++k) {
                double a = (double)A[AIX(i, k)];
                if (!tb) {
                    const float *brow = B + (int64_t)k * ldb;
                    for (int j = 0; j < N; ++j) o[j] += a * (double)brow[j];
                } else {
                    for (int j = 0; j < N; ++j) o[j] += a * (double)B[BIX(k, j)];
                }
            }
        }
        for (int j = 0; j < N; ++j) {
            double v = alpha * o[j];
            if (beta != 0.0) v += beta * (double)C0[(int64_t)i * ldc + j];
            o[j] = v;
        }
    }
}

/* Same for FP64 inputs: a plain double i-k-j accumulation. */
void orc_gemm_f64_d(int M, int N, int K, double alpha, const double *A, int lda, int ta,
                    const double *B, int ldb, int tb, double beta, const double *C0, int ldc,
                    double *out, int ldo, int row0, int rows) {
    (void)M;
#pragma omp parallel for schedule(dynamic, 1)
    for (int ii = 0; ii < rows; ++ii) {
        int i = row0 + ii;
        double *o = out + (int64_t)ii * ldo;
        for (int j = 0; j < N; ++j) o[j] = 0.0;
        for (int k = 0; k < K; ++k) {
            double a = A[AIX(i, k)];
            for (int j = 0; j < N; ++j) o[j] += a * B[BIX(k, j)];
        }
        for (int j = 0; j < N; ++j) {
            double v = alpha * o[j];
            if (beta != 0.0) v += beta * C0[(int64_t)i * ldc + j];
            o[j] = v;
        }
    }
}

/* Long-double oracle for FP64 inputs (x87 80-bit, 64-bit significand) with
 * Neumaier-compensated summation over k for each element; the result is
 * rounded to double once at the end.  Rows [row0, row0+rows). */
void orc_gemm_ld(int M, int N, int K, double alpha, const double *A, int lda, int ta,
                 const double *B, int ldb, int tb, double beta, const double *C0, int ldc,
                 double *out, int ldo, int row0, int rows) {
    (void)M;
#pragma omp parallel for schedule(dynamic, 1)
    for (int ii = 0; ii < rows; ++ii) {
        int i = row0 + ii;
        long double *s = (long double *)calloc((size_t)N, sizeof(long double));
        long double *c = (long double *)calloc((size_t)N, sizeof(long double));
        for (int k = 0; k < K; ++k) {
            long double a = (long double)A[AIX(i, k)];
            for (int j = 0; j < N; ++j) {
                long double p = a * (long double)B[BIX(k, j)];
                long double t = s[j] + p;
                if (fabsl(s[j]) >= fabsl(p))
                    c[j] += (s[j] - t) + p;
                else
                    c[j] += (p - t) + s[j];
                s[j] = t;
            }
        }
        double *o = out + (int64_t)ii * ldo;
        for (int j = 0; j < N; ++j) {
            long double v = (long double)alpha * (s[j] + c[j]);
            if (beta != 0.0) v += (long double)beta * (long double)C0[(int64_t)i * ldc + j];
            o[j] = (double)v;
        }
        free(s);
        free(c);
    }
}

/* ------------------------------------------------------------------------ */
/* Reference kernel's per-element arithmetic (row-major A and B).            */
/* ------------------------------------------------------------------------ */

/* One element's K slice as cosmaSgemm_kernel walks it: the slice of split z is
 * k in [z*tbk, z*tbk + tbk) clipped to K; chunks of load_k start at
 * K_START = (ceil(tbk/load_k) - 1) * load_k and step down to 0; inside a chunk
 * k ascends; every step is acc = fmaf(a, b, acc) from acc = 0. */
static float refsim_slice(int i, int j, int K, const float *A, int lda, const float *B,
                          int ldb, int z, int tbk, int load_k) {
    float acc = 0.0f;
    int kstart = ((tbk + load_k - 1) / load_k - 1) * load_k;
    for (int c = kstart; c >= 0; c -= load_k) {
        for (int kk = 0; kk < load_k; ++kk) {
            int kl = c + kk;
            int k = z * tbk + kl;
            if (kl >= tbk || k >= K) continue; /* zero-filled in shared memory */
            acc = fmaf(A[(int64_t)i * lda + k], B[(int64_t)k * ldb + j], acc);
        }
    }
    return acc;
}

/* C <- refsim(alpha*A*B + beta*C).  split_k == 1 requires tbk >= K (the
 * reference sets THREADBLOCK_TILE_K = K then).  fused_beta selects whether
 * acc + BETA*c is evaluated as one fma (nvcc/hipcc contraction) or as a
 * rounded product then a rounded sum.  beta_is_double models the split-K
 * reduce kernel where a double BETA literal promotes BETA*c + ALPHA*val to
 * double (Kernel/cuda/src/cosmaSplitKReduce.cuh:38-40). */
void orc_refsim_sgemm(int M, int N, int K, float alpha, float beta, const float *A, int lda,
                      const float *B, int ldb, float *C, int ldc, int split_k, int load_k,
                      int tbk, int fused_beta, int beta_is_double) {
    if (alpha == 0.0f) { /* pre-scale only (cuCOSMAV100.cuh:2782-2797) */
        if (beta != 1.0f)
            for (int i = 0; i < M; ++i)
                for (int j = 0; j < N; ++j) C[(int64_t)i * ldc + j] *= beta;
        return;
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < M; ++i) {
        for (int j = 0; j < N; ++j) {
            float *cp = C + (int64_t)i * ldc + j;
            if (split_k == 1) {
                float acc = refsim_slice(i, j, K, A, lda, B, ldb, 0, tbk, load_k);
                if (alpha != 1.0f) acc *= alpha;
                if (beta != 0.0f) {
                    if (fused_beta)
                        acc = fmaf(beta, *cp, acc);
                    else {
                        volatile float p = beta * *cp;
                        acc = acc + p;
                    }
                }
                *cp = acc;
            } else {
                float val = 0.0f;
                for (int z = 0; z < split_k; ++z) {
                    volatile float part = refsim_slice(i, j, K, A, lda, B, ldb, z, tbk, load_k);
                    val = val + part;
                }
                if (beta_is_double) {
                    double r = (double)beta * (double)*cp + (double)alpha * (double)val;
                    *cp = (float)r;
                } else if (fused_beta) {
                    volatile float av = alpha * val;
                    *cp = fmaf(beta, *cp, av);
                } else {
                    volatile float bc = beta * *cp;
                    volatile float av = alpha * val;
                    *cp = bc + av;
                }
            }
        }
    }
}

/* Matrix<T>::operator* (Kernel/cuda/src/Matrix.h:183-201): naive i-j-k, float
 * accumulator, B read at stride N, no fma (this file is built with
 * -ffp-contract=off).  Rows [row0, row0+rows) of C = A*B, C is rows x N.
 * nthreads > 1 splits rows over OpenMP threads; every element keeps its own
 * sequential k loop, so the result is bit-identical to nthreads == 1. */
void orc_cpu_naive_sgemm(int M, int N, int K, const float *A, const float *B, float *C,
                         int row0, int rows, int nthreads) {
    (void)M;
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : 1;
#pragma omp parallel for num_threads(nt) schedule(static)
#endif
    for (int ii = 0; ii < rows; ++ii) {
        int i = row0 + ii;
        for (int j = 0; j < N; ++j) {
            float acc = 0.0f;
            for (int k = 0; k < K; ++k) acc += A[(int64_t)i * K + k] * B[(int64_t)k * N + j];
            C[(int64_t)ii * N + j] = acc;
        }
    }
}

int orc_omp_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* BF16 split arithmetic.                                                    */
/* ------------------------------------------------------------------------ */

/* float -> bfloat16, round to nearest even; NaN stays a quiet NaN. */
uint16_t orc_bf16_rne(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x0040u);
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}
float orc_bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
/* hi = RNE_bf16(x), lo = RNE_bf16(x - hi); x - hi is exact in float. */
void orc_split_bf16(int64_t n, const float *x, uint16_t *hi, uint16_t *lo) {
    for (int64_t i = 0; i < n; ++i) {
        uint16_t h = orc_bf16_rne(x[i]);
        float r = x[i] - orc_bf16_to_f32(h);
        hi[i] = h;
        lo[i] = orc_bf16_rne(r);
    }
}

/* What the split-precision GEMM computes, accumulated exactly enough
 * (double) that only the split itself is modelled:
 * out = alpha * sum_k (ah*bh + ah*bl + al*bh [+ al*bl when nprod == 4]) + beta*C0.
 * Each bf16 product is exact in float and in double.  Rows [row0, row0+rows). */
void orc_gemm_bf16xn_f64(int M, int N, int K, double alpha, const float *A, int lda, int ta,
                         const float *B, int ldb, int tb, double beta, const float *C0,
                         int ldc, double *out, int ldo, int row0, int rows, int nprod) {
    (void)M;
    float *bh = (float *)malloc(sizeof(float) * (size_t)K * N);
    float *bl = (float *)malloc(sizeof(float) * (size_t)K * N);
    for (int k = 0; k < K; ++k)
        for (int j = 0; j < N; ++j) {
            float x = B[BIX(k, j)];
            float h = orc_bf16_to_f32(orc_bf16_rne(x));
            bh[(int64_t)k * N + j] = h;
            bl[(int64_t)k * N + j] = orc_bf16_to_f32(orc_bf16_rne(x - h));
        }
    const double w4 = nprod == 4 ? 1.0 : 0.0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int ii = 0; ii < rows; ++ii) {
        int i = row0 + ii;
        double *o = out + (int64_t)ii * ldo;
        for (int j = 0; j < N; ++j) o[j] = 0.0;
        for (int k = 0; k < K; ++k) {
            float x = A[AIX(i, k)];
            float h = orc_bf16_to_f32(orc_bf16_rne(x));
            float l = orc_bf16_to_f32(orc_bf16_rne(x - h));
            const float *bhr = bh + (int64_t)k * N, *blr = bl + (int64_t)k * N;
            for (int j = 0; j < N; ++j)
                o[j] += (double)h * bhr[j] + (double)h * blr[j] + (double)l * bhr[j] +
                        w4 * ((double)l * blr[j]);
        }
        for (int j = 0; j < N; ++j) {
            double v = alpha * o[j];
            if (beta != 0.0) v += beta * (double)C0[(int64_t)i * ldc + j];
            o[j] = v;
        }
    }
    free(bh);
    free(bl);
}

void orc_gemm_bf16x3_f64(int M, int N, int K, double alpha, const float *A, int lda, int ta,
                         const float *B, int ldb, int tb, double beta, const float *C0,
                         int ldc, double *out, int ldo, int row0, int rows) {
    orc_gemm_bf16xn_f64(M, N, K, alpha, A, lda, ta, B, ldb, tb, beta, C0, ldc, out, ldo, row0,
                        rows, 3);
}
void orc_gemm_bf16x4_f64(int M, int N, int K, double alpha, const float *A, int lda, int ta,
                         const float *B, int ldb, int tb, double beta, const float *C0,
                         int ldc, double *out, int ldo, int row0, int rows) {
    orc_gemm_bf16xn_f64(M, N, K, alpha, A, lda, ta, B, ldb, tb, beta, C0, ldc, out, ldo, row0,
                        rows, 4);
}

/* ------------------------------------------------------------------------ */
/* Error metrics.                                                            */
/* ------------------------------------------------------------------------ */

/* stats[0] = sum (C - R)^2, stats[1] = sum R^2, stats[2] = max |C - R|,
 * stats[3] = max over R != 0 of |C - R| / |R|, stats[4] = max |R|.
 * C is float (ldc), R double (ldr), rows x cols. */
void orc_err_f32(int rows, int cols, const float *C, int ldc, const double *R, int ldr,
                 double *stats) {
    double s2 = 0, r2 = 0, mx = 0, mr = 0, rmax = 0;
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) {
            double r = R[(int64_t)i * ldr + j];
            double d = fabs((double)C[(int64_t)i * ldc + j] - r);
            s2 += d * d;
            r2 += r * r;
            if (d > mx) mx = d;
            if (r != 0.0 && d / fabs(r) > mr) mr = d / fabs(r);
            if (fabs(r) > rmax) rmax = fabs(r);
        }
    stats[0] = s2;
    stats[1] = r2;
    stats[2] = mx;
    stats[3] = mr;
    stats[4] = rmax;
}
void orc_err_f64(int rows, int cols, const double *C, int ldc, const double *R, int ldr,
                 double *stats) {
    double s2 = 0, r2 = 0, mx = 0, mr = 0, rmax = 0;
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) {
            double r = R[(int64_t)i * ldr + j];
            double d = fabs(C[(int64_t)i * ldc + j] - r);
            s2 += d * d;
            r2 += r * r;
            if (d > mx) mx = d;
            if (r != 0.0 && d / fabs(r) > mr) mr = d / fabs(r);
            if (fabs(r) > rmax) rmax = fabs(r);
        }
    stats[0] = s2;
    stats[1] = r2;
    stats[2] = mx;
    stats[3] = mr;
    stats[4] = rmax;
}
/* Sum of squares in double (Frobenius norm squared) of a float / double array. */
double orc_sumsq_f32(int64_t n, const float *x) {
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += (double)x[i] * (double)x[i];
    return s;
}
double orc_sumsq_f64(int64_t n, const double *x) {
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
    return s;
}

/* Output activations of the fused epilogue, in double: act 1 is ReLU
 * (fmax(x, 0), so NaN -> 0 as with fmaxf), act 2 the logistic sigmoid
 * 1 / (1 + exp(-x)); act 0 leaves x unchanged.  The reference applies them to
 * the thread tile just before the store (Kernel/cuda/src/cuCOSMAV100.cuh:2466-2476)
 * or as a separate pass over C (Kernel/cuda/src/sigmoid4.cuh:13-27). */
void orc_activate_f64(int64_t n, double *x, int act) {
    for (int64_t i = 0; i < n; ++i) {
        if (act == 1) x[i] = fmax(x[i], 0.0);
        else if (act == 2) x[i] = 1.0 / (1.0 + exp(-x[i]));
    }
}
