/*
 * cosma.h -- C ABI of the MI355X (gfx950) COSMA-tiled GEMM library
 * (libcosma_amd.so).
 *
 * Drop-in boundary.  The reference exposes one call per build,
 *
 *     cublasStatus_t cosmaSgemm(const TYPE* A, int lda, const TYPE* B, int ldb,
 *                               TYPE* C, int ldc);
 *
 * (Kernel/cuda/src/cuCOSMAV100.cuh:2774-2776; the HIP copy returns bool,
 * Kernel/hip/src/cuCOSMAV100.cuh:2787-2789) with M, N, K, ALPHA, BETA and the
 * tile sizes fixed by the macros of a generated config.h
 * (Kernel/cuda/src/config.h:9-31).  C = ALPHA*A*B + BETA*C, all operands
 * row-major device buffers owned by the caller, ld* in elements.
 *
 * This header keeps that contract (row-major, caller-owned device buffers,
 * leading dimensions in elements, C = alpha*A*B + beta*C) and makes alpha and
 * beta arguments.  Differences from the reference, all deliberate:
 *   - beta == 0 never reads C (BLAS convention; the reference's split-K
 *     reduce reads C even then, Kernel/cuda/src/cosmaSplitKReduce.cuh:38-40).
 *   - alpha == 0 scales only the M x N window of C by beta (the reference
 *     scales ldc*M elements, padding included, cuCOSMAV100.cuh:2782-2796).
 *   - Split-K partials are always reduced in a fixed order; there is no
 *     atomic variant, so results are bitwise repeatable.
 *   - Errors are returned, never abort the process.
 *   - Calls take an optional stream; NULL is the default stream.
 *   - Threads: calls on different streams may be made from different threads
 *     at once; calls that use one stream must not overlap on the host (a call
 *     that grows the stream's workspace frees the previous buffer).
 *
 * Status codes: 0 success; negative = contract violation detected on the host
 * before any GPU work; positive = the hipError_t reported by the HIP runtime.
 *
 * Operand contract: A, B and C are aligned to their element size.  When A's
 * and B's rows start on 16-byte boundaries (base 16-byte aligned, lda/ldb a
 * multiple of 4 floats or 2 doubles -- the reference's implicit float4
 * contract, cuCOSMAV100.cuh:143-144, 315-317) the kernels use 16-byte vector
 * loads; any other leading dimension is accepted and served by the
 * bounds-checked instantiation with element loads.
 */
#ifndef COSMA_COSMA_H_
#define COSMA_COSMA_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define COSMA_STATUS_SUCCESS 0
#define COSMA_STATUS_INVALID_VALUE (-1) /* negative size, null pointer, ld too small, bad split/config */
#define COSMA_STATUS_MISALIGNED (-2)    /* operand base not aligned to its element size */
#define COSMA_STATUS_NOT_SUPPORTED (-3) /* combination not built (e.g. both operands transposed) */
#define COSMA_STATUS_WORKSPACE (-4)     /* split-K workspace missing while the stream is capturing */
#define COSMA_STATUS_SCRATCH (-5)       /* cosma_gemm_ex_static: needs more than the static scratch */

#define COSMA_MODE_FP32 0   /* FP32 in, FP32 MFMA, FP32 out */
#define COSMA_MODE_FP64 1   /* FP64 in, FP64 MFMA, FP64 out */
#define COSMA_MODE_BF16X3 2 /* FP32 in, bf16 hi/lo split, 3 bf16 MFMA products, FP32 out */
#define COSMA_MODE_BF16X4 3 /* FP32 in, bf16 hi/lo split, 4 products (adds lo*lo), FP32 out */

/* Fused element-wise epilogues: C = act(alpha * op(A) op(B) + beta * C).
 * Replace the ReLU / sigmoid loops of store_C (Kernel/cuda/src/cuCOSMAV100.cuh:2466-2476)
 * and the separate sigmoid pass (Kernel/cuda/src/sigmoid4.cuh:13-27).  With split-K
 * the activation runs once, after the fixed-order reduction. */
#define COSMA_EPILOGUE_NONE 0
#define COSMA_EPILOGUE_RELU 1    /* fmax(x, 0); NaN -> 0 */
#define COSMA_EPILOGUE_SIGMOID 2 /* 1 / (1 + exp(-x)) in the element type */

/* Full description of one GEMM: C = alpha * op(A) * op(B) + beta * C with
 * op(A) M x K and op(B) K x N.  trans_a == 0: A(i,k) = a[i*lda + k];
 * trans_a == 1: A(i,k) = a[k*lda + i].  trans_b == 0: B(k,j) = b[k*ldb + j];
 * trans_b == 1: B(k,j) = b[j*ldb + k].  trans_a == trans_b == 1 is not built.
 * Element type: float for FP32, BF16X3 and BF16X4, double for FP64. */
typedef struct cosma_gemm_desc {
    int mode;
    int m, n, k;
    int trans_a, trans_b;
    const void *a;
    int64_t lda;
    const void *b;
    int64_t ldb;
    void *c;
    int64_t ldc;
    double alpha, beta;
    int split_k;            /* <= 0: library heuristic; else 1..4096 */
    int tile_cfg;           /* < 0: library heuristic; else index < cosma_num_tile_configs(mode) */
    int group_m;            /* <= 0: default (8) tile rows per grouped-order group */
    void *workspace;        /* split-K partials and planes; NULL: installed or library-owned buffer */
    size_t workspace_bytes; /* size of workspace */
    void *stream;           /* hipStream_t; NULL: default stream */
    int epilogue;           /* COSMA_EPILOGUE_*: activation applied to the final C */
    /* BF16X3 / BF16X4 only: operands already split by cosma_split_operand
     * (NULL: split inside the call when the tile configuration works on
     * pre-split operands).  Setting either selects such a configuration. */
    const void *a_planes;   /* op(A) as M rows of K, split */
    const void *b_planes;   /* op(B) as N rows of K (i.e. B transposed), split */
} cosma_gemm_desc;

/* Fill *d with defaults (heuristic split and tile, default stream). */
void cosma_gemm_desc_init(cosma_gemm_desc *d);

/* Run the GEMM described by *d.  Replaces the body of cosmaSgemm
 * (Kernel/cuda/src/cuCOSMAV100.cuh:2774-2848) for every mode and layout. */
int cosma_gemm_ex(const cosma_gemm_desc *d);

/* cosma_gemm_ex for callers that cannot allocate (code under hipGraph
 * capture, serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  Allocates nothing: split-K partials and the planes of
 * operands split inside the call live in a static device array of
 * cosma_static_scratch_bytes() bytes (128 MiB) per GPU, held until the
 * process ends; its address on a device is looked up on the first call on
 * that device.  d->workspace must be NULL (an installed workspace is not
 * used).  Calls that need more return COSMA_STATUS_SCRATCH, except that the
 * library's pre-split BF16 configuration gives way to the on-the-fly tile,
 * which computes the same products in the same order (bitwise the same C).
 * Nothing in the array has to start at zero: every call writes all it
 * reads.  All calls on one GPU share its array: calls that overlap on one
 * GPU (other streams, other threads) race and must be ordered by the
 * caller. */
int cosma_gemm_ex_static(const cosma_gemm_desc *d);
/* Size of the static scratch array of cosma_gemm_ex_static. */
size_t cosma_static_scratch_bytes(void);

/* Bytes of workspace the call *d needs: split-K partials, then the planes of
 * operands split inside the call (0 when neither applies).  When the library
 * chose the pre-split tile (tile_cfg < 0) and d->workspace is smaller than
 * this, the call runs the on-the-fly 256x256 tile at the same split instead,
 * which needs only the partials and gives the same bits. */
int cosma_workspace_query(const cosma_gemm_desc *d, size_t *bytes);

/* Install a caller-owned device buffer used for split-K partials (and the
 * planes of operands split inside the call) by calls whose descriptor brings
 * no workspace.  NULL/0 reverts to the library-owned buffers.  Calls that
 * share an installed workspace must be ordered on one stream.
 *
 * The reference allocates and frees the partial buffer inside every call
 * (cuCOSMAV100.cuh:2826-2838).  Here, without a caller or installed buffer,
 * each (device, stream) pair owns one library buffer, allocated on first need
 * and grown only when the stream is not capturing (COSMA_STATUS_WORKSPACE
 * then): calls on different streams never share one.  Growing waits for the
 * stream.  A buffer given to a call while its stream was capturing is held
 * until the process ends, because the captured graph may be replayed later;
 * replay such a graph ordered with other calls on the stream it was captured
 * on. */
int cosma_workspace_set(void *ptr, size_t bytes);

/* Free the current device's library-owned workspace buffers, after
 * synchronising the device (the memory cudaFree returns at the end of every
 * reference call, cuCOSMAV100.cuh:2838).  Buffers a captured graph uses are
 * kept.  *freed_bytes (may be NULL) receives the bytes freed.  Not while any
 * stream is capturing, nor while another thread may be inside a GEMM call on
 * this device (a call that has picked its buffer but not yet launched would
 * then use freed memory). */
int cosma_workspace_release(size_t *freed_bytes);

/* Pre-split operands (BF16X3 / BF16X4).  The split modes can split each FP32
 * operand once into bf16 hi/lo "planes" -- rows of 2 * Kp bf16, Kp = K
 * rounded up to 32, each 32-deep K chunk stored as 32 hi then 32 lo parts --
 * and run the GEMM on planes copied straight into LDS.  Large GEMMs do this
 * inside the call (the tile configuration with pipeline schedule 5); a
 * caller that multiplies one operand several times can split it once with
 * cosma_split_operand and pass the planes in the descriptor.
 *   element (r, k) of the operand is X[r * ld + k] when kmajor != 0, else
 *   X[k * ld + r]; planes must hold cosma_planes_bytes(rows, k) bytes,
 *   16-byte aligned.  For A pass rows = M (kmajor = !trans_a); for B pass
 *   rows = N (kmajor = trans_b); k <= 2^31 - 32 (Kp must fit an int).
 *   Asynchronous on stream. */
size_t cosma_planes_bytes(int rows, int k);
int cosma_split_operand(int mode, int kmajor, int rows, int k, const float *X, int64_t ld, void *planes,
                        void *stream);

/* Fixed-shape entry points: the shape is fixed when the library is built
 * (COSMA_M, COSMA_N, COSMA_K; default 1024 x 1024 x 1024), as the reference
 * fixes it in config.h.  Same operand order as cosmaSgemm, with alpha/beta
 * passed instead of taken from ALPHA/BETA macros.  Default stream. */
int cosma_sgemm(float alpha, const float *A, int lda, const float *B, int ldb, float beta,
                float *C, int ldc);
int cosma_dgemm(double alpha, const double *A, int lda, const double *B, int ldb, double beta,
                double *C, int ldc);
/* Shape the fixed-shape entry points were built for. */
void cosma_fixed_shape(int *m, int *n, int *k);

/* Compile-time specialisations for the benchmark shapes: the shape, tile
 * configuration and split-K factor are constants of the instantiation, as one
 * reference build per config.h row (Kernel/cuda/run.sh:185-257). */
int cosma_sgemm_1024x1024x1024(float alpha, const float *A, int lda, const float *B, int ldb,
                               float beta, float *C, int ldc);
int cosma_sgemm_8192x8192x8192(float alpha, const float *A, int lda, const float *B, int ldb,
                               float beta, float *C, int ldc);
int cosma_sgemm_256x256x1048576(float alpha, const float *A, int lda, const float *B, int ldb,
                                float beta, float *C, int ldc);
int cosma_sgemm_bf16x3_8192x8192x8192(float alpha, const float *A, int lda, const float *B,
                                      int ldb, float beta, float *C, int ldc);
int cosma_sgemm_bf16x3_256x256x1048576(float alpha, const float *A, int lda, const float *B,
                                       int ldb, float beta, float *C, int ldc);
int cosma_dgemm_32768x1024x1024(double alpha, const double *A, int lda, const double *B, int ldb,
                                double beta, double *C, int ldc);
int cosma_dgemm_1024x65536x1024(double alpha, const double *A, int lda, const double *B, int ldb,
                                double beta, double *C, int ldc);

/* Runtime-shape variants (row-major, no transposes), on an explicit stream. */
int cosma_sgemm_mnk(int m, int n, int k, float alpha, const float *A, int lda, const float *B,
                    int ldb, float beta, float *C, int ldc, void *stream);
int cosma_dgemm_mnk(int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                    int ldb, double beta, double *C, int ldc, void *stream);
int cosma_sgemm_bf16x3_mnk(int m, int n, int k, float alpha, const float *A, int lda,
                           const float *B, int ldb, float beta, float *C, int ldc, void *stream);

/* Tile configurations built into the library.  out[0..7] = BM, BN, BK, WM,
 * WN, threads per workgroup, LDS bytes, pipeline schedule (0: next-stage copy
 * after the MFMAs, 1: after half of them, 2: after three quarters, 3:
 * register double-buffered, 4: ping-pong -- half the waves copy before their
 * MFMAs, half after, 5: pre-split operands copied with direct-to-LDS loads,
 * 7: every wave copies first and spreads its next loads among its MFMAs).
 * Replaces the tile
 * macros of config.h (THREADBLOCK_TILE_*, WARP_TILE_*, LOAD_K). */
int cosma_num_tile_configs(int mode);
int cosma_tile_config(int mode, int cfg, int *out8);
/* The (tile_cfg, split_k) the library heuristic picks for a shape on the
 * current device (its compute-unit count; without a device, a whole MI355X:
 * 256 CUs). */
int cosma_heuristic(int mode, int m, int n, int k, int *tile_cfg, int *split_k);
/* The same for op(A) stored transposed or not (trans_a) on a device of `cus`
 * compute units.  The reference's generator takes the SM count, shared
 * memory and registers of the target as inputs
 * (ScheduleGenerator/ParametersV100.h:27-36); here only the CU count enters
 * the choice. */
int cosma_heuristic_topo(int mode, int m, int n, int k, int trans_a, int cus, int *tile_cfg,
                         int *split_k);
/* Compute units, XCDs and peak engine clock (kHz) of the current device,
 * queried once per device.  Returns COSMA_STATUS_NOT_SUPPORTED (and the
 * nominal MI355X figures 256, 8, 2400000) when no device answers. */
int cosma_device_topology(int *cus, int *xcds, int *clock_khz);

/* FP64 residual statistics computed on the GPU with a deterministic
 * reduction.  out5: [0] sum (C-R)^2, [1] sum R^2, [2] max |C-R|,
 * [3] max over R != 0 of |C-R|/|R|, [4] max |R|.  C is float (f32) or
 * double (f64), R double; rows x cols windows with leading dimensions.
 * Synchronises the stream and writes out5 on the host. */
int cosma_error_stats_f32(int rows, int cols, const float *C, int64_t ldc, const double *R,
                          int64_t ldr, double *out5, void *stream);
int cosma_error_stats_f64(int rows, int cols, const double *C, int64_t ldc, const double *R,
                          int64_t ldr, double *out5, void *stream);
/* The same, with out5[3] the max of |C-R|/|R| over the entries with
 * |R| >= rel_floor (> 0): a relative error that near-zero reference
 * entries cannot blow up.  rel_floor = 0 is the plain statistic. */
int cosma_error_stats_ex_f32(int rows, int cols, const float *C, int64_t ldc, const double *R,
                             int64_t ldr, double rel_floor, double *out5, void *stream);
int cosma_error_stats_ex_f64(int rows, int cols, const double *C, int64_t ldc, const double *R,
                             int64_t ldr, double rel_floor, double *out5, void *stream);
/* out5[0] = sum x^2 (FP64), out5[2] = max |x|; others 0.  Synchronises. */
int cosma_sumsq_f32(int rows, int cols, const float *X, int64_t ldx, double *out5, void *stream);
int cosma_sumsq_f64(int rows, int cols, const double *X, int64_t ldx, double *out5, void *stream);
/* Y = (double)X, element-wise over a rows x cols window. */
int cosma_convert_f32_f64(int rows, int cols, const float *X, int64_t ldx, double *Y, int64_t ldy,
                          void *stream);

/* ---- Autotuner: replaces ScheduleGenerator's communication-volume model
 * (ScheduleGenerator/ScheduleGeneratorv4.h:13-391, main.cpp:22-342) with
 * measurement.  Candidates are every built tile configuration x split-K
 * factor x XCD group size that fits the shape; each is timed with HIP events
 * on the descriptor's stream (caches flushed before every timed call) and
 * reported with its occupancy and achieved fraction of the mode's peak.  The
 * fastest is the choice.  MFMA busy and L2 hit rates are hardware counters,
 * read by the Python autotuner (cucosma_amd.autotune.counters) under
 * rocprofv3 for chosen candidates. ---- */
typedef struct cosma_tune_result {
    int tile_cfg, split_k, group_m;
    int status;           /* 0, or the status of the failed launch */
    double ms;            /* mean time per GEMM (HIP events, one pair per timed call) */
    double tflops;        /* 2*M*N*K / ms */
    double achieved_frac; /* tflops / the mode's dense MFMA peak on this device (CUs x clock x
                             per-CU rate: FP32 256, FP64 128, BF16 4096 flops per clock;
                             BF16X3 / BF16X4 a third / a quarter of the BF16 rate) */
    int blocks;           /* workgroups of the main kernel */
    int occupancy;        /* workgroups per CU the kernel can hold */
    int waves_per_simd;
    size_t workspace_bytes;
} cosma_tune_result;

/* Time the candidates for the GEMM *d (its operands are used and C is
 * overwritten; d->tile_cfg / split_k / group_m >= 0 pin that parameter).
 * warmup untimed calls, then reps timed calls per candidate, each after
 * overwriting flush_bytes of device memory (0: no flush).  Fills up to
 * max_results entries sorted fastest first (failed candidates last), sets
 * *n_results and *best (0 when a candidate succeeded, else -1).  Allocates
 * split-K workspace and the flush buffer as needed: do not call while
 * capturing. */
int cosma_autotune_ex(const cosma_gemm_desc *d, int warmup, int reps, size_t flush_bytes,
                      cosma_tune_result *results, int max_results, int *n_results, int *best);
/* cosma_autotune_ex with a 512 MiB flush (L2 and the 256 MB Infinity Cache). */
int cosma_autotune(const cosma_gemm_desc *d, int warmup, int reps, cosma_tune_result *results,
                   int max_results, int *n_results, int *best);
/* Tuned table consulted by every entry point whose descriptor leaves tile_cfg
 * and split_k to the library: an exact (mode, m, n, k, trans_a, trans_b) match
 * overrides the built-in heuristic. */
int cosma_tune_set(int mode, int m, int n, int k, int trans_a, int trans_b, int tile_cfg,
                   int split_k, int group_m);
int cosma_tune_lookup(int mode, int m, int n, int k, int trans_a, int trans_b, int *tile_cfg,
                      int *split_k, int *group_m);
int cosma_tune_clear(void);

/* Benchmark protocol of the reference harness (Kernel/hip/src/benchmark.h:39-91,
 * Cacheflusher.h:18-49): `warmup` untimed calls, then `runs` calls each
 * bracketed by a HIP event pair on the descriptor's stream, with the caches
 * flushed between runs by overwriting flush_bytes of device memory (0: no
 * flush; the library sizes 512 MiB so that both L2 and the 256 MB Infinity
 * Cache are evicted, where the reference flushed an L2-sized buffer).
 * ms_out[i] receives run i's event time.  Allocates the flush buffer on first
 * use. */
int cosma_benchmark(const cosma_gemm_desc *d, int warmup, int runs, size_t flush_bytes,
                    float *ms_out);

/* ---- Grouped GEMM: many independent row-major problems, one persistent
 * launch per tile-configuration bucket, replayable as one hipGraph.  No
 * reference counterpart: the reference builds one binary per shape
 * (Kernel/cuda/run.sh:185-313). ---- */
typedef struct cosma_group_problem {
    int m, n, k;
    const void *a;
    int64_t lda;
    const void *b;
    int64_t ldb;
    void *c;
    int64_t ldc;
    double alpha, beta;
} cosma_group_problem;

typedef struct cosma_grouped_plan_s *cosma_grouped_plan;

/* Validate the problems (same contract as cosma_gemm_ex, row-major only;
 * alpha == 0 or K == 0 never reads A or B and gives C = beta * C), choose a
 * tile configuration and a split-K factor per problem, and upload, per tile
 * family (see cosma_group_families), one device table, one workgroup
 * schedule (longest tiles first, to the least-loaded workgroup) and the
 * split members' partial buffer.  Problem i takes tile_cfgs[i] and
 * splits[i] when given (tile_cfgs NULL or tile_cfgs[i] < 0, splits NULL or
 * splits[i] <= 0: the library's choice -- each problem on the variant of the
 * mode's first family with the least weighted padding, or one small tile for
 * batches of small outputs; K split only for tiles that would be long poles).
 * Allocates and synchronises; do not call while capturing.  blocks_per_cu <=
 * 0 uses the kernel's occupancy. */
int cosma_grouped_plan_create_ex(int mode, int count, const cosma_group_problem *problems,
                                 const int *tile_cfgs, const int *splits, int blocks_per_cu,
                                 cosma_grouped_plan *plan);
/* cosma_grouped_plan_create_ex with the library's split choice. */
int cosma_grouped_plan_create(int mode, int count, const cosma_group_problem *problems,
                              const int *tile_cfgs, int blocks_per_cu, cosma_grouped_plan *plan);
/* One persistent-kernel launch per tile family on stream, each followed by
 * the fixed-order reduction of its split members (no allocation, no copy:
 * capturable). */
int cosma_grouped_plan_run(cosma_grouped_plan plan, void *stream);
/* Launch the plan as a hipGraph: captured once on first use, then
 * hipGraphLaunch on stream. */
int cosma_grouped_plan_run_graph(cosma_grouped_plan plan, void *stream);
/* buckets, total tiles, sum of 2*M*N*K over the problems. */
int cosma_grouped_plan_info(cosma_grouped_plan plan, int *buckets, int *total_tiles, double *flops);
/* Bucket (tile family) b: the family's first tile config, problem count,
 * tiles over all K slices, persistent workgroups. */
int cosma_grouped_plan_bucket(cosma_grouped_plan plan, int b, int *tile_cfg, int *count,
                              int *tiles, int *blocks);
/* Tile configuration and split-K factor problem i runs with. */
int cosma_grouped_plan_choice(cosma_grouped_plan plan, int i, int *tile_cfg, int *split_k);
int cosma_grouped_plan_destroy(cosma_grouped_plan plan);
/* Grouped tile families of a mode: the tile configurations one persistent
 * kernel runs side by side (one per problem).  fam < 0 returns the number of
 * families; otherwise cfgs4 receives family fam's configurations (-1 past
 * the last) and the return value is their number.  Every tile configuration
 * except the pre-split ones belongs to exactly one family. */
int cosma_group_families(int mode, int fam, int *cfgs4);

/* Grouped autotuner: starting from the library's choices, the problems are
 * put in size buckets (default tile configuration x K class: K <= 256,
 * <= 1024, above); bucket by bucket, in decreasing order of flops, every
 * configuration of the mode's first tile family (and the bucket's default)
 * with K split 1, 2 or 4 ways is timed as a whole plan (graph replays, HIP
 * events, warmup + reps on the null stream), in pairs with the plan of the
 * choices kept so far (kept, candidate, candidate, kept), and kept when both
 * candidate timings beat their neighbours by at least 1%.
 * tile_cfgs / splits (count entries) receive the choices, for
 * cosma_grouped_plan_create_ex.  Every timed plan runs on the problems'
 * buffers (C is overwritten).  results (up to max_results) list the timed
 * candidates in order: bucket -1 is the starting plan. */
typedef struct cosma_group_tune_result {
    int bucket;   /* size bucket (default tile config * 3 + K class), -1: the starting plan */
    int tile_cfg; /* configuration tried for the bucket */
    int split_k;  /* split tried for the bucket */
    int status;
    double ms;    /* time per pass of the whole plan (mean of the candidate's two timings) */
} cosma_group_tune_result;
int cosma_grouped_autotune(int mode, int count, const cosma_group_problem *problems, int warmup,
                           int reps, int *tile_cfgs, int *splits, cosma_group_tune_result *results,
                           int max_results, int *n_results);

/* Human-readable text for a status code (static storage). */
const char *cosma_status_string(int status);
/* Library version, e.g. "0.1.0". */
const char *cosma_version(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

/* Literal drop-in for the reference call site: a translation unit that
 * defines the reference's config.h macros M, N, K, ALPHA and BETA before
 * including this header, and COSMA_REFERENCE_MACROS, can keep
 * `cosmaSgemm(A, lda, B, ldb, C, ldc)` unchanged (test_correctness.h:96,
 * benchmark.h:216).  The return value is 0 on success instead of
 * CUBLAS_STATUS_SUCCESS. */
#if defined(COSMA_REFERENCE_MACROS) && defined(M) && defined(N) && defined(K) && \
    defined(ALPHA) && defined(BETA)
#define cosmaSgemm(A_, lda_, B_, ldb_, C_, ldc_)                                                  \
    cosma_sgemm_mnk((M), (N), (K), (float)(ALPHA), (A_), (lda_), (B_), (ldb_), (float)(BETA), \
                    (C_), (ldc_), (void *)0)
#endif

#endif /* COSMA_COSMA_H_ */
