// cosma_aux.hip -- split-K reduction, beta scaling, FP64 residual statistics
// and the tile-configuration table.
#include "cosma_launch.hpp"

namespace cosma {

// Fixed-order split-K reduction: C = act(alpha * sum_{z=0..S-1} ws[z] (+ beta*C)).
// Every element sums its partials in ascending z, so the result does not
// depend on launch order or timing (Kernel/cuda/src/cosmaSplitKReduce.cuh:20-42
// sums in the same order).  beta == 0 does not read C.  The activation runs
// here, on the reduced value, not on each slice's partial.
//
// One element per thread is a chain of S loads.  With U > 1 a thread issues
// the loads of U slices before adding them (still one slice after another,
// from zero), so U loads per thread are in flight; the last, partial batch
// loads and adds only the slices that exist.  The U per-slice offsets cost
// scalar set-up in every wave, so the host picks U by S (U = 1 is the plain
// loop, for few slices).
template <typename E, int U>
__global__ void __launch_bounds__(256)
    splitk_reduce_kernel(const E* __restrict__ ws, E* __restrict__ C, int64_t ldc, int M, int N,
                         int S, E alpha, E beta, int act) {
  const int64_t total = (int64_t)M * N;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const E* p = ws + e;
    E v = E(0);
    int s = 0;
    for (; s + U <= S; s += U) {
      E t[U];
#pragma unroll
      for (int u = 0; u < U; ++u) t[u] = p[(int64_t)(s + u) * total];
#pragma unroll
      for (int u = 0; u < U; ++u) v += t[u];
    }
    if (U > 1 && s < S) {
      E t[U];
#pragma unroll
      for (int u = 0; u < U; ++u) t[u] = s + u < S ? p[(int64_t)(s + u) * total] : E(0);
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (s + u < S) v += t[u];
    }
    const int row = (int)(e / N);
    const int col = (int)(e - (int64_t)row * N);
    E* cp = C + (int64_t)row * ldc + col;
    E out = alpha * v;
    if (beta != E(0)) out = fma_e(beta, *cp, out);
    if (act != kActNone) out = activate(out, act);
    *cp = out;
  }
}

// C = act(beta * C) over the M x N window (alpha == 0 or K == 0); beta == 0
// writes act(0) without reading C.
template <typename E>
__global__ void __launch_bounds__(256)
    scale_kernel(E* __restrict__ C, int64_t ldc, int M, int N, E beta, int act) {
  const int64_t total = (int64_t)M * N;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int row = (int)(e / N);
    const int col = (int)(e - (int64_t)row * N);
    E* cp = C + (int64_t)row * ldc + col;
    const E v = beta == E(0) ? E(0) : beta * *cp;
    *cp = act != kActNone ? activate(v, act) : v;
  }
}

// ---------------------------------------------------------------------------
// FP64 residual statistics on the GPU.  Deterministic: the grid size is fixed,
// each thread walks a fixed element sequence, each block reduces in a fixed
// tree order, and one block reduces the block partials in index order.
// Statistics: [0] sum (C-R)^2, [1] sum R^2, [2] max |C-R|,
// [3] max over R != 0 of |C-R|/|R|, [4] max |R|.
// ---------------------------------------------------------------------------
constexpr int kStatBlocks = 1024;
constexpr int kStatThreads = 256;
constexpr int kNStat = 5;

__device__ __forceinline__ void stat_combine(double* a, const double* b) {
  a[0] += b[0];
  a[1] += b[1];
  a[2] = fmax(a[2], b[2]);
  a[3] = fmax(a[3], b[3]);
  a[4] = fmax(a[4], b[4]);
}

__device__ __forceinline__ void block_stat_reduce(double (&mine)[kNStat], double* out) {
  __shared__ double red[kStatThreads][kNStat];
#pragma unroll
  for (int s = 0; s < kNStat; ++s) red[threadIdx.x][s] = mine[s];
  __syncthreads();
  for (int w = kStatThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) stat_combine(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int s = 0; s < kNStat; ++s) out[s] = red[0][s];
}

template <typename TC>
__global__ void __launch_bounds__(kStatThreads)
    err_stats_partial(int rows, int cols, const TC* __restrict__ C, int64_t ldc,
                      const double* __restrict__ R, int64_t ldr, double rel_floor, double* __restrict__ part) {
  double mine[kNStat] = {0, 0, 0, 0, 0};
  const int64_t total = (int64_t)rows * cols;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int i = (int)(e / cols);
    const int j = (int)(e - (int64_t)i * cols);
    const double r = R[(int64_t)i * ldr + j];
    const double d = fabs((double)C[(int64_t)i * ldc + j] - r);
    mine[0] += d * d;
    mine[1] += r * r;
    mine[2] = fmax(mine[2], d);
    if (r != 0.0 && fabs(r) >= rel_floor) mine[3] = fmax(mine[3], d / fabs(r));
    mine[4] = fmax(mine[4], fabs(r));
  }
  block_stat_reduce(mine, part + (int64_t)blockIdx.x * kNStat);
}

template <typename T>
__global__ void __launch_bounds__(kStatThreads)
    sumsq_partial(int rows, int cols, const T* __restrict__ X, int64_t ldx,
                  double* __restrict__ part) {
  double mine[kNStat] = {0, 0, 0, 0, 0};
  const int64_t total = (int64_t)rows * cols;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int i = (int)(e / cols);
    const int j = (int)(e - (int64_t)i * cols);
    const double x = (double)X[(int64_t)i * ldx + j];
    mine[0] += x * x;
    mine[2] = fmax(mine[2], fabs(x));
  }
  block_stat_reduce(mine, part + (int64_t)blockIdx.x * kNStat);
}

__global__ void __launch_bounds__(kStatThreads)
    stats_final(const double* __restrict__ part, int nparts, double* __restrict__ out) {
  double mine[kNStat] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) stat_combine(mine, part + (int64_t)b * kNStat);
  block_stat_reduce(mine, out);
}

template <typename TI, typename TO>
__global__ void __launch_bounds__(256)
    convert_kernel(int rows, int cols, const TI* __restrict__ X, int64_t ldx, TO* __restrict__ Y,
                   int64_t ldy) {
  const int64_t total = (int64_t)rows * cols;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int i = (int)(e / cols);
    const int j = (int)(e - (int64_t)i * cols);
    Y[(int64_t)i * ldy + j] = (TO)X[(int64_t)i * ldx + j];
  }
}


const ModeOps* ops_f32();
const ModeOps* ops_f64();
const ModeOps* ops_bf16x3();
const ModeOps* ops_bf16x4();

static const ModeOps* ops(int mode) {
  switch (mode) {
    case kF32: return ops_f32();
    case kF64: return ops_f64();
    case kBF16x3: return ops_bf16x3();
    case kBF16x4: return ops_bf16x4();
    default: return nullptr;
  }
}

hipError_t run_gemm(const Problem& p) {
  const ModeOps* o = ops(p.mode);
  if (!o || p.cfg < 0 || p.cfg >= o->ncfg) return hipErrorInvalidValue;
  return o->run(p.cfg, p);
}

int num_tile_cfgs(int mode) {
  const ModeOps* o = ops(mode);
  return o ? o->ncfg : 0;
}

const TileInfo* tile_info(int mode, int cfg) {
  const ModeOps* o = ops(mode);
  if (!o || cfg < 0 || cfg >= o->ncfg) return nullptr;
  return &o->infos()[cfg];
}

int gemm_occupancy(int mode, int cfg, int ta, int tb, bool edge) {
  const ModeOps* o = ops(mode);
  if (!o || cfg < 0 || cfg >= o->ncfg) return 0;
  return o->occupancy(cfg, ta, tb, edge);
}

const GroupFamily* group_families(int mode, int* count) {
  const ModeOps* o = ops(mode);
  *count = o ? o->nfam : 0;
  return o ? o->families() : nullptr;
}

hipError_t run_grouped(int mode, int fam, const void* table, const int* sched, int blocks, hipStream_t s) {
  const ModeOps* o = ops(mode);
  if (!o || fam < 0 || fam >= o->nfam || blocks < 1) return hipErrorInvalidValue;
  return o->grouped(fam, table, sched, blocks, s);
}

int grouped_occupancy(int mode, int fam) {
  const ModeOps* o = ops(mode);
  if (!o || fam < 0 || fam >= o->nfam) return 1;
  return o->grouped_occupancy(fam);
}

hipError_t run_grouped_reduce(int mode, const void* tab, int count, int blocks, hipStream_t s) {
  if (count < 1 || blocks < 1) return hipSuccess;
  if (mode == kF64)
    hipLaunchKernelGGL((grouped_reduce_kernel<double, double>), dim3(blocks), dim3(256), 0, s,
                       static_cast<const GroupReduce<double>*>(tab), count);
  else
    hipLaunchKernelGGL((grouped_reduce_kernel<float, float>), dim3(blocks), dim3(256), 0, s,
                       static_cast<const GroupReduce<float>*>(tab), count);
  return hipGetLastError();
}

static int grid_for(int64_t total, int threads, int cap) {
  int64_t b = (total + threads - 1) / threads;
  if (b < 1) b = 1;
  return (int)(b < cap ? b : cap);
}

hipError_t run_scale_f32(float* C, int64_t ldc, int M, int N, float beta, int act, hipStream_t s) {
  hipLaunchKernelGGL(scale_kernel<float>, dim3(grid_for((int64_t)M * N, 256, 8192)), dim3(256), 0,
                     s, C, ldc, M, N, beta, act);
  return hipGetLastError();
}
hipError_t run_scale_f64(double* C, int64_t ldc, int M, int N, double beta, int act,
                         hipStream_t s) {
  hipLaunchKernelGGL(scale_kernel<double>, dim3(grid_for((int64_t)M * N, 256, 8192)), dim3(256), 0,
                     s, C, ldc, M, N, beta, act);
  return hipGetLastError();
}

// Loads in flight per thread: one (the plain loop) when the slices are few
// or the elements alone give every CU its 32 resident waves (524288 elements
// and up: 1024^2 with 16 slices ran 14.7 us against 15.8 with U = 8; 1024^2
// with 3 slices 5.7 us against 11.4 with U = 32); else 8, or 32 (FP32) / 16
// (FP64) from 64 slices (256x256 with 256 slices: 12 us against 75 with one
// load at a time).
template <typename E>
static hipError_t launch_reduce(const E* ws, E* C, int64_t ldc, int M, int N, int S, E alpha, E beta,
                                int act, hipStream_t s) {
  const dim3 grid(grid_for((int64_t)M * N, 256, 8192)), block(256);
  if (S < 8 || (int64_t)M * N >= (int64_t)256 * 32 * 64)
    hipLaunchKernelGGL((splitk_reduce_kernel<E, 1>), grid, block, 0, s, ws, C, ldc, M, N, S, alpha, beta, act);
  else if (S < 64)
    hipLaunchKernelGGL((splitk_reduce_kernel<E, 8>), grid, block, 0, s, ws, C, ldc, M, N, S, alpha, beta, act);
  else
    hipLaunchKernelGGL((splitk_reduce_kernel<E, sizeof(E) == 4 ? 32 : 16>), grid, block, 0, s, ws, C, ldc, M,
                       N, S, alpha, beta, act);
  return hipGetLastError();
}
hipError_t run_reduce_f32(const float* ws, float* C, int64_t ldc, int M, int N, int S, float alpha,
                          float beta, int act, hipStream_t s) {
  return launch_reduce<float>(ws, C, ldc, M, N, S, alpha, beta, act, s);
}
hipError_t run_reduce_f64(const double* ws, double* C, int64_t ldc, int M, int N, int S,
                          double alpha, double beta, int act, hipStream_t s) {
  return launch_reduce<double>(ws, C, ldc, M, N, S, alpha, beta, act, s);
}

// Cache flush: overwrite a buffer larger than L2 + Infinity Cache, then read
// it back.  The writes evict whatever the caches held; the reads evict the
// writes' own dirty lines, so their write-back happens here and not inside
// the timed call after the flush.  Overwriting alone left up to L2 + Infinity
// Cache of dirty lines for the timed call to write back on its first
// evictions: 57 us (12%) on the 256x256x1048576 BF16x3 kernel, which streams
// 2 GiB (0.510 against 0.453 ms), 1% at 4096^3.
__global__ void __launch_bounds__(256) flush_kernel(float4* __restrict__ buf, int64_t n, float v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    buf[i] = make_float4(v, v, v, v);
}

// Reads the buffer flush_kernel filled with v > 0; the store is never taken
// (a sum of positive values), it only keeps the loads.
__global__ void __launch_bounds__(256) flush_read_kernel(float4* __restrict__ buf, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  float acc = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float4 w = buf[i];
    acc += w.x + w.w;
  }
  if (acc < 0.f && (int64_t)threadIdx.x < n) buf[threadIdx.x] = make_float4(acc, acc, acc, acc);
}

hipError_t run_flush(void* buf, size_t bytes, hipStream_t s) {
  const int64_t n = (int64_t)(bytes / sizeof(float4));
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(flush_kernel, dim3(4096), dim3(256), 0, s, static_cast<float4*>(buf), n, 1.0f);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(flush_read_kernel, dim3(4096), dim3(256), 0, s, static_cast<float4*>(buf), n);
  return hipGetLastError();
}

static_assert(kStatParts == kStatBlocks, "partial buffer size must match the stats grid");

hipError_t run_err_stats_f32(int rows, int cols, const float* C, int64_t ldc, const double* R,
                             int64_t ldr, double rel_floor, double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL(err_stats_partial<float>, dim3(kStatBlocks), dim3(kStatThreads), 0, s, rows,
                     cols, C, ldc, R, ldr, rel_floor, part);
  hipLaunchKernelGGL(stats_final, dim3(1), dim3(kStatThreads), 0, s, part, kStatBlocks, out);
  return hipGetLastError();
}
hipError_t run_err_stats_f64(int rows, int cols, const double* C, int64_t ldc, const double* R,
                             int64_t ldr, double rel_floor, double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL(err_stats_partial<double>, dim3(kStatBlocks), dim3(kStatThreads), 0, s, rows,
                     cols, C, ldc, R, ldr, rel_floor, part);
  hipLaunchKernelGGL(stats_final, dim3(1), dim3(kStatThreads), 0, s, part, kStatBlocks, out);
  return hipGetLastError();
}
hipError_t run_sumsq_f32(int rows, int cols, const float* X, int64_t ldx, double* part, double* out,
                         hipStream_t s) {
  hipLaunchKernelGGL(sumsq_partial<float>, dim3(kStatBlocks), dim3(kStatThreads), 0, s, rows, cols,
                     X, ldx, part);
  hipLaunchKernelGGL(stats_final, dim3(1), dim3(kStatThreads), 0, s, part, kStatBlocks, out);
  return hipGetLastError();
}
hipError_t run_sumsq_f64(int rows, int cols, const double* X, int64_t ldx, double* part,
                         double* out, hipStream_t s) {
  hipLaunchKernelGGL(sumsq_partial<double>, dim3(kStatBlocks), dim3(kStatThreads), 0, s, rows,
                     cols, X, ldx, part);
  hipLaunchKernelGGL(stats_final, dim3(1), dim3(kStatThreads), 0, s, part, kStatBlocks, out);
  return hipGetLastError();
}
hipError_t run_convert_f32_f64(int rows, int cols, const float* X, int64_t ldx, double* Y,
                               int64_t ldy, hipStream_t s) {
  hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for((int64_t)rows * cols, 256, 8192)),
                     dim3(256), 0, s, rows, cols, X, ldx, Y, ldy);
  return hipGetLastError();
}

}  // namespace cosma
