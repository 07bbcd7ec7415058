// cosma_internal.hpp -- host-side interface between the C ABI (cosma_api.cpp)
// and the per-mode kernel launchers (cosma_f32.hip, cosma_f64.hip,
// cosma_bf16x3.hip, cosma_aux.hip).  No HIP device code here.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

namespace cosma {

// One fully validated GEMM launch.  Shapes are non-negative, pointers are
// non-null where used, leading dimensions and alignment satisfy the contract
// checked in cosma_api.cpp, split >= 1, and ws holds split*M*N accumulator
// elements whenever split > 1.
struct Problem {
  int mode;  // 0 fp32, 1 fp64, 2 bf16x3, 3 bf16x4
  int M, N, K;
  int ta, tb;
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  void* C;
  int64_t ldc;
  double alpha, beta;
  int split;
  int cfg;
  int group_m;
  int xcds;  // XCDs of the device (workgroups are dealt to them round-robin by id)
  int act;  // output activation: 0 none, 1 ReLU, 2 sigmoid
  void* ws;
  void* planes;  // pre-split tile configurations: room for the planes not given below
  const void* a_planes;  // caller's pre-split op(A) (M rows), or nullptr
  const void* b_planes;  // caller's pre-split op(B) (N rows), or nullptr
  hipStream_t stream;
};

// Tile configuration of one kernel instantiation family.
struct TileInfo {
  int BM, BN, BK, WM, WN;
  int threads;
  int lds_bytes;
  // Next-stage copy placement: 0 after the MFMAs, 1 after half of them, 2
  // after three quarters, 3 register double-buffered, 4 ping-pong (half the
  // waves copy before their MFMAs, half after), 5 pre-split operands copied
  // with direct-to-LDS loads (no grouped launch), 7 every wave copies first
  // and issues its next loads one by one among its MFMAs.
  int pipe;
};

// One problem of a grouped launch, as the grouped kernel reads it from
// device memory.  variant: index of the problem's tile in the launch's tile
// family.  split > 1: the tiles of K slice z write raw partials to
// ws + z*M*N (accumulator type, row length N), reduced after the launch.  K
// is 0 for alpha == 0 (the tiles then only scale C by beta).  vec_a / vec_b:
// the operand's rows start on 16-byte boundaries.
template <typename G>
struct GroupEntry {
  const G* A;
  const G* B;
  G* C;
  void* ws;
  int64_t lda, ldb, ldc;
  int M, N, K;
  int variant;
  int tiles_m, tiles_n;
  int split, kslice;
  int vec_a, vec_b;
  double alpha, beta;
};

// One split member of a grouped launch for the fixed-order reduction:
// C = alpha * sum_z ws[z] (+ beta * C).  block0: first reduction block of the
// member (256 elements per block, ascending over the table).
template <typename G>
struct GroupReduce {
  const void* ws;
  G* C;
  int64_t ldc;
  int M, N, S;
  int block0;
  double alpha, beta;
};

// A grouped tile family: tile configurations (indices of the mode's gemm
// tile table) that one persistent kernel runs side by side, one per problem.
struct GroupFamily {
  int ncfg;
  int cfgs[4];
};

// Bytes of bf16 planes a pre-split tile configuration needs: (M + N) rows of
// 2 * Kp elements, Kp = K rounded up to 32.
inline size_t planar_bytes(int M, int N, int K) {
  const int64_t kp = ((int64_t)K + 31) / 32 * 32;
  return (size_t)((int64_t)M + N) * 2 * (size_t)kp * 2;
}
// Split an FP32 operand into planes (kmajor: element (r, k) at X[r*ld + k]).
hipError_t run_split_operand(int mode, bool kmajor, int rows, int K, const float* X, int64_t ld, void* planes,
                             hipStream_t s);
constexpr int kPipePlanar = 5;

int num_tile_cfgs(int mode);
const TileInfo* tile_info(int mode, int cfg);  // nullptr when out of range

// Launch the main kernel of tile configuration p.cfg of p.mode (and, when
// split > 1, the fixed-order reduction).
hipError_t run_gemm(const Problem& p);

// Tile families of a mode's grouped launches (every non-pre-split tile
// configuration belongs to exactly one); *count receives their number.
const GroupFamily* group_families(int mode, int* count);
// Grouped persistent launch of family fam: `blocks` workgroups over the
// table (device memory); sched (device memory) holds blocks + 1 offsets and
// then, for each workgroup's tiles in order, the pair (table index of the
// problem, tile index within the problem).
hipError_t run_grouped(int mode, int fam, const void* table, const int* sched, int blocks, hipStream_t s);
// Workgroups of family fam's grouped kernel that fit on one CU.
int grouped_occupancy(int mode, int fam);
// Fixed-order reduction of the split members of a grouped launch: `count`
// GroupReduce entries (device memory), `blocks` blocks of 256 threads.
hipError_t run_grouped_reduce(int mode, const void* tab, int count, int blocks, hipStream_t s);
// Workgroups of one gemm_kernel instantiation that fit on one CU (0 if the
// query fails or the layout is not built).
int gemm_occupancy(int mode, int cfg, int ta, int tb, bool edge);

// C = act(beta*C) over the M x N window (alpha == 0 or K == 0).
hipError_t run_scale_f32(float* C, int64_t ldc, int M, int N, float beta, int act, hipStream_t s);
hipError_t run_scale_f64(double* C, int64_t ldc, int M, int N, double beta, int act,
                         hipStream_t s);

// Fixed-order split-K reduction, then the activation.
hipError_t run_reduce_f32(const float* ws, float* C, int64_t ldc, int M, int N, int S, float alpha,
                          float beta, int act, hipStream_t s);
hipError_t run_reduce_f64(const double* ws, double* C, int64_t ldc, int M, int N, int S,
                          double alpha, double beta, int act, hipStream_t s);

// Overwrite `bytes` of buf (cache flush between timed runs).
hipError_t run_flush(void* buf, size_t bytes, hipStream_t s);

// FP64 statistics; part must hold kStatParts * 5 doubles, out 5 doubles (device).
constexpr int kStatParts = 1024;
hipError_t run_err_stats_f32(int rows, int cols, const float* C, int64_t ldc, const double* R,
                             int64_t ldr, double rel_floor, double* part, double* out, hipStream_t s);
hipError_t run_err_stats_f64(int rows, int cols, const double* C, int64_t ldc, const double* R,
                             int64_t ldr, double rel_floor, double* part, double* out, hipStream_t s);
hipError_t run_sumsq_f32(int rows, int cols, const float* X, int64_t ldx, double* part, double* out,
                         hipStream_t s);
hipError_t run_sumsq_f64(int rows, int cols, const double* X, int64_t ldx, double* part,
                         double* out, hipStream_t s);
hipError_t run_convert_f32_f64(int rows, int cols, const float* X, int64_t ldx, double* Y,
                               int64_t ldy, hipStream_t s);

}  // namespace cosma
