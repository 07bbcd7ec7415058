// cosma_kernels.hpp -- gfx950 device code of the COSMA-tiled GEMM.
//
// One kernel template covers three arithmetic modes that share the tiling:
//   kF32     C = alpha*A*B + beta*C on v_mfma_f32_16x16x4_f32 (FP32 in, FP32 accumulate)
//   kF64     the same on v_mfma_f64_16x16x4_f64
//   kBF16x3  FP32 operands split on the fly into bf16 hi + lo parts while the
//            tile is written to LDS; hi*hi + hi*lo + lo*hi on
//            v_mfma_f32_16x16x32_bf16 with FP32 accumulation
//   kBF16x4  the same split with the fourth product lo*lo added
//
// Reference behaviour this replaces (Kernel/cuda/src/cuCOSMAV100.cuh):
//   threadblock / warp / thread tiles     :2505-2565   -> workgroup / wave / MFMA sub-tiles
//   double-buffered shared-memory staging :2567-2581, 2658-2718 -> two LDS stages, one barrier per K chunk
//   split-K over blockIdx.z               :59,125,298,2814-2831 -> K slice per split index, partials to a workspace
//   split-K reduce                        Kernel/cuda/src/cosmaSplitKReduce.cuh:20-42 -> splitk_reduce_kernel
//   ALPHA / BETA epilogue                 :1575-1595, 876-950  -> epilogue(): alpha*acc, then fma(beta, C, .)
//   block swizzle                         :2583-2603  -> map_block(): XCD-aware remap + grouped tile order
//
// Layout conventions.  A is logically M x K and B is K x N, both addressed
// through a leading dimension.  TA == false: A(i,k) = A[i*lda + k] (row-major,
// the reference's only layout); TA == true: A(i,k) = A[k*lda + i].  TB ==
// false: B(k,j) = B[k*ldb + j]; TB == true: B(k,j) = B[j*ldb + k].  In LDS both
// operands are stored "k-contiguous": sA[i][k], sB[j][k], rows of BK elements
// without padding, the 16-byte chunks of each row permuted (LdsLayout) so
// that fragment reads and tile stores are free of bank conflicts.  An operand
// whose global layout is k-contiguous is copied with 16-byte loads and
// stores; the other one is loaded as 4x4 (2x2 for FP64) element blocks and
// transposed in registers before its LDS stores.
//
// MFMA operand mapping (16x16 output sub-tile, 64 lanes, g = lane/16,
// r16 = lane%16): lane supplies A(row r16, k-group g) and B(k-group g, col
// r16); for FP32/BF16 products the accumulator holds C(4g + r, r16) in
// element r, for FP64 C(g + 4r, r16) (acc_row()).  Each lane reads
// KPL consecutive k of its row with one 16-byte LDS read, and step s of the
// MFMA sequence consumes element s of that read for both operands, so A and B
// always agree on which k an MFMA slot carries.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cosma_internal.hpp"

namespace cosma {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

enum Mode : int { kF32 = 0, kF64 = 1, kBF16x3 = 2, kBF16x4 = 3 };

// Modes that split FP32 operands into bf16 hi/lo planes in LDS.
constexpr bool is_split(int mode) { return mode == kBF16x3 || mode == kBF16x4; }

template <int MODE>
struct Traits;

template <>
struct Traits<kF32> {
  using G = float;    // element in global memory (A, B, C)
  using S = float;    // element in LDS
  using GV = f32x4;   // one 16-byte global vector
  using Frag = f32x4; // one 16-byte LDS fragment read
  using Acc = f32x4;  // accumulator of a 16x16 sub-tile, per lane
  using E = float;    // accumulator / epilogue scalar
  static constexpr int GVEC = 4;  // elements per global vector
  static constexpr int KPL = 4;   // consecutive k per lane per fragment read
  static constexpr int STEPS = 4; // MFMAs per fragment read
  static constexpr int PLANES = 1;
};
template <>
struct Traits<kF64> {
  using G = double;
  using S = double;
  using GV = f64x2;
  using Frag = f64x2;
  using Acc = f64x4;
  using E = double;
  static constexpr int GVEC = 2;
  static constexpr int KPL = 2;
  static constexpr int STEPS = 2;
  static constexpr int PLANES = 1;
};
template <>
struct Traits<kBF16x3> {
  using G = float;
  using S = __bf16;
  using GV = f32x4;
  using Frag = bf16x8;
  using Acc = f32x4;
  using E = float;
  static constexpr int GVEC = 4;
  static constexpr int KPL = 8;
  static constexpr int STEPS = 1;
  static constexpr int PLANES = 2;  // hi and lo
};
// Four products: the three of kBF16x3 plus lo*lo.
template <>
struct Traits<kBF16x4> : Traits<kBF16x3> {};

// Global-memory (address space 1) views of operand pointers.  A pointer the
// kernel reads from memory -- the grouped launch's problem table -- is a
// generic pointer to the compiler, which then emits flat loads and stores.
// Flat accesses count against lgkmcnt as well as vmcnt, so every wait for an
// LDS fragment read would also wait for all operand loads in flight and the
// prefetch would stop overlapping the MFMAs.  Loads and stores through these
// views are global instructions whatever the pointer's origin (kernel
// arguments are inferred global already).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
  return (gptr<T>)p;
}

template <int MODE>
struct GemmParams {
  using G = typename Traits<MODE>::G;
  using E = typename Traits<MODE>::E;
  const G* A;
  const G* B;
  G* C;
  E* ws;  // split-K partials [split][M][N]; unused when split == 1
  int64_t lda, ldb, ldc;
  int M, N, K;
  int tiles_m, tiles_n;
  int split;   // number of K slices
  int kslice;  // K per slice, a multiple of the tile's BK
  int group_m; // tile rows per group in the grouped tile order (1..tiles_m)
  int xcds;    // XCDs of the device: workgroups are dealt to them round-robin by id
  int vec_a;   // 1: A's rows start on 16-byte boundaries (vector loads allowed)
  int vec_b;   // 1: same for B
  int act;     // element-wise activation applied to the final C values (kActNone/ReLU/Sigmoid)
  E alpha, beta;
};

// Element-wise activations of the output (fused epilogue).  The reference has
// them as commented-out loops over the thread tile in store_C
// (Kernel/cuda/src/cuCOSMAV100.cuh:2466-2476: fmaxf(x, 0) and
// 1 / (1 + expf(-x))) and as a separate pass (Kernel/cuda/src/sigmoid4.cuh:13-27).
// Here they act on alpha*AB + beta*C, after the split-K reduction.
constexpr int kActNone = 0;
constexpr int kActReLU = 1;
constexpr int kActSigmoid = 2;
template <typename E>
__device__ __forceinline__ E activate(E v, int act) {
  if (act == kActReLU) return fmax(v, E(0));  // NaN -> 0, as fmaxf
  if (act == kActSigmoid) return E(1) / (E(1) + exp(-v));
  return v;
}

// LDS placement of one operand tile: row r (M or N index) holds BK elements
// in NC 16-byte chunks, chunk c of row r stored at 16-byte unit phys(r, c).
// gfx950's LDS has 64 banks of 4 bytes; a 16-byte read by 64 lanes is served
// 16 lanes at a time, lanes {0-7, 16-23}, {8-15, 24-31}, {32-39, 48-55},
// {40-47, 56-63} (measured with SQ_LDS_BANK_CONFLICT: padded rows of 80 or
// 144 bytes took 8 cycles per read instead of 4).  A fragment read puts rows
// r0..r0+7 and two adjacent chunks in one such group, so the 16 chunks must
// land in 16 distinct bank quads:
//   NC = 16 (256-byte rows, every row starts at bank 0): c ^ 2*(r & 7);
//   NC = 8 (two rows per 64 banks): c ^ 2*((r >> 1) & 3);
//   NC = 4 (four rows per 64 banks): c ^ 2*((r >> 2) & 1), and rows r and r+4
//   of a block of four rows also swap bank quarters ((r & 3) ^ ((r >> 2) & 3))
//   -- without that, the transposing store, which writes rows 4j + t for
//   eight values of j in one instruction, puts all of them in one quarter.
// Stores of 8 or 16 bytes per lane are conflict-free in every case (measured
// on the loaders' exact lane patterns).  phys maps each block of four rows
// (NC = 4) or each row (NC = 8, 16) onto its own units, so a tile of R rows
// (R a multiple of 4) occupies exactly R * BK elements.
template <int MODE, int BK>
struct LdsLayout {
  using S = typename Traits<MODE>::S;
  static constexpr int EPC = 16 / (int)sizeof(S);  // elements per 16-byte chunk
  static constexpr int NC = BK / EPC;              // chunks per row
  static_assert(NC == 4 || NC == 8 || NC == 16, "LDS rows of 64, 128 or 256 bytes");
  // Element offset of element e (of a run of at most EPC - e % EPC
  // contiguous elements) of row r.
  __device__ __forceinline__ static int off(int r, int e) {
    const int c = e / EPC;
    int u;
    if constexpr (NC == 4)
      u = ((r >> 2) << 4) + (((r & 3) ^ ((r >> 2) & 3)) << 2) + (c ^ (((r >> 2) & 1) << 1));
    else if constexpr (NC == 8)
      u = (r << 3) + (c ^ (((r >> 1) & 3) << 1));
    else
      u = (r << 4) + (c ^ ((r & 7) << 1));
    return u * EPC + (e % EPC);
  }
};

// Split four floats into RNE bf16 hi parts and RNE bf16 lo = x - hi parts
// (x - hi is exact in FP32).  v_cvt_pk_bf16_f32 does the rounding.
__device__ __forceinline__ void split_bf16x4(f32x4 v, bf16x4& hi, bf16x4& lo) {
  f32x2 v01 = {v[0], v[1]};
  f32x2 v23 = {v[2], v[3]};
  bf16x2 h01 = __builtin_convertvector(v01, bf16x2);
  bf16x2 h23 = __builtin_convertvector(v23, bf16x2);
  f32x2 r01 = v01 - __builtin_convertvector(h01, f32x2);
  f32x2 r23 = v23 - __builtin_convertvector(h23, f32x2);
  bf16x2 l01 = __builtin_convertvector(r01, bf16x2);
  bf16x2 l23 = __builtin_convertvector(r23, bf16x2);
  hi = bf16x4{h01[0], h01[1], h23[0], h23[1]};
  lo = bf16x4{l01[0], l01[1], l23[0], l23[1]};
}

// Linear block id -> (tile row, tile col, split index).
// Workgroups are handed to the device's xcds XCDs round-robin by id, so ids
// b, b + xcds, ... share an L2.  The first remap gives each XCD a contiguous
// range of logical ids; the grouped order then walks group_m tile rows
// column by column, so the workgroups resident on one XCD at a time share A
// row panels and B column panels.  Split index is the slowest-varying
// coordinate.  group_m is at most tiles_m (the host clamps it).
__device__ __forceinline__ void map_block(int bid, int nblk, int tiles_m, int tiles_n,
                                          int group_m, int xcds, int& tm, int& tn, int& z) {
  const int full = xcds > 1 ? (nblk / xcds) * xcds : 0;
  int lid = bid;
  if (bid < full) lid = (bid % xcds) * (full / xcds) + bid / xcds;
  const int tiles = tiles_m * tiles_n;
  z = lid / tiles;
  const int t = lid - z * tiles;
  const int per_group = group_m * tiles_n;
  const int grp = t / per_group;
  const int first_m = grp * group_m;
  const int gsz = min(tiles_m - first_m, group_m);
  const int r = t - grp * per_group;
  tm = first_m + r % gsz;
  tn = r / gsz;
}

// Global -> register -> LDS copy of one operand tile of R rows (M or N
// extent) by BK (k extent).  KMAJ: global memory is k-contiguous for a row.
template <int MODE, int R, int BK, int NT, bool KMAJ, bool EDGE>
struct TileLoader {
  using T = Traits<MODE>;
  using G = typename T::G;
  using GV = typename T::GV;
  using S = typename T::S;
  static constexpr int GVEC = T::GVEC;
  using L = LdsLayout<MODE, BK>;
  // KMAJ: units are 16-byte vectors along k.  Otherwise units are GVEC x GVEC
  // blocks (GVEC k rows of one GVEC-wide vector along the row index); the k
  // block varies fastest across lanes, so the transposed LDS stores of
  // neighbouring lanes fill one LDS row contiguously (no bank conflicts) while
  // each global k row is still read in whole 128-byte lines.
  static constexpr int UNITS = KMAJ ? (R * BK / GVEC) : ((R / GVEC) * (BK / GVEC));
  static constexpr int PER = (UNITS + NT - 1) / NT;
  static constexpr int NREG = KMAJ ? PER : PER * GVEC;
  // When the tile has fewer units than threads, the surplus threads idle.
  __device__ __forceinline__ static bool active(int idx) { return UNITS % NT == 0 || idx < UNITS; }
  static_assert(BK % GVEC == 0 && R % GVEC == 0, "tile extents must be whole vectors");

  GV v[NREG];

  // A whole 16-byte vector when it is in range and aligned; otherwise guarded
  // element loads (EDGE only: the unchecked instantiation is launched only
  // for aligned, tile-divisible problems).
  __device__ __forceinline__ static GV load_vec(gptr<const G> p, bool row_ok, int e_valid, bool vec_ok) {
    if (!EDGE || (vec_ok && row_ok && e_valid >= GVEC)) return *reinterpret_cast<gptr<const GV>>(p);
    GV out;
#pragma unroll
    for (int e = 0; e < GVEC; ++e) out[e] = (row_ok && e < e_valid) ? p[e] : G(0);
    return out;
  }

  // A thread's units are UNIT_STEP rows (KMAJ) or columns (otherwise) apart,
  // a compile-time stride, so one 64-bit pointer to its first unit plus
  // scalar offsets (multiples of ld) address all of them.
  static constexpr int VPR = BK / GVEC;  // KMAJ: vectors per row; otherwise k blocks per tile
  static constexpr int UNIT_STEP = KMAJ ? NT / VPR : (NT / VPR) * GVEC;
  static_assert(NT % VPR == 0, "unit stride must be uniform across the workgroup");
  gptr<const G> p0;
  int r0_;  // KMAJ: row of the first unit; otherwise its first column
  int k0_;  // first k of every unit of this thread
  int64_t ld_;

  // base: operand pointer; ld: its leading dimension; r0: first tile row;
  // k0: first k of the tile's K slice.
  __device__ __forceinline__ void init(const G* __restrict__ base, int64_t ld, int r0, int k0, int tid) {
    ld_ = ld;
    const int a = tid / VPR;
    const int b = tid - a * VPR;
    if constexpr (KMAJ) {
      r0_ = r0 + a;
      k0_ = k0 + b * GVEC;
      p0 = to_global(base) + (int64_t)r0_ * ld + k0_;
    } else {
      r0_ = r0 + a * GVEC;
      k0_ = k0 + b * GVEC;
      p0 = to_global(base) + (int64_t)k0_ * ld + r0_;
    }
  }

  // Loads the chunk that starts kofs past init()'s k0.  rmax: rows (KMAJ) or
  // columns of the operand; kend: end of the valid k range of this split.
  __device__ __forceinline__ void load(int kofs, int rmax, int kend, int tid, bool vec_ok) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (!active(tid + i * NT)) continue;
      const int rr = r0_ + i * UNIT_STEP;
      const int gk = k0_ + kofs;
      if constexpr (KMAJ) {
        gptr<const G> p = p0 + (int64_t)(i * UNIT_STEP) * ld_ + kofs;
        v[i] = load_vec(p, rr < rmax, kend - gk, vec_ok);
      } else {
        gptr<const G> p = p0 + (int64_t)kofs * ld_ + i * UNIT_STEP;
#pragma unroll
        for (int u = 0; u < GVEC; ++u)
          v[i * GVEC + u] = load_vec(p + (int64_t)u * ld_, gk + u < kend, rmax - rr, vec_ok);
      }
    }
  }

  // The same loads one instruction at a time: load_one(j, ...) for j in
  // [0, NREG) issues the j-th of load()'s vector loads, so a caller can spread
  // them among its MFMAs.
  __device__ __forceinline__ void load_one(int j, int kofs, int rmax, int kend, int tid, bool vec_ok) {
    const int i = KMAJ ? j : j / GVEC;
    if (!active(tid + i * NT)) return;
    const int rr = r0_ + i * UNIT_STEP;
    const int gk = k0_ + kofs;
    if constexpr (KMAJ) {
      gptr<const G> p = p0 + (int64_t)(i * UNIT_STEP) * ld_ + kofs;
      v[i] = load_vec(p, rr < rmax, kend - gk, vec_ok);
    } else {
      const int u = j - i * GVEC;
      gptr<const G> p = p0 + (int64_t)(kofs + u) * ld_ + i * UNIT_STEP;
      v[j] = load_vec(p, gk + u < kend, rmax - rr, vec_ok);
    }
  }

  // Writes the registers to LDS plane(s) sh (and sl for the bf16 lo part).
  __device__ __forceinline__ void store(S* __restrict__ sh, S* __restrict__ sl, int tid) const {
    if constexpr (KMAJ) {
      constexpr int VPR = BK / GVEC;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int idx = tid + i * NT;
        if (!active(idx)) continue;
        const int r = idx / VPR;
        const int kv = idx - r * VPR;
        put(sh, sl, L::off(r, kv * GVEC), v[i]);
      }
    } else {
      constexpr int KB = BK / GVEC;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int idx = tid + i * NT;
        if (!active(idx)) continue;
        const int rb = idx / KB;
        const int kb = idx - rb * KB;
#pragma unroll
        for (int t = 0; t < GVEC; ++t) {
          GV w;
#pragma unroll
          for (int u = 0; u < GVEC; ++u) w[u] = v[i * GVEC + u][t];
          put(sh, sl, L::off(rb * GVEC + t, kb * GVEC), w);
        }
      }
    }
  }

  __device__ __forceinline__ static void put(S* sh, S* sl, int off, GV w) {
    if constexpr (is_split(MODE)) {
      bf16x4 h, l;
      split_bf16x4(w, h, l);
      *reinterpret_cast<bf16x4*>(sh + off) = h;
      *reinterpret_cast<bf16x4*>(sl + off) = l;
    } else {
      (void)sl;
      *reinterpret_cast<GV*>(sh + off) = w;
    }
  }
};

__device__ __forceinline__ float fma_e(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_e(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <int MODE>
__device__ __forceinline__ typename Traits<MODE>::Acc zero_acc() {
  typename Traits<MODE>::Acc a;
#pragma unroll
  for (int r = 0; r < 4; ++r) a[r] = 0;
  return a;
}

// Row of accumulator element r held by lane group g of a 16x16 sub-tile.
// v_mfma_f32_16x16x4_f32 and v_mfma_f32_16x16x32_bf16 give lane group g rows
// 4g..4g+3; v_mfma_f64_16x16x4_f64 gives it rows g, g+4, g+8, g+12.
template <int MODE>
__device__ __forceinline__ int acc_row(int g, int r) {
  return MODE == kF64 ? g + 4 * r : 4 * g + r;
}

template <int MODE, int BM, int BN, int BK, int WM, int WN>
struct TileShape {
  static constexpr int WAVES_M = BM / WM;
  static constexpr int WAVES_N = BN / WN;
  static constexpr int NT = 64 * WAVES_M * WAVES_N;
  static constexpr int TM = WM / 16;
  static constexpr int TN = WN / 16;
  static constexpr int KR = 4 * Traits<MODE>::KPL;  // k covered per fragment read round
  static constexpr int PLANE = (BM + BN) * BK;       // elements of one hi or lo plane
  static constexpr int STAGE = PLANE * Traits<MODE>::PLANES;
  static_assert(BM % WM == 0 && BN % WN == 0 && WM % 16 == 0 && WN % 16 == 0, "wave tiling");
  static_assert(BK % KR == 0, "BK must cover whole fragment rounds");
};

// The products of one sub-tile row of the split modes for one 32-deep
// k-step: smallest terms first, then the leading hi*hi term, the product
// index outermost so TN independent MFMAs separate two updates of one
// accumulator.  kBF16x3 chains the three products into the running
// accumulator; kBF16x4 sums its four products of this k-step in a zeroed
// temporary and adds that with one round-to-nearest VALU add, so the large
// running sum is rounded once per k-step.
template <int MODE, int TN>
__device__ __forceinline__ void split_products(f32x4 (&row)[TN], const bf16x8& ah, const bf16x8& al,
                                               const bf16x8 (&bh)[TN], const bf16x8 (&bl)[TN]) {
  if constexpr (MODE == kBF16x4) {
    f32x4 t[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) t[j] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < TN; ++j) t[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bl[j], t[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < TN; ++j) t[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], t[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < TN; ++j) t[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], t[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < TN; ++j) t[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], t[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) row[j][r] += t[j][r];
  } else {
#pragma unroll
    for (int j = 0; j < TN; ++j) row[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], row[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < TN; ++j) row[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], row[j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < TN; ++j) row[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], row[j], 0, 0, 0);
  }
}

// The products of one sub-tile (one accumulator) for one 32-deep k-step, in
// the order split_products gives that accumulator: kBF16x3 lo*hi, hi*lo,
// hi*hi into the running sum; kBF16x4 lo*lo, lo*hi, hi*lo, hi*hi into a
// zeroed temporary, then one add.
template <int MODE>
__device__ __forceinline__ void split_products_one(f32x4& acc, const bf16x8& ah, const bf16x8& al,
                                                   const bf16x8& bh, const bf16x8& bl) {
  if constexpr (MODE == kBF16x4) {
    f32x4 t = f32x4{0.f, 0.f, 0.f, 0.f};
    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bl, t, 0, 0, 0);
    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, t, 0, 0, 0);
    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, t, 0, 0, 0);
    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, t, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] += t[r];
  } else {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
  }
}

// MFMA sequence over one LDS stage for one wave's TM x TN sub-tiles.
//
// PIPE selects how the copy of the next stage (mid(): split + LDS writes of
// the registers loaded at the top of the iteration) is placed:
//   0  all fragments of a k-round are read first and mid() runs after every
//      MFMA has been issued -- the longest time for the global loads to land,
//      and the widest distance between MFMAs on one accumulator;
//   1  B fragments are read per k-round, A fragments per sub-tile row, and
//      mid() runs between the two halves of the rows of the last k-round, so
//      its VALU and LDS work overlaps MFMAs still queued;
//   2  as 1, with mid() after three quarters of the rows (more time for the
//      global loads to land);
//   3  as 1, with mid() after the first quarter of the rows; the caller
//      double-buffers the loader registers so mid()'s data arrived an
//      iteration earlier (gemm_tile);
//   4  ping-pong (gemm_tile): half the waves copy before their MFMAs, half
//      after, each calling this with the PIPE 0 order;
//   7  spread loads (gemm_tile): every wave copies first, then calls this
//      with the PIPE 0 order and an after_row hook that issues the next
//      loads one scheduling group per sub-tile row.
// Which is faster depends on the tile (register headroom, MFMA latency of the
// type); each tile configuration fixes its choice from measurement.
template <int MODE, int BM, int BN, int BK, int WM, int WN, int PIPE, typename Mid, typename Row = void (*)(int)>
__device__ __forceinline__ void compute_stage(const typename Traits<MODE>::S* __restrict__ st,
                                              typename Traits<MODE>::Acc (&acc)[WM / 16][WN / 16],
                                              int wm0, int wn0, int lane, Mid&& mid,
                                              Row&& after_row = [](int) {}) {
  using Sh = TileShape<MODE, BM, BN, BK, WM, WN>;
  using T = Traits<MODE>;
  using Frag = typename T::Frag;
  using Acc = typename T::Acc;
  constexpr int TM = Sh::TM, TN = Sh::TN;
  using L = LdsLayout<MODE, BK>;
  constexpr int NQ = BK / Sh::KR;
  constexpr int kMidRow = PIPE == 1 ? TM / 2 : (PIPE == 2 ? (3 * TM) / 4 : TM / 4);  // row before which mid() runs
  const int g = lane >> 4;
  const int r16 = lane & 15;
  const typename T::S* sa = st;
  const typename T::S* sb = st + BM * BK;

  // One sub-tile row of the split modes (split_products).  (Generic lambda:
  // its body is instantiated only for the modes that call it.)
  auto split_row = [&](auto& row, const auto& ah, const auto& al, const auto& bh, const auto& bl) {
    split_products<MODE, TN>(row, ah, al, bh, bl);
  };
  auto mfma1 = [&](Acc& c, auto a, auto b) {
    if constexpr (MODE == kF32)
      c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    else
      c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  };

#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int kof = q * Sh::KR + g * T::KPL;
    if constexpr (is_split(MODE)) {
      const typename T::S* sal = st + Sh::PLANE;
      const typename T::S* sbl = sal + BM * BK;
      Frag bh[TN], bl[TN];
      if constexpr (PIPE != 0) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          const int off = L::off(wn0 + j * 16 + r16, kof);
          bh[j] = *reinterpret_cast<const Frag*>(sb + off);
          bl[j] = *reinterpret_cast<const Frag*>(sbl + off);
        }
      }
      if constexpr (PIPE == 0) {
        // Sub-tile by sub-tile, as in planar_tile: A fragments per row, the B
        // fragments of a column just before its first products, each
        // sub-tile's products kept together by a scheduling group.
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const int off = L::off(wm0 + i * 16 + r16, kof);
          const Frag ah = *reinterpret_cast<const Frag*>(sa + off);
          const Frag al = *reinterpret_cast<const Frag*>(sal + off);
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            if (i == 0) {
              const int offb = L::off(wn0 + j * 16 + r16, kof);
              bh[j] = *reinterpret_cast<const Frag*>(sb + offb);
              bl[j] = *reinterpret_cast<const Frag*>(sbl + offb);
            }
            split_products_one<MODE>(acc[i][j], ah, al, bh[j], bl[j]);
            __builtin_amdgcn_sched_group_barrier(0x008, MODE == kBF16x4 ? 4 : 3, 0);
          }
          if (q == NQ - 1) after_row(i);
        }
      } else {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          if (q == NQ - 1 && i == kMidRow) mid();
          const int off = L::off(wm0 + i * 16 + r16, kof);
          const Frag ah = *reinterpret_cast<const Frag*>(sa + off);
          const Frag al = *reinterpret_cast<const Frag*>(sal + off);
          split_row(acc[i], ah, al, bh, bl);
        }
      }
    } else {
      Frag b[TN];
#pragma unroll
      for (int j = 0; j < TN; ++j)
        b[j] = *reinterpret_cast<const Frag*>(sb + L::off(wn0 + j * 16 + r16, kof));
      if constexpr (PIPE == 0) {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          const Frag a = *reinterpret_cast<const Frag*>(sa + L::off(wm0 + i * 16 + r16, kof));
#pragma unroll
          for (int s = 0; s < T::STEPS; ++s)
#pragma unroll
            for (int j = 0; j < TN; ++j) mfma1(acc[i][j], a[s], b[j][s]);
          __builtin_amdgcn_sched_group_barrier(0x008, T::STEPS * TN, 0);
        }
      } else {
#pragma unroll
        for (int i = 0; i < TM; ++i) {
          if (q == NQ - 1 && i == kMidRow) mid();
          const Frag a = *reinterpret_cast<const Frag*>(sa + L::off(wm0 + i * 16 + r16, kof));
#pragma unroll
          for (int s = 0; s < T::STEPS; ++s)
#pragma unroll
            for (int j = 0; j < TN; ++j) mfma1(acc[i][j], a[s], b[j][s]);
        }
      }
    }
  }
  if constexpr (PIPE == 0) mid();
}

// 4x4 transpose inside each quad of lanes (lanes 4c .. 4c+3): on entry lane
// 4c+b holds x[r] = element (row r, column b) of the quad's 4x4 block; on
// exit it holds x[j] = element (row b, column j).  Two butterfly stages swap
// bit 0, then bit 1, of the (lane, element) indices with DPP quad
// permutations (xor 1: quad_perm [1,0,3,2]; xor 2: [2,3,0,1]).
__device__ __forceinline__ float dpp_xor1(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_xor2(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false));
}
__device__ __forceinline__ f32x4 quad_transpose(f32x4 x, int b) {
  const bool o0 = (b & 1) != 0;
  float s0 = o0 ? x[0] : x[1], s1 = o0 ? x[2] : x[3];
  float r0 = dpp_xor1(s0), r1 = dpp_xor1(s1);
  if (o0) {
    x[0] = r0;
    x[2] = r1;
  } else {
    x[1] = r0;
    x[3] = r1;
  }
  const bool o1 = (b & 2) != 0;
  s0 = o1 ? x[0] : x[2];
  s1 = o1 ? x[1] : x[3];
  r0 = dpp_xor2(s0);
  r1 = dpp_xor2(s1);
  if (o1) {
    x[0] = r0;
    x[1] = r1;
  } else {
    x[2] = r0;
    x[3] = r1;
  }
  return x;
}

// Epilogue of the whole wave tile.  split == 1: C = alpha*acc, plus beta*C
// when beta != 0 (C is not read when beta == 0); split > 1: the raw partial
// goes to ws[z] (scale 1, exact).  Destination, leading dimension, scale and
// whether to read C are uniform and chosen once; one loop serves all cases
// (separate loops per case made the compiler keep extra copies of the
// accumulators, raising the kernel's register count).
//
// The split modes' FP32 accumulators hold 4 consecutive rows of one column
// per lane; when the destination's rows start on 16-byte boundaries the 4x4
// blocks of each lane quad are transposed first (quad_transpose), so every
// lane owns 4 consecutive columns of one row and C is read and written with
// 16-byte accesses -- a quarter of the memory instructions of element-wise
// stores (a 256x256 tile's stores cost a grouped launch 4-7 us per tile
// element-wise; with 16-byte stores the 512-problem BF16x3 batch ran 5%
// faster, the pre-split BF16x4 8192^3 kernel 3%).  Each element's arithmetic
// is unchanged, so the values are bitwise those of the element-wise path,
// which serves unaligned destinations, FP64, and FP32 (whose 16-wave tile
// sits at 128 registers: the transpose made it spill, and its grouped launch
// ran 3% slower).  With beta != 0 the C values of one sub-tile row
// are loaded together before any of them is stored: a load cannot move above
// an earlier store that may alias it (ld is a run-time value).
template <int MODE, bool EDGE, int TM, int TN>
__device__ __forceinline__ void epilogue(const GemmParams<MODE>& p, int z, int row0, int col0,
                                         int lane, const typename Traits<MODE>::Acc (&acc)[TM][TN]) {
  using E = typename Traits<MODE>::E;
  const int g = lane >> 4;
  const bool partial = p.split > 1;
  const int64_t ld = partial ? (int64_t)p.N : p.ldc;
  E* const dst = partial ? p.ws + (int64_t)z * p.M * p.N : p.C;
  const E scale = partial ? E(1) : p.alpha;
  const bool read_c = !partial && p.beta != E(0);
  const int act = partial ? kActNone : p.act;
  const E beta = p.beta;
  if constexpr (is_split(MODE)) {
    if (((uintptr_t)dst % 16 == 0) && (ld % 4 == 0)) {
      const int b = lane & 3;
      const int trow = row0 + 4 * g + b;     // row of this lane after the transpose (sub-tile row 0)
      const int tcol = col0 + (lane & 12);   // first of its 4 columns (sub-tile column 0)
      gptr<E> base = to_global(dst + (int64_t)trow * ld + tcol);
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        const int row = trow + i * 16;
        gptr<E> q = base + (int64_t)(i * 16) * ld;
        const bool row_ok = !EDGE || row < p.M;
        f32x4 cv[TN];
        if (read_c) {
#pragma unroll
          for (int j = 0; j < TN; ++j) {
            const int col = tcol + j * 16;
            if (!EDGE || (row_ok && col + 4 <= p.N)) {
              cv[j] = *reinterpret_cast<gptr<const f32x4>>(q + j * 16);
            } else {
#pragma unroll
              for (int e = 0; e < 4; ++e) cv[j][e] = (row_ok && col + e < p.N) ? q[j * 16 + e] : E(0);
            }
          }
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          f32x4 v = quad_transpose(acc[i][j], b);
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            E out = scale * v[e];
            if (read_c) out = fma_e(beta, cv[j][e], out);
            if (act != kActNone) out = activate(out, act);
            v[e] = out;
          }
          const int col = tcol + j * 16;
          if (!EDGE || (row_ok && col + 4 <= p.N)) {
            *reinterpret_cast<gptr<f32x4>>(q + j * 16) = v;
          } else if (row_ok) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if (col + e < p.N) q[j * 16 + e] = v[e];
          }
        }
      }
      return;
    }
  }
  const int lrow = row0 + (MODE == kF64 ? g : 4 * g);  // row of element 0 of this lane
  const int lcol = col0 + (lane & 15);
  gptr<E> base = to_global(dst + (int64_t)lrow * ld + lcol);
#pragma unroll
  for (int i = 0; i < TM; ++i) {
    E cv[4][TN];
    if (read_c) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ro = i * 16 + (MODE == kF64 ? 4 * r : r);  // row offset of element r from lrow
        gptr<E> q = base + (int64_t)ro * ld;
#pragma unroll
        for (int j = 0; j < TN; ++j)
          cv[r][j] = (!EDGE || (lrow + ro < p.M && lcol + j * 16 < p.N)) ? q[j * 16] : E(0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ro = i * 16 + (MODE == kF64 ? 4 * r : r);
      gptr<E> q = base + (int64_t)ro * ld;
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        if (EDGE && (lrow + ro >= p.M || lcol + j * 16 >= p.N)) continue;
        E out = scale * acc[i][j][r];
        if (read_c) out = fma_e(beta, cv[r][j], out);
        if (act != kActNone) out = activate(out, act);
        q[j * 16] = out;
      }
    }
  }
}

// One output tile (tm, tn) of split z: the K-slice main loop and the
// epilogue.  lds holds two stages.  Ends with every wave past a barrier, so
// a caller may start another tile on the same LDS right away.
// EDGE == false requires M % BM == 0, N % BN == 0 and K % BK == 0.
template <int MODE, int BM, int BN, int BK, int WM, int WN, int PIPE, bool TA, bool TB, bool EDGE>
__device__ __forceinline__ void gemm_tile(const GemmParams<MODE>& p, int tm, int tn, int z,
                                          typename Traits<MODE>::S* lds) {
  using Sh = TileShape<MODE, BM, BN, BK, WM, WN>;
  using T = Traits<MODE>;
  using S = typename T::S;
  constexpr int NT = Sh::NT;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / Sh::WAVES_N) * WM;
  const int wn0 = (wave % Sh::WAVES_N) * WN;

  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const int kbeg = z * p.kslice;
  const int kend = min(p.K, kbeg + p.kslice);
  const int nk = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;

  TileLoader<MODE, BM, BK, NT, !TA, EDGE> la;
  TileLoader<MODE, BN, BK, NT, TB, EDGE> lb;
  la.init(p.A, p.lda, m0, kbeg, tid);
  lb.init(p.B, p.ldb, n0, kbeg, tid);

  typename T::Acc acc[Sh::TM][Sh::TN];
#pragma unroll
  for (int i = 0; i < Sh::TM; ++i)
#pragma unroll
    for (int j = 0; j < Sh::TN; ++j) acc[i][j] = zero_acc<MODE>();

  auto stage_ptr = [&](int s) { return lds + s * Sh::STAGE; };

  auto store_to = [&](const auto& a, const auto& b, S* st) {
    a.store(st, st + Sh::PLANE, tid);
    b.store(st + BM * BK, st + Sh::PLANE + BM * BK, tid);
  };

  // BF16x4's guarded loads spread among the MFMAs (EDGE, PIPE 7) need about
  // 400 registers and spill with 256 per wave; that instantiation runs the
  // ping-pong schedule, which fits (the two compute the same products in
  // the same order).
  constexpr int SCHED = (PIPE == 7 && EDGE && MODE == kBF16x4) ? 4 : PIPE;
  if constexpr (SCHED == 3) {
    // Two register sets: chunk it+2 is loaded while chunk it is computed and
    // chunk it+1 (loaded one iteration earlier) is written to LDS among the
    // MFMAs, so that copy never waits on memory.  The loop is unrolled by two
    // so each set is named statically.
    TileLoader<MODE, BM, BK, NT, !TA, EDGE> la1 = la;
    TileLoader<MODE, BN, BK, NT, TB, EDGE> lb1 = lb;
    if (nk > 0) {
      la.load(0, p.M, kend, tid, p.vec_a);
      lb.load(0, p.N, kend, tid, p.vec_b);
      store_to(la, lb, stage_ptr(0));
      const int k1 = (nk > 1 ? 1 : 0) * BK;
      la1.load(k1, p.M, kend, tid, p.vec_a);
      lb1.load(k1, p.N, kend, tid, p.vec_b);
      __syncthreads();
    }
    // One iteration: the set `ld` was written to LDS last iteration and is
    // free to receive chunk it+2 (clamped to the last chunk: always in
    // range); the set `st` holds chunk it+1.
    auto step = [&](int it, auto& lda_, auto& ldb_, const auto& sta, const auto& stb) {
      const int cur = it & 1;
      const int kofs = min(it + 2, nk - 1) * BK;
      lda_.load(kofs, p.M, kend, tid, p.vec_a);
      ldb_.load(kofs, p.N, kend, tid, p.vec_b);
      S* nxt = stage_ptr(cur ^ 1);
      compute_stage<MODE, BM, BN, BK, WM, WN, PIPE>(stage_ptr(cur), acc, wm0, wn0, lane,
                                                    [&]() { store_to(sta, stb, nxt); });
      __syncthreads();
    };
    int it = 0;
    for (; it + 1 < nk; it += 2) {
      step(it, la, lb, la1, lb1);
      step(it + 1, la1, lb1, la, lb);
    }
    if (it < nk) step(it, la, lb, la1, lb1);
  } else if constexpr (SCHED == 4) {
    // Ping-pong: waves in the upper half of the workgroup (one per SIMD for
    // 8 waves) copy first -- they write chunk it+1, loaded during the previous
    // iteration, then load chunk it+2 into the freed registers, then run the
    // MFMAs of chunk it; the lower half loads chunk it+1, runs its MFMAs and
    // copies last.  The two kinds of wave on a SIMD do their copies at
    // opposite ends of the iteration, so one wave's VALU/LDS work overlaps
    // the other's MFMAs.  Both write stage it+1 before the barrier and read
    // stage it after the previous one.
    const bool copy_first = wave >= Sh::WAVES_M * Sh::WAVES_N / 2;
    if (nk > 0) {
      la.load(0, p.M, kend, tid, p.vec_a);
      lb.load(0, p.N, kend, tid, p.vec_b);
      store_to(la, lb, stage_ptr(0));
      if (copy_first) {
        const int k1 = (nk > 1 ? 1 : 0) * BK;
        la.load(k1, p.M, kend, tid, p.vec_a);
        lb.load(k1, p.N, kend, tid, p.vec_b);
      }
      __syncthreads();
    }
    for (int it = 0; it < nk; ++it) {
      const int cur = it & 1;
      S* nxt = stage_ptr(cur ^ 1);
      if (copy_first) {
        store_to(la, lb, nxt);
        const int kofs = min(it + 2, nk - 1) * BK;
        la.load(kofs, p.M, kend, tid, p.vec_a);
        lb.load(kofs, p.N, kend, tid, p.vec_b);
        compute_stage<MODE, BM, BN, BK, WM, WN, 0>(stage_ptr(cur), acc, wm0, wn0, lane, []() {});
      } else {
        const int kofs = min(it + 1, nk - 1) * BK;
        la.load(kofs, p.M, kend, tid, p.vec_a);
        lb.load(kofs, p.N, kend, tid, p.vec_b);
        compute_stage<MODE, BM, BN, BK, WM, WN, 0>(stage_ptr(cur), acc, wm0, wn0, lane,
                                                   [&]() { store_to(la, lb, nxt); });
      }
      __syncthreads();
    }
  } else if constexpr (SCHED == 7) {
    // Every wave copies first: it writes chunk it+1 (loaded during the
    // previous iteration) to the free stage, then runs the MFMAs of chunk it
    // with the loads of chunk it+2 spread among them, one group after each
    // sub-tile row.  A wave's loads are capped by the memory pipeline's
    // queue: issued together before the MFMAs, the later ones wait for the
    // earlier ones to drain and hold back the MFMAs behind them.
    constexpr int NLA = decltype(la)::NREG, NLB = decltype(lb)::NREG, NL = NLA + NLB;
    if (nk > 0) {
      la.load(0, p.M, kend, tid, p.vec_a);
      lb.load(0, p.N, kend, tid, p.vec_b);
      store_to(la, lb, stage_ptr(0));
      const int k1 = (nk > 1 ? 1 : 0) * BK;
      la.load(k1, p.M, kend, tid, p.vec_a);
      lb.load(k1, p.N, kend, tid, p.vec_b);
      __syncthreads();
    }
    for (int it = 0; it < nk; ++it) {
      const int cur = it & 1;
      store_to(la, lb, stage_ptr(cur ^ 1));
      const int kofs = min(it + 2, nk - 1) * BK;
      compute_stage<MODE, BM, BN, BK, WM, WN, 0>(stage_ptr(cur), acc, wm0, wn0, lane, []() {}, [&](int i) {
#pragma unroll
        for (int j = i * NL / Sh::TM; j < (i + 1) * NL / Sh::TM; ++j) {
          if (j < NLA)
            la.load_one(j, kofs, p.M, kend, tid, p.vec_a);
          else
            lb.load_one(j - NLA, kofs, p.N, kend, tid, p.vec_b);
        }
        constexpr int kPerRow = NL / Sh::TM > 0 ? NL / Sh::TM : 1;
        __builtin_amdgcn_sched_group_barrier(0x020, kPerRow, 0);  // these loads, after this row's MFMAs
      });
      __syncthreads();
    }
  } else {
    if (nk > 0) {
      la.load(0, p.M, kend, tid, p.vec_a);
      lb.load(0, p.N, kend, tid, p.vec_b);
      store_to(la, lb, stage_ptr(0));
      __syncthreads();
    }
    // Each iteration loads chunk it+1 (the last one reloads its own chunk, in
    // range and never read) and writes it to the other stage during or after
    // the MFMAs of chunk it (PIPE); no branch splits the loop body.
    for (int it = 0; it < nk; ++it) {
      const int cur = it & 1;
      const int kofs = (it + 1 < nk ? it + 1 : it) * BK;
      la.load(kofs, p.M, kend, tid, p.vec_a);
      lb.load(kofs, p.N, kend, tid, p.vec_b);
      S* nxt = stage_ptr(cur ^ 1);
      compute_stage<MODE, BM, BN, BK, WM, WN, PIPE>(stage_ptr(cur), acc, wm0, wn0, lane,
                                                    [&]() { store_to(la, lb, nxt); });
      __syncthreads();
    }
  }

  epilogue<MODE, EDGE, Sh::TM, Sh::TN>(p, z, m0 + wm0, n0 + wn0, lane, acc);
}

// The GEMM kernel.  Grid: tiles_m * tiles_n * split workgroups (1-D).
// OCC: minimum waves per SIMD the register allocation must allow (1: no
// constraint beyond the workgroup size).
template <int MODE, int BM, int BN, int BK, int WM, int WN, int PIPE, int OCC, bool TA, bool TB,
          bool EDGE>
__global__ void __launch_bounds__(64 * (BM / WM) * (BN / WN)) __attribute__((amdgpu_waves_per_eu(OCC)))
    gemm_kernel(GemmParams<MODE> p) {
  using Sh = TileShape<MODE, BM, BN, BK, WM, WN>;
  __shared__ __attribute__((aligned(16))) typename Traits<MODE>::S lds[2 * Sh::STAGE];
  int tm, tn, z;
  map_block(blockIdx.x, gridDim.x, p.tiles_m, p.tiles_n, p.group_m, p.xcds, tm, tn, z);
  gemm_tile<MODE, BM, BN, BK, WM, WN, PIPE, TA, TB, EDGE>(p, tm, tn, z, lds);
}

// ---------------------------------------------------------------------------
// Grouped GEMM: many independent row-major problems in one persistent launch.
// A launch runs one tile family: tile variants (Vs) with one workgroup size,
// each problem of the table using one of them (GroupEntry::variant), so thin
// outputs take thin tiles inside the same launch and under the same
// schedule.  sched holds the host-computed schedule: sched[w] .. sched[w + 1]
// index the tiles of workgroup w, each a pair (problem, tile within the
// problem) at sched[gridDim.x + 1 + 2i] and the next int (a longest-first
// assignment that balances the work per workgroup).  A problem's tiles are
// numbered z * tiles_m * tiles_n + tm * tiles_n + tn over its split-K slices
// z.  Each tile is computed with the same gemm_tile as the single-problem
// kernel, so a problem's result is bitwise the single launch's of the same
// tile configuration and split: slices of a split member write their raw
// partials to the member's workspace (ws[z][M][N]), which
// grouped_reduce_kernel sums in ascending z after the launch.  In the split
// modes a tile that lies inside its problem (and whose operands allow 16-byte
// loads and whose K is a whole number of chunks) runs the unchecked loaders,
// the others the bounds-checked ones -- the same values either way.
// ---------------------------------------------------------------------------
template <int BM_, int BN_, int BK_, int WM_, int WN_, int PIPE_>
struct TileV {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_, PIPE = PIPE_;
  static constexpr int NT = 64 * (BM / WM) * (BN / WN);
};

template <int MODE, class V>
__device__ __forceinline__ void grouped_tile(const GroupEntry<typename Traits<MODE>::G>& e, int local,
                                             typename Traits<MODE>::S* lds) {
  using E = typename Traits<MODE>::E;
  GemmParams<MODE> p;
  p.A = e.A;
  p.B = e.B;
  p.C = e.C;
  p.ws = static_cast<E*>(e.ws);
  p.act = kActNone;
  p.lda = e.lda;
  p.ldb = e.ldb;
  p.ldc = e.ldc;
  p.M = e.M;
  p.N = e.N;
  p.K = e.K;
  p.tiles_m = e.tiles_m;
  p.tiles_n = e.tiles_n;
  p.split = e.split;
  p.kslice = e.kslice;
  p.group_m = 1;
  p.xcds = 1;
  p.vec_a = e.vec_a;
  p.vec_b = e.vec_b;
  p.alpha = (E)e.alpha;
  p.beta = (E)e.beta;
  const int per = e.tiles_m * e.tiles_n;
  const int z = local / per;
  const int t = local - z * per;
  const int tm = t / e.tiles_n;
  const int tn = t - tm * e.tiles_n;
  // The split modes' kernels have 256 registers per lane and run interior
  // tiles on the unchecked loaders (BF16x3 512-problem batch 6% faster);
  // FP32 and FP64 families cap registers at 128, and a second copy of the
  // tile body made them spill (FP32 3% slower), so they keep one.
  const bool inside = is_split(MODE) && e.vec_a && e.vec_b && (tm + 1) * V::BM <= e.M &&
                      (tn + 1) * V::BN <= e.N && e.K % V::BK == 0;
  if constexpr (!is_split(MODE)) {
    (void)inside;
    gemm_tile<MODE, V::BM, V::BN, V::BK, V::WM, V::WN, V::PIPE, false, false, true>(p, tm, tn, z, lds);
  } else if (inside)
    gemm_tile<MODE, V::BM, V::BN, V::BK, V::WM, V::WN, V::PIPE, false, false, false>(p, tm, tn, z, lds);
  else
    gemm_tile<MODE, V::BM, V::BN, V::BK, V::WM, V::WN, V::PIPE, false, false, true>(p, tm, tn, z, lds);
}

template <int MODE, class V0, class... Vs>
__device__ __forceinline__ void grouped_dispatch(int v, const GroupEntry<typename Traits<MODE>::G>& e,
                                                 int local, typename Traits<MODE>::S* lds) {
  if constexpr (sizeof...(Vs) == 0) {
    (void)v;
    grouped_tile<MODE, V0>(e, local, lds);
  } else {
    if (v == 0)
      grouped_tile<MODE, V0>(e, local, lds);
    else
      grouped_dispatch<MODE, Vs...>(v - 1, e, local, lds);
  }
}

template <class V0, class... Vs>
struct FamilyNT {
  static constexpr int value = V0::NT;
  static_assert(((Vs::NT == V0::NT) && ...), "every variant of a family has one workgroup size");
};

// LDS elements of two stages of the largest variant.
template <int MODE, class... Vs>
constexpr int family_lds_elems() {
  int m = 0;
  ((m = m > 2 * TileShape<MODE, Vs::BM, Vs::BN, Vs::BK, Vs::WM, Vs::WN>::STAGE
            ? m
            : 2 * TileShape<MODE, Vs::BM, Vs::BN, Vs::BK, Vs::WM, Vs::WN>::STAGE),
   ...);
  return m;
}

template <int MODE, int OCC, class... Vs>
__global__ void __launch_bounds__(FamilyNT<Vs...>::value) __attribute__((amdgpu_waves_per_eu(OCC)))
    grouped_kernel(const GroupEntry<typename Traits<MODE>::G>* __restrict__ table,
                   const int* __restrict__ sched) {
  __shared__ __attribute__((aligned(16))) typename Traits<MODE>::S lds[family_lds_elems<MODE, Vs...>()];
  const int* order = sched + gridDim.x + 1;
  const int end = sched[blockIdx.x + 1];
  for (int i = sched[blockIdx.x]; i < end; ++i) {
    const int prob = order[2 * (int64_t)i];
    const int local = order[2 * (int64_t)i + 1];
    const GroupEntry<typename Traits<MODE>::G>& e = table[prob];
    grouped_dispatch<MODE, Vs...>(e.variant, e, local, lds);
  }
}

// Fixed-order reduction of the split members of a grouped launch: for each
// member, C = alpha * sum_{z ascending} ws[z] (+ beta * C), the arithmetic of
// splitk_reduce_kernel (no activation).  Block b serves member i with
// tab[i].block0 <= b < tab[i + 1].block0 (256 elements per block).
template <typename G, typename E>
__global__ void __launch_bounds__(256)
    grouped_reduce_kernel(const GroupReduce<G>* __restrict__ tab, int count) {
  const int b = blockIdx.x;
  int lo = 0, hi = count - 1;  // last member whose block0 <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].block0 <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  const GroupReduce<G>& r = tab[lo];
  const int64_t total = (int64_t)r.M * r.N;
  const int64_t e = (int64_t)(b - r.block0) * 256 + threadIdx.x;
  if (e >= total) return;
  const E* p = static_cast<const E*>(r.ws) + e;
  E v = E(0);
  for (int z = 0; z < r.S; ++z) v += p[(int64_t)z * total];
  const int row = (int)(e / r.N);
  const int col = (int)(e - (int64_t)row * r.N);
  G* cp = r.C + (int64_t)row * r.ldc + col;
  E out = (E)r.alpha * v;
  if ((E)r.beta != E(0)) out = fma_e((E)r.beta, *cp, out);
  *cp = out;
}

// ---------------------------------------------------------------------------
// Pre-split BF16 path.  A pass (split_pair / split_one) writes each FP32
// operand once as bf16 "planes": row r (an M or N index) of 2*Kp elements,
// each 32-deep K chunk stored as its 32 hi parts followed by its 32 lo parts
// (hi = RNE(x), lo = RNE(x - hi), the split the on-the-fly path does), K
// zero-padded to Kp, a multiple of 32.  Both operands are then K-major, so the
// GEMM (planar_kernel) copies 16-byte chunks straight from global memory into
// LDS with global_load_lds_dwordx4: no VALU split, no staging registers, no
// LDS store instructions.  Products and their order are those of the
// on-the-fly kernel with the same tile, so the results are bitwise equal.
// ---------------------------------------------------------------------------

// One operand's split: element (r, k) at X[r * ld + k] when kmajor (k-contiguous
// rows), else at X[k * ld + r].  blocks: workgroups of 256 threads it takes.
struct SplitJob {
  const float* X;
  int64_t ld;
  __bf16* out;
  int R, K, Kp;
  int kmajor;
  int vec;     // rows start on 16-byte boundaries and ld % 4 == 0
  int nrb;     // mn-major: ceil(R / 256) row blocks
  int blocks;
};

// k-major operand: one thread per (row, 8 consecutive k); t is the thread's
// global index.
__device__ __forceinline__ void split_kmajor_body(const SplitJob& j, int64_t t) {
  const int cpr = j.Kp / 8;
  if (t >= (int64_t)j.R * cpr) return;
  const int r = (int)(t / cpr);
  const int k0 = (int)(t - (int64_t)r * cpr) * 8;
  const float* x = j.X + (int64_t)r * j.ld + k0;
  f32x4 v0, v1;
  if (j.vec && k0 + 8 <= j.K) {
    v0 = *reinterpret_cast<const f32x4*>(x);
    v1 = *reinterpret_cast<const f32x4*>(x + 4);
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      v0[e] = k0 + e < j.K ? x[e] : 0.f;
      v1[e] = k0 + 4 + e < j.K ? x[4 + e] : 0.f;
    }
  }
  bf16x4 h0, l0, h1, l1;
  split_bf16x4(v0, h0, l0);
  split_bf16x4(v1, h1, l1);
  __bf16* o = j.out + (int64_t)r * 2 * j.Kp + (int64_t)(k0 / 32) * 64 + (k0 % 32);
  *reinterpret_cast<bf16x8*>(o) = bf16x8{h0[0], h0[1], h0[2], h0[3], h1[0], h1[1], h1[2], h1[3]};
  *reinterpret_cast<bf16x8*>(o + 32) = bf16x8{l0[0], l0[1], l0[2], l0[3], l1[0], l1[1], l1[2], l1[3]};
}

// mn-major operand (r-contiguous), transposed in registers.  A thread loads
// 4 consecutive r from each of 8 consecutive k rows (one 16-byte load per row
// when the rows allow it), so for each of its r it holds the 8 k of one hi
// and one lo 16-byte chunk of the plane row.  Lanes: the 4 k groups of a
// 32-deep chunk fastest, then 16 groups of 4 r, so each load instruction
// reads 256 contiguous bytes of 4 k rows and each store instruction fills
// 64-byte halves of 16 plane lines (hi, then lo).  Workgroup b covers 256 r
// x 32 k: row block b % nrb, k chunk b / nrb (row blocks fastest: a
// chunk-fastest order measured 7-16% slower).
__device__ __forceinline__ void split_mnmajor_body(const SplitJob& j, int b) {
  const int lane = threadIdx.x & 63;
  const int r0 = (b % j.nrb) * 256 + (int)(threadIdx.x >> 6) * 64 + (lane >> 2) * 4;
  const int chunk = b / j.nrb;
  const int k0 = chunk * 32 + (lane & 3) * 8;
  if (r0 >= j.R) return;
  const bool full = j.vec && r0 + 4 <= j.R;
  f32x4 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = k0 + e;
    const float* x = j.X + (int64_t)k * j.ld + r0;
    if (full && k < j.K) {
      v[e] = *reinterpret_cast<const f32x4*>(x);
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[e][q] = (k < j.K && r0 + q < j.R) ? x[q] : 0.f;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (r0 + q >= j.R) break;
    const f32x4 a{v[0][q], v[1][q], v[2][q], v[3][q]};
    const f32x4 c{v[4][q], v[5][q], v[6][q], v[7][q]};
    bf16x4 h0, l0, h1, l1;
    split_bf16x4(a, h0, l0);
    split_bf16x4(c, h1, l1);
    __bf16* o = j.out + (int64_t)(r0 + q) * 2 * j.Kp + (int64_t)chunk * 64 + (lane & 3) * 8;
    *reinterpret_cast<bf16x8*>(o) = bf16x8{h0[0], h0[1], h0[2], h0[3], h1[0], h1[1], h1[2], h1[3]};
    *reinterpret_cast<bf16x8*>(o + 32) = bf16x8{l0[0], l0[1], l0[2], l0[3], l1[0], l1[1], l1[2], l1[3]};
  }
}

// Workgroup b of one operand's split.
__device__ __forceinline__ void split_block(const SplitJob& j, int b) {
  if (j.kmajor)
    split_kmajor_body(j, (int64_t)b * 256 + threadIdx.x);
  else
    split_mnmajor_body(j, b);
}

// One operand: j.blocks workgroups.
template <int MODE>
__global__ void __launch_bounds__(256) split_one(SplitJob j) {
  split_block(j, blockIdx.x);
}

// Both operands in one launch: workgroups [0, a.blocks) split a, the rest b.
// One launch instead of two saves a kernel boundary and lets the second
// operand's first workgroups fill the first one's tail (8192^3, A k-major and
// B mn-major, the two orders alternated in one process: 0.176-0.209 against
// 0.179-0.214 ms for two launches, 2-4% less, same planes).
template <int MODE>
__global__ void __launch_bounds__(256) split_pair(SplitJob a, SplitJob b) {
  const int bid = blockIdx.x;
  if (bid < a.blocks)
    split_block(a, bid);
  else
    split_block(b, bid - a.blocks);
}

template <int MODE>
struct PlanarParams {
  GemmParams<MODE> g;  // C, ws, M, N, K, tiles, split, kslice, group_m, alpha, beta, act
  const __bf16* Ap;    // [M][2*Kp]
  const __bf16* Bp;    // [N][2*Kp]
  int Kp;
};

// GEMM over planes.  Workgroup tile BM x BN, waves of WM x WN, two LDS stages
// of (BM + BN) rows x 64 bf16 (hi and lo of a 32-deep chunk) in the 128-byte
// row layout of LdsLayout.  Each global_load_lds_dwordx4 of a wave fills one
// 1-KiB block of 8 LDS rows: lane L writes 16-byte unit L of the block, so it
// loads the chunk that the layout places there.  Rows beyond M (N) load the
// last valid row and are never stored.  Loads for chunk it+1 are issued
// before the MFMAs of chunk it and waited for (vmcnt) before the barrier that
// ends the iteration.
template <int MODE, int BM, int BN, int WM, int WN, bool EDGE>
__device__ __forceinline__ void planar_tile(const PlanarParams<MODE>& pp, __bf16* lds) {
  constexpr int BK = 32;
  constexpr int ROWE = 2 * BK;              // elements per LDS row
  constexpr int STAGE = (BM + BN) * ROWE;
  constexpr int NT = 64 * (BM / WM) * (BN / WN);
  constexpr int NW = NT / 64;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int BLK_A = BM / 8, BLK_B = BN / 8;  // 1-KiB blocks per operand and stage
  static_assert(BLK_A % NW == 0 && BLK_B % NW == 0, "whole blocks per wave");
  constexpr int PA = BLK_A / NW, PB = BLK_B / NW;
  using L = LdsLayout<MODE, ROWE>;
  static_assert(L::NC == 8, "128-byte LDS rows");
  const GemmParams<MODE>& p = pp.g;
  int tm, tn, z;
  map_block(blockIdx.x, gridDim.x, p.tiles_m, p.tiles_n, p.group_m, p.xcds, tm, tn, z);
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave / (BN / WN)) * WM;
  const int wn0 = (wave % (BN / WN)) * WN;
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const int64_t rowlen = 2 * (int64_t)pp.Kp;

  // Source of this lane's unit in each of its blocks (chunk 0 of the slice).
  const int cbeg = z * (p.kslice / BK);
  const int cend = min(pp.Kp / BK, cbeg + p.kslice / BK);
  const int nk = cend > cbeg ? cend - cbeg : 0;
  // Row r = 8 * block + lane / 8 of a block loads chunk c of the layout; the
  // chunk permutation depends on (r / 2) % 4 = (lane / 16) % 4 only, so all
  // blocks of a lane share c.  Without edges every block's source is a
  // wave-uniform row base (scalar registers) plus one per-lane byte offset
  // shared by all blocks of both operands (rowlen * 2 * 8 < 2^32 is checked
  // by the launcher); 64-bit per-lane pointers per block would push the
  // 16-wave BF16x4 kernel past its 128 registers.  With edges each block
  // clamps its rows to the operand, so each keeps its own pointer.
  const int rin = lane >> 3;
  const int c = (lane & 7) ^ (((rin >> 1) & 3) << 1);  // inverse of LdsLayout (an involution)
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  const uint32_t loff = (uint32_t)(((int64_t)rin * rowlen + c * 8) * (int64_t)sizeof(__bf16));
  const __bf16* srcA[EDGE ? PA : 1];
  const __bf16* srcB[EDGE ? PB : 1];
  if constexpr (EDGE) {
#pragma unroll
    for (int j = 0; j < PA; ++j)
      srcA[j] = pp.Ap + (int64_t)min(m0 + 8 * (wv * PA + j) + rin, p.M - 1) * rowlen + (int64_t)cbeg * ROWE + c * 8;
#pragma unroll
    for (int j = 0; j < PB; ++j)
      srcB[j] = pp.Bp + (int64_t)min(n0 + 8 * (wv * PB + j) + rin, p.N - 1) * rowlen + (int64_t)cbeg * ROWE + c * 8;
  }
  auto issue = [&](int it, int s) {
    __bf16* st = lds + s * STAGE;
#pragma unroll
    for (int j = 0; j < PA; ++j) {
      const __bf16* g;
      if constexpr (EDGE) {
        g = srcA[j] + it * ROWE;
      } else {
        const char* base = reinterpret_cast<const char*>(pp.Ap + (int64_t)(m0 + 8 * (wv * PA + j)) * rowlen +
                                                         (int64_t)(cbeg + it) * ROWE);
        g = reinterpret_cast<const __bf16*>(base + loff);
      }
      __builtin_amdgcn_global_load_lds(g, st + (wv * PA + j) * 8 * ROWE, 16, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < PB; ++j) {
      const __bf16* g;
      if constexpr (EDGE) {
        g = srcB[j] + it * ROWE;
      } else {
        const char* base = reinterpret_cast<const char*>(pp.Bp + (int64_t)(n0 + 8 * (wv * PB + j)) * rowlen +
                                                         (int64_t)(cbeg + it) * ROWE);
        g = reinterpret_cast<const __bf16*>(base + loff);
      }
      __builtin_amdgcn_global_load_lds(g, st + BM * ROWE + (wv * PB + j) * 8 * ROWE, 16, 0, 0);
    }
  };

  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int g = lane >> 4;
  const int r16 = lane & 15;
  if (nk > 0) {
    issue(0, 0);
    // The direct loads write LDS behind the compiler's back: wait for them
    // and fence compiler reordering before the barrier publishes the stage.
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
  for (int it = 0; it < nk; ++it) {
    const int cur = it & 1;
    if (it + 1 < nk) issue(it + 1, cur ^ 1);
    const __bf16* sa = lds + cur * STAGE;
    const __bf16* sb = sa + BM * ROWE;
    // Sub-tile by sub-tile, row-major: the first row reads each B column's
    // fragments just before its products, so a wave's first MFMAs wait for
    // four fragment reads instead of all of them (every wave reads at once
    // after the barrier); each sub-tile's products are kept together in
    // program order by a scheduling group.  Each accumulator receives its
    // products in split_products' order, so results are unchanged.
    bf16x8 bh[TN], bl[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int row = wm0 + i * 16 + r16;
      const bf16x8 ah = *reinterpret_cast<const bf16x8*>(sa + L::off(row, g * 8));
      const bf16x8 al = *reinterpret_cast<const bf16x8*>(sa + L::off(row, BK + g * 8));
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        if (i == 0) {
          const int rowb = wn0 + j * 16 + r16;
          bh[j] = *reinterpret_cast<const bf16x8*>(sb + L::off(rowb, g * 8));
          bl[j] = *reinterpret_cast<const bf16x8*>(sb + L::off(rowb, BK + g * 8));
        }
        split_products_one<MODE>(acc[i][j], ah, al, bh[j], bl[j]);
        __builtin_amdgcn_sched_group_barrier(0x008, MODE == kBF16x4 ? 4 : 3, 0);  // this sub-tile's MFMAs
      }
    }
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  }
  epilogue<MODE, EDGE, TM, TN>(p, z, m0 + wm0, n0 + wn0, lane, acc);
}

// The pre-split GEMM kernel: grid tiles_m * tiles_n * split (1-D), two LDS
// stages of (BM + BN) x 64 bf16.
template <int MODE, int BM, int BN, int WM, int WN, bool EDGE>
__global__ void __launch_bounds__(64 * (BM / WM) * (BN / WN)) planar_kernel(PlanarParams<MODE> pp) {
  __shared__ __attribute__((aligned(16))) __bf16 lds[2 * (BM + BN) * 64];
  planar_tile<MODE, BM, BN, WM, WN, EDGE>(pp, lds);
}

}  // namespace cosma
