// cosma_api.cpp -- the C ABI declared in include/cosma/cosma.h.
//
// Host side of the reference's cosmaSgemm (Kernel/cuda/src/cuCOSMAV100.cuh:2774-2848):
// argument checks, the alpha == 0 / beta pre-scale branch (:2782-2797), tile
// and split-K choice (:2797-2822, one build per config.h row in the
// reference), split-K workspace (:2824-2838) and the launches.  Every check
// that can fail runs before the first HIP call, so a contract violation
// returns a negative status without touching the GPU.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <functional>
#include <map>
#include <mutex>
#include <queue>
#include <tuple>
#include <vector>

#include "../../include/cosma/cosma.h"
#include "cosma_internal.hpp"

#ifndef COSMA_M
#define COSMA_M 1024
#endif
#ifndef COSMA_N
#define COSMA_N 1024
#endif
#ifndef COSMA_K
#define COSMA_K 1024
#endif

namespace {

using cosma::Problem;

constexpr int kMaxSplit = 4096;
constexpr int kCUs = 256;  // MI355X compute units

// Tile extents of the built configurations, mirrored as constants so the
// fixed-shape entry points can pick them at compile time.  Must match the
// tables in cosma_f32.hip / cosma_f64.hip / cosma_bf16x3.hip (checked at
// first use by check_tables()).
struct TileDims {
  int bm, bn, bk;
};
constexpr int kMaxCfg = 9;
constexpr TileDims kDims[4][kMaxCfg] = {
    // FP32
    {{128, 128, 32}, {64, 64, 32}, {128, 64, 32}, {256, 128, 32}, {128, 128, 64}, {256, 128, 32},
     {256, 128, 32}, {256, 128, 32}},
    // FP64
    {{128, 128, 16}, {64, 64, 16}, {128, 64, 16}, {128, 128, 16}, {64, 64, 32}, {128, 128, 32},
     {256, 128, 16}, {128, 128, 16}},
    // BF16X3 (cfg 7: pre-split operands)
    {{128, 128, 32}, {64, 64, 32}, {256, 128, 32}, {256, 256, 32}, {256, 128, 32}, {128, 128, 64},
     {256, 256, 32}, {256, 256, 32}, {256, 256, 32}},
    // BF16X4 (same tiles as BF16X3)
    {{128, 128, 32}, {64, 64, 32}, {256, 128, 32}, {256, 256, 32}, {256, 128, 32}, {128, 128, 64},
     {256, 256, 32}, {256, 256, 32}, {256, 256, 32}},
};
constexpr int kNumModes = 4;
constexpr int kNumCfg[kNumModes] = {8, 8, 9, 9};

constexpr int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

constexpr int tiles_of(int mode, int cfg, int m, int n) {
  return cdiv(m, kDims[mode][cfg].bm) * cdiv(n, kDims[mode][cfg].bn);
}

// Heuristic tile / split choice, used when neither the caller nor the tuned
// table (cosma_tune_set) fixes them.  Derived from autotuner sweeps on one
// MI355X:
//  - BF16x3 / BF16x4: the large tile -- 256x256, 8 waves of 128x64 with
//    ping-pong copies (for BF16x4 0-14% above the copy-between-MFMAs version
//    on eight shapes, within 3% below on two) -- whenever the output fills
//    at least half of its tiles and the tiles, with K split into slices of at
//    least 512, give at least half the CUs a workgroup.  K is split only when
//    the tiles alone leave more than half the CUs idle, until every CU holds
//    one workgroup (256x256x1048576: 256 slices of 4096, so every input byte
//    is read from HBM once).  It beat every smaller tile from 2048^3 up, and
//    for 4096x4096x16384 (the weight gradient of a 16384x4096x4096 layer) ran
//    at twice the speed of 128x128.  The 256x256 tile runs over pre-split
//    operands (planar_cfg) when its rounds of workgroups (ceil(tiles / CUs))
//    times 5504 reach M + N: the split pass costs (M+N)*K*8 bytes of HBM
//    traffic, the GEMM without the in-loop split saves 16-22% of each round,
//    and both scale with K (8192^3: 2.35 ms in total against 2.64 for cfg 6;
//    4608^2x4096, 5120^2x4096, 4352^3, 4096x6144x4096 and 4096x12288x4096,
//    with partly filled or few rounds, 6-8% faster in both split modes;
//    4096^3, 2048x8192x2048 and 4096x8192x4096 tie or lose, and keep cfg 6;
//    5632^2x2048 lost 1% and is just outside).  5504 >= 2^24 / 3072, so every
//    output with M*N/(M+N) >= 3072 also qualifies.  The pre-split tile is
//    also taken when its 256x256 tiles number between half the CUs and all
//    of them (one workgroup per CU, some CUs idle): there the 16-wave kernel's
//    faster K loop outweighs the split pass (BF16x3 +2% to +17% over nine
//    shapes from 2048x4096x4096 to 3584^2x4096, row-major and A transposed;
//    BF16x4 -3% to +3% row-major, +5% to +13% with A transposed);
//  - FP32 / FP64 large outputs (at least two workgroups per CU): FP32 256x128
//    in 16 waves of 64x32 with the next stage held in registers (four waves
//    per SIMD; 3-5% above the 8-wave version, which itself matched the
//    copy-between-MFMAs variant with A row-major and beat it by 12% with B
//    stored transposed), FP64 128x128 (8 waves of 64x32, 4 waves per SIMD,
//    ping-pong copies; a 16-wave version of it ran 12% slower).  FP32 splits K in two when
//    the last round of workgroups would be at most half full (6144^3: 4.5
//    rounds of 256 become 9; +7%);
//  - deep-K outputs (few tiles, K >= 16384): FP32 splits K over the 16-wave
//    256x128 tile until every CU holds one workgroup (0-6% above the 8-wave
//    ping-pong version of the tile on four deep shapes); FP64, and outputs
//    smaller than the large tile, split a 128-row tile with a long K chunk
//    (128x64x32, or 128x128x64 for BF16) until every CU holds two (128x64) or
//    one (128x128x64) workgroups;
//  - medium outputs (autotuner comparisons on one MI355X): BF16 splits K
//    over the large tile unless that leaves CUs idle with slices of at
//    most 512, then over 128x128 tiles toward two workgroups per CU
//    (1024x1024x4096: 213 TFLOP/s against 160 over the large tile; at
//    1536^2x4096 the large tile with 1024-deep slices stays ahead); FP32
//    runs the large tile with K split in two when that makes about one
//    round of workgroups (2048^3: 112 against 108 for 128x64); FP64 takes
//    the 128x128 tile when it fills the CUs, K split in two unless its
//    tiles make whole rounds and K < 4096 (2048^3: 65 against 54;
//    3072x2048x2048: 65 against 50 unsplit); FP32 128x64 tiles split K in
//    two when the last round would be at most half full (3072^3: 119
//    against 110);
//  - otherwise 128x128 (BF16) / 128x64 tiles while they give every CU a
//    workgroup, else 64x64; split K when the tiles give the CUs fewer than
//    two workgroups each, aiming at three, with K slices at least 256 deep
//    (1024^3 with 64x64 tiles: FP32 split 3 ran in 0.038 ms against 0.050
//    unsplit, BF16x3 split 3 in 0.031 against 0.035 with split 2; cache
//    flushed, beta 0.5).
struct Choice {
  int cfg, split;
};
constexpr bool split_mode(int mode) { return mode == 2 || mode == 3; }
// Tile configurations over pre-split operands (cfg 7: 8 waves of 128x64,
// cfg 8: 16 waves of 64x64; no grouped launch) and the one the heuristic
// uses: 16 waves, four per SIMD (BF16x3 124 VGPRs; BF16x4 128, its per-k-step
// temporary kept for one sub-tile at a time; 8192^3 BF16x4 371 TFLOP/s per
// call against 352 with 8 waves, 3072^3 and 6144^3 +12-16%).
constexpr bool is_planar(int mode, int cfg) { return split_mode(mode) && (cfg == 7 || cfg == 8); }
constexpr int planar_cfg(int mode) { (void)mode; return 8; }
constexpr int large_cfg(int mode) { return mode == 0 ? 7 : (mode == 1 ? 7 : 6); }
constexpr int medium_cfg(int mode) { return split_mode(mode) ? 0 : 2; }
// Deep-K tile: the large tile when the output fills it, else a 128-row tile
// with a long K chunk (128x64x32 for FP32/FP64, 128x128x64 for BF16).
constexpr int deep_cfg(int mode, int m, int n) {
  const int big = mode == 0 ? 7 : (mode == 1 ? 2 : 6);
  if (m >= kDims[mode][big].bm && n >= kDims[mode][big].bn) return big;
  return split_mode(mode) ? 5 : 2;
}
constexpr int clamp_split(int split, int k, int min_depth, int max_split = 64) {
  int max_by_depth = k / min_depth;
  if (split > max_by_depth) split = max_by_depth;
  if (split > max_split) split = max_split;
  return split < 1 ? 1 : split;
}
// ta: op(A) is stored transposed (K x M).  The on-the-fly tiles transpose A
// in their loaders; the pre-split pass does it once, so with a transposed A the
// pre-split tile wins even at one round of workgroups (BF16x3 4096x4096x16384,
// the weight gradient of a 16384x4096x4096 layer: 1.185 ms against 1.353 for
// cfg 6; BF16x4 1.562 against 1.606; 4096^2x1024 0.91x, 6144^2x1024 0.77x,
// 2048x8192x2048 0.89x the cfg 6 time).  Not for K < 512, where the split
// pass and its launches outweigh it (4096^2x256: 1.15x).
constexpr Choice heuristic(int mode, int m, int n, int k, int ta = 0) {
  int cfg = large_cfg(mode);
  const int lt = tiles_of(mode, cfg, m, n);
  if (split_mode(mode)) {
    const int64_t covered = (int64_t)lt * kDims[mode][cfg].bm * kDims[mode][cfg].bn;
    const int split = lt >= kCUs / 2 ? 1 : clamp_split(kCUs / lt, k, 512, kCUs);
    // K is split over the large tile while that gives half the CUs a
    // workgroup, unless the slices end up at most 512 deep and the CUs are
    // still not all busy (then 128x128 tiles below do better).
    const bool thin = split > 1 && (int64_t)lt * split < kCUs && k / split <= 512;
    if (2 * (int64_t)m * n >= covered && (int64_t)lt * split >= kCUs / 2 && !thin) {
      // Large outputs, and outputs whose tiles leave CUs idle: split the
      // operands once and copy planes into LDS.
      if (split == 1 && (lt < kCUs || (ta && k >= 512) || (int64_t)cdiv(lt, kCUs) * 5504 >= (int64_t)m + n))
        return Choice{planar_cfg(mode), 1};
      return Choice{cfg, split};
    }
  } else if (lt >= 2 * kCUs) {
    const int rem = lt % kCUs;
    const bool even_out = mode == 0 && rem > 0 && rem <= kCUs / 2 && lt < 8 * kCUs && k >= 2048;
    // FP64's large tile runs two workgroups per CU: K split in two when the
    // last round of 2 * kCUs would be at most a quarter full (cache flushed,
    // beta 0.5: 576 tiles 0.86-0.97x the unsplit time, 640 tiles 0.91x; 648
    // and 672 tiles 1.04-1.07x, so they stay unsplit).
    const int rem64 = lt % (2 * kCUs);
    const bool even_out64 = mode == 1 && rem64 > 0 && rem64 <= kCUs / 2 && lt < 8 * kCUs && k >= 1024;
    return Choice{cfg, (even_out || even_out64) ? 2 : 1};
  } else if (mode == 0 && lt >= kCUs / 2 && 2 * lt <= kCUs + kCUs / 4 && k >= 2048) {
    return Choice{cfg, 2};  // the large FP32 tile, K split in two, about one round
  }
  const int dc = deep_cfg(mode, m, n);
  const int dtiles = tiles_of(mode, dc, m, n);
  if (k >= 16384 && dtiles < kCUs) {
    const int per_cu = (mode == 1 || dc == 2) ? 2 : 1;
    return Choice{dc, clamp_split(cdiv(per_cu * kCUs, dtiles), k, 512, kCUs)};
  }
  if (mode == 1 && tiles_of(mode, 3, m, n) >= kCUs) {
    // FP64 128x128 (copy after the MFMAs); K split in two unless the tiles
    // make whole rounds of CUs and K is short.
    const int t3 = tiles_of(mode, 3, m, n);
    return Choice{3, (t3 % kCUs == 0 && k < 4096) ? 1 : 2};
  }
  cfg = medium_cfg(mode);
  if (tiles_of(mode, cfg, m, n) < kCUs) {
    // BF16: 128x128 with K split (a power of two) toward two workgroups per
    // CU when that fills every CU with slices at least 512 deep; otherwise
    // 64x64.
    const int t0 = tiles_of(mode, cfg, m, n);
    int want = 1;
    while (want * 2 * t0 <= 2 * kCUs) want *= 2;
    const int s0 = clamp_split(want, k, 512);
    if (split_mode(mode) && (int64_t)t0 * s0 >= kCUs) return Choice{cfg, s0};
    cfg = 1;  // 64x64
  }
  const int tiles = tiles_of(mode, cfg, m, n);
  int split = 1;
  if (tiles < 2 * kCUs && k >= 1024) {
    split = clamp_split(cdiv(3 * kCUs, tiles), k, 256);
    if (mode == 1) {
      // FP64 fits two 128x64 or four 64x64 workgroups per CU, not three: take
      // the smallest split whose last round of workgroups is at least three
      // quarters full (cache flushed, beta 0.5: 1024x2048x2048 185 us at
      // split 2 against 239 at 3, 2048x1024x1024 108 against 139, 1792^3 287
      // at split 1 against 307 at 2; the other shapes tried keep their split).
      const int64_t slots = (int64_t)(cfg == 1 ? 4 : 2) * kCUs;
      const int smax = clamp_split(64, k, 256);
      for (int s = 1; s <= smax; ++s) {
        const int64_t w = (int64_t)tiles * s;
        const int64_t rounds = (w + slots - 1) / slots;
        if (4 * w >= 3 * rounds * slots) {
          split = s;
          break;
        }
      }
    }
  } else if (mode == 0 && k >= 2048) {
    const int rem = tiles % kCUs;  // FP32: even out a half-empty last round
    if (rem > 0 && rem <= kCUs / 2 && tiles < 8 * kCUs) split = 2;
  }
  return Choice{cfg, split};
}

size_t elem_size(int mode) { return mode == COSMA_MODE_FP64 ? 8 : 4; }

// Workspace of one call: split-K partials [split][M][N] at offset 0, then (for
// a pre-split configuration) room for the planes of the operands the caller
// did not split, 256-byte aligned.
size_t workspace_need(const cosma_gemm_desc* d, int cfg, int split, size_t* planes_off) {
  size_t bytes = split > 1 ? (size_t)split * (size_t)d->m * (size_t)d->n * elem_size(d->mode) : 0;
  *planes_off = (bytes + 255) / 256 * 256;
  if (is_planar(d->mode, cfg)) {
    const int rows_a = d->a_planes ? 0 : d->m, rows_b = d->b_planes ? 0 : d->n;
    if (rows_a + rows_b > 0) bytes = *planes_off + cosma::planar_bytes(rows_a, rows_b, d->k);
  }
  return bytes;
}

std::mutex g_mu;
struct DevState {
  double* stat_part = nullptr;  // kStatParts * 5 doubles + 5 doubles of output
};
std::vector<DevState> g_dev;
void* g_user_ws = nullptr;
size_t g_user_ws_bytes = 0;

// Library-owned workspace: one buffer per (device, stream), so calls on
// different streams never share partials or planes.  A buffer handed to a
// call while its stream was capturing is referenced by the captured graph,
// which may be replayed at any later time: it is never freed (growth retires
// it to g_ws_retired).
struct WsBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool captured = false;
};
std::map<std::pair<int, hipStream_t>, WsBuf> g_ws;
std::vector<void*> g_ws_retired;

// Serialises the residual-statistics calls: they share one partial buffer per
// device and return only after their stream has finished with it.
std::mutex g_stats_mu;

DevState* dev_state(int* status) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *status = (int)e;
    return nullptr;
  }
  if ((int)g_dev.size() <= dev) g_dev.resize(dev + 1);
  return &g_dev[dev];
}

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

// Resolve the split-K buffer: the descriptor's, the installed one, or the
// library-owned one of (current device, s), allocated or grown only when s
// is not being captured.
int resolve_workspace(const cosma_gemm_desc* d, size_t need, hipStream_t s, void** out) {
  if (d->workspace) {
    if (d->workspace_bytes < need) return COSMA_STATUS_INVALID_VALUE;
    *out = d->workspace;
    return 0;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_user_ws && g_user_ws_bytes >= need) {
    *out = g_user_ws;
    return 0;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  const bool cap = capturing(s);
  WsBuf& b = g_ws[{dev, s}];
  if (b.bytes < need) {
    if (cap) return COSMA_STATUS_WORKSPACE;
    if (b.p) {
      if (b.captured) {
        g_ws_retired.push_back(b.p);
      } else {
        // only work queued on s uses this buffer
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return (int)e;
        e = hipFree(b.p);
      }
      b = WsBuf{};
      if (e != hipSuccess) return (int)e;
    }
    e = hipMalloc(&b.p, need);
    if (e != hipSuccess) {
      b.p = nullptr;
      return (int)e;
    }
    b.bytes = need;
  }
  if (cap) b.captured = true;
  *out = b.p;
  return 0;
}

// Tuned choices installed by cosma_tune_set (or the Python autotuner's DB).
using TuneKey = std::tuple<int, int, int, int, int, int>;
struct TuneVal {
  int cfg, split, group_m;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

bool tuned(int mode, int m, int n, int k, int ta, int tb, TuneVal* out) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(TuneKey{mode, m, n, k, ta, tb});
  if (it == g_tune.end()) return false;
  *out = it->second;
  return true;
}

// Host-side contract check; returns 0, or a negative status.
int validate(const cosma_gemm_desc* d) {
  if (!d) return COSMA_STATUS_INVALID_VALUE;
  if (d->mode < 0 || d->mode >= kNumModes) return COSMA_STATUS_INVALID_VALUE;
  if (d->m < 0 || d->n < 0 || d->k < 0) return COSMA_STATUS_INVALID_VALUE;
  if ((d->trans_a != 0 && d->trans_a != 1) || (d->trans_b != 0 && d->trans_b != 1))
    return COSMA_STATUS_INVALID_VALUE;
  if (d->split_k > kMaxSplit) return COSMA_STATUS_INVALID_VALUE;
  if (d->tile_cfg >= kNumCfg[d->mode]) return COSMA_STATUS_INVALID_VALUE;
  if (d->epilogue < COSMA_EPILOGUE_NONE || d->epilogue > COSMA_EPILOGUE_SIGMOID)
    return COSMA_STATUS_INVALID_VALUE;
  if (d->a_planes || d->b_planes) {
    if (!split_mode(d->mode)) return COSMA_STATUS_INVALID_VALUE;
    if (d->tile_cfg >= 0 && !is_planar(d->mode, d->tile_cfg)) return COSMA_STATUS_INVALID_VALUE;
    if (((uintptr_t)d->a_planes % 16) || ((uintptr_t)d->b_planes % 16)) return COSMA_STATUS_MISALIGNED;
  }
  if (d->m == 0 || d->n == 0) return 0;
  if (!d->c) return COSMA_STATUS_INVALID_VALUE;
  if (d->ldc < (d->n > 1 ? d->n : 1)) return COSMA_STATUS_INVALID_VALUE;
  const bool uses_ab = d->k > 0 && d->alpha != 0.0;
  if (!uses_ab) return 0;
  if (d->trans_a && d->trans_b) return COSMA_STATUS_NOT_SUPPORTED;
  if (!d->a || !d->b) return COSMA_STATUS_INVALID_VALUE;
  const int64_t a_cols = d->trans_a ? d->m : d->k;
  const int64_t b_cols = d->trans_b ? d->k : d->n;
  if (d->lda < (a_cols > 1 ? a_cols : 1) || d->ldb < (b_cols > 1 ? b_cols : 1))
    return COSMA_STATUS_INVALID_VALUE;
  const uintptr_t es = elem_size(d->mode);
  if (((uintptr_t)d->a % es) || ((uintptr_t)d->b % es)) return COSMA_STATUS_MISALIGNED;
  return 0;
}

// Tile configuration, split factor and group size of a validated call with
// K > 0: the caller's, else the tuned table's, else the heuristic's.
void resolve_choice(const cosma_gemm_desc* d, int* cfg, int* split, int* group_m) {
  Choice ch = heuristic(d->mode, d->m, d->n, d->k, d->trans_a);
  *group_m = d->group_m;
  TuneVal tv;
  if (d->tile_cfg < 0 && d->split_k <= 0 &&
      tuned(d->mode, d->m, d->n, d->k, d->trans_a, d->trans_b, &tv)) {
    ch.cfg = tv.cfg;
    ch.split = tv.split;
    if (*group_m <= 0) *group_m = tv.group_m;
  }
  *cfg = d->tile_cfg >= 0 ? d->tile_cfg : ch.cfg;
  if ((d->a_planes || d->b_planes) && !is_planar(d->mode, *cfg)) *cfg = planar_cfg(d->mode);
  *split = d->split_k > 0 ? d->split_k : (d->tile_cfg >= 0 ? 1 : ch.split);
  if (*split > d->k) *split = d->k > 0 ? d->k : 1;
}

int run(const cosma_gemm_desc* d) {
  int st = validate(d);
  if (st != 0) return st;
  if (d->m == 0 || d->n == 0) return 0;
  hipStream_t s = static_cast<hipStream_t>(d->stream);
  const bool uses_ab = d->k > 0 && d->alpha != 0.0;
  if (!uses_ab) {  // C = act(beta*C) (reference: pre-scale only, cuCOSMAV100.cuh:2782-2797)
    if (d->beta == 1.0 && d->epilogue == COSMA_EPILOGUE_NONE) return 0;
    hipError_t e = d->mode == COSMA_MODE_FP64
                       ? cosma::run_scale_f64(static_cast<double*>(d->c), d->ldc, d->m, d->n, d->beta,
                                              d->epilogue, s)
                       : cosma::run_scale_f32(static_cast<float*>(d->c), d->ldc, d->m, d->n,
                                              (float)d->beta, d->epilogue, s);
    return (int)e;
  }
  int cfg, split, group_m;
  resolve_choice(d, &cfg, &split, &group_m);
  Problem p;
  p.mode = d->mode;
  p.M = d->m;
  p.N = d->n;
  p.K = d->k;
  p.ta = d->trans_a;
  p.tb = d->trans_b;
  p.A = d->a;
  p.lda = d->lda;
  p.B = d->b;
  p.ldb = d->ldb;
  p.C = d->c;
  p.ldc = d->ldc;
  p.alpha = d->alpha;
  p.beta = d->beta;
  p.cfg = cfg;
  p.split = split;
  p.group_m = group_m;
  p.act = d->epilogue;
  p.ws = nullptr;
  p.planes = nullptr;
  p.a_planes = d->a_planes;
  p.b_planes = d->b_planes;
  p.stream = s;
  size_t planes_off = 0;
  const size_t need = workspace_need(d, cfg, split, &planes_off);
  if (need > 0) {
    void* ws = nullptr;
    st = resolve_workspace(d, need, s, &ws);
    if (st != 0) return st;
    if (split > 1) p.ws = ws;
    if (is_planar(d->mode, cfg)) p.planes = static_cast<char*>(ws) + planes_off;
  }
  hipError_t e;
  switch (d->mode) {
    case COSMA_MODE_FP32: e = cosma::run_gemm_f32(p); break;
    case COSMA_MODE_FP64: e = cosma::run_gemm_f64(p); break;
    case COSMA_MODE_BF16X3: e = cosma::run_gemm_bf16x3(p); break;
    default: e = cosma::run_gemm_bf16x4(p); break;
  }
  return (int)e;
}

cosma_gemm_desc make_desc(int mode, int m, int n, int k, double alpha, const void* A, int lda,
                          const void* B, int ldb, double beta, void* C, int ldc, void* stream) {
  cosma_gemm_desc d;
  cosma_gemm_desc_init(&d);
  d.mode = mode;
  d.m = m;
  d.n = n;
  d.k = k;
  d.a = A;
  d.lda = lda;
  d.b = B;
  d.ldb = ldb;
  d.c = C;
  d.ldc = ldc;
  d.alpha = alpha;
  d.beta = beta;
  d.stream = stream;
  return d;
}

// Fixed-shape call: shape, tile configuration and split factor are
// compile-time constants of the instantiation.
template <int MODE, int M_, int N_, int K_>
int fixed_gemm(double alpha, const void* A, int lda, const void* B, int ldb, double beta, void* C,
               int ldc) {
  static_assert(M_ > 0 && N_ > 0 && K_ > 0, "fixed shapes are positive");
  constexpr Choice ch = heuristic(MODE, M_, N_, K_);
  static_assert(ch.cfg >= 0 && ch.split >= 1, "heuristic");
  cosma_gemm_desc d = make_desc(MODE, M_, N_, K_, alpha, A, lda, B, ldb, beta, C, ldc, nullptr);
  d.tile_cfg = ch.cfg;
  d.split_k = ch.split;
  return run(&d);
}

int stats_scratch(double** part, double** out) {
  std::lock_guard<std::mutex> lk(g_mu);
  int status = 0;
  DevState* ds = dev_state(&status);
  if (!ds) return status;
  if (!ds->stat_part) {
    hipError_t e = hipMalloc(&ds->stat_part, sizeof(double) * (cosma::kStatParts * 5 + 5));
    if (e != hipSuccess) {
      ds->stat_part = nullptr;
      return (int)e;
    }
  }
  *part = ds->stat_part;
  *out = ds->stat_part + cosma::kStatParts * 5;
  return 0;
}

int finish_stats(hipError_t e, const double* dev_out, double* out5, hipStream_t s) {
  if (e != hipSuccess) return (int)e;
  e = hipMemcpyAsync(out5, dev_out, 5 * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return (int)e;
  return (int)hipStreamSynchronize(s);
}

// Per-tile cost beyond its K-steps in the grouped schedule (prologue load
// and C store), in K-steps.
constexpr int kTileOverheadSteps = 2;

// Default tile of a grouped plan: one configuration for the whole batch, so
// the plan is a single persistent launch balanced by one schedule (separate
// buckets run as concurrent launches that compete for CUs without
// coordination).  The large tile when outputs with a side of at least 256
// hold half the batch's flops, else 128x128 when outputs with a side of at
// least 128 do, else 64x64.  On 512 problems with sides log-uniform in
// [32, 2048] the one large-tile launch ran at 76 / 44 / 118 / 110 TFLOP/s
// (FP32 / FP64 / BF16x3 / BF16x4) against 59 / 22 / 94 / 88 for per-size
// buckets of 64x64, 128x128 and large tiles, and within 2% of the best of
// every combination of those tiles over three size classes.
int group_default_cfg(int mode, int count, const cosma_group_problem* problems) {
  double all = 0.0, from256 = 0.0, from128 = 0.0;
  for (int i = 0; i < count; ++i) {
    const cosma_group_problem& q = problems[i];
    const double f = (double)q.m * q.n * (double)q.k;
    const int side = std::max(q.m, q.n);
    all += f;
    if (side >= 256) from256 += f;
    if (side >= 128) from128 += f;
  }
  // (BF16x4: the copy-between-MFMAs version of the 256x256 tile, 5% above
  // the ping-pong one in one launch over the 512 problems.)
  if (2.0 * from256 >= all) return mode == 3 ? 3 : large_cfg(mode);
  return 2.0 * from128 >= all ? 0 : 1;
}

}  // namespace

struct cosma_grouped_plan_s {
  int mode;
  struct Bucket {
    int cfg;
    int count;
    int tiles;
    int blocks;
    void* table;        // device GroupEntry array, then the schedule
    const int* sched;   // blocks + 1 offsets, then (problem, tile) pairs (inside table's allocation)
  };
  std::vector<Bucket> buckets;
  double flops = 0.0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// Tiles are numbered problem by problem, problems ordered by decreasing K
// (the cost of each of their tiles).  The schedule hands them out
// longest-processing-time first: each tile, in numbering order, goes to the
// workgroup with the least work so far (K-steps plus a fixed per-tile cost
// for the prologue and the C store), ties to the lowest logical workgroup.
// Logical workgroup L runs as workgroup (L % (B/8)) * 8 + L / (B/8) when B,
// the grid, is a multiple of 8: workgroups are dealt to the 8 XCDs
// round-robin by id, so logical neighbours -- which receive the neighbouring
// tiles of one problem in each round -- share one XCD's L2.  On 512 problems
// with sides in [32, 2048] (256x256 tiles) this cost model puts the largest
// workgroup load at 1.01x the mean, against 1.19x for a round-robin walk.
template <typename G>
int build_bucket(int mode, int cfg, const std::vector<const cosma_group_problem*>& probs_in,
                 int blocks_per_cu, cosma_grouped_plan_s::Bucket* out) {
  const TileDims td = kDims[mode][cfg];
  std::vector<const cosma_group_problem*> probs(probs_in);
  std::stable_sort(probs.begin(), probs.end(),
                   [](const cosma_group_problem* a, const cosma_group_problem* b) { return a->k > b->k; });
  std::vector<cosma::GroupEntry<G>> table;
  table.reserve(probs.size());
  std::vector<int64_t> cost;  // per problem: cost of each of its tiles
  int64_t tiles = 0;
  for (const cosma_group_problem* q : probs) {
    cosma::GroupEntry<G> e{};
    e.A = static_cast<const G*>(q->a);
    e.B = static_cast<const G*>(q->b);
    e.C = static_cast<G*>(q->c);
    e.lda = q->lda;
    e.ldb = q->ldb;
    e.ldc = q->ldc;
    e.M = q->m;
    e.N = q->n;
    e.K = q->k;
    e.tiles_n = cdiv(q->n, td.bn);
    e.tile_start = (int)tiles;
    e.alpha = q->alpha;
    e.beta = q->beta;
    tiles += (int64_t)cdiv(q->m, td.bm) * e.tiles_n;
    if (tiles > 0x3fffffff) return COSMA_STATUS_INVALID_VALUE;  // two schedule ints per tile
    table.push_back(e);
    cost.push_back(cdiv(q->k, td.bk) + kTileOverheadSteps);
  }
  out->cfg = cfg;
  out->count = (int)table.size();
  out->tiles = (int)tiles;
  const int occ = blocks_per_cu > 0 ? blocks_per_cu : cosma::grouped_occupancy(mode, cfg);
  int64_t blocks64 = (int64_t)occ * kCUs;
  if (blocks64 > tiles) blocks64 = tiles;
  const int B = (int)(blocks64 > 0 ? blocks64 : 1);
  out->blocks = B;

  // Longest-processing-time-first schedule (see above).
  std::vector<std::vector<std::pair<int, int>>> lists(B);  // (problem, tile within it)
  using Slot = std::pair<int64_t, int>;  // (load, logical workgroup)
  std::priority_queue<Slot, std::vector<Slot>, std::greater<Slot>> heap;
  for (int l = 0; l < B; ++l) heap.push(Slot{0, l});
  for (size_t i = 0; i < table.size(); ++i) {
    const int64_t first = table[i].tile_start;
    const int64_t last = i + 1 < table.size() ? table[i + 1].tile_start : tiles;
    for (int64_t t = first; t < last; ++t) {
      Slot s = heap.top();
      heap.pop();
      lists[s.second].push_back({(int)i, (int)(t - first)});
      heap.push(Slot{s.first + cost[i], s.second});
    }
  }
  const int per_xcd = B % 8 == 0 ? B / 8 : 0;
  std::vector<int> sched((size_t)B + 1 + 2 * (size_t)tiles);
  std::vector<int> of_phys(B);  // logical workgroup run by physical id
  for (int l = 0; l < B; ++l) of_phys[per_xcd ? (l % per_xcd) * 8 + l / per_xcd : l] = l;
  int pos = 0;
  for (int w = 0; w < B; ++w) {
    sched[w] = pos;
    for (const auto& pt : lists[of_phys[w]]) {
      sched[(size_t)B + 1 + 2 * (size_t)pos] = pt.first;
      sched[(size_t)B + 1 + 2 * (size_t)pos + 1] = pt.second;
      ++pos;
    }
  }
  sched[B] = pos;

  const size_t tbytes = (table.size() * sizeof(cosma::GroupEntry<G>) + 255) / 256 * 256;
  const size_t sbytes = sched.size() * sizeof(int);
  out->table = nullptr;
  out->sched = nullptr;
  hipError_t e = hipMalloc(&out->table, tbytes + sbytes);
  if (e != hipSuccess) {
    out->table = nullptr;
    return (int)e;
  }
  out->sched = reinterpret_cast<const int*>(static_cast<char*>(out->table) + tbytes);
  e = hipMemcpy(out->table, table.data(), table.size() * sizeof(cosma::GroupEntry<G>), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(const_cast<int*>(out->sched), sched.data(), sbytes, hipMemcpyHostToDevice);
  return (int)e;
}

void free_plan(cosma_grouped_plan_s* p) {
  if (!p) return;
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  for (auto& b : p->buckets)
    if (b.table) hipFree(b.table);
  delete p;
}

int launch_plan(cosma_grouped_plan_s* p, hipStream_t s) {
  for (auto& b : p->buckets) {
    if (b.tiles == 0) continue;
    hipError_t e = cosma::run_grouped(p->mode, b.cfg, b.table, b.count, b.sched, b.blocks, s);
    if (e != hipSuccess) return (int)e;
  }
  return 0;
}

}  // namespace

extern "C" {

int cosma_grouped_plan_create(int mode, int count, const cosma_group_problem* problems,
                              const int* tile_cfgs, int blocks_per_cu, cosma_grouped_plan* plan) {
  if (!plan) return COSMA_STATUS_INVALID_VALUE;
  *plan = nullptr;
  if (mode < 0 || mode >= kNumModes || count < 0 || (count > 0 && !problems))
    return COSMA_STATUS_INVALID_VALUE;
  std::vector<std::vector<const cosma_group_problem*>> by_cfg(kNumCfg[mode]);
  const int default_cfg = group_default_cfg(mode, count, problems);
  double flops = 0.0;
  for (int i = 0; i < count; ++i) {
    const cosma_group_problem& q = problems[i];
    cosma_gemm_desc d = make_desc(mode, q.m, q.n, q.k, q.alpha, q.a, (int)0, q.b, 0, q.beta, q.c,
                                  0, nullptr);
    d.lda = q.lda;
    d.ldb = q.ldb;
    d.ldc = q.ldc;
    int st = validate(&d);
    if (st != 0) return st;
    if (q.m == 0 || q.n == 0) continue;
    if (q.k > 0 && q.alpha != 0.0 && ((uintptr_t)q.c == 0)) return COSMA_STATUS_INVALID_VALUE;
    int cfg = (tile_cfgs && tile_cfgs[i] >= 0) ? tile_cfgs[i] : default_cfg;
    if (cfg >= kNumCfg[mode]) return COSMA_STATUS_INVALID_VALUE;
    if (is_planar(mode, cfg)) return COSMA_STATUS_NOT_SUPPORTED;
    by_cfg[cfg].push_back(&q);
    flops += 2.0 * q.m * q.n * (double)q.k;
  }
  cosma_grouped_plan_s* p = new cosma_grouped_plan_s();
  p->mode = mode;
  p->flops = flops;
  for (int cfg = 0; cfg < kNumCfg[mode]; ++cfg) {
    if (by_cfg[cfg].empty()) continue;
    cosma_grouped_plan_s::Bucket b{};
    int st = mode == COSMA_MODE_FP64 ? build_bucket<double>(mode, cfg, by_cfg[cfg], blocks_per_cu, &b)
                                     : build_bucket<float>(mode, cfg, by_cfg[cfg], blocks_per_cu, &b);
    p->buckets.push_back(b);
    if (st != 0) {
      free_plan(p);
      return st;
    }
  }
  *plan = p;
  return 0;
}

int cosma_grouped_plan_run(cosma_grouped_plan plan, void* stream) {
  if (!plan) return COSMA_STATUS_INVALID_VALUE;
  return launch_plan(plan, static_cast<hipStream_t>(stream));
}

int cosma_grouped_plan_run_graph(cosma_grouped_plan plan, void* stream) {
  if (!plan) return COSMA_STATUS_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!plan->exec) {
    // Captured as one graph whose bucket launches are independent branches
    // (fork/join through events), so a later bucket's workgroups start on CUs
    // that an earlier bucket's tail leaves idle.
    hipStream_t cs = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return (int)e;
    const int nb = (int)plan->buckets.size();
    std::vector<hipStream_t> side(nb > 1 ? nb - 1 : 0, nullptr);
    std::vector<hipEvent_t> ev(nb > 0 ? nb : 1, nullptr);
    int st = 0;
    for (auto& x : side)
      if (st == 0) st = (int)hipStreamCreateWithFlags(&x, hipStreamNonBlocking);
    for (auto& x : ev)
      if (st == 0) st = (int)hipEventCreateWithFlags(&x, hipEventDisableTiming);
    e = st == 0 ? hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) : hipSuccess;
    if (st == 0) st = (int)e;
    const bool capturing_ok = st == 0;
    if (st == 0 && nb > 1) st = (int)hipEventRecord(ev[0], cs);  // fork
    for (int b = 0; b < nb && st == 0; ++b) {
      hipStream_t bs = b == 0 ? cs : side[b - 1];
      if (b > 0) st = (int)hipStreamWaitEvent(bs, ev[0], 0);
      const auto& k = plan->buckets[b];
      if (st == 0 && k.tiles > 0)
        st = (int)cosma::run_grouped(plan->mode, k.cfg, k.table, k.count, k.sched, k.blocks, bs);
      if (st == 0 && b > 0) st = (int)hipEventRecord(ev[b], bs);
    }
    for (int b = 1; b < nb && st == 0; ++b) st = (int)hipStreamWaitEvent(cs, ev[b], 0);  // join
    hipGraph_t g = nullptr;
    hipError_t e2 = capturing_ok ? hipStreamEndCapture(cs, &g) : hipSuccess;
    for (auto& x : side)
      if (x) hipStreamDestroy(x);
    for (auto& x : ev)
      if (x) hipEventDestroy(x);
    hipStreamDestroy(cs);
    if (st != 0) {
      if (g) hipGraphDestroy(g);
      return st;
    }
    if (e2 != hipSuccess) return (int)e2;
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      hipGraphDestroy(g);
      return (int)e;
    }
    plan->graph = g;
    plan->exec = x;
  }
  return (int)hipGraphLaunch(plan->exec, s);
}

int cosma_grouped_plan_info(cosma_grouped_plan plan, int* buckets, int* total_tiles, double* flops) {
  if (!plan) return COSMA_STATUS_INVALID_VALUE;
  int tiles = 0;
  for (auto& b : plan->buckets) tiles += b.tiles;
  if (buckets) *buckets = (int)plan->buckets.size();
  if (total_tiles) *total_tiles = tiles;
  if (flops) *flops = plan->flops;
  return 0;
}

int cosma_grouped_plan_bucket(cosma_grouped_plan plan, int b, int* tile_cfg, int* count, int* tiles,
                              int* blocks) {
  if (!plan || b < 0 || b >= (int)plan->buckets.size()) return COSMA_STATUS_INVALID_VALUE;
  const auto& k = plan->buckets[b];
  if (tile_cfg) *tile_cfg = k.cfg;
  if (count) *count = k.count;
  if (tiles) *tiles = k.tiles;
  if (blocks) *blocks = k.blocks;
  return 0;
}

int cosma_grouped_plan_destroy(cosma_grouped_plan plan) {
  free_plan(plan);
  return 0;
}


void cosma_gemm_desc_init(cosma_gemm_desc* d) {
  if (!d) return;
  d->mode = COSMA_MODE_FP32;
  d->m = d->n = d->k = 0;
  d->trans_a = d->trans_b = 0;
  d->a = d->b = nullptr;
  d->c = nullptr;
  d->lda = d->ldb = d->ldc = 0;
  d->alpha = 1.0;
  d->beta = 0.0;
  d->split_k = 0;
  d->tile_cfg = -1;
  d->group_m = 0;
  d->workspace = nullptr;
  d->workspace_bytes = 0;
  d->stream = nullptr;
  d->epilogue = COSMA_EPILOGUE_NONE;
  d->a_planes = nullptr;
  d->b_planes = nullptr;
}

int cosma_gemm_ex(const cosma_gemm_desc* d) { return run(d); }

int cosma_workspace_query(const cosma_gemm_desc* d, size_t* bytes) {
  int st = validate(d);
  if (st != 0) return st;
  if (!bytes) return COSMA_STATUS_INVALID_VALUE;
  *bytes = 0;
  if (d->m == 0 || d->n == 0 || d->k == 0 || d->alpha == 0.0) return 0;
  int cfg, split, group_m;
  resolve_choice(d, &cfg, &split, &group_m);
  size_t planes_off = 0;
  *bytes = workspace_need(d, cfg, split, &planes_off);
  return 0;
}

size_t cosma_planes_bytes(int rows, int k) {
  if (rows <= 0 || k <= 0) return 0;
  return cosma::planar_bytes(rows, 0, k);
}

int cosma_split_operand(int mode, int kmajor, int rows, int k, const float* X, int64_t ld, void* planes,
                        void* stream) {
  if (!split_mode(mode) || rows < 0 || k < 0) return COSMA_STATUS_INVALID_VALUE;
  if (rows == 0 || k == 0) return 0;
  if (!X || !planes) return COSMA_STATUS_INVALID_VALUE;
  const int64_t cols = kmajor ? k : rows;
  if (ld < cols) return COSMA_STATUS_INVALID_VALUE;
  if (((uintptr_t)X % 4) || ((uintptr_t)planes % 16)) return COSMA_STATUS_MISALIGNED;
  return (int)cosma::run_split_operand(mode, kmajor != 0, rows, k, X, ld, planes,
                                       static_cast<hipStream_t>(stream));
}

int cosma_workspace_set(void* ptr, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_user_ws = bytes ? ptr : nullptr;
  g_user_ws_bytes = ptr ? bytes : 0;
  return 0;
}

int cosma_workspace_release(size_t* freed_bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (freed_bytes) *freed_bytes = 0;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  bool synced = false;
  for (auto it = g_ws.begin(); it != g_ws.end();) {
    if (it->first.first != dev || it->second.captured) {
      ++it;
      continue;
    }
    if (!synced) {
      e = hipDeviceSynchronize();
      if (e != hipSuccess) return (int)e;
      synced = true;
    }
    e = hipFree(it->second.p);
    if (e != hipSuccess) return (int)e;
    if (freed_bytes) *freed_bytes += it->second.bytes;
    it = g_ws.erase(it);
  }
  return 0;
}

int cosma_sgemm(float alpha, const float* A, int lda, const float* B, int ldb, float beta,
                float* C, int ldc) {
  return fixed_gemm<COSMA_MODE_FP32, COSMA_M, COSMA_N, COSMA_K>(alpha, A, lda, B, ldb, beta, C, ldc);
}
int cosma_dgemm(double alpha, const double* A, int lda, const double* B, int ldb, double beta,
                double* C, int ldc) {
  return fixed_gemm<COSMA_MODE_FP64, COSMA_M, COSMA_N, COSMA_K>(alpha, A, lda, B, ldb, beta, C, ldc);
}
void cosma_fixed_shape(int* m, int* n, int* k) {
  if (m) *m = COSMA_M;
  if (n) *n = COSMA_N;
  if (k) *k = COSMA_K;
}

int cosma_sgemm_1024x1024x1024(float alpha, const float* A, int lda, const float* B, int ldb,
                               float beta, float* C, int ldc) {
  return fixed_gemm<COSMA_MODE_FP32, 1024, 1024, 1024>(alpha, A, lda, B, ldb, beta, C, ldc);
}
int cosma_sgemm_8192x8192x8192(float alpha, const float* A, int lda, const float* B, int ldb,
                               float beta, float* C, int ldc) {
  return fixed_gemm<COSMA_MODE_FP32, 8192, 8192, 8192>(alpha, A, lda, B, ldb, beta, C, ldc);
}
int cosma_sgemm_256x256x1048576(float alpha, const float* A, int lda, const float* B, int ldb,
                                float beta, float* C, int ldc) {
  return fixed_gemm<COSMA_MODE_FP32, 256, 256, 1048576>(alpha, A, lda, B, ldb, beta, C, ldc);
}
int cosma_sgemm_bf16x3_8192x8192x8192(float alpha, const float* A, int lda, const float* B,
                                      int ldb, float beta, float* C, int ldc) {
  return fixed_gemm<COSMA_MODE_BF16X3, 8192, 8192, 8192>(alpha, A, lda, B, ldb, beta, C, ldc);
}
int cosma_sgemm_bf16x3_256x256x1048576(float alpha, const float* A, int lda, const float* B,
                                       int ldb, float beta, float* C, int ldc) {
  return fixed_gemm<COSMA_MODE_BF16X3, 256, 256, 1048576>(alpha, A, lda, B, ldb, beta, C, ldc);
}
int cosma_dgemm_32768x1024x1024(double alpha, const double* A, int lda, const double* B, int ldb,
                                double beta, double* C, int ldc) {
  return fixed_gemm<COSMA_MODE_FP64, 32768, 1024, 1024>(alpha, A, lda, B, ldb, beta, C, ldc);
}
int cosma_dgemm_1024x65536x1024(double alpha, const double* A, int lda, const double* B, int ldb,
                                double beta, double* C, int ldc) {
  return fixed_gemm<COSMA_MODE_FP64, 1024, 65536, 1024>(alpha, A, lda, B, ldb, beta, C, ldc);
}

int cosma_sgemm_mnk(int m, int n, int k, float alpha, const float* A, int lda, const float* B,
                    int ldb, float beta, float* C, int ldc, void* stream) {
  cosma_gemm_desc d = make_desc(COSMA_MODE_FP32, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream);
  return run(&d);
}
int cosma_dgemm_mnk(int m, int n, int k, double alpha, const double* A, int lda, const double* B,
                    int ldb, double beta, double* C, int ldc, void* stream) {
  cosma_gemm_desc d = make_desc(COSMA_MODE_FP64, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream);
  return run(&d);
}
int cosma_sgemm_bf16x3_mnk(int m, int n, int k, float alpha, const float* A, int lda,
                           const float* B, int ldb, float beta, float* C, int ldc, void* stream) {
  cosma_gemm_desc d =
      make_desc(COSMA_MODE_BF16X3, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream);
  return run(&d);
}

int cosma_num_tile_configs(int mode) {
  if (mode < 0 || mode >= kNumModes) return 0;
  return kNumCfg[mode];
}

int cosma_tile_config(int mode, int cfg, int* out8) {
  if (!out8 || mode < 0 || mode >= kNumModes || cfg < 0 || cfg >= kNumCfg[mode])
    return COSMA_STATUS_INVALID_VALUE;
  const cosma::TileInfo* t = cosma::tile_info(mode, cfg);
  if (!t || t->BM != kDims[mode][cfg].bm || t->BN != kDims[mode][cfg].bn ||
      t->BK != kDims[mode][cfg].bk)
    return COSMA_STATUS_INVALID_VALUE;
  out8[0] = t->BM;
  out8[1] = t->BN;
  out8[2] = t->BK;
  out8[3] = t->WM;
  out8[4] = t->WN;
  out8[5] = t->threads;
  out8[6] = t->lds_bytes;
  out8[7] = t->pipe;
  return 0;
}

int cosma_heuristic(int mode, int m, int n, int k, int* tile_cfg, int* split_k) {
  if (mode < 0 || mode >= kNumModes || m < 0 || n < 0 || k < 0 || !tile_cfg || !split_k)
    return COSMA_STATUS_INVALID_VALUE;
  Choice ch = heuristic(mode, m, n, k);
  *tile_cfg = ch.cfg;
  *split_k = ch.split;
  return 0;
}

int cosma_error_stats_f32(int rows, int cols, const float* C, int64_t ldc, const double* R,
                          int64_t ldr, double* out5, void* stream) {
  if (rows < 0 || cols < 0 || !out5 || ((rows && cols) && (!C || !R)) || ldc < cols || ldr < cols)
    return COSMA_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_stats_mu);
  double *part, *dout;
  int st = stats_scratch(&part, &dout);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return finish_stats(cosma::run_err_stats_f32(rows, cols, C, ldc, R, ldr, part, dout, s), dout, out5, s);
}
int cosma_error_stats_f64(int rows, int cols, const double* C, int64_t ldc, const double* R,
                          int64_t ldr, double* out5, void* stream) {
  if (rows < 0 || cols < 0 || !out5 || ((rows && cols) && (!C || !R)) || ldc < cols || ldr < cols)
    return COSMA_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_stats_mu);
  double *part, *dout;
  int st = stats_scratch(&part, &dout);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return finish_stats(cosma::run_err_stats_f64(rows, cols, C, ldc, R, ldr, part, dout, s), dout, out5, s);
}
int cosma_sumsq_f32(int rows, int cols, const float* X, int64_t ldx, double* out5, void* stream) {
  if (rows < 0 || cols < 0 || !out5 || ((rows && cols) && !X) || ldx < cols)
    return COSMA_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_stats_mu);
  double *part, *dout;
  int st = stats_scratch(&part, &dout);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return finish_stats(cosma::run_sumsq_f32(rows, cols, X, ldx, part, dout, s), dout, out5, s);
}
int cosma_sumsq_f64(int rows, int cols, const double* X, int64_t ldx, double* out5, void* stream) {
  if (rows < 0 || cols < 0 || !out5 || ((rows && cols) && !X) || ldx < cols)
    return COSMA_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_stats_mu);
  double *part, *dout;
  int st = stats_scratch(&part, &dout);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return finish_stats(cosma::run_sumsq_f64(rows, cols, X, ldx, part, dout, s), dout, out5, s);
}
int cosma_convert_f32_f64(int rows, int cols, const float* X, int64_t ldx, double* Y, int64_t ldy,
                          void* stream) {
  if (rows < 0 || cols < 0 || ((rows && cols) && (!X || !Y)) || ldx < cols || ldy < cols)
    return COSMA_STATUS_INVALID_VALUE;
  if (rows == 0 || cols == 0) return 0;
  return (int)cosma::run_convert_f32_f64(rows, cols, X, ldx, Y, ldy,
                                         static_cast<hipStream_t>(stream));
}

int cosma_tune_set(int mode, int m, int n, int k, int trans_a, int trans_b, int tile_cfg,
                   int split_k, int group_m) {
  if (mode < 0 || mode >= kNumModes || m < 0 || n < 0 || k < 0 || (trans_a != 0 && trans_a != 1) ||
      (trans_b != 0 && trans_b != 1) || tile_cfg < 0 || tile_cfg >= kNumCfg[mode] || split_k < 1 ||
      split_k > kMaxSplit || group_m < 0)
    return COSMA_STATUS_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[TuneKey{mode, m, n, k, trans_a, trans_b}] = TuneVal{tile_cfg, split_k, group_m};
  return 0;
}

int cosma_tune_lookup(int mode, int m, int n, int k, int trans_a, int trans_b, int* tile_cfg,
                      int* split_k, int* group_m) {
  if (!tile_cfg || !split_k || !group_m) return COSMA_STATUS_INVALID_VALUE;
  TuneVal tv;
  if (!tuned(mode, m, n, k, trans_a, trans_b, &tv)) return COSMA_STATUS_NOT_SUPPORTED;
  *tile_cfg = tv.cfg;
  *split_k = tv.split;
  *group_m = tv.group_m;
  return 0;
}

int cosma_tune_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  return 0;
}

int cosma_autotune(const cosma_gemm_desc* d, int warmup, int reps, cosma_tune_result* results,
                   int max_results, int* n_results, int* best) {
  int st = validate(d);
  if (st != 0) return st;
  if (!n_results || !best || (max_results > 0 && !results) || warmup < 0 || reps < 1)
    return COSMA_STATUS_INVALID_VALUE;
  *n_results = 0;
  *best = -1;
  if (d->m == 0 || d->n == 0 || d->k == 0 || d->alpha == 0.0) return 0;
  const double peak[kNumModes] = {157.3, 78.6, 2516.6 / 3.0, 2516.6 / 4.0};
  hipStream_t s = static_cast<hipStream_t>(d->stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    if (e0) hipEventDestroy(e0);
    return (int)he;
  }
  const double flops = 2.0 * d->m * (double)d->n * d->k;
  std::vector<cosma_tune_result> all;
  auto try_one = [&](int cfg, int split, int gm) {
    cosma_gemm_desc c = *d;
    c.tile_cfg = cfg;
    c.split_k = split;
    c.group_m = gm;
    cosma_tune_result r{};
    r.tile_cfg = cfg;
    r.split_k = split;
    r.group_m = gm;
    size_t ws = 0;
    cosma_workspace_query(&c, &ws);
    r.workspace_bytes = ws;
    const TileDims td = kDims[d->mode][cfg];
    r.blocks = tiles_of(d->mode, cfg, d->m, d->n) * split;
    const bool edge = (d->m % td.bm) || (d->n % td.bn) || (d->k % td.bk);
    r.occupancy = cosma::gemm_occupancy(d->mode, cfg, d->trans_a, d->trans_b, edge);
    const cosma::TileInfo* ti = cosma::tile_info(d->mode, cfg);
    r.waves_per_simd = ti ? r.occupancy * (ti->threads / 64) / 4 : 0;
    int status = 0;
    for (int i = 0; i < warmup && status == 0; ++i) status = run(&c);
    if (status == 0) status = (int)hipEventRecord(e0, s);
    for (int i = 0; i < reps && status == 0; ++i) status = run(&c);
    if (status == 0) status = (int)hipEventRecord(e1, s);
    if (status == 0) status = (int)hipEventSynchronize(e1);
    float ms = 0.f;
    if (status == 0) status = (int)hipEventElapsedTime(&ms, e0, e1);
    r.status = status;
    if (status == 0) {
      r.ms = ms / reps;
      r.tflops = r.ms > 0 ? flops / (r.ms * 1e-3) / 1e12 : 0.0;
      r.mfma_util = r.tflops / peak[d->mode];
    }
    all.push_back(r);
  };
  // Pass 1: tile configuration x split-K at the default group size.
  const int cfg_lo = d->tile_cfg >= 0 ? d->tile_cfg : 0;
  const int cfg_hi = d->tile_cfg >= 0 ? d->tile_cfg + 1 : kNumCfg[d->mode];
  const int gm0 = d->group_m > 0 ? d->group_m : 8;
  for (int cfg = cfg_lo; cfg < cfg_hi; ++cfg) {
    const TileDims td = kDims[d->mode][cfg];
    const int tiles = tiles_of(d->mode, cfg, d->m, d->n);
    for (int split = 1; split <= 1024; split *= 2) {
      if (d->split_k > 0 && split != d->split_k) continue;
      if (split > 1 && (split * 2 * td.bk > d->k || (int64_t)tiles * split > 16 * kCUs)) continue;
      try_one(cfg, split, gm0);
    }
    if (d->split_k > 1024) try_one(cfg, d->split_k, gm0);
  }
  auto best_of = [&]() {
    int b = -1;
    for (int i = 0; i < (int)all.size(); ++i)
      if (all[i].status == 0 && (b < 0 || all[i].ms < all[b].ms)) b = i;
    return b;
  };
  // Pass 2: XCD group size for the best (cfg, split) when the grid has
  // enough tile rows for grouping to matter.
  int b = best_of();
  if (b >= 0 && d->group_m <= 0) {
    const int cfg = all[b].tile_cfg, split = all[b].split_k;
    if (cdiv(d->m, kDims[d->mode][cfg].bm) > 8)
      for (int gm : {2, 4, 16, 32}) try_one(cfg, split, gm);
  }
  he = hipEventDestroy(e0);
  hipEventDestroy(e1);
  *best = best_of();
  const int n = (int)all.size() < max_results ? (int)all.size() : max_results;
  for (int i = 0; i < n; ++i) results[i] = all[i];
  *n_results = n;
  if (*best >= n) *best = -1;
  return 0;
}

int cosma_benchmark(const cosma_gemm_desc* d, int warmup, int runs, size_t flush_bytes,
                    float* ms_out) {
  int st = validate(d);
  if (st != 0) return st;
  if (warmup < 0 || runs < 0 || (runs > 0 && !ms_out)) return COSMA_STATUS_INVALID_VALUE;
  hipStream_t s = static_cast<hipStream_t>(d->stream);
  void* flush = nullptr;
  if (flush_bytes > 0) {
    static std::mutex fmu;
    static void* fbuf = nullptr;
    static size_t fbytes = 0;
    std::lock_guard<std::mutex> lk(fmu);
    if (fbytes < flush_bytes) {
      if (capturing(s)) return COSMA_STATUS_WORKSPACE;
      if (fbuf) hipFree(fbuf);
      fbuf = nullptr;
      fbytes = 0;
      hipError_t e = hipMalloc(&fbuf, flush_bytes);
      if (e != hipSuccess) return (int)e;
      fbytes = flush_bytes;
    }
    flush = fbuf;
  }
  for (int i = 0; i < warmup; ++i) {
    st = run(d);
    if (st) return st;
  }
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return (int)e;
  hipEvent_t e0, e1;
  if ((e = hipEventCreate(&e0)) != hipSuccess) return (int)e;
  if ((e = hipEventCreate(&e1)) != hipSuccess) {
    hipEventDestroy(e0);
    return (int)e;
  }
  for (int i = 0; i < runs && st == 0; ++i) {
    if (flush) st = (int)cosma::run_flush(flush, flush_bytes, s);
    if (st == 0) st = (int)hipEventRecord(e0, s);
    if (st == 0) st = run(d);
    if (st == 0) st = (int)hipEventRecord(e1, s);
    if (st == 0) st = (int)hipEventSynchronize(e1);
    if (st == 0) st = (int)hipEventElapsedTime(&ms_out[i], e0, e1);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return st;
}

const char* cosma_status_string(int status) {
  switch (status) {
    case COSMA_STATUS_SUCCESS: return "success";
    case COSMA_STATUS_INVALID_VALUE: return "invalid value (size, pointer, leading dimension, split or config)";
    case COSMA_STATUS_MISALIGNED: return "operand base not aligned to its element size";
    case COSMA_STATUS_NOT_SUPPORTED: return "combination not supported";
    case COSMA_STATUS_WORKSPACE: return "split-K workspace unavailable while the stream is capturing";
    default: return status > 0 ? hipGetErrorString((hipError_t)status) : "unknown status";
  }
}

const char* cosma_version(void) { return "0.1.0"; }

}  // extern "C"
