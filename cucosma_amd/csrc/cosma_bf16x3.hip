// cosma_bf16x3.hip -- split-precision SGEMM (bf16 hi/lo, three products on
// v_mfma_f32_16x16x32_bf16, FP32 accumulation): tile configurations.
// Tile<mode, BM, BN, BK, WM, WN, PIPE, OCC> as in cosma_f32.hip.
// cfg 0: 128x128, 4 waves of 64x64
// cfg 1: 64x64, 4 waves of 32x32
// cfg 2: 256x128, 8 waves of 64x64
// cfg 3: 256x256, 8 waves of 128x64 (160 KiB of LDS)
// cfg 4: 256x128, 4 waves of 128x64
// cfg 5: 128x128 with a 64-deep K chunk (deep-K split)
// cfg 6: cfg 3 with ping-pong copies (PIPE 4): the large-shape tile
// cfg 7: 256x256, 8 waves of 128x64 over pre-split operands (split pass, then
//        direct-to-LDS loads; Planar in cosma_launch.hpp); not for grouped plans
// cfg 8: cfg 7 with 16 waves of 64x64 (four waves per SIMD): the pre-split
//        configuration the heuristic uses (one box, 8192^3: planar kernels 504
//        and 507 TFLOP/s for cfg 8 and 7; 3072^3: 317 against 301 per call)
#include "cosma_launch.hpp"

namespace cosma {

using B3C0 = Tile<kBF16x3, 128, 128, 32, 64, 64>;
using B3C1 = Tile<kBF16x3, 64, 64, 32, 32, 32>;
using B3C2 = Tile<kBF16x3, 256, 128, 32, 64, 64>;
using B3C3 = Tile<kBF16x3, 256, 256, 32, 128, 64>;
using B3C4 = Tile<kBF16x3, 256, 128, 32, 128, 64>;
using B3C5 = Tile<kBF16x3, 128, 128, 64, 64, 64>;
using B3C6 = Tile<kBF16x3, 256, 256, 32, 128, 64, 4>;
using B3C7 = Planar<kBF16x3, 256, 256, 128, 64>;
using B3C8 = Planar<kBF16x3, 256, 256, 64, 64>;

extern const TileInfo kBF16x3Tiles[] = {B3C0::info(), B3C1::info(), B3C2::info(), B3C3::info(), B3C4::info(), B3C5::info(), B3C6::info(), B3C7::info(), B3C8::info()};
extern const int kBF16x3NumTiles = 9;

hipError_t run_gemm_bf16x3(const Problem& p) {
  switch (p.cfg) {
    case 0: return B3C0::run(p);
    case 1: return B3C1::run(p);
    case 2: return B3C2::run(p);
    case 3: return B3C3::run(p);
    case 4: return B3C4::run(p);
    case 5: return B3C5::run(p);
    case 6: return B3C6::run(p);
    case 7: return B3C7::run(p);
    case 8: return B3C8::run(p);
    default: return hipErrorInvalidValue;
  }
}

hipError_t run_grouped_bf16x3(int cfg, const void* table, int count, const int* sched, int blocks,
                         hipStream_t s) {
  switch (cfg) {
    case 0: return B3C0::grouped(table, count, sched, blocks, s);
    case 1: return B3C1::grouped(table, count, sched, blocks, s);
    case 2: return B3C2::grouped(table, count, sched, blocks, s);
    case 3: return B3C3::grouped(table, count, sched, blocks, s);
    case 4: return B3C4::grouped(table, count, sched, blocks, s);
    case 5: return B3C5::grouped(table, count, sched, blocks, s);
    case 6: return B3C6::grouped(table, count, sched, blocks, s);
    case 7: return hipErrorNotSupported;  // pre-split: no grouped launch
    case 8: return hipErrorNotSupported;
    default: return hipErrorInvalidValue;
  }
}

int grouped_occupancy_bf16x3(int cfg) {
  switch (cfg) {
    case 0: return B3C0::grouped_occupancy();
    case 1: return B3C1::grouped_occupancy();
    case 2: return B3C2::grouped_occupancy();
    case 3: return B3C3::grouped_occupancy();
    case 4: return B3C4::grouped_occupancy();
    case 5: return B3C5::grouped_occupancy();
    case 6: return B3C6::grouped_occupancy();
    default: return 1;
  }
}

int gemm_occupancy_bf16x3(int cfg, int ta, int tb, bool edge) {
  switch (cfg) {
    case 0: return B3C0::occupancy(ta, tb, edge);
    case 1: return B3C1::occupancy(ta, tb, edge);
    case 2: return B3C2::occupancy(ta, tb, edge);
    case 3: return B3C3::occupancy(ta, tb, edge);
    case 4: return B3C4::occupancy(ta, tb, edge);
    case 5: return B3C5::occupancy(ta, tb, edge);
    case 6: return B3C6::occupancy(ta, tb, edge);
    case 7: return B3C7::occupancy(ta, tb, edge);
    case 8: return B3C8::occupancy(ta, tb, edge);
    default: return 0;
  }
}

hipError_t run_split_operand(int mode, bool kmajor, int rows, int K, const float* X, int64_t ld, void* planes,
                             hipStream_t s) {
  (void)mode;  // the split is the same for three and four products
  const int Kp = (int)(((int64_t)K + 31) / 32 * 32);
  return B3C7::split(X, ld, kmajor, rows, K, Kp, static_cast<__bf16*>(planes), s);
}

}  // namespace cosma
