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
// cfg 9: 256x128, 8 waves of 64x64, ping-pong: cfg 6's thin-N variant
// cfg 10: 128x256, 8 waves of 64x64, ping-pong: the thin-M variant
// cfg 11: 128x128, 8 waves of 64x32, ping-pong: the small variant
// Grouped tile families: {6, 9, 10, 11} in one launch; cfgs 0-5 each on
// their own.
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
using B3C9 = Tile<kBF16x3, 256, 128, 32, 64, 64, 4>;
using B3C10 = Tile<kBF16x3, 128, 256, 32, 64, 64, 4>;
using B3C11 = Tile<kBF16x3, 128, 128, 32, 64, 32, 4>;
using B3C12 = Tile<kBF16x3, 256, 256, 32, 128, 64, 7>;

using B3Tiles = TileTable<B3C0, B3C1, B3C2, B3C3, B3C4, B3C5, B3C6, B3C7, B3C8, B3C9, B3C10, B3C11, B3C12>;
using B3Fams = FamilyTable<Family<kBF16x3, B3Tiles, 6, 9, 10, 11>, Family<kBF16x3, B3Tiles, 0>,
                           Family<kBF16x3, B3Tiles, 1>, Family<kBF16x3, B3Tiles, 2>, Family<kBF16x3, B3Tiles, 3>,
                           Family<kBF16x3, B3Tiles, 4>, Family<kBF16x3, B3Tiles, 5>,
                           Family<kBF16x3, B3Tiles, 12>>;

// Host-side dispatch of this mode (a function-local object: a namespace-scope
// constant would also be emitted for the device).
const ModeOps* ops_bf16x3() {
  static const ModeOps ops = make_ops<B3Tiles, B3Fams>();
  return &ops;
}

hipError_t run_split_operand(int mode, bool kmajor, int rows, int K, const float* X, int64_t ld, void* planes,
                             hipStream_t s) {
  (void)mode;  // the split is the same for three and four products
  if (K > 0x7fffffe0) return hipErrorInvalidValue;  // Kp, K rounded up to 32, must fit an int
  const int Kp = (int)(((int64_t)K + 31) / 32 * 32);
  return B3C7::split(X, ld, kmajor, rows, K, Kp, static_cast<__bf16*>(planes), s);
}

}  // namespace cosma
