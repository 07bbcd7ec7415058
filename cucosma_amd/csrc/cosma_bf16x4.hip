// cosma_bf16x4.hip -- split-precision SGEMM with four products (bf16 hi/lo,
// hi*hi + hi*lo + lo*hi + lo*lo on v_mfma_f32_16x16x32_bf16, FP32
// accumulation): the tiles of cosma_bf16x3.hip; cfg 3 places the next-stage
// copy between the MFMAs (PIPE 1); cfg 6 (ping-pong) is the large and
// deep-K tile (its copy-after schedule reads fragments per sub-tile and keeps
// the per-k-step temporaries of one sub-tile live, so it fits without
// spilling); cfg 7 is the pre-split 256x256 tile (as in cosma_bf16x3.hip);
// cfg 8 is cfg 7 with 16 waves of 64x64 (four waves per SIMD; its per-k-step
// temporary is kept for one sub-tile at a time, which fits in 128
// registers); cfgs 9-11 are cfg 6's thin and small variants, as for BF16x3.
// Grouped tile families: {6, 9, 10, 11} in one launch; cfgs 0-5 each on
// their own.
#include "cosma_launch.hpp"

namespace cosma {

using B4C0 = Tile<kBF16x4, 128, 128, 32, 64, 64>;
using B4C1 = Tile<kBF16x4, 64, 64, 32, 32, 32>;
using B4C2 = Tile<kBF16x4, 256, 128, 32, 64, 64>;
using B4C3 = Tile<kBF16x4, 256, 256, 32, 128, 64, 1>;
using B4C4 = Tile<kBF16x4, 256, 128, 32, 128, 64>;
using B4C5 = Tile<kBF16x4, 128, 128, 64, 64, 64>;
using B4C6 = Tile<kBF16x4, 256, 256, 32, 128, 64, 4>;
using B4C7 = Planar<kBF16x4, 256, 256, 128, 64>;
using B4C8 = Planar<kBF16x4, 256, 256, 64, 64>;
using B4C9 = Tile<kBF16x4, 256, 128, 32, 64, 64, 4>;
using B4C10 = Tile<kBF16x4, 128, 256, 32, 64, 64, 4>;
using B4C11 = Tile<kBF16x4, 128, 128, 32, 64, 32, 4>;
using B4C12 = Tile<kBF16x4, 256, 256, 32, 128, 64, 7>;

using B4Tiles = TileTable<B4C0, B4C1, B4C2, B4C3, B4C4, B4C5, B4C6, B4C7, B4C8, B4C9, B4C10, B4C11, B4C12>;
using B4Fams = FamilyTable<Family<kBF16x4, B4Tiles, 6, 9, 10, 11>, Family<kBF16x4, B4Tiles, 0>,
                           Family<kBF16x4, B4Tiles, 1>, Family<kBF16x4, B4Tiles, 2>, Family<kBF16x4, B4Tiles, 3>,
                           Family<kBF16x4, B4Tiles, 4>, Family<kBF16x4, B4Tiles, 5>,
                           Family<kBF16x4, B4Tiles, 12>>;

// Host-side dispatch of this mode (a function-local object: a namespace-scope
// constant would also be emitted for the device).
const ModeOps* ops_bf16x4() {
  static const ModeOps ops = make_ops<B4Tiles, B4Fams>();
  return &ops;
}

}  // namespace cosma
