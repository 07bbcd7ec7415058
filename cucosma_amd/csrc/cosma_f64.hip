// cosma_f64.hip -- FP64 GEMM on v_mfma_f64_16x16x4_f64: tile configurations.
// Tile<mode, BM, BN, BK, WM, WN, PIPE, OCC> as in cosma_f32.hip.
// cfg 0: 128x128, 4 waves of 64x64
// cfg 1: 64x64, 4 waves of 32x32
// cfg 2: 128x64, 4 waves of 64x32
// cfg 3: 128x128, 8 waves of 64x32, registers capped for 4 waves per SIMD
//        (two workgroups per CU)
// cfg 4: 64x64 with a 32-deep K chunk
// cfg 5: 128x128 with a 32-deep K chunk, 8 waves of 64x32
// cfg 6: 256x128, 8 waves of 64x64
// cfg 7: cfg 3 with ping-pong copies (PIPE 4): the large-shape tile
// cfg 8: 128x64, 8 waves of 32x32 (PIPE 4, 4 waves per SIMD): cfg 7's thin-N variant
// cfg 9: 64x128, 8 waves of 32x32 (PIPE 4): the thin-M variant
// cfg 10: 64x64, 8 waves of 32x16 (PIPE 4): the small variant
// Grouped tile families: {7, 8, 9, 10} in one launch; every other
// configuration on its own.
#include "cosma_launch.hpp"

namespace cosma {

using F64C0 = Tile<kF64, 128, 128, 16, 64, 64>;
using F64C1 = Tile<kF64, 64, 64, 16, 32, 32>;
using F64C2 = Tile<kF64, 128, 64, 16, 64, 32>;
using F64C3 = Tile<kF64, 128, 128, 16, 64, 32, 0, 4>;
using F64C4 = Tile<kF64, 64, 64, 32, 32, 32>;
using F64C5 = Tile<kF64, 128, 128, 32, 64, 32>;
using F64C6 = Tile<kF64, 256, 128, 16, 64, 64>;
using F64C7 = Tile<kF64, 128, 128, 16, 64, 32, 4, 4>;
using F64C8 = Tile<kF64, 128, 64, 16, 32, 32, 4, 4>;
using F64C9 = Tile<kF64, 64, 128, 16, 32, 32, 4, 4>;
using F64C10 = Tile<kF64, 64, 64, 16, 32, 16, 4, 4>;

using F64Tiles = TileTable<F64C0, F64C1, F64C2, F64C3, F64C4, F64C5, F64C6, F64C7, F64C8, F64C9, F64C10>;
using F64Fams = FamilyTable<Family<kF64, F64Tiles, 7, 8, 9, 10>, Family<kF64, F64Tiles, 0>, Family<kF64, F64Tiles, 1>,
                            Family<kF64, F64Tiles, 2>, Family<kF64, F64Tiles, 3>, Family<kF64, F64Tiles, 4>,
                            Family<kF64, F64Tiles, 5>, Family<kF64, F64Tiles, 6>>;

// Host-side dispatch of this mode (a function-local object: a namespace-scope
// constant would also be emitted for the device).
const ModeOps* ops_f64() {
  static const ModeOps ops = make_ops<F64Tiles, F64Fams>();
  return &ops;
}

}  // namespace cosma
