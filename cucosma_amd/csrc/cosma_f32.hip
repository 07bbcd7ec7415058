// cosma_f32.hip -- FP32 GEMM on v_mfma_f32_16x16x4_f32: tile configurations.
// Tile<mode, BM, BN, BK, WM, WN, PIPE, OCC>: workgroup tile BM x BN with a K
// chunk of BK, waves of WM x WN, next-stage copy placement PIPE (see
// compute_stage), minimum waves per SIMD OCC.
// cfg 0: 128x128, 4 waves of 64x64
// cfg 1: 64x64, 4 waves of 32x32 (small shapes, split-K slices)
// cfg 2: 128x64, 4 waves of 64x32 (deep-K split)
// cfg 3: 256x128, 8 waves of 64x64
// cfg 4: 128x128 with a 64-deep K chunk
// cfg 5: cfg 3 with register double-buffered prefetch (PIPE 3) (135 TFLOP/s
//        at 8192^3, 145 with B stored transposed)
// cfg 6: cfg 3 with ping-pong copies (PIPE 4) (deep-K split)
// cfg 7: cfg 5 with 16 waves of 64x32 (128 VGPRs, four waves per SIMD): the
//        large-shape tile (139 TFLOP/s at 8192^3, 3-5% above cfg 5 from 4096^3 up)
// cfg 8: 128x256, 16 waves of 32x64 (PIPE 3): cfg 7's thin-M variant
// cfg 9: 128x128, 16 waves of 32x32 (PIPE 3): the small variant
// Grouped tile families: {7, 9} in one launch (the 128x256 variant took no
// problem of the 512-problem batch, and without it the family kernel holds
// 88 instead of 148 bytes of spills); every other configuration on its own.
#include "cosma_launch.hpp"

namespace cosma {

using F32C0 = Tile<kF32, 128, 128, 32, 64, 64, 1>;
using F32C1 = Tile<kF32, 64, 64, 32, 32, 32, 1>;
using F32C2 = Tile<kF32, 128, 64, 32, 64, 32, 1>;
using F32C3 = Tile<kF32, 256, 128, 32, 64, 64, 1>;
using F32C4 = Tile<kF32, 128, 128, 64, 64, 64, 1>;
using F32C5 = Tile<kF32, 256, 128, 32, 64, 64, 3>;
using F32C6 = Tile<kF32, 256, 128, 32, 64, 64, 4>;
using F32C7 = Tile<kF32, 256, 128, 32, 64, 32, 3>;
using F32C8 = Tile<kF32, 128, 256, 32, 32, 64, 3>;
using F32C9 = Tile<kF32, 128, 128, 32, 32, 32, 3>;

using F32Tiles = TileTable<F32C0, F32C1, F32C2, F32C3, F32C4, F32C5, F32C6, F32C7, F32C8, F32C9>;
using F32Fams = FamilyTable<Family<kF32, F32Tiles, 7, 9>, Family<kF32, F32Tiles, 0>, Family<kF32, F32Tiles, 1>,
                            Family<kF32, F32Tiles, 2>, Family<kF32, F32Tiles, 3>, Family<kF32, F32Tiles, 4>,
                            Family<kF32, F32Tiles, 5>, Family<kF32, F32Tiles, 6>, Family<kF32, F32Tiles, 8>>;

// Host-side dispatch of this mode (a function-local object: a namespace-scope
// constant would also be emitted for the device).
const ModeOps* ops_f32() {
  static const ModeOps ops = make_ops<F32Tiles, F32Fams>();
  return &ops;
}

}  // namespace cosma
