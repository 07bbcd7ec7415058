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

extern const TileInfo kF32Tiles[] = {F32C0::info(), F32C1::info(), F32C2::info(), F32C3::info(), F32C4::info(), F32C5::info(), F32C6::info(), F32C7::info()};
extern const int kF32NumTiles = 8;

hipError_t run_gemm_f32(const Problem& p) {
  switch (p.cfg) {
    case 0: return F32C0::run(p);
    case 1: return F32C1::run(p);
    case 2: return F32C2::run(p);
    case 3: return F32C3::run(p);
    case 4: return F32C4::run(p);
    case 5: return F32C5::run(p);
    case 6: return F32C6::run(p);
    case 7: return F32C7::run(p);
    default: return hipErrorInvalidValue;
  }
}

hipError_t run_grouped_f32(int cfg, const void* table, int count, const int* sched, int blocks,
                         hipStream_t s) {
  switch (cfg) {
    case 0: return F32C0::grouped(table, count, sched, blocks, s);
    case 1: return F32C1::grouped(table, count, sched, blocks, s);
    case 2: return F32C2::grouped(table, count, sched, blocks, s);
    case 3: return F32C3::grouped(table, count, sched, blocks, s);
    case 4: return F32C4::grouped(table, count, sched, blocks, s);
    case 5: return F32C5::grouped(table, count, sched, blocks, s);
    case 6: return F32C6::grouped(table, count, sched, blocks, s);
    case 7: return F32C7::grouped(table, count, sched, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

int grouped_occupancy_f32(int cfg) {
  switch (cfg) {
    case 0: return F32C0::grouped_occupancy();
    case 1: return F32C1::grouped_occupancy();
    case 2: return F32C2::grouped_occupancy();
    case 3: return F32C3::grouped_occupancy();
    case 4: return F32C4::grouped_occupancy();
    case 5: return F32C5::grouped_occupancy();
    case 6: return F32C6::grouped_occupancy();
    case 7: return F32C7::grouped_occupancy();
    default: return 1;
  }
}

int gemm_occupancy_f32(int cfg, int ta, int tb, bool edge) {
  switch (cfg) {
    case 0: return F32C0::occupancy(ta, tb, edge);
    case 1: return F32C1::occupancy(ta, tb, edge);
    case 2: return F32C2::occupancy(ta, tb, edge);
    case 3: return F32C3::occupancy(ta, tb, edge);
    case 4: return F32C4::occupancy(ta, tb, edge);
    case 5: return F32C5::occupancy(ta, tb, edge);
    case 6: return F32C6::occupancy(ta, tb, edge);
    case 7: return F32C7::occupancy(ta, tb, edge);
    default: return 0;
  }
}

}  // namespace cosma
