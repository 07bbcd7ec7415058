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

extern const TileInfo kF64Tiles[] = {F64C0::info(), F64C1::info(), F64C2::info(), F64C3::info(), F64C4::info(), F64C5::info(), F64C6::info(), F64C7::info()};
extern const int kF64NumTiles = 8;

hipError_t run_gemm_f64(const Problem& p) {
  switch (p.cfg) {
    case 0: return F64C0::run(p);
    case 1: return F64C1::run(p);
    case 2: return F64C2::run(p);
    case 3: return F64C3::run(p);
    case 4: return F64C4::run(p);
    case 5: return F64C5::run(p);
    case 6: return F64C6::run(p);
    case 7: return F64C7::run(p);
    default: return hipErrorInvalidValue;
  }
}

hipError_t run_grouped_f64(int cfg, const void* table, int count, const int* sched, int blocks,
                         hipStream_t s) {
  switch (cfg) {
    case 0: return F64C0::grouped(table, count, sched, blocks, s);
    case 1: return F64C1::grouped(table, count, sched, blocks, s);
    case 2: return F64C2::grouped(table, count, sched, blocks, s);
    case 3: return F64C3::grouped(table, count, sched, blocks, s);
    case 4: return F64C4::grouped(table, count, sched, blocks, s);
    case 5: return F64C5::grouped(table, count, sched, blocks, s);
    case 6: return F64C6::grouped(table, count, sched, blocks, s);
    case 7: return F64C7::grouped(table, count, sched, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

int grouped_occupancy_f64(int cfg) {
  switch (cfg) {
    case 0: return F64C0::grouped_occupancy();
    case 1: return F64C1::grouped_occupancy();
    case 2: return F64C2::grouped_occupancy();
    case 3: return F64C3::grouped_occupancy();
    case 4: return F64C4::grouped_occupancy();
    case 5: return F64C5::grouped_occupancy();
    case 6: return F64C6::grouped_occupancy();
    case 7: return F64C7::grouped_occupancy();
    default: return 1;
  }
}

int gemm_occupancy_f64(int cfg, int ta, int tb, bool edge) {
  switch (cfg) {
    case 0: return F64C0::occupancy(ta, tb, edge);
    case 1: return F64C1::occupancy(ta, tb, edge);
    case 2: return F64C2::occupancy(ta, tb, edge);
    case 3: return F64C3::occupancy(ta, tb, edge);
    case 4: return F64C4::occupancy(ta, tb, edge);
    case 5: return F64C5::occupancy(ta, tb, edge);
    case 6: return F64C6::occupancy(ta, tb, edge);
    case 7: return F64C7::occupancy(ta, tb, edge);
    default: return 0;
  }
}

}  // namespace cosma
