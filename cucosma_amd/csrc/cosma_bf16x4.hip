// cosma_bf16x4.hip -- split-precision SGEMM with four products (bf16 hi/lo,
// hi*hi + hi*lo + lo*hi + lo*lo on v_mfma_f32_16x16x32_bf16, FP32
// accumulation): the tiles of cosma_bf16x3.hip; cfg 3 places the next-stage
// copy between the MFMAs (PIPE 1) and is the grouped default; cfg 6
// (ping-pong) is the large and deep-K tile (its copy-after schedule reads
// fragments per sub-tile and keeps the per-k-step temporaries of one
// sub-tile live, so it fits without spilling); cfg 7 is the pre-split 256x256 tile (as in
// cosma_bf16x3.hip); cfg 8 is cfg 7 with 16 waves of 64x64 (four waves per
// SIMD; its per-k-step temporary is kept for one sub-tile at a time, which
// fits in 128 registers).
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

extern const TileInfo kBF16x4Tiles[] = {B4C0::info(), B4C1::info(), B4C2::info(), B4C3::info(), B4C4::info(), B4C5::info(), B4C6::info(), B4C7::info(), B4C8::info()};
extern const int kBF16x4NumTiles = 9;

hipError_t run_gemm_bf16x4(const Problem& p) {
  switch (p.cfg) {
    case 0: return B4C0::run(p);
    case 1: return B4C1::run(p);
    case 2: return B4C2::run(p);
    case 3: return B4C3::run(p);
    case 4: return B4C4::run(p);
    case 5: return B4C5::run(p);
    case 6: return B4C6::run(p);
    case 7: return B4C7::run(p);
    case 8: return B4C8::run(p);
    default: return hipErrorInvalidValue;
  }
}

hipError_t run_grouped_bf16x4(int cfg, const void* table, int count, const int* sched, int blocks,
                         hipStream_t s) {
  switch (cfg) {
    case 0: return B4C0::grouped(table, count, sched, blocks, s);
    case 1: return B4C1::grouped(table, count, sched, blocks, s);
    case 2: return B4C2::grouped(table, count, sched, blocks, s);
    case 3: return B4C3::grouped(table, count, sched, blocks, s);
    case 4: return B4C4::grouped(table, count, sched, blocks, s);
    case 5: return B4C5::grouped(table, count, sched, blocks, s);
    case 6: return B4C6::grouped(table, count, sched, blocks, s);
    case 7: return hipErrorNotSupported;  // pre-split: no grouped launch
    case 8: return hipErrorNotSupported;
    default: return hipErrorInvalidValue;
  }
}

int grouped_occupancy_bf16x4(int cfg) {
  switch (cfg) {
    case 0: return B4C0::grouped_occupancy();
    case 1: return B4C1::grouped_occupancy();
    case 2: return B4C2::grouped_occupancy();
    case 3: return B4C3::grouped_occupancy();
    case 4: return B4C4::grouped_occupancy();
    case 5: return B4C5::grouped_occupancy();
    case 6: return B4C6::grouped_occupancy();
    default: return 1;
  }
}

int gemm_occupancy_bf16x4(int cfg, int ta, int tb, bool edge) {
  switch (cfg) {
    case 0: return B4C0::occupancy(ta, tb, edge);
    case 1: return B4C1::occupancy(ta, tb, edge);
    case 2: return B4C2::occupancy(ta, tb, edge);
    case 3: return B4C3::occupancy(ta, tb, edge);
    case 4: return B4C4::occupancy(ta, tb, edge);
    case 5: return B4C5::occupancy(ta, tb, edge);
    case 6: return B4C6::occupancy(ta, tb, edge);
    case 7: return B4C7::occupancy(ta, tb, edge);
    case 8: return B4C8::occupancy(ta, tb, edge);
    default: return 0;
  }
}

}  // namespace cosma
