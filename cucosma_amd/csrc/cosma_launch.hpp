// cosma_launch.hpp -- instantiation and launch of gemm_kernel for one tile
// configuration.  Included by the per-mode translation units only.
#pragma once
#include <algorithm>
#include <tuple>

#include "cosma_internal.hpp"
#include "cosma_kernels.hpp"

namespace cosma {

template <int MODE, int BM, int BN, int BK, int WM, int WN, int PIPE = 0, int OCC = 1>
struct Tile {
  using Sh = TileShape<MODE, BM, BN, BK, WM, WN>;
  using V = TileV<BM, BN, BK, WM, WN, PIPE>;  // as a variant of a grouped tile family
  static constexpr int kOcc = OCC;
  static constexpr bool kGrouped = true;
  static constexpr TileInfo info() {
    return TileInfo{BM, BN, BK, WM, WN, Sh::NT,
                    (int)(2 * Sh::STAGE * sizeof(typename Traits<MODE>::S)), PIPE};
  }

  template <bool TA, bool TB, bool EDGE>
  static hipError_t go(const GemmParams<MODE>& gp, int nblk, hipStream_t s) {
    hipLaunchKernelGGL((gemm_kernel<MODE, BM, BN, BK, WM, WN, PIPE, OCC, TA, TB, EDGE>), dim3(nblk),
                       dim3(Sh::NT), 0, s, gp);
    return hipGetLastError();
  }

  // Workgroups of one gemm_kernel instantiation that fit on a CU.
  template <bool TA, bool TB, bool EDGE>
  static int occ() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &n, reinterpret_cast<const void*>(&gemm_kernel<MODE, BM, BN, BK, WM, WN, PIPE, OCC, TA, TB, EDGE>),
            Sh::NT, 0) != hipSuccess)
      return 0;
    return n;
  }
  static int occupancy(int ta, int tb, bool edge) {
    if (ta == 0 && tb == 0) return edge ? occ<false, false, true>() : occ<false, false, false>();
    if (ta == 0 && tb == 1) return edge ? occ<false, true, true>() : occ<false, true, false>();
    if (ta == 1 && tb == 0) return edge ? occ<true, false, true>() : occ<true, false, false>();
    return 0;
  }

  static hipError_t run(const Problem& pr) {
    using G = typename Traits<MODE>::G;
    using E = typename Traits<MODE>::E;
    GemmParams<MODE> gp;
    gp.A = static_cast<const G*>(pr.A);
    gp.B = static_cast<const G*>(pr.B);
    gp.C = static_cast<G*>(pr.C);
    gp.ws = static_cast<E*>(pr.ws);
    gp.lda = pr.lda;
    gp.ldb = pr.ldb;
    gp.ldc = pr.ldc;
    gp.M = pr.M;
    gp.N = pr.N;
    gp.K = pr.K;
    gp.tiles_m = (pr.M + BM - 1) / BM;
    gp.tiles_n = (pr.N + BN - 1) / BN;
    gp.split = pr.split;
    const int per = (pr.K + pr.split - 1) / pr.split;
    gp.kslice = ((per + BK - 1) / BK) * BK;
    gp.group_m = std::min(pr.group_m > 0 ? pr.group_m : 8, gp.tiles_m);
    gp.xcds = pr.xcds > 0 ? pr.xcds : 1;
    gp.act = pr.act;
    gp.alpha = (E)pr.alpha;
    gp.beta = (E)pr.beta;
    const int64_t nblk64 = (int64_t)gp.tiles_m * gp.tiles_n * pr.split;
    if (nblk64 <= 0 || nblk64 > 0x7fffffff) return hipErrorInvalidValue;
    const int nblk = (int)nblk64;
    constexpr int GVEC = Traits<MODE>::GVEC;
    gp.vec_a = ((uintptr_t)pr.A % 16 == 0) && (pr.lda % GVEC == 0);
    gp.vec_b = ((uintptr_t)pr.B % 16 == 0) && (pr.ldb % GVEC == 0);
    // Unaligned rows take the bounds-checked instantiation with element loads.
    const bool edge = (pr.M % BM) || (pr.N % BN) || (pr.K % BK) || !gp.vec_a || !gp.vec_b;
    hipError_t e;
    if (pr.ta == 0 && pr.tb == 0)
      e = edge ? go<false, false, true>(gp, nblk, pr.stream) : go<false, false, false>(gp, nblk, pr.stream);
    else if (pr.ta == 0 && pr.tb == 1)
      e = edge ? go<false, true, true>(gp, nblk, pr.stream) : go<false, true, false>(gp, nblk, pr.stream);
    else if (pr.ta == 1 && pr.tb == 0)
      e = edge ? go<true, false, true>(gp, nblk, pr.stream) : go<true, false, false>(gp, nblk, pr.stream);
    else
      return hipErrorNotSupported;
    if (e != hipSuccess || pr.split == 1) return e;
    if constexpr (MODE == kF64)
      return run_reduce_f64(static_cast<const double*>(pr.ws), static_cast<double*>(pr.C), pr.ldc,
                            pr.M, pr.N, pr.split, pr.alpha, pr.beta, pr.act, pr.stream);
    else
      return run_reduce_f32(static_cast<const float*>(pr.ws), static_cast<float*>(pr.C), pr.ldc,
                            pr.M, pr.N, pr.split, (float)pr.alpha, (float)pr.beta, pr.act, pr.stream);
  }
};

// Pre-split BF16 tile configuration: split pass into pr.planes, then
// planar_kernel (and the reduction when split > 1).
template <int MODE, int BM, int BN, int WM, int WN>
struct Planar {
  static constexpr bool kGrouped = false;  // no grouped launch over pre-split operands
  static constexpr int kThreads = 64 * (BM / WM) * (BN / WN);
  static constexpr TileInfo info() {
    return TileInfo{BM, BN, 32, WM, WN, kThreads, (int)(2 * (BM + BN) * 64 * sizeof(__bf16)), kPipePlanar};
  }
  template <bool EDGE>
  static int occ() {
    int n = 0;
    void (*fn)(PlanarParams<MODE>) = planar_kernel<MODE, BM, BN, WM, WN, EDGE>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(fn), kThreads, 0) != hipSuccess)
      return 0;
    return n;
  }
  static int occupancy(int, int, bool edge) { return edge ? occ<true>() : occ<false>(); }

  // The split of operand X into planes: kmajor when element (r, k) is
  // X[r*ld + k].  False when the grid would not fit one launch.
  static bool split_job(const float* X, int64_t ld, bool kmajor, int R, int K, int Kp, __bf16* out, SplitJob* j) {
    j->X = X;
    j->ld = ld;
    j->out = out;
    j->R = R;
    j->K = K;
    j->Kp = Kp;
    j->kmajor = kmajor ? 1 : 0;
    j->vec = ((uintptr_t)X % 16 == 0) && (ld % 4 == 0);
    j->nrb = (R + 255) / 256;
    const int64_t blocks = kmajor ? ((int64_t)R * (Kp / 8) + 255) / 256 : (int64_t)j->nrb * (Kp / 32);
    j->blocks = (int)blocks;
    return blocks <= 0x7fffffff;
  }

  static hipError_t split(const float* X, int64_t ld, bool kmajor, int R, int K, int Kp, __bf16* out,
                          hipStream_t s) {
    SplitJob j;
    if (!split_job(X, ld, kmajor, R, K, Kp, out, &j)) return hipErrorInvalidValue;
    if (j.blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(split_one<MODE>, dim3((unsigned)j.blocks), dim3(256), 0, s, j);
    return hipGetLastError();
  }

  static hipError_t run(const Problem& pr) {
    if (pr.K > 0x7fffffe0) return hipErrorInvalidValue;  // Kp, K rounded up to 32, must fit an int
    const int Kp = (int)(((int64_t)pr.K + 31) / 32 * 32);
    __bf16* room = static_cast<__bf16*>(pr.planes);
    const __bf16* Ap = static_cast<const __bf16*>(pr.a_planes);
    const __bf16* Bp = static_cast<const __bf16*>(pr.b_planes);
    hipError_t e = hipSuccess;
    if (!Ap && !Bp) {
      // Both operands: one launch.
      __bf16* pa = room;
      __bf16* pb = room + (int64_t)pr.M * 2 * Kp;
      SplitJob ja, jb;
      if (!split_job(static_cast<const float*>(pr.A), pr.lda, pr.ta == 0, pr.M, pr.K, Kp, pa, &ja) ||
          !split_job(static_cast<const float*>(pr.B), pr.ldb, pr.tb == 1, pr.N, pr.K, Kp, pb, &jb) ||
          (int64_t)ja.blocks + jb.blocks > 0x7fffffff)
        return hipErrorInvalidValue;
      if (ja.blocks + jb.blocks > 0) {
        hipLaunchKernelGGL(split_pair<MODE>, dim3((unsigned)(ja.blocks + jb.blocks)), dim3(256), 0, pr.stream, ja,
                           jb);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
      }
      Ap = pa;
      Bp = pb;
    } else if (!Ap) {
      e = split(static_cast<const float*>(pr.A), pr.lda, pr.ta == 0, pr.M, pr.K, Kp, room, pr.stream);
      if (e != hipSuccess) return e;
      Ap = room;
    } else if (!Bp) {
      e = split(static_cast<const float*>(pr.B), pr.ldb, pr.tb == 1, pr.N, pr.K, Kp, room, pr.stream);
      if (e != hipSuccess) return e;
      Bp = room;
    }
    PlanarParams<MODE> pp;
    GemmParams<MODE>& gp = pp.g;
    gp.A = static_cast<const float*>(pr.A);
    gp.B = static_cast<const float*>(pr.B);
    gp.C = static_cast<float*>(pr.C);
    gp.ws = static_cast<float*>(pr.ws);
    gp.lda = pr.lda;
    gp.ldb = pr.ldb;
    gp.ldc = pr.ldc;
    gp.M = pr.M;
    gp.N = pr.N;
    gp.K = pr.K;
    gp.tiles_m = (pr.M + BM - 1) / BM;
    gp.tiles_n = (pr.N + BN - 1) / BN;
    gp.split = pr.split;
    const int per = (pr.K + pr.split - 1) / pr.split;
    gp.kslice = ((per + 31) / 32) * 32;
    gp.group_m = std::min(pr.group_m > 0 ? pr.group_m : 8, gp.tiles_m);
    gp.xcds = pr.xcds > 0 ? pr.xcds : 1;
    gp.vec_a = 1;
    gp.vec_b = 1;
    gp.act = pr.act;
    gp.alpha = (float)pr.alpha;
    gp.beta = (float)pr.beta;
    pp.Ap = Ap;
    pp.Bp = Bp;
    pp.Kp = Kp;
    const int64_t nblk64 = (int64_t)gp.tiles_m * gp.tiles_n * pr.split;
    if (nblk64 <= 0 || nblk64 > 0x7fffffff) return hipErrorInvalidValue;
    // The unchecked instantiation addresses a lane's rows with a 32-bit byte
    // offset of up to 8 plane rows.
    const bool edge = (pr.M % BM) || (pr.N % BN) || (int64_t)Kp * 2 * 2 * 8 >= ((int64_t)1 << 32);
    if (edge)
      hipLaunchKernelGGL((planar_kernel<MODE, BM, BN, WM, WN, true>), dim3((unsigned)nblk64), dim3(kThreads), 0,
                         pr.stream, pp);
    else
      hipLaunchKernelGGL((planar_kernel<MODE, BM, BN, WM, WN, false>), dim3((unsigned)nblk64), dim3(kThreads), 0,
                         pr.stream, pp);
    e = hipGetLastError();
    if (e != hipSuccess || pr.split == 1) return e;
    return run_reduce_f32(static_cast<const float*>(pr.ws), static_cast<float*>(pr.C), pr.ldc, pr.M, pr.N,
                          pr.split, (float)pr.alpha, (float)pr.beta, pr.act, pr.stream);
  }
};

// A mode's tile table: configuration cfg is the cfg-th type.  Dispatch by
// index folds over the types.
template <class... Ts>
struct TileTable {
  static constexpr int kCount = sizeof...(Ts);
  template <int I>
  using at = std::tuple_element_t<I, std::tuple<Ts...>>;
  static const TileInfo* infos() {
    static const TileInfo t[] = {Ts::info()...};
    return t;
  }
  static hipError_t run(int cfg, const Problem& p) {
    hipError_t e = hipErrorInvalidValue;
    int i = 0;
    (void)((i++ == cfg && ((e = Ts::run(p)), true)) || ...);
    return e;
  }
  static int occupancy(int cfg, int ta, int tb, bool edge) {
    int r = 0, i = 0;
    (void)((i++ == cfg && ((r = Ts::occupancy(ta, tb, edge)), true)) || ...);
    return r;
  }
};

// Grouped tile family: the configurations Is of Table that one persistent
// grouped_kernel runs side by side (one per problem).  They share the
// workgroup size; the family takes the first one's register cap (OCC).
template <int MODE, class Table, int... Is>
struct Family {
  static_assert(sizeof...(Is) >= 1 && sizeof...(Is) <= 4, "one to four variants");
  using First = typename Table::template at<std::get<0>(std::make_tuple(Is...))>;
  static constexpr int kOcc = First::kOcc;
  static constexpr int kThreads = FamilyNT<typename Table::template at<Is>::V...>::value;
  static_assert((Table::template at<Is>::kGrouped && ...), "no pre-split configuration in a family");
  static constexpr GroupFamily info() { return GroupFamily{(int)sizeof...(Is), {Is...}}; }
  static auto kernel() { return &grouped_kernel<MODE, kOcc, typename Table::template at<Is>::V...>; }
  static hipError_t launch(const void* table, const int* sched, int blocks, hipStream_t s) {
    using G = typename Traits<MODE>::G;
    hipLaunchKernelGGL((grouped_kernel<MODE, kOcc, typename Table::template at<Is>::V...>), dim3(blocks),
                       dim3(kThreads), 0, s, static_cast<const GroupEntry<G>*>(table), sched);
    return hipGetLastError();
  }
  static int occupancy() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(kernel()), kThreads, 0) !=
        hipSuccess)
      return 1;
    return n > 0 ? n : 1;
  }
};

template <class... Fs>
struct FamilyTable {
  static constexpr int kCount = sizeof...(Fs);
  static const GroupFamily* infos() {
    static const GroupFamily t[] = {Fs::info()...};
    return t;
  }
  static hipError_t launch(int fam, const void* table, const int* sched, int blocks, hipStream_t s) {
    hipError_t e = hipErrorInvalidValue;
    int i = 0;
    (void)((i++ == fam && ((e = Fs::launch(table, sched, blocks, s)), true)) || ...);
    return e;
  }
  static int occupancy(int fam) {
    int r = 1, i = 0;
    (void)((i++ == fam && ((r = Fs::occupancy()), true)) || ...);
    return r;
  }
};

// The entry points one mode's translation unit exports to cosma_aux.hip.
struct ModeOps {
  int ncfg;
  const TileInfo* (*infos)();
  hipError_t (*run)(int, const Problem&);
  int (*occupancy)(int, int, int, bool);
  int nfam;
  const GroupFamily* (*families)();
  hipError_t (*grouped)(int, const void*, const int*, int, hipStream_t);
  int (*grouped_occupancy)(int);
};

template <class Tiles, class Fams>
ModeOps make_ops() {
  return ModeOps{Tiles::kCount, &Tiles::infos, &Tiles::run, &Tiles::occupancy,
                 Fams::kCount, &Fams::infos, &Fams::launch, &Fams::occupancy};
}

}  // namespace cosma
