// ttg_plan.h -- contraction plan: device-built block list + grouped-GEMM work.
#pragma once
#include "ttg_internal.h"

namespace ttg {

// One matched block pair C_c += A_a * B_b (qtensor.cpp:719-721).  Operands
// keep their stored column-major layout; each of the M, N and K ranges is
// an (at most) two-component mixed-radix index with one stride per
// component:  x -> (x % e0) * s0 + (x / e0) * s1.
struct GemmTask {
  int64_t a_off, b_off;
  int64_t sam0, sam1, sak0, sak1;
  int64_t sbk0, sbk1, sbn0, sbn1;
  int em0, ek0, en0, K;
  int kt0;            // first k-tile of this task in its block's k walk
  int nk1;            // K / ek0: steps of the second k component
};

// Output block c: its K-contributions are tasks [t0, t1) (nkt k-tiles in
// total), written once.
// most pieces one output tile's k walk is split into (split-K)
constexpr int kMaxSplit = 16;

struct CBlock {
  int64_t c_off;
  int M, N;
  int t0, t1;
  int nkt, pad;
};

// One workgroup work item: rows [m0, m0+TM) x cols [n0, n0+TN) of block c
// over the k-tiles [kt0, kt1) of its walk.  slot < 0: the item owns the
// whole walk and writes C; slot >= 0: it writes a TM x TN partial that a
// reduction pass adds up (split-K, fixed order).
struct Tile {
  int c, m0, n0, kt0, kt1, slot;
};
// reduction group of a split tile: partial slots [slot0, slot0 + ns)
struct SplitGroup {
  int c, m0, n0, slot0, ns, pad;
};

inline int tk_of(int prec) { return prec == TTG_BF16X3 ? 32 : 16; }

struct PairSpec {
  int np, nfa, nfb;
  int pa[MAXR], pb[MAXR], fa[MAXR], fb[MAXR];
};

// tile shapes the planner chooses between (MFMA 16x16 building block); the
// x16 shapes serve contractions whose output blocks are at most 16 wide.
// TILE_SKINNY is the vector-ALU kernel for contractions whose every task has
// N <= kSkinnyN and K <= kSkinnyK (one output row per thread, the K x N
// operand block in LDS): there an MFMA tile would be mostly zero padding.
// The 96-wide shapes fit sector dims up to 96 with one tile per dimension.
enum TileCfg { TILE_64x64 = 0, TILE_32x32 = 1, TILE_64x32 = 2, TILE_32x64 = 3, TILE_128x64 = 4,
               TILE_256x16 = 5, TILE_128x16 = 6, TILE_64x16 = 7, TILE_96x96 = 8, TILE_96x32 = 9,
               TILE_32x96 = 10, TILE_NCFG = 11,
               TILE_SKINNY = 11 };
constexpr int kSkinnyN = 16, kSkinnyK = 32, kSkinnyRows = 256;
__host__ __device__ inline int cfg_tm(int c) {
  const int t[] = {64, 32, 64, 32, 128, 256, 128, 64, 96, 96, 32, kSkinnyRows};
  return t[c];
}
__host__ __device__ inline int cfg_tn(int c) {
  const int t[] = {64, 32, 32, 64, 64, 16, 16, 16, 96, 32, 96, kSkinnyN};
  return t[c];
}

// device header filled by the planning kernels
struct PlanHdr {
  int64_t nuL, nuR, nuM, nQ, nblocksC;
  int64_t totalC, ntasks;
  double flops;
  int64_t tiles[TILE_NCFG];
  double padded[TILE_NCFG];      // sum over tiles of tile area x padded K
  int64_t kmax;                  // largest padded K walk of one output block
  int64_t votes[4];              // tasks with unit stride: A m, A k, B k, B n
  int64_t nmax, ktmax;           // widest output block with tasks; largest task K
  int64_t opmax;                 // largest operand block (elements) of any task
  int64_t err;
  int64_t ncand;                 // (result block, mid) candidates: an upper bound of ntasks
};

}  // namespace ttg

struct ttg_plan_s {
  ttg::MetaPtr A, B, C;          // structures the plan was built on / produces
  ttg::PairSpec ps;
  int prec = TTG_FP64;
  int nblocksC = 0;
  int64_t ntasks = 0;
  double flops = 0;
  int cfg = ttg::TILE_64x64;
  int ntiles = 0;
  bool a_mcontig = true, b_kcontig = true;
  // split-BF16 operand format: packed_in, the MFMA stage reads both operands
  // as pre-split words (bf16 hi << 16 | bf16 lo, see split_word); split_out,
  // the vector-ALU stage writes its result as such words
  bool packed_in = false, split_out = false;
  ttg::GemmTask* d_tasks = nullptr;
  ttg::CBlock* d_cblk = nullptr;
  ttg::Tile* d_tiles = nullptr;
  ttg::SplitGroup* d_groups = nullptr;   // split-K reduction groups (inside d_tiles' allocation)
  int ngroups = 0;
  void* d_part = nullptr;                // split-K partial tiles (accumulator type)
  int64_t nslots = 0;
  // 6-tuples (c, a, b, M, N, K) for the bookkeeping checks
  int* d_tuples = nullptr;
  // host copy of d_cblk as planned (the fields a work list needs: M, N and
  // k-tile counts; c_off is the planned one, before any remap on the device),
  // so a work list can be rebuilt without a read-back
  std::vector<ttg::CBlock> hcb;
  // the work list is not built yet (plan_build with defer_tiles: the caller
  // chooses the tile shape first, tune_stage)
  bool tiles_pending = false;
  ~ttg_plan_s();
};

namespace ttg {
// defer_tiles: leave the work list unbuilt (tiles_pending) for a caller that
// picks the tile shape itself before the first launch (tune_stage)
int plan_build(const MetaPtr& A, const MetaPtr& B, int np, const int* pa, const int* pb, int prec,
               int force_cfg, ttg_plan* out, bool defer_tiles = false);
// launch the grouped GEMM of a plan.  Operand element types follow prec:
// FP64 reads double A/B; FP32 and BF16X3 read float A/B, or split-BF16
// words (unsigned) for a BF16X3 plan with packed_in.  out_f64 selects a
// double or float C; a BF16X3 vector-ALU plan with split_out writes words.
int gemm_launch(const ttg_plan_s& p, const void* A, const void* B, void* C, bool out_f64,
                hipStream_t s);
// split-BF16 words of n FP64 values: hi = bf16 of the value rounded to FP32,
// lo = bf16 of the FP32 remainder, word = hi << 16 | lo (the operands the
// split-BF16 kernel forms from FP32 values when it stages them)
void split_words(const double* src, unsigned* dst, size_t n, hipStream_t s);
// workgroups per CU the GEMM kernel of (precision, tile shape, operand
// layout) can keep resident
int gemm_occupancy(int prec, int cfg, bool a_mcontig, bool b_kcontig);
// work list of a plan from its output blocks (host copy): tiles of the
// chosen shape, split along K when the tile count alone cannot fill the GPU
int build_tiles(ttg_plan_s& p, const std::vector<CBlock>& cb);
// the same plan's work list rebuilt for another MFMA tile shape
int rebuild_tiles(ttg_plan_s& p, int cfg);
// a plan for precision prec on src's output blocks and task list (copied on
// the device, no QN bookkeeping), with src's tile shape: the k-tile offsets
// are recomputed when prec's k-tile depth differs
int plan_clone(const ttg_plan_s& src, int prec, ttg_plan* out);
// measured tile-shape choice of one stage of a multi-stage operator (the
// operands are the stage's own; C is overwritten); `stage` keys the cache of
// choices together with the stage's precision, size and operand layout
int tune_stage(ttg_plan p, int stage, const void* A, const void* B, void* C, bool out_f64,
               hipStream_t s);
}  // namespace ttg
