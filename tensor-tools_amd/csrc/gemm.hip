// gemm.hip -- persistent grouped block GEMM over a plan's tile list.
//
// Every uneven block pair of a contraction runs in ONE launch.  A work item
// is a TM x TN piece of one output block and a range of its k walk (all of
// the block's tasks, the K contributions in the reference's mid order).  An
// item owning the whole walk writes the tile once, so no atomics and no
// zero-fill pass are needed (blocks without tasks are written as zeros);
// long walks are split into pieces that write partial tiles, added in fixed
// order by k_split_reduce (deterministic split-K).  Operand tiles are staged
// through LDS; the next k-slice is prefetched into registers while the MFMA
// pipe works on the current one.
//
// K walk: a task's contraction index is k = k0 + e0*k1 (two stride
// components).  k-tiles never straddle a k1 boundary: for each k1 the
// k0 range is cut into chunks of TK (the last one partial, zero-filled), so
// an operand element's address is base + row_offset + (c + kl)*s0 + k1*s1
// with no division inside the loop; row offsets (the M or N mixed-radix
// decomposition) are computed once per task.
//
// MFMA building block: 16x16 output fragment per wave-instruction.
//   FP64:   v_mfma_f64_16x16x4_f64      (A,B one double per lane)
//   FP32:   v_mfma_f32_16x16x4_f32      (A,B one float per lane)
//   BF16X3: v_mfma_f32_16x16x32_bf16 x3 (lo*hi + hi*lo + hi*hi, FP32 acc)
// Operand lane layout (k-slice of 4):  lane l holds A[l%16][l/16] and
// B[l/16][l%16]; result register r of lane l holds D[row][l%16] with
// row = l/16 + 4r for f64 and 4*(l/16) + r for the f32-accumulating ops.
// The kernels pass the B fragment as the first operand and the A fragment as
// the second, so D is the transposed tile (B^T A^T): lane l holds output
// row l%16, and a column of C is stored by 16 consecutive lanes.
#include <algorithm>
#include <unordered_map>

#include "ttg_internal.h"
#include "ttg_plan.h"

namespace ttg {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <typename T>
struct MfmaOp;
template <>
struct MfmaOp<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct MfmaOp<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

__device__ __forceinline__ int64_t off2(int x, int e0, int64_t s0, int64_t s1) {
  return (int64_t)(x % e0) * s0 + (int64_t)(x / e0) * s1;
}
// the same in 32 bits: operand blocks are below 2^28 elements (checked by the
// planner), so element offsets and their byte offsets fit an unsigned
__device__ __forceinline__ unsigned off2u(int x, int e0, unsigned s0, unsigned s1) {
  return (unsigned)(x % e0) * s0 + (unsigned)(x / e0) * s1;
}

constexpr int kThreads = 256;

// k-walk state of one tile: task t, second k component k1, chunk start c
struct KWalk {
  int t, k1, c;
};
__device__ __forceinline__ void kwalk_next(KWalk& w, const GemmTask& g, int TK) {
  w.c += TK;
  if (w.c >= g.ek0) {
    w.c = 0;
    if (++w.k1 >= g.nk1) {
      w.k1 = 0;
      w.t++;
    }
  }
}
// position of k-tile kt of block cb's walk: the last task whose first k-tile
// is <= kt (tasks[t0..t1) hold ascending kt0), then (k1, chunk) inside it
__device__ __forceinline__ KWalk kwalk_start(const GemmTask* __restrict__ tasks, const CBlock& cb,
                                             int kt, int TK) {
  if (kt == 0) return KWalk{cb.t0, 0, 0};   // the walk's start (every item without split-K)
  int lo = cb.t0, hi = cb.t1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tasks[mid].kt0 <= kt) lo = mid;
    else hi = mid;
  }
  const int ek0 = tasks[lo].ek0, local = kt - tasks[lo].kt0;
  const int nch = (ek0 + TK - 1) / TK;
  return KWalk{lo, local / nch, (local % nch) * TK};
}

// Work items are taken with stride gridDim.x starting from a permuted block
// id: workgroups are dispatched round-robin over the 8 XCDs, so block b runs
// on XCD b % 8; the permutation gives each XCD a run of consecutive items
// (neighbouring tiles of one output block share their A rows / B columns).
__device__ __forceinline__ int xcd_first_item() {
  const int G = gridDim.x, b = blockIdx.x;
  return (G & 7) ? b : (b & 7) * (G >> 3) + (b >> 3);
}

// ------------------------------------------------------------------ operand loads
// Raw buffer loads over one operand block: the resource's base is the block
// start, offsets are bytes, and an offset at or above 2^31 reads as zero.
// Row padding and the ragged end of a k chunk become such offsets, so the
// staging loop has no branches.  Blocks are below 2^31 bytes (checked by the
// planner), so a padding offset plus a k offset cannot wrap into range.
constexpr unsigned kOOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t block_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7FFFFFFF,
                                           0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, unsigned off);
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <>
__device__ __forceinline__ unsigned buf_load<unsigned>(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}

// ------------------------------------------------------------------ FP64/FP32
// WM x (4/WM) waves; each wave owns (TM/WM) x (TN/WN) of the tile.
//
// Staging: element e = tid + 256 i of the TM x TK A tile (TN x TK B tile)
// is slot i of thread tid, laid along the operand's contiguous HBM
// direction (AM: rows, else k; BK: k, else columns) so that a wave's loads
// coalesce.  Per task each slot keeps one 32-bit byte offset into the
// operand block; per k-tile only the uniform (chunk, k1) offset is added.
// LDS holds [row][k] tiles (row pitch LDK), double-buffered: one barrier per
// k-tile.  MFMA j of a 16-deep k-tile takes tile k = 4q + j from lane group
// q = lane / 16, so each lane reads its four k values of a row with 16-byte
// LDS loads.
template <typename T, int TM, int TN, int WM, bool AM, bool BK, typename OutT>
__global__ void __launch_bounds__(kThreads)
k_gemm(const Tile* __restrict__ tiles, int ntiles, const CBlock* __restrict__ cblk,
       const GemmTask* __restrict__ tasks, const T* __restrict__ A, const T* __restrict__ B,
       OutT* __restrict__ C, T* __restrict__ part) {
  constexpr int TK = 16;
  constexpr int LDK = sizeof(T) == 8 ? TK + 2 : TK + 4;   // 16-row bank spread
  constexpr int ESZ = (int)sizeof(T);
  constexpr int WN = 4 / WM;
  constexpr int FM = TM / (16 * WM), FN = TN / (16 * WN);
  constexpr int NA = TM * TK / kThreads, NB = TN * TK / kThreads;
  static_assert(FM >= 1 && FN >= 1 && NA >= 1 && NB >= 1, "tile shape");
  static_assert(TM * TK % kThreads == 0 && TN * TK % kThreads == 0, "staging slots");
  typedef typename MfmaOp<T>::acc_t acc_t;
  typedef T v2t __attribute__((ext_vector_type(2)));
  __shared__ __attribute__((aligned(16))) T As[2][TM * LDK];
  __shared__ __attribute__((aligned(16))) T Bs[2][TN * LDK];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int q4 = 4 * (lane >> 4), r16 = lane & 15;
  int ar[NA], ak[NA], br[NB], bk[NB];
#pragma unroll
  for (int i = 0; i < NA; i++) {
    const int e = tid + kThreads * i;
    ar[i] = AM ? e % TM : e / TK;
    ak[i] = AM ? e / TM : e % TK;
  }
#pragma unroll
  for (int i = 0; i < NB; i++) {
    const int e = tid + kThreads * i;
    br[i] = BK ? e / TK : e % TN;
    bk[i] = BK ? e % TK : e / TN;
  }

  for (int ti = xcd_first_item(); ti < ntiles; ti += gridDim.x) {
    const Tile tl = tiles[ti];
    const CBlock cb = cblk[tl.c];
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++) acc[i][j] = acc_t{0, 0, 0, 0};

    T ra[NA], rb[NB];
    unsigned aoff[NA], boff[NB];
    __amdgpu_buffer_rsrc_t rsA, rsB;
    int cur_task = -1;
    GemmTask g;
    auto prepare = [&](int tt) {
      g = tasks[tt];
      cur_task = tt;
      rsA = block_rsrc(A + g.a_off);
      rsB = block_rsrc(B + g.b_off);
      // a tile whose rows (columns) all lie in the first folded component
      // (uniform test) needs no division to place them
      const unsigned sam0 = (unsigned)g.sam0, sam1 = (unsigned)g.sam1, sak0 = (unsigned)g.sak0;
      const unsigned sbn0 = (unsigned)g.sbn0, sbn1 = (unsigned)g.sbn1, sbk0 = (unsigned)g.sbk0;
      if (tl.m0 + TM <= g.em0) {
#pragma unroll
        for (int i = 0; i < NA; i++) {
          const int mg = tl.m0 + ar[i];
          aoff[i] = mg < cb.M ? ((unsigned)mg * sam0 + (unsigned)ak[i] * sak0) * ESZ : kOOB;
        }
      } else {
#pragma unroll
        for (int i = 0; i < NA; i++) {
          const int mg = tl.m0 + ar[i];
          aoff[i] = mg < cb.M ? (off2u(mg, g.em0, sam0, sam1) + (unsigned)ak[i] * sak0) * ESZ : kOOB;
        }
      }
      if (tl.n0 + TN <= g.en0) {
#pragma unroll
        for (int i = 0; i < NB; i++) {
          const int ng = tl.n0 + br[i];
          boff[i] = ng < cb.N ? ((unsigned)ng * sbn0 + (unsigned)bk[i] * sbk0) * ESZ : kOOB;
        }
      } else {
#pragma unroll
        for (int i = 0; i < NB; i++) {
          const int ng = tl.n0 + br[i];
          boff[i] = ng < cb.N ? (off2u(ng, g.en0, sbn0, sbn1) + (unsigned)bk[i] * sbk0) * ESZ : kOOB;
        }
      }
    };
    auto load = [&](const KWalk& w) {
      if (w.t != cur_task) prepare(w.t);
      const unsigned ua = (unsigned)(((int64_t)w.c * g.sak0 + (int64_t)w.k1 * g.sak1) * ESZ);
      const unsigned ub = (unsigned)(((int64_t)w.c * g.sbk0 + (int64_t)w.k1 * g.sbk1) * ESZ);
      const int klim = g.ek0 - w.c;
      if (klim >= TK) {   // the whole k-tile lies inside the chunk (uniform): no per-element test
#pragma unroll
        for (int i = 0; i < NA; i++) ra[i] = buf_load<T>(rsA, aoff[i] + ua);
#pragma unroll
        for (int i = 0; i < NB; i++) rb[i] = buf_load<T>(rsB, boff[i] + ub);
      } else {
#pragma unroll
        for (int i = 0; i < NA; i++) ra[i] = buf_load<T>(rsA, ak[i] < klim ? aoff[i] + ua : kOOB);
#pragma unroll
        for (int i = 0; i < NB; i++) rb[i] = buf_load<T>(rsB, bk[i] < klim ? boff[i] + ub : kOOB);
      }
    };
    auto store = [&](int buf) {
#pragma unroll
      for (int i = 0; i < NA; i++) As[buf][ar[i] * LDK + ak[i]] = ra[i];
#pragma unroll
      for (int i = 0; i < NB; i++) Bs[buf][br[i] * LDK + bk[i]] = rb[i];
    };
    int left = tl.kt1 - tl.kt0;   // k-tiles still to compute
    if (left > 0) {
      KWalk w = kwalk_start(tasks, cb, tl.kt0, TK);
      load(w);
      store(0);
      __syncthreads();
      int buf = 0;
      while (left > 0) {
        kwalk_next(w, g, TK);
        if (--left > 0) load(w);   // next k-tile in flight during the MFMAs
        v2t af[FM][2], bf[FN][2];
#pragma unroll
        for (int i = 0; i < FM; i++) {
          const T* p = &As[buf][(wm * (TM / WM) + i * 16 + r16) * LDK + q4];
          af[i][0] = *(const v2t*)p;
          af[i][1] = *(const v2t*)(p + 2);
        }
#pragma unroll
        for (int j = 0; j < FN; j++) {
          const T* p = &Bs[buf][(wn * (TN / WN) + j * 16 + r16) * LDK + q4];
          bf[j][0] = *(const v2t*)p;
          bf[j][1] = *(const v2t*)(p + 2);
        }
#pragma unroll
        for (int kk = 0; kk < 4; kk++)
#pragma unroll
          for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++)
              acc[i][j] = MfmaOp<T>::mac(bf[j][kk >> 1][kk & 1], af[i][kk >> 1][kk & 1], acc[i][j]);
        if (left > 0) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
      }
    }
    // epilogue: the fragments hold the transposed tile (B^T A^T: the B
    // fragment goes in as the MFMA's first operand), so lane l has rows
    // m = l % 16 of output columns n = row(l, r) and the 16 lanes of a column
    // store 128 contiguous bytes of the column-major M x N C block; a split
    // piece writes its whole TM x TN partial (column-major) to its slot
    T* pt = tl.slot >= 0 ? part + (int64_t)tl.slot * (TM * TN) : nullptr;
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++) {
        const int ml = wm * (TM / WM) + i * 16 + r16, m = tl.m0 + ml;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int nl = wn * (TN / WN) + j * 16 + MfmaOp<T>::row(lane, r), n = tl.n0 + nl;
          if (pt) pt[ml + TM * nl] = acc[i][j][r];
          else if (m < cb.M && n < cb.N) C[cb.c_off + m + (int64_t)cb.M * n] = (OutT)acc[i][j][r];
        }
      }
  }
}

// ------------------------------------------------------------------ BF16X3
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// split-BF16 operand word of an FP32 value: bf16 hi (round to nearest even)
// in the upper half, bf16 of the remainder x - hi in the lower half
__device__ __forceinline__ unsigned split_word(float x) {
  __bf16 hi, lo;
  split_bf16(x, hi, lo);
  return ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16) | __builtin_bit_cast(unsigned short, lo);
}

__global__ void k_split_words(const double* __restrict__ s, unsigned* __restrict__ d, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    d[i] = split_word((float)s[i]);
}

void split_words(const double* src, unsigned* dst, size_t n, hipStream_t s) {
  const unsigned g = (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(k_split_words, dim3(g), dim3(256), 0, s, src, dst, n);
}

// eight staged operand values -> their hi and lo bf16 vectors.  FP32 values
// are split here; pre-split words (split_word) only have their halves
// regrouped, which gives the same bits.
__device__ __forceinline__ void split8(const float (&v)[8], v8bf& h, v8bf& l) {
#pragma unroll
  for (int u = 0; u < 8; u++) {
    __bf16 a, b;
    split_bf16(v[u], a, b);
    h[u] = a;
    l[u] = b;
  }
}
__device__ __forceinline__ void split8(const unsigned (&w)[8], v8bf& h, v8bf& l) {
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  v4u hv, lv;
#pragma unroll
  for (int j = 0; j < 4; j++) {   // byte selectors 0-3 pick from the second operand, 4-7 from the first
    hv[j] = __builtin_amdgcn_perm(w[2 * j + 1], w[2 * j], 0x07060302u);   // hi(w0) | hi(w1) << 16
    lv[j] = __builtin_amdgcn_perm(w[2 * j + 1], w[2 * j], 0x05040100u);   // lo(w0) | lo(w1) << 16
  }
  h = __builtin_bit_cast(v8bf, hv);
  l = __builtin_bit_cast(v8bf, lv);
}

// Operands are FP32 in HBM (In = float), split into hi/lo bf16 planes when
// staged into LDS ([row][k] rows of 32 k, pitch 40 to spread banks),
// double-buffered with one barrier per k-tile; or pre-split words
// (In = unsigned, split_word: the same bits, only regrouped when staged).
// A staging slot is one row's 8-element k group; per task each slot keeps a
// 32-bit byte offset into the operand block (buffer loads, out-of-range
// offsets read zero as in k_gemm), per k-tile the uniform chunk offset and
// the element's k stride are added.  MFMA operands: lane l holds
// A[l%16][8*(l/16) .. +7] and B[8*(l/16) .. +7][l%16].
template <int TM, int TN, int WM, bool AM, bool BK, typename OutT, typename In>
__global__ void __launch_bounds__(kThreads)
k_gemm_bf16x3(const Tile* __restrict__ tiles, int ntiles, const CBlock* __restrict__ cblk,
              const GemmTask* __restrict__ tasks, const In* __restrict__ A,
              const In* __restrict__ B, OutT* __restrict__ C, float* __restrict__ part) {
  constexpr int TK = 32;
  constexpr int WN = 4 / WM;
  constexpr int FM = TM / (16 * WM), FN = TN / (16 * WN);
  constexpr int KP = TK + 8;
  constexpr int NGA = TM * TK / 8, NGB = TN * TK / 8;   // 8-element k-groups per tile
  constexpr int GA = (NGA + kThreads - 1) / kThreads;
  constexpr int GB = (NGB + kThreads - 1) / kThreads;
  static_assert(FM >= 1 && FN >= 1, "tile shape");
  __shared__ __attribute__((aligned(16))) __bf16 AsH[2][TM * KP];
  __shared__ __attribute__((aligned(16))) __bf16 AsL[2][TM * KP];
  __shared__ __attribute__((aligned(16))) __bf16 BsH[2][TN * KP];
  __shared__ __attribute__((aligned(16))) __bf16 BsL[2][TN * KP];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  // slot i of this thread: row / column and k group; row-contiguous operands
  // put lanes along the rows, k-contiguous ones put 4 lanes on a row's 32 k
  int am[GA], akg[GA], bn[GB], bkg[GB];
#pragma unroll
  for (int i = 0; i < GA; i++) {
    const int e = tid + kThreads * i;
    am[i] = AM ? e % TM : e / 4;
    akg[i] = AM ? e / TM : e % 4;
  }
#pragma unroll
  for (int i = 0; i < GB; i++) {
    const int e = tid + kThreads * i;
    bn[i] = BK ? e / 4 : e % TN;
    bkg[i] = BK ? e % 4 : e / TN;
  }

  for (int ti = xcd_first_item(); ti < ntiles; ti += gridDim.x) {
    const Tile tl = tiles[ti];
    const CBlock cb = cblk[tl.c];
    v4f acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++) acc[i][j] = v4f{0, 0, 0, 0};
    In ra[GA][8], rb[GB][8];
    unsigned aoff[GA], boff[GB];
    __amdgpu_buffer_rsrc_t rsA, rsB;
    int cur_task = -1;
    GemmTask g;
    auto prepare = [&](int tt) {
      g = tasks[tt];
      cur_task = tt;
      rsA = block_rsrc(A + g.a_off);
      rsB = block_rsrc(B + g.b_off);
#pragma unroll
      for (int i = 0; i < GA; i++) {
        const int mg = tl.m0 + am[i];
        aoff[i] = (tid + kThreads * i < NGA && mg < cb.M)
                      ? (unsigned)((off2(mg, g.em0, g.sam0, g.sam1) + (int64_t)(8 * akg[i]) * g.sak0) * 4)
                      : kOOB;
      }
#pragma unroll
      for (int i = 0; i < GB; i++) {
        const int ng = tl.n0 + bn[i];
        boff[i] = (tid + kThreads * i < NGB && ng < cb.N)
                      ? (unsigned)((off2(ng, g.en0, g.sbn0, g.sbn1) + (int64_t)(8 * bkg[i]) * g.sbk0) * 4)
                      : kOOB;
      }
    };
    auto load = [&](const KWalk& w) {
      if (w.t != cur_task) prepare(w.t);
      const unsigned ua = (unsigned)(((int64_t)w.c * g.sak0 + (int64_t)w.k1 * g.sak1) * 4);
      const unsigned ub = (unsigned)(((int64_t)w.c * g.sbk0 + (int64_t)w.k1 * g.sbk1) * 4);
      const unsigned sa = (unsigned)(g.sak0 * 4), sb = (unsigned)(g.sbk0 * 4);
      const int klim = g.ek0 - w.c;
      if (klim >= TK) {   // the whole k-tile lies inside the chunk (uniform): no per-element test
#pragma unroll
        for (int i = 0; i < GA; i++)
#pragma unroll
          for (int u = 0; u < 8; u++) ra[i][u] = buf_load<In>(rsA, aoff[i] + (ua + u * sa));
#pragma unroll
        for (int i = 0; i < GB; i++)
#pragma unroll
          for (int u = 0; u < 8; u++) rb[i][u] = buf_load<In>(rsB, boff[i] + (ub + u * sb));
      } else {
#pragma unroll
        for (int i = 0; i < GA; i++) {
          const int kl = klim - 8 * akg[i];
#pragma unroll
          for (int u = 0; u < 8; u++)
            ra[i][u] = buf_load<In>(rsA, u < kl ? aoff[i] + (ua + u * sa) : kOOB);
        }
#pragma unroll
        for (int i = 0; i < GB; i++) {
          const int kl = klim - 8 * bkg[i];
#pragma unroll
          for (int u = 0; u < 8; u++)
            rb[i][u] = buf_load<In>(rsB, u < kl ? boff[i] + (ub + u * sb) : kOOB);
        }
      }
    };
    auto store = [&](int buf) {
#pragma unroll
      for (int i = 0; i < GA; i++) {
        if (tid + kThreads * i >= NGA) break;
        v8bf h, l;
        split8(ra[i], h, l);
        *(v8bf*)&AsH[buf][am[i] * KP + akg[i] * 8] = h;
        *(v8bf*)&AsL[buf][am[i] * KP + akg[i] * 8] = l;
      }
#pragma unroll
      for (int i = 0; i < GB; i++) {
        if (tid + kThreads * i >= NGB) break;
        v8bf h, l;
        split8(rb[i], h, l);
        *(v8bf*)&BsH[buf][bn[i] * KP + bkg[i] * 8] = h;
        *(v8bf*)&BsL[buf][bn[i] * KP + bkg[i] * 8] = l;
      }
    };
    int left = tl.kt1 - tl.kt0;
    if (left > 0) {
      KWalk w = kwalk_start(tasks, cb, tl.kt0, TK);
      load(w);
      store(0);
      __syncthreads();
      int buf = 0;
      const int kq = 8 * (lane >> 4);
      while (left > 0) {
        kwalk_next(w, g, TK);
        if (--left > 0) load(w);
        v8bf ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
        for (int i = 0; i < FM; i++) {
          const int row = wm * (TM / WM) + i * 16 + (lane & 15);
          ah[i] = *(const v8bf*)&AsH[buf][row * KP + kq];
          al[i] = *(const v8bf*)&AsL[buf][row * KP + kq];
        }
#pragma unroll
        for (int j = 0; j < FN; j++) {
          const int col = wn * (TN / WN) + j * 16 + (lane & 15);
          bh[j] = *(const v8bf*)&BsH[buf][col * KP + kq];
          bl[j] = *(const v8bf*)&BsL[buf][col * KP + kq];
        }
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int j = 0; j < FN; j++) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
          }
        if (left > 0) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
      }
    }
    // transposed fragments (see k_gemm): lane l holds rows m = l % 16
    float* pt = tl.slot >= 0 ? part + (int64_t)tl.slot * (TM * TN) : nullptr;
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++) {
        const int ml = wm * (TM / WM) + i * 16 + (lane & 15), m = tl.m0 + ml;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int nl = wn * (TN / WN) + j * 16 + 4 * (lane >> 4) + r, n = tl.n0 + nl;
          if (pt) pt[ml + TM * nl] = acc[i][j][r];
          else if (m < cb.M && n < cb.N) C[cb.c_off + m + (int64_t)cb.M * n] = (OutT)acc[i][j][r];
        }
      }
  }
}

// ------------------------------------------------------------------ skinny
// output element conversion: a plain cast, or a split-BF16 word (OutT = unsigned)
template <typename OutT, typename T>
__device__ __forceinline__ OutT to_out(T v) {
  return (OutT)v;
}
template <>
__device__ __forceinline__ unsigned to_out<unsigned, float>(float v) {
  return split_word(v);
}

// Contractions whose tasks all have N <= kSkinnyN and K <= kSkinnyK (the MPO
// stage of H_eff and of the environment update: K and N are MPO-link sector
// dims).  One work item = kSkinnyRows rows of one output block, one row per
// thread; per task the K x N operand block and the K offsets of A go to LDS
// (zero-padded to kSkinnyK x NB, NB = 4, 8 or 16, the smallest that holds the
// block's N), each thread loads its row's K values (coalesced across the wave
// when A's row stride is 1) and accumulates its NB outputs with FMA in T.  A
// and C are streamed once per task / once.
//
// One output block's rows for a compile-time column count NB (>= the
// block's N): the FMAs run over NB columns only.
template <typename T, typename OutT, int NB>
__device__ __forceinline__ void skinny_rows(const Tile& tl, const CBlock& cb, const GemmTask* __restrict__ tasks,
                                            const T* __restrict__ A, const T* __restrict__ B,
                                            OutT* __restrict__ C, T* Bs, unsigned* koff) {
  const int m = tl.m0 + (int)threadIdx.x;
  const bool row_ok = m < cb.M;
  T acc[NB];
#pragma unroll
  for (int n = 0; n < NB; n++) acc[n] = T(0);
  for (int t = cb.t0; t < cb.t1; t++) {
    const GemmTask g = tasks[t];
    __syncthreads();   // previous task's LDS reads are done
    // the K x N operand block, zero-padded to kSkinnyK x NB; a single folded
    // component (uniform test) is placed without division
    const bool k1 = g.K <= g.ek0, n1 = cb.N <= g.en0;
    for (int e = threadIdx.x; e < kSkinnyK * NB; e += kThreads) {
      const int k = e / NB, n = e % NB;
      T v = T(0);
      if (k < g.K && n < cb.N) {
        const unsigned ko = k1 ? (unsigned)k * (unsigned)g.sbk0 : off2u(k, g.ek0, (unsigned)g.sbk0, (unsigned)g.sbk1);
        const unsigned no = n1 ? (unsigned)n * (unsigned)g.sbn0 : off2u(n, g.en0, (unsigned)g.sbn0, (unsigned)g.sbn1);
        v = B[g.b_off + ko + no];
      }
      Bs[e] = v;
    }
    if (threadIdx.x < kSkinnyK) {
      const int k = threadIdx.x;
      koff[k] = k < g.K ? (k1 ? (unsigned)k * (unsigned)g.sak0 : off2u(k, g.ek0, (unsigned)g.sak0, (unsigned)g.sak1))
                        : 0u;
    }
    __syncthreads();
    if (!row_ok) continue;
    const T* ar = A + g.a_off + (m < g.em0 ? (int64_t)m * g.sam0 : off2(m, g.em0, g.sam0, g.sam1));
    for (int k0 = 0; k0 < g.K; k0 += 8) {
      T a[8];
      if (k0 + 8 <= g.K) {   // uniform: no per-element test
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] = ar[koff[k0 + u]];
      } else {
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] = k0 + u < g.K ? ar[koff[k0 + u]] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
#pragma unroll
        for (int n = 0; n < NB; n++) acc[n] += a[u] * Bs[(k0 + u) * NB + n];
    }
  }
  if (!row_ok) return;
  OutT* cr = C + cb.c_off + m;
#pragma unroll
  for (int n = 0; n < NB; n++)
    if (n < cb.N) cr[(int64_t)cb.M * n] = to_out<OutT>(acc[n]);
  for (int n = NB; n < cb.N; n++) cr[(int64_t)cb.M * n] = OutT(0);   // blocks without tasks
}

template <typename T, typename OutT>
__global__ void __launch_bounds__(kThreads)
k_skinny(const Tile* __restrict__ tiles, const CBlock* __restrict__ cblk,
         const GemmTask* __restrict__ tasks, const T* __restrict__ A, const T* __restrict__ B,
         OutT* __restrict__ C) {
  static_assert(kSkinnyRows == kThreads, "one row per thread");
  __shared__ T Bs[kSkinnyK * kSkinnyN];
  __shared__ unsigned koff[kSkinnyK];
  const Tile tl = tiles[blockIdx.x];
  const CBlock cb = cblk[tl.c];
  // the column count picks the accumulator width (uniform); blocks without
  // tasks are written as zeros by every form
  if (cb.N <= 4) skinny_rows<T, OutT, 4>(tl, cb, tasks, A, B, C, Bs, koff);
  else if (cb.N <= 8) skinny_rows<T, OutT, 8>(tl, cb, tasks, A, B, C, Bs, koff);
  else skinny_rows<T, OutT, kSkinnyN>(tl, cb, tasks, A, B, C, Bs, koff);
}

template <typename T, typename OutT>
static void launch_skinny(const ttg_plan_s& p, const void* A, const void* B, void* C,
                          hipStream_t s) {
  hipLaunchKernelGGL((k_skinny<T, OutT>), dim3(p.ntiles), dim3(kThreads), 0, s, p.d_tiles,
                     p.d_cblk, p.d_tasks, (const T*)A, (const T*)B, (OutT*)C);
}

// ------------------------------------------------------------------ split-K
// C tile of a split group = sum of its partial slots, added in slot order.
// An element's slot values are all loaded before the first add (a loop with
// one load per slot would wait one memory latency per slot).
template <typename AccT, typename OutT>
__global__ void __launch_bounds__(256)
k_split_reduce(const SplitGroup* __restrict__ groups, const CBlock* __restrict__ cblk,
               const AccT* __restrict__ part, int TM, int TN, OutT* __restrict__ C) {
  const SplitGroup gr = groups[blockIdx.x];
  const CBlock cb = cblk[gr.c];
  const int64_t tile = (int64_t)TM * TN;
  for (int e = threadIdx.x; e < TM * TN; e += blockDim.x) {
    const int m = gr.m0 + e % TM, n = gr.n0 + e / TM;
    if (m >= cb.M || n >= cb.N) continue;
    const AccT* src = part + (int64_t)gr.slot0 * tile + e;
    AccT v[kMaxSplit];
#pragma unroll
    for (int k = 0; k < kMaxSplit; k++) v[k] = k < gr.ns ? src[k * tile] : AccT(0);
    AccT sum = v[0];
#pragma unroll
    for (int k = 1; k < kMaxSplit; k++)
      if (k < gr.ns) sum += v[k];
    C[cb.c_off + m + (int64_t)cb.M * n] = (OutT)sum;
  }
}

template <typename AccT, typename OutT>
static void launch_reduce(const ttg_plan_s& p, void* C, hipStream_t s) {
  if (p.ngroups == 0) return;
  hipLaunchKernelGGL((k_split_reduce<AccT, OutT>), dim3(p.ngroups), dim3(256), 0, s, p.d_groups,
                     p.d_cblk, (const AccT*)p.d_part, cfg_tm(p.cfg), cfg_tn(p.cfg), (OutT*)C);
}

// ------------------------------------------------------------------ dispatch
// Tile shapes: (config, TM, TN, waves along M).
#define TTG_TILE_SHAPES(X)                                                                       \
  X(TILE_64x64, 64, 64, 2) X(TILE_32x32, 32, 32, 2) X(TILE_64x32, 64, 32, 2)                     \
  X(TILE_32x64, 32, 64, 2) X(TILE_128x64, 128, 64, 2) X(TILE_256x16, 256, 16, 4)                 \
  X(TILE_128x16, 128, 16, 4) X(TILE_64x16, 64, 16, 4) X(TILE_96x96, 96, 96, 2)                   \
  X(TILE_96x32, 96, 32, 2) X(TILE_32x96, 32, 96, 2)

// the kernel instantiation that serves (precision, shape, operand layout)
template <typename T, int TM, int TN, int WM, typename OutT>
static const void* fp_kernel(bool am, bool bk) {
  if (am && bk) return (const void*)k_gemm<T, TM, TN, WM, true, true, OutT>;
  if (am) return (const void*)k_gemm<T, TM, TN, WM, true, false, OutT>;
  if (bk) return (const void*)k_gemm<T, TM, TN, WM, false, true, OutT>;
  return (const void*)k_gemm<T, TM, TN, WM, false, false, OutT>;
}
template <int TM, int TN, int WM, typename OutT, typename In = float>
static const void* bf_kernel(bool am, bool bk) {
  if (am && bk) return (const void*)k_gemm_bf16x3<TM, TN, WM, true, true, OutT, In>;
  if (am) return (const void*)k_gemm_bf16x3<TM, TN, WM, true, false, OutT, In>;
  if (bk) return (const void*)k_gemm_bf16x3<TM, TN, WM, false, true, OutT, In>;
  return (const void*)k_gemm_bf16x3<TM, TN, WM, false, false, OutT, In>;
}
template <typename OutT>
static const void* kernel_of(int prec, int cfg, bool am, bool bk) {
  switch (cfg) {
#define TTG_K(C, TM, TN, WM)                                                                     \
  case C:                                                                                        \
    return prec == TTG_FP64   ? fp_kernel<double, TM, TN, WM, OutT>(am, bk)                      \
           : prec == TTG_FP32 ? fp_kernel<float, TM, TN, WM, OutT>(am, bk)                       \
                              : bf_kernel<TM, TN, WM, OutT>(am, bk);
    TTG_TILE_SHAPES(TTG_K)
#undef TTG_K
    default: return nullptr;
  }
}

// workgroups of a kernel one CU holds at once (register / LDS occupancy),
// cached per instantiation
static int occupancy_of(const void* kernel) {
  static std::unordered_map<const void*, int> per_cu;
  auto it = per_cu.find(kernel);
  if (it == per_cu.end()) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kThreads, 0) != hipSuccess || n < 1)
      n = 4;
    it = per_cu.emplace(kernel, std::min(n, 8)).first;
  }
  return it->second;
}

int gemm_occupancy(int prec, int cfg, bool a_mcontig, bool b_kcontig) {
  const void* k = prec == TTG_FP64 ? kernel_of<double>(prec, cfg, a_mcontig, b_kcontig)
                                   : kernel_of<float>(prec, cfg, a_mcontig, b_kcontig);
  return k ? occupancy_of(k) : 1;
}

// persistent grid: every workgroup the CUs can hold at once for this kernel,
// at most one per work item
static dim3 persistent_grid(const ttg_plan_s& p, const void* kernel) {
  const int64_t g = (int64_t)occupancy_of(kernel) * rt().num_cu;
  return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(g, p.ntiles)));
}

template <typename T, int TM, int TN, int WM, typename OutT>
static void launch_mode(const ttg_plan_s& p, const void* A, const void* B, void* C, hipStream_t s) {
  const dim3 grid = persistent_grid(p, fp_kernel<T, TM, TN, WM, OutT>(p.a_mcontig, p.b_kcontig));
#define TTG_L(AMv, BKv)                                                                           \
  hipLaunchKernelGGL((k_gemm<T, TM, TN, WM, AMv, BKv, OutT>), grid, dim3(kThreads), 0, s,        \
                     p.d_tiles, p.ntiles, p.d_cblk, p.d_tasks, (const T*)A, (const T*)B,        \
                     (OutT*)C, (T*)p.d_part)
  if (p.a_mcontig && p.b_kcontig) TTG_L(true, true);
  else if (p.a_mcontig) TTG_L(true, false);
  else if (p.b_kcontig) TTG_L(false, true);
  else TTG_L(false, false);
#undef TTG_L
}

template <int TM, int TN, int WM, typename OutT, typename In>
static void launch_bf16(const ttg_plan_s& p, const void* A, const void* B, void* C, hipStream_t s) {
  const dim3 grid = persistent_grid(p, bf_kernel<TM, TN, WM, OutT, In>(p.a_mcontig, p.b_kcontig));
#define TTG_L(AMv, BKv)                                                                             \
  hipLaunchKernelGGL((k_gemm_bf16x3<TM, TN, WM, AMv, BKv, OutT, In>), grid, dim3(kThreads), 0, s, \
                     p.d_tiles, p.ntiles, p.d_cblk, p.d_tasks, (const In*)A, (const In*)B,       \
                     (OutT*)C, (float*)p.d_part)
  if (p.a_mcontig && p.b_kcontig) TTG_L(true, true);
  else if (p.a_mcontig) TTG_L(true, false);
  else if (p.b_kcontig) TTG_L(false, true);
  else TTG_L(false, false);
#undef TTG_L
}

template <typename T, typename OutT>
static void by_cfg(const ttg_plan_s& p, const void* A, const void* B, void* C, hipStream_t s) {
  switch (p.cfg) {
#define TTG_C(C_, TM, TN, WM) \
  case C_: launch_mode<T, TM, TN, WM, OutT>(p, A, B, C, s); break;
    TTG_TILE_SHAPES(TTG_C)
#undef TTG_C
    default: break;
  }
}

template <typename OutT, typename In>
static void by_cfg_bf16(const ttg_plan_s& p, const void* A, const void* B, void* C, hipStream_t s) {
  switch (p.cfg) {
#define TTG_C(C_, TM, TN, WM) \
  case C_: launch_bf16<TM, TN, WM, OutT, In>(p, A, B, C, s); break;
    TTG_TILE_SHAPES(TTG_C)
#undef TTG_C
    default: break;
  }
}

int gemm_launch(const ttg_plan_s& p, const void* A, const void* B, void* C, bool out_f64,
                hipStream_t s) {
  TTG_REQUIRE(!p.tiles_pending, "gemm: the plan's work list is not built");
  if (p.ntiles == 0) return 0;
  TTG_REQUIRE(!(p.packed_in || p.split_out) || p.prec == TTG_BF16X3,
              "gemm: split-BF16 words are operands of the split-BF16 form only");
  if (p.cfg == TILE_SKINNY) {
    TTG_REQUIRE(!p.packed_in, "gemm: the vector-ALU kernel does not read split words");
    if (p.split_out) {
      TTG_REQUIRE(!out_f64, "gemm: a split-word output is not FP64");
      launch_skinny<float, unsigned>(p, A, B, C, s);
    } else if (p.prec == TTG_FP64) {
      if (out_f64) launch_skinny<double, double>(p, A, B, C, s);
      else launch_skinny<double, float>(p, A, B, C, s);
    } else {
      if (out_f64) launch_skinny<float, double>(p, A, B, C, s);
      else launch_skinny<float, float>(p, A, B, C, s);
    }
    TTG_HIP(hipGetLastError());
    return 0;
  }
  if (p.prec == TTG_FP64) {
    if (out_f64) by_cfg<double, double>(p, A, B, C, s);
    else by_cfg<double, float>(p, A, B, C, s);
  } else if (p.prec == TTG_FP32) {
    if (out_f64) by_cfg<float, double>(p, A, B, C, s);
    else by_cfg<float, float>(p, A, B, C, s);
  } else {
    TTG_REQUIRE(!p.split_out, "gemm: the MFMA kernels do not write split words");
    if (p.packed_in) {
      if (out_f64) by_cfg_bf16<double, unsigned>(p, A, B, C, s);
      else by_cfg_bf16<float, unsigned>(p, A, B, C, s);
    } else {
      if (out_f64) by_cfg_bf16<double, float>(p, A, B, C, s);
      else by_cfg_bf16<float, float>(p, A, B, C, s);
    }
  }
  if (p.prec == TTG_FP64) {
    if (out_f64) launch_reduce<double, double>(p, C, s);
    else launch_reduce<double, float>(p, C, s);
  } else {
    if (out_f64) launch_reduce<float, double>(p, C, s);
    else launch_reduce<float, float>(p, C, s);
  }
  TTG_HIP(hipGetLastError());
  return 0;
}

}  // namespace ttg
