// gemm.hip -- persistent grouped block GEMM over a plan's tile list.
//
// Every uneven block pair of a contraction runs in ONE launch.  A tile is a
// TM x TN piece of one output block; the workgroup owning it walks all of
// the block's tasks (the K contributions, in the reference's mid order) and
// writes the tile once, so no atomics and no zero-fill pass are needed
// (blocks without tasks are written as zeros).  Operand tiles are staged
// through LDS; the next k-slice is prefetched into registers while the MFMA
// pipe works on the current one.
//
// MFMA building block: 16x16 output fragment per wave-instruction.
//   FP64:   v_mfma_f64_16x16x4_f64      (A,B one double per lane)
//   FP32:   v_mfma_f32_16x16x4_f32      (A,B one float per lane)
//   BF16X3: v_mfma_f32_16x16x32_bf16 x3 (hi*hi + hi*lo + lo*hi, FP32 acc)
// Operand lane layout (k-slice of 4):  lane l holds A[l%16][l/16] and
// B[l/16][l%16]; result register r of lane l holds D[row][l%16] with
// row = l/16 + 4r for f64 and 4*(l/16) + r for the f32-accumulating ops.
#include <algorithm>

#include "ttg_internal.h"
#include "ttg_plan.h"

namespace ttg {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <typename T>
struct MfmaOp;
// row of result register r in lane l of a 16x16 fragment (column = l%16):
// the f64 MFMA interleaves rows across registers, the f32 ones keep each
// lane's four rows contiguous (both checked by an in-bounds layout probe).
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

constexpr int kThreads = 256;

// ------------------------------------------------------------------ FP64/FP32
template <typename T, int TM, int TN, bool AM, bool BK, typename OutT>
__global__ void __launch_bounds__(kThreads)
k_gemm(const Tile* __restrict__ tiles, int ntiles, const CBlock* __restrict__ cblk,
       const GemmTask* __restrict__ tasks, const T* __restrict__ A, const T* __restrict__ B,
       OutT* __restrict__ C) {
  constexpr int TK = 16;
  constexpr int FM = TM / 32, FN = TN / 32;   // 16x16 fragments per wave (2x2 waves)
  constexpr int NA = TM * TK / kThreads, NB = TN * TK / kThreads;
  constexpr int LDA = TM + 1, LDB = TN + 1;
  typedef typename MfmaOp<T>::acc_t acc_t;
  __shared__ T As[TK * LDA];
  __shared__ T Bs[TK * LDB];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
    const Tile tl = tiles[ti];
    const CBlock cb = cblk[tl.c];
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++) acc[i][j] = acc_t{0, 0, 0, 0};

    T ra[NA], rb[NB];
    int t = cb.t0, kk = 0;
    auto load = [&](int tt, int k0) {
      const GemmTask g = tasks[tt];
#pragma unroll
      for (int i = 0; i < NA; i++) {
        int m, k;
        if (AM) { m = tid % TM; k = tid / TM + (kThreads / TM) * i; }
        else { k = tid % TK; m = tid / TK + (kThreads / TK) * i; }
        const int mg = tl.m0 + m, kg = k0 + k;
        T v = T(0);
        if (mg < cb.M && kg < g.K)
          v = A[g.a_off + off2(mg, g.em0, g.sam0, g.sam1) + off2(kg, g.ek0, g.sak0, g.sak1)];
        ra[i] = v;
      }
#pragma unroll
      for (int i = 0; i < NB; i++) {
        int n, k;
        if (BK) { k = tid % TK; n = tid / TK + (kThreads / TK) * i; }
        else { n = tid % TN; k = tid / TN + (kThreads / TN) * i; }
        const int ng = tl.n0 + n, kg = k0 + k;
        T v = T(0);
        if (ng < cb.N && kg < g.K)
          v = B[g.b_off + off2(kg, g.ek0, g.sbk0, g.sbk1) + off2(ng, g.en0, g.sbn0, g.sbn1)];
        rb[i] = v;
      }
    };
    bool have = t < cb.t1;
    if (have) load(t, kk);
    while (have) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < NA; i++) {
        int m, k;
        if (AM) { m = tid % TM; k = tid / TM + (kThreads / TM) * i; }
        else { k = tid % TK; m = tid / TK + (kThreads / TK) * i; }
        As[k * LDA + m] = ra[i];
      }
#pragma unroll
      for (int i = 0; i < NB; i++) {
        int n, k;
        if (BK) { k = tid % TK; n = tid / TK + (kThreads / TK) * i; }
        else { n = tid % TN; k = tid / TN + (kThreads / TN) * i; }
        Bs[k * LDB + n] = rb[i];
      }
      __syncthreads();
      kk += TK;
      if (kk >= tasks[t].K) {
        t++;
        kk = 0;
      }
      have = t < cb.t1;
      if (have) load(t, kk);
#pragma unroll
      for (int k4 = 0; k4 < TK; k4 += 4) {
        const int kr = k4 + (lane >> 4);
        T af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; i++) af[i] = As[kr * LDA + wm * (TM / 2) + i * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; j++) bf[j] = Bs[kr * LDB + wn * (TN / 2) + j * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
          for (int j = 0; j < FN; j++) acc[i][j] = MfmaOp<T>::mac(af[i], bf[j], acc[i][j]);
      }
    }
    // epilogue: C block is column-major M x N
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++) {
        const int n = tl.n0 + wn * (TN / 2) + j * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int m = tl.m0 + wm * (TM / 2) + i * 16 + MfmaOp<T>::row(lane, r);
          if (m < cb.M && n < cb.N) C[cb.c_off + m + (int64_t)cb.M * n] = (OutT)acc[i][j][r];
        }
      }
  }
}

// ------------------------------------------------------------------ BF16X3
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <int TM, int TN, bool AM, bool BK, typename OutT>
__global__ void __launch_bounds__(kThreads)
k_gemm_bf16x3(const Tile* __restrict__ tiles, int ntiles, const CBlock* __restrict__ cblk,
              const GemmTask* __restrict__ tasks, const float* __restrict__ A,
              const float* __restrict__ B, OutT* __restrict__ C) {
  constexpr int TK = 32;
  constexpr int FM = TM / 32, FN = TN / 32;
  constexpr int KP = TK + 8;                     // padded k-row of bf16 (80 B)
  constexpr int NGA = TM * TK / 8, NGB = TN * TK / 8;   // 8-element k-groups per tile
  constexpr int GA = (NGA + kThreads - 1) / kThreads;
  constexpr int GB = (NGB + kThreads - 1) / kThreads;
  __shared__ __attribute__((aligned(16))) __bf16 AsH[TM * KP];
  __shared__ __attribute__((aligned(16))) __bf16 AsL[TM * KP];
  __shared__ __attribute__((aligned(16))) __bf16 BsH[TN * KP];
  __shared__ __attribute__((aligned(16))) __bf16 BsL[TN * KP];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
    const Tile tl = tiles[ti];
    const CBlock cb = cblk[tl.c];
    v4f acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++) acc[i][j] = v4f{0, 0, 0, 0};
    float ra[GA][8], rb[GB][8];
    int t = cb.t0, kk = 0;
    // thread -> (row, k-group): row-contiguous operands put lanes along the
    // row; k-contiguous ones put 4 consecutive lanes on one row's 32 k.
    auto a_map = [&](int i, int& m, int& kg) {
      const int e = tid + kThreads * i;
      if (AM) { m = e % TM; kg = e / TM; } else { kg = e % 4; m = e / 4; }
    };
    auto b_map = [&](int i, int& n, int& kg) {
      const int e = tid + kThreads * i;
      if (BK) { kg = e % 4; n = e / 4; } else { n = e % TN; kg = e / TN; }
    };
    auto load = [&](int tt, int k0) {
      const GemmTask g = tasks[tt];
#pragma unroll
      for (int i = 0; i < GA; i++) {
        int m, kg;
        a_map(i, m, kg);
        const int mg = tl.m0 + m;
        const bool mok = mg < cb.M && tid + kThreads * i < NGA;
        const int64_t mo = mok ? off2(mg, g.em0, g.sam0, g.sam1) : 0;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int kq = k0 + kg * 8 + u;
          ra[i][u] = (mok && kq < g.K) ? A[g.a_off + mo + off2(kq, g.ek0, g.sak0, g.sak1)] : 0.f;
        }
      }
#pragma unroll
      for (int i = 0; i < GB; i++) {
        int n, kg;
        b_map(i, n, kg);
        const int ng = tl.n0 + n;
        const bool nok = ng < cb.N && tid + kThreads * i < NGB;
        const int64_t no = nok ? off2(ng, g.en0, g.sbn0, g.sbn1) : 0;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int kq = k0 + kg * 8 + u;
          rb[i][u] = (nok && kq < g.K) ? B[g.b_off + no + off2(kq, g.ek0, g.sbk0, g.sbk1)] : 0.f;
        }
      }
    };
    bool have = t < cb.t1;
    if (have) load(t, kk);
    while (have) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < GA; i++) {
        if (tid + kThreads * i >= NGA) break;
        int m, kg;
        a_map(i, m, kg);
        v8bf h, l;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          __bf16 a, b;
          split_bf16(ra[i][u], a, b);
          h[u] = a;
          l[u] = b;
        }
        *(v8bf*)&AsH[m * KP + kg * 8] = h;
        *(v8bf*)&AsL[m * KP + kg * 8] = l;
      }
#pragma unroll
      for (int i = 0; i < GB; i++) {
        if (tid + kThreads * i >= NGB) break;
        int n, kg;
        b_map(i, n, kg);
        v8bf h, l;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          __bf16 a, b;
          split_bf16(rb[i][u], a, b);
          h[u] = a;
          l[u] = b;
        }
        *(v8bf*)&BsH[n * KP + kg * 8] = h;
        *(v8bf*)&BsL[n * KP + kg * 8] = l;
      }
      __syncthreads();
      kk += TK;
      if (kk >= tasks[t].K) {
        t++;
        kk = 0;
      }
      have = t < cb.t1;
      if (have) load(t, kk);
      const int kq = 8 * (lane >> 4);
      v8bf ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; i++) {
        const int row = wm * (TM / 2) + i * 16 + (lane & 15);
        ah[i] = *(const v8bf*)&AsH[row * KP + kq];
        al[i] = *(const v8bf*)&AsL[row * KP + kq];
      }
#pragma unroll
      for (int j = 0; j < FN; j++) {
        const int col = wn * (TN / 2) + j * 16 + (lane & 15);
        bh[j] = *(const v8bf*)&BsH[col * KP + kq];
        bl[j] = *(const v8bf*)&BsL[col * KP + kq];
      }
#pragma unroll
      for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++) {
        const int n = tl.n0 + wn * (TN / 2) + j * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int m = tl.m0 + wm * (TM / 2) + i * 16 + 4 * (lane >> 4) + r;
          if (m < cb.M && n < cb.N) C[cb.c_off + m + (int64_t)cb.M * n] = (OutT)acc[i][j][r];
        }
      }
  }
}

// ------------------------------------------------------------------ dispatch
int gemm_grid(const ttg_plan_s& p) {
  int g = rt().num_cu * 4;
  return std::max(1, std::min(g, p.ntiles));
}

template <typename T, int TM, int TN, typename OutT>
static void launch_mode(const ttg_plan_s& p, const void* A, const void* B, void* C, hipStream_t s) {
  dim3 grid(gemm_grid(p)), block(kThreads);
#define TTG_L(AMv, BKv)                                                                         \
  hipLaunchKernelGGL((k_gemm<T, TM, TN, AMv, BKv, OutT>), grid, block, 0, s, p.d_tiles, p.ntiles, \
                     p.d_cblk, p.d_tasks, (const T*)A, (const T*)B, (OutT*)C)
  if (p.a_mcontig && p.b_kcontig) TTG_L(true, true);
  else if (p.a_mcontig) TTG_L(true, false);
  else if (p.b_kcontig) TTG_L(false, true);
  else TTG_L(false, false);
#undef TTG_L
}

template <int TM, int TN, typename OutT>
static void launch_bf16(const ttg_plan_s& p, const void* A, const void* B, void* C, hipStream_t s) {
  dim3 grid(gemm_grid(p)), block(kThreads);
#define TTG_L(AMv, BKv)                                                                       \
  hipLaunchKernelGGL((k_gemm_bf16x3<TM, TN, AMv, BKv, OutT>), grid, block, 0, s, p.d_tiles,    \
                     p.ntiles, p.d_cblk, p.d_tasks, (const float*)A, (const float*)B, (OutT*)C)
  if (p.a_mcontig && p.b_kcontig) TTG_L(true, true);
  else if (p.a_mcontig) TTG_L(true, false);
  else if (p.b_kcontig) TTG_L(false, true);
  else TTG_L(false, false);
#undef TTG_L
}

template <typename OutT>
static int dispatch(const ttg_plan_s& p, const void* A, const void* B, void* C, hipStream_t s) {
  switch (p.prec) {
    case TTG_FP64:
      switch (p.cfg) {
        case TILE_64x64: launch_mode<double, 64, 64, OutT>(p, A, B, C, s); break;
        case TILE_32x32: launch_mode<double, 32, 32, OutT>(p, A, B, C, s); break;
        case TILE_64x32: launch_mode<double, 64, 32, OutT>(p, A, B, C, s); break;
        case TILE_32x64: launch_mode<double, 32, 64, OutT>(p, A, B, C, s); break;
        default: launch_mode<double, 128, 64, OutT>(p, A, B, C, s); break;
      }
      break;
    case TTG_FP32:
      switch (p.cfg) {
        case TILE_64x64: launch_mode<float, 64, 64, OutT>(p, A, B, C, s); break;
        case TILE_32x32: launch_mode<float, 32, 32, OutT>(p, A, B, C, s); break;
        case TILE_64x32: launch_mode<float, 64, 32, OutT>(p, A, B, C, s); break;
        case TILE_32x64: launch_mode<float, 32, 64, OutT>(p, A, B, C, s); break;
        default: launch_mode<float, 128, 64, OutT>(p, A, B, C, s); break;
      }
      break;
    default:
      switch (p.cfg) {
        case TILE_64x64: launch_bf16<64, 64, OutT>(p, A, B, C, s); break;
        case TILE_32x32: launch_bf16<32, 32, OutT>(p, A, B, C, s); break;
        case TILE_64x32: launch_bf16<64, 32, OutT>(p, A, B, C, s); break;
        case TILE_32x64: launch_bf16<32, 64, OutT>(p, A, B, C, s); break;
        default: launch_bf16<128, 64, OutT>(p, A, B, C, s); break;
      }
  }
  return 0;
}

int gemm_launch(const ttg_plan_s& p, const void* A, const void* B, void* C, bool out_f64,
                hipStream_t s) {
  if (p.ntiles == 0) return 0;
  if (out_f64) dispatch<double>(p, A, B, C, s);
  else dispatch<float>(p, A, B, C, s);
  TTG_HIP(hipGetLastError());
  return 0;
}

}  // namespace ttg
