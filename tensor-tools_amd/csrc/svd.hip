// svd.hip -- sector SVD with global truncation (qtensor_op.cpp:1097-1453).
//
// Per mid-QN sector q (flux of the left indices, ascending QN order) the
// reference assembles a dense matrix from the blocks (rows: left position
// tuples in lexicographic order, columns: right tuples), takes its SVD,
// absorbs S into V (MoveFromLeft) or U (MoveFromRight), sorts all singular
// values globally (descending), keeps them while 1 - cumsum(s^2/sum s^2)
// >= cutoff (stopping at exact zeros, capping at K) and slices U, S, V.
//
// Device form (all FP64):
//   1. bookkeeping kernels: per-block sector keys, sorted unique row / column
//      tuple lists, sector sizes and row / column offsets;
//   2. gather of the blocks into packed sector matrices;
//   3. Gram matrix per sector on the grouped MFMA GEMM: A*A^T for
//      MoveFromLeft (the left factor U is what must be orthonormal),
//      A^T*A for MoveFromRight;
//   4. one workgroup per sector: parallel cyclic Jacobi eigensolver
//      (round-robin pairs, packed upper triangle in LDS when it fits) with
//      the eigenvectors accumulated in HBM;  s = sqrt(max(lambda, 0));
//   5. truncation on the device: per-sector rank-K cap, sum of s^2 and the
//      sequential cumsum rule in one thread, in the reference's order;
//   6. the absorbed factor as a second grouped GEMM (U^T A or A V), then the
//      slicing of U, V and S into blocks.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "ttg_device.h"
#include "ttg_internal.h"
#include "ttg_plan.h"

namespace ttg {

__device__ inline uint32_t qcode_s(int q0, int q1) {
  return ((uint32_t)(q0 + 32768) << 16) | (uint32_t)(q1 + 32768);
}

// 1a. keys: (q << 32 | left tuple), (q << 32 | right tuple)
__global__ void k_svd_keys(MetaView A, int nleft, uint64_t* kl, uint64_t* kr, int* ids) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= A.nblocks) return;
  const int* qi = A.qi + (size_t)b * A.rank;
  int q0 = 0, q1 = 0;
  uint64_t l = 0, r = 0;
  for (int j = 0; j < A.rank; j++) {
    if (j < nleft) {
      const int* e = A.idx + 3 * (A.idx_off[j] + qi[j]);
      int sg = A.arrow[j] == 0 ? 1 : -1;   // qtensor_op.cpp:1129-1133
      q0 += sg * e[0];
      q1 += sg * e[1];
      l = l * A.nsec[j] + qi[j];
    } else {
      r = r * A.nsec[j] + qi[j];
    }
  }
  uint64_t qc = (uint64_t)qcode_s(q0, q1) << 32;
  kl[b] = qc | l;
  kr[b] = qc | r;
  ids[b] = b;
}

// size (product of sector dims) of a tuple over indices [j0, j1)
__device__ inline int64_t tuple_size(const MetaView& A, int j0, int j1, uint64_t key) {
  int64_t s = 1;
  for (int j = j1 - 1; j >= j0; j--) {
    int p = (int)(key % A.nsec[j]);
    key /= A.nsec[j];
    s *= A.idx[3 * (A.idx_off[j] + p) + 2];
  }
  return s;
}

struct SvdTab {
  uint64_t *uL, *uR;          // unique sorted keys
  int64_t *lsz, *loff;        // per uL entry: size, row offset inside its sector
  int64_t *rsz, *roff;        // per uR entry
  int *qL0, *qL1, *qR0, *qR1; // per sector ranges
  int64_t *rows, *cols;       // per sector
  uint32_t* qc;               // per sector QN code
  int64_t* hdr;               // [nuL, nuR, nQ]
};

// 1b. unique lists, sector table, offsets (one workgroup)
__global__ void __launch_bounds__(1024) k_svd_unique(MetaView A, int nleft, const uint64_t* sl,
                                                     const uint64_t* sr, int n, SvdTab t) {
  __shared__ int64_t lds[32];
  int64_t carry = 0;
  for (int base = 0; base < n; base += 1024) {   // unique left
    int i = base + threadIdx.x;
    int f = i < n && (i == 0 || sl[i] != sl[i - 1]);
    int64_t tot;
    int64_t pos = block_excl_scan<1024>((int64_t)f, lds, &tot);
    if (f) t.uL[carry + pos] = sl[i];
    carry += tot;
  }
  __syncthreads();
  const int nuL = (int)carry;
  carry = 0;
  for (int base = 0; base < n; base += 1024) {   // unique right
    int i = base + threadIdx.x;
    int f = i < n && (i == 0 || sr[i] != sr[i - 1]);
    int64_t tot;
    int64_t pos = block_excl_scan<1024>((int64_t)f, lds, &tot);
    if (f) t.uR[carry + pos] = sr[i];
    carry += tot;
  }
  __syncthreads();
  const int nuR = (int)carry;
  carry = 0;
  for (int base = 0; base < nuL; base += 1024) {  // sectors from uL
    int i = base + threadIdx.x;
    int f = i < nuL && (i == 0 || (t.uL[i] >> 32) != (t.uL[i - 1] >> 32));
    int64_t tot;
    int64_t pos = block_excl_scan<1024>((int64_t)f, lds, &tot);
    if (f) t.qL0[carry + pos] = i;
    carry += tot;
  }
  __syncthreads();
  const int nQ = (int)carry;
  for (int q = threadIdx.x; q < nQ; q += 1024) {
    int l0 = t.qL0[q], l1 = q + 1 < nQ ? t.qL0[q + 1] : nuL;
    t.qL1[q] = l1;
    uint64_t qc = t.uL[l0] >> 32;
    t.qc[q] = (uint32_t)qc;
    t.qR0[q] = lower_bound_u64(t.uR, nuR, qc << 32);
    t.qR1[q] = lower_bound_u64(t.uR, nuR, (qc + 1) << 32);
  }
  __syncthreads();
  // per-entry sizes and offsets inside the sector (sequential per sector)
  for (int q = threadIdx.x; q < nQ; q += 1024) {
    int64_t o = 0;
    for (int i = t.qL0[q]; i < t.qL1[q]; i++) {
      int64_t s = tuple_size(A, 0, nleft, t.uL[i] & 0xffffffffull);
      t.lsz[i] = s;
      t.loff[i] = o;
      o += s;
    }
    t.rows[q] = o;
    o = 0;
    for (int i = t.qR0[q]; i < t.qR1[q]; i++) {
      int64_t s = tuple_size(A, nleft, A.rank, t.uR[i] & 0xffffffffull);
      t.rsz[i] = s;
      t.roff[i] = o;
      o += s;
    }
    t.cols[q] = o;
  }
  if (threadIdx.x == 0) {
    t.hdr[0] = nuL;
    t.hdr[1] = nuR;
    t.hdr[2] = nQ;
  }
}

// 2. gather blocks into sector matrices (column-major rows_q x cols_q)
__global__ void k_svd_gather(MetaView A, int nleft, const double* __restrict__ a, SvdTab t,
                             int nuL, int nuR, int nQ, const int64_t* moff,
                             double* __restrict__ mat) {
  for (int b = blockIdx.x; b < A.nblocks; b += gridDim.x) {
    const int* qi = A.qi + (size_t)b * A.rank;
    int q0 = 0, q1 = 0;
    uint64_t l = 0, r = 0;
    for (int j = 0; j < A.rank; j++) {
      if (j < nleft) {
        const int* e = A.idx + 3 * (A.idx_off[j] + qi[j]);
        int sg = A.arrow[j] == 0 ? 1 : -1;
        q0 += sg * e[0];
        q1 += sg * e[1];
        l = l * A.nsec[j] + qi[j];
      } else {
        r = r * A.nsec[j] + qi[j];
      }
    }
    const uint64_t qc = (uint64_t)qcode_s(q0, q1) << 32;
    const int il = lower_bound_u64(t.uL, nuL, qc | l);
    const int ir = lower_bound_u64(t.uR, nuR, qc | r);
    // sector of il
    int lo = 0, hi = nQ;
    while (lo < hi) {
      int m = (lo + hi) >> 1;
      if (t.qL0[m] <= il) lo = m + 1;
      else hi = m;
    }
    const int q = lo - 1;
    const int64_t rows = t.rows[q];
    const int64_t ls = t.lsz[il], rs = t.rsz[ir];
    double* dst = mat + moff[q] + t.loff[il] + rows * t.roff[ir];
    const double* src = a + A.off[b];
    const int64_t n = ls * rs;
#pragma unroll 8
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
      const int64_t i = e % ls, j = e / ls;
      dst[i + rows * j] = src[e];
    }
  }
}

// 4. parallel cyclic Jacobi, one workgroup per sector.  n even (padded with
// a zero row/column when odd; the pad never rotates).  Packed upper triangle
// (i <= j at j(j+1)/2 + i).  Dynamic LDS holds, when they fit: the triangle,
// the eigenvector matrix V and the table of 2x2 block pairs of a round;
// otherwise V, then also the triangle, live in HBM (scratch / output).
// A sweep converges when no rotation is significant: |a_pq| above 1e-12 of
// sqrt(a_pp a_qq) (an off-diagonal d moves eigenvalues by ~d^2 / gap, so
// the eigenvalues are exact to ~1e-24 relative) and above 1e-14 of the
// largest diagonal.  Below that floor
// a rotation moves no eigenvalue by more than 1e-14 of the largest (singular
// values below 1e-7 of the largest, under the Gram matrix's own resolution of
// ~sqrt(eps)), and the rounding noise in the null space of a rank-deficient
// sector's Gram matrix never passes the relative test alone.
constexpr int kJacThreads = 1024;
constexpr double kJacFloor = 1e-14;   // convergence floor relative to the largest diagonal
constexpr int kLdsDoubles = 20000;   // dynamic LDS (160 000 B of the CU's 160 KiB)

__device__ inline int pk(int i, int j) { return i <= j ? j * (j + 1) / 2 + i : i * (i + 1) / 2 + j; }

// LDS placement of one sector of size n (doubles): rotation arrays always;
// then the packed triangle, V and the pair table in that priority.  Bit 0 of
// *where: triangle in LDS, bit 1: V in LDS, bit 2: pair table in LDS.
struct JacLayout {
  int64_t rot, tri, v, tab, total;
  int where;
};
__host__ __device__ inline JacLayout jac_layout(int n) {
  const int64_t np = n + (n & 1), npair = np / 2;
  JacLayout L;
  L.rot = 3 * npair;   // c, s (doubles) and the (p, q) pairs (two ints)
  L.tri = np * (np + 1) / 2;
  L.v = (int64_t)n * n;
  L.tab = (npair * (npair + 1) / 2 + 1) / 2;
  L.where = 0;
  int64_t used = L.rot;
  if (used + L.tri <= kLdsDoubles) { used += L.tri; L.where |= 1; }
  if (used + L.v <= kLdsDoubles) { used += L.v; L.where |= 2; }
  if (used + L.tab <= kLdsDoubles) { used += L.tab; L.where |= 4; }
  L.total = used;
  return L;
}

__device__ inline void tri_index(int64_t w, int& i, int& j) {   // w -> (i <= j)
  j = (int)((sqrt(8.0 * (double)w + 1.0) - 1.0) / 2.0);
  while ((int64_t)j * (j + 1) / 2 > w) j--;
  while ((int64_t)(j + 1) * (j + 2) / 2 <= w) j++;
  i = (int)(w - (int64_t)j * (j + 1) / 2);
}

__global__ void __launch_bounds__(kJacThreads)
k_jacobi(const int* __restrict__ nq, const int64_t* __restrict__ goff,
         const double* __restrict__ Gin, double* __restrict__ gscratch,
         const int64_t* __restrict__ poff, double* __restrict__ Vout,
         double* __restrict__ evals, const int64_t* __restrict__ eoff, int max_sweeps,
         int nmax_here, double floor_rel) {
  extern __shared__ double lds_g[];
  __shared__ int any_rot;
  __shared__ double red[kJacThreads / 64];
  const int q = blockIdx.x;
  const int n = nq[q];                 // true size
  if (n == 0 || n > nmax_here) return;   // larger sectors: block Jacobi
  const int np = n + (n & 1);          // padded even size
  const int npair = np / 2;
  const int64_t psize = (int64_t)np * (np + 1) / 2;
  const int nblk = npair * (npair + 1) / 2;
  const JacLayout lay = jac_layout(n);
  double* cs_c = lds_g;
  double* cs_s = lds_g + npair;
  int* pp = (int*)(lds_g + 2 * npair);
  int* qq = pp + npair;
  double* next = lds_g + lay.rot;
  double* g = (lay.where & 1) ? next : gscratch + poff[q];
  if (lay.where & 1) next += psize;
  double* Vg = Vout + goff[q];         // n x n column-major eigenvectors (output)
  double* V = (lay.where & 2) ? next : Vg;
  if (lay.where & 2) next += (int64_t)n * n;
  int* tab = (lay.where & 4) ? (int*)next : nullptr;
  const double* G0 = Gin + goff[q];    // full n x n column-major
  // load packed triangle (pad row/col zero), V = I, block-pair table
#pragma unroll 8
  for (int64_t e = threadIdx.x; e < psize; e += blockDim.x) {
    int i, j;
    tri_index(e, i, j);
    g[e] = (i < n && j < n) ? 0.5 * (G0[i + (int64_t)n * j] + G0[j + (int64_t)n * i]) : 0.0;
  }
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += blockDim.x)
    V[e] = (e % n == e / n) ? 1.0 : 0.0;
  if (tab)
    for (int w = threadIdx.x; w < nblk; w += blockDim.x) {
      int k1, k2;
      tri_index(w, k1, k2);
      tab[w] = k1 | (k2 << 16);
    }
  // largest diagonal element: the scale of the convergence floor
  double dmax = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    dmax = fmax(dmax, fabs(G0[i + (int64_t)n * i]));
  for (int o = 32; o > 0; o >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dmax;
  __syncthreads();
  dmax = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); i++) dmax = fmax(dmax, red[i]);
  const double floor_abs = floor_rel * dmax;
  const int m1 = np - 1;
  for (int sweep = 0; sweep < max_sweeps; sweep++) {
    if (threadIdx.x == 0) any_rot = 0;
    __syncthreads();
    for (int round = 0; round < m1; round++) {
      // circle method: player m1 fixed, pairs (round, m1) and (round+k, round-k)
      for (int k = threadIdx.x; k < npair; k += blockDim.x) {
        int a, b;
        if (k == 0) { a = round; b = m1; }
        else { a = (round + k) % m1; b = (round - k + m1) % m1; }
        if (a > b) { int t = a; a = b; b = t; }
        pp[k] = a;
        qq[k] = b;
        const double app = g[pk(a, a)], aqq = g[pk(b, b)], apq = g[pk(a, b)];
        double c = 1.0, s = 0.0;
        if (apq != 0.0) {
          if (fabs(apq) > 1e-12 * sqrt(fabs(app * aqq)) && fabs(apq) > floor_abs) any_rot = 1;
          const double theta = (aqq - app) / (2.0 * apq);
          const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          c = 1.0 / sqrt(t * t + 1.0);
          s = t * c;
        }
        cs_c[k] = c;
        cs_s[k] = s;
      }
      __syncthreads();
      // G <- J^T G J over 2x2 blocks (k1 <= k2) of the packed triangle
      for (int w = threadIdx.x; w < nblk; w += blockDim.x) {
        int k1, k2;
        if (tab) { k1 = tab[w] & 0xffff; k2 = tab[w] >> 16; }
        else tri_index(w, k1, k2);
        const int a = pp[k1], b = qq[k1], c2i = pp[k2], d = qq[k2];
        const double c1 = cs_c[k1], s1 = cs_s[k1], c2 = cs_c[k2], s2 = cs_s[k2];
        if (k1 == k2) {
          const double gaa = g[pk(a, a)], gab = g[pk(a, b)], gbb = g[pk(b, b)];
          // rows then columns of the symmetric 2x2 block
          const double ra0 = c1 * gaa - s1 * gab, ra1 = c1 * gab - s1 * gbb;
          const double rb0 = s1 * gaa + c1 * gab, rb1 = s1 * gab + c1 * gbb;
          g[pk(a, a)] = c1 * ra0 - s1 * ra1;
          g[pk(a, b)] = s1 * ra0 + c1 * ra1;
          g[pk(b, b)] = s1 * rb0 + c1 * rb1;
        } else {
          const int iac = pk(a, c2i), iad = pk(a, d), ibc = pk(b, c2i), ibd = pk(b, d);
          const double gac = g[iac], gad = g[iad], gbc = g[ibc], gbd = g[ibd];
          const double ra_c = c1 * gac - s1 * gbc, ra_d = c1 * gad - s1 * gbd;
          const double rb_c = s1 * gac + c1 * gbc, rb_d = s1 * gad + c1 * gbd;
          g[iac] = c2 * ra_c - s2 * ra_d;
          g[iad] = s2 * ra_c + c2 * ra_d;
          g[ibc] = c2 * rb_c - s2 * rb_d;
          g[ibd] = s2 * rb_c + c2 * rb_d;
        }
      }
      // V <- V J (columns of real indices only)
      for (int w = threadIdx.x; w < npair * n; w += blockDim.x) {
        const int k = w / n, i = w % n;
        const int a = pp[k], b = qq[k];
        if (b >= n) continue;   // pad never rotates (apq == 0)
        const double c = cs_c[k], s = cs_s[k];
        const double va = V[i + (int64_t)n * a], vb = V[i + (int64_t)n * b];
        V[i + (int64_t)n * a] = c * va - s * vb;
        V[i + (int64_t)n * b] = s * va + c * vb;
      }
      __syncthreads();
    }
    // every thread reads the flag before thread 0 may clear it for the next
    // sweep: the barrier between the read and the reset keeps the decision
    // uniform over the workgroup
    const int rotated = any_rot;
    __syncthreads();
    if (!rotated) break;
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) evals[eoff[q] + i] = g[pk(i, i)];
  if (lay.where & 2)
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += blockDim.x) Vg[e] = V[e];
}

// 4b. Sectors larger than kJacBlockMin: two-sided block Jacobi.  The padded
// Gram matrix (N = n rounded up to 64, zero padding never rotates) is cut into
// 32-wide blocks; a round pairs the blocks (circle method) and
//   k_bjac_solve  diagonalises every 64 x 64 subproblem G[I+J, I+J] in LDS
//                 (one inner cyclic sweep) into Q_IJ, its eigenpairs sorted
//                 in descending order, and
//   k_bjac_apply  updates G <- P^T G P (upper tiles, mirrored: G stays
//                 exactly symmetric) and V <- V P, P = diag(Q_IJ), as 64 x 64
//                 tile products on the FP64 matrix cores,
// so a round moves G and V through memory once instead of once per scalar
// rotation.  A sweep (nbk - 1 rounds) leaves a sector active while any
// subproblem held an off-diagonal element above 1e-12 sqrt(g_ii g_jj) and
// above the sector's floor, 1e-14 of the largest diagonal.  A subproblem
// with nothing to rotate leaves Q_IJ = I and marks its pair unrotated; the
// tile products of a round skip the tiles whose pairs were both unrotated.
constexpr int kJacBlockMin = 64;
constexpr double kBJRel = 1e-12;
constexpr int kBJ = 32;                // block width; subproblems are 2 kBJ
constexpr int kSub = 2 * kBJ;

typedef double bj_v4d __attribute__((ext_vector_type(4)));

struct BJSector {
  int n, N, nbk, q;                    // size, padded size, blocks, sector id
  int64_t gp, qp;                      // offsets of padded G / V and of the pair Q's
  double dmax;                         // largest diagonal of the Gram matrix
  double floor;                        // off-diagonal size below which nothing rotates
  int pb, pad;                         // index of the sector's first pair in the pair list
};

// circle method over m = 2 npair indices: pair k of round r (a < b)
__device__ inline void circle_pair(int m, int r, int k, int& a, int& b) {
  const int m1 = m - 1;
  if (k == 0) { a = r % m1; b = m1; }
  else { a = (r + k) % m1; b = (r - k + m1) % m1; }
  if (a > b) { const int t = a; a = b; b = t; }
}
// subproblem entry t (0..63) of pair k of round r is global index
// (t < kBJ ? a : b) * kBJ + t % kBJ, (a, b) = circle_pair(nbk, r, k)

// sec[li].dmax = max_i |G_ii| of the sector, sec[li].floor = floor_rel * dmax
__global__ void __launch_bounds__(256) k_bjac_floor(BJSector* sec, int nsec, const int64_t* goff, const double* G,
                                                    double floor_rel) {
  __shared__ double red[4];
  const int li = blockIdx.x;
  if (li >= nsec) return;
  const int q = sec[li].q, n = sec[li].n;
  const double* g = G + goff[q];
  double mx = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) mx = fmax(mx, fabs(g[i + (int64_t)n * i]));
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    sec[li].dmax = m;
    sec[li].floor = floor_rel * m;
  }
}

__global__ void k_bjac_init(const BJSector* sec, int nsec, const int64_t* goff, const double* G,
                            double* Gp, double* Vp) {
  const int li = blockIdx.y;
  if (li >= nsec) return;
  const BJSector S = sec[li];
  const double* g = G + goff[S.q];
  const int64_t NN = (int64_t)S.N * S.N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < NN;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % S.N), j = (int)(e / S.N);
    Gp[S.gp + e] = (i < S.n && j < S.n) ? 0.5 * (g[i + (int64_t)S.n * j] + g[j + (int64_t)S.n * i]) : 0.0;
    Vp[S.gp + e] = i == j ? 1.0 : 0.0;
  }
}

// Jacobi rotation zeroing a_pq of [[app, apq], [apq, aqq]]: rows/columns
// (p, q) <- (c x_p - s x_q, s x_p + c x_q), t = s/c the smaller root of
// t^2 + 2 theta t - 1 = 0, theta = (aqq - app) / (2 apq).  Reciprocal and
// reciprocal square root from the hardware estimates with two Newton steps
// each (full double precision; c^2 + s^2 = 1 to rounding).
__device__ inline double rcp_nr(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  return fma(r, e, r);
}
__device__ inline double rsq_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}
__device__ inline void jac_rotation(double app, double aqq, double apq, double& c, double& s) {
  c = 1.0;
  s = 0.0;
  if (!(fabs(apq) > 1e-290)) return;   // nothing to rotate (and no overflow below)
  const double theta = (aqq - app) * (0.5 * rcp_nr(apq));
  const double at = fabs(theta);
  double t;
  if (at < 1e150) {
    const double h = fma(theta, theta, 1.0);
    t = rcp_nr(at + h * rsq_nr(h));
  } else {
    t = 0.5 * rcp_nr(at);
  }
  if (theta < 0) t = -t;
  c = rsq_nr(fma(t, t, 1.0));
  s = t * c;
}

// one workgroup per (sector, pair): the 64 x 64 subproblem, one cyclic
// sweep of 63 rounds of 32 disjoint rotations (circle method), in slot
// order: slot k holds the pair of round r at positions (2k, 2k+1), so every
// rotation acts on adjacent rows and columns.  Between rounds the indices
// move by the fixed permutation slot_next (period 63); a round reads S and
// Q^T from one buffer and writes the rotated values to their next-round
// positions in the other.  A round, one barrier:
//   1. every wave: lane k < 32 computes the rotation of slot k from its 2x2
//      diagonal block (the four waves compute the same values from the same
//      LDS words), and each thread takes the rotations it needs from those
//      lanes by cross-lane shuffles, so no rotation table passes through LDS;
//   2. all threads: S <- M S M^T (thread t: rows 2k, 2k+1, k = t/8, and the
//      column pairs k' = t%8 + 8u) and Q^T <- M Q^T (the same rows).
constexpr int kSP = 72;                    // LDS row pitch (doubles) of S and Q^T
constexpr int kSlotMat = kSub * kSP;
constexpr size_t kSolveLds = 4 * kSlotMat * sizeof(double) + (2 * kSub + 1) * sizeof(int);
__device__ inline int slot_of(int x) { return x < kBJ ? 2 * x : 2 * (kSub - 1 - x) + 1; }
__device__ inline int slot_next(int p) {
  const int k = p >> 1;
  if (!(p & 1)) return k ? 2 * (k - 1) : 3;
  if (k == 0) return 1;
  return k < kBJ - 1 ? 2 * k + 3 : 2 * kBJ - 2;
}

constexpr int kSolveThreads = 256;
__global__ void __launch_bounds__(kSolveThreads)
k_bjac_solve(const BJSector* sec, const int2* items, int nitems, int r, const int* active, int* flag,
             int* rotated, const double* Gp, double* Qp) {
  extern __shared__ __attribute__((aligned(16))) double bj_lds[];
  int* perm = (int*)(bj_lds + 4 * kSlotMat);
  int* cnt = perm + kSub;
  int* sig = cnt + kSub;
  const int2 it = items[blockIdx.x];
  const int li = it.x, k = it.y;
  const BJSector S = sec[li];
  if (!active[li] || r >= S.nbk - 1) return;
  const int tid = threadIdx.x;
  // the pair's two blocks (uniform): subproblem entry t is global index
  // (t < 32 ? ba : bb) * 32 + t % 32; the sector's own indices (global index
  // < n) are entries [0, na) and [32, 32 + nbb)
  int ba, bb;
  circle_pair(S.nbk, r, k, ba, bb);
  const int na = min(max(S.n - kBJ * ba, 0), kBJ), nbb = min(max(S.n - kBJ * bb, 0), kBJ);
  auto gidx = [&](int t) { return (t >= kBJ ? bb : ba) * kBJ + (t & (kBJ - 1)); };
  auto own = [&](int t) { return t < kBJ ? t < na : t - kBJ < nbb; };
  if (tid == 0) *sig = 0;
  if (tid < kSub) cnt[tid] = 0;
  {
    // all of a thread's loads are issued before the first LDS store (a loop
    // with one load per iteration would wait one memory latency per entry)
    constexpr int kPer = kSub * kSub / kSolveThreads;
    double* S0 = bj_lds;
    double* Q0 = bj_lds + 2 * kSlotMat;
    double gv[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int e = tid + kSolveThreads * u, i = e % kSub, j = e / kSub;   // G is exactly symmetric
      gv[u] = Gp[S.gp + gidx(i) + (int64_t)S.N * gidx(j)];
    }
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int e = tid + kSolveThreads * u, i = e % kSub, j = e / kSub;
      S0[slot_of(i) * kSP + slot_of(j)] = gv[u];
      Q0[slot_of(j) * kSP + i] = i == j ? 1.0 : 0.0;
    }
  }
  __syncthreads();
  {
    // significance test of this subproblem (before rotating):
    // |v| > 1e-12 sqrt(g_pp g_qq), compared squared; thread t's rows p are
    // all t % 64 (so one diagonal load), its columns q = t / 64 + 4 i
    const double* S0 = bj_lds;
    const int p = tid % kSub;
    const double gpp = S0[p * kSP + p];
    bool any = false;
    for (int q = tid / kSub; q < kSub; q += kSolveThreads / kSub) {
      if (p == q) continue;
      const double v = S0[p * kSP + q];
      any |= v * v > (kBJRel * kBJRel) * fabs(gpp * S0[q * kSP + q]) && fabs(v) > S.floor;
    }
    if (any) *sig = 1;
  }
  __syncthreads();
  double* Qg = Qp + S.qp + (int64_t)k * kSub * kSub;   // column-major 64 x 64
  if (!*sig) {   // already diagonal: Q = I
    for (int e = tid; e < kSub * kSub; e += kSolveThreads) Qg[e] = (e % kSub == e / kSub) ? 1.0 : 0.0;
    if (tid == 0) rotated[blockIdx.x] = 0;
    return;
  }
  if (tid == 0) {
    atomicOr(&flag[li], 1);
    rotated[blockIdx.x] = 1;
  }
  typedef double d2 __attribute__((ext_vector_type(2)));
  const int kk = tid >> 3, c8 = tid & 7, lane = tid & 63;
  const int r0 = 2 * kk, r1 = 2 * kk + 1;
  const int nr0 = slot_next(r0) * kSP, nr1 = slot_next(r1) * kSP;
  int ncol[4][2];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    ncol[u][0] = slot_next(2 * c8 + 16 * u);
    ncol[u][1] = slot_next(2 * c8 + 16 * u + 1);
  }
  for (int round = 0; round < kSub - 1; round++) {
    const double* Sc = bj_lds + (round & 1) * kSlotMat;
    double* Sn = bj_lds + ((round & 1) ^ 1) * kSlotMat;
    const double* Qc = bj_lds + (2 + (round & 1)) * kSlotMat;
    double* Qn = bj_lds + (2 + ((round & 1) ^ 1)) * kSlotMat;
    // lane k (k < 32) of every wave: the rotation of slot k from its 2x2
    // diagonal block; lanes 32..63 compute slot k - 32 again (same words, no
    // divergence).  The block's three words are read before the round's S and
    // Q^T rows (the scheduling barriers keep that order), so the rotation's
    // arithmetic runs while those 16 reads stream in.
    d2 x0[4], x1[4], q0[4], q1[4], cp[4];
    const int pd = 2 * (lane & (kBJ - 1));
    const double app = Sc[pd * kSP + pd], aqq = Sc[(pd + 1) * kSP + pd + 1], apq = Sc[pd * kSP + pd + 1];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      x0[u] = *(const d2*)&Sc[r0 * kSP + 2 * c8 + 16 * u];
      x1[u] = *(const d2*)&Sc[r1 * kSP + 2 * c8 + 16 * u];
      q0[u] = *(const d2*)&Qc[r0 * kSP + 2 * c8 + 16 * u];
      q1[u] = *(const d2*)&Qc[r1 * kSP + 2 * c8 + 16 * u];
    }
    __builtin_amdgcn_sched_barrier(0);
    double cl, sl;
    jac_rotation(app, aqq, apq, cl, sl);
#pragma unroll
    for (int u = 0; u < 4; u++) cp[u] = d2{__shfl(cl, c8 + 8 * u), __shfl(sl, c8 + 8 * u)};
    const double c1 = __shfl(cl, kk), s1 = __shfl(sl, kk);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      // rows 2k, 2k+1 of the 2x2 block (k, k' = c8 + 8u), then its columns
      const double y00 = c1 * x0[u].x - s1 * x1[u].x, y01 = c1 * x0[u].y - s1 * x1[u].y;
      const double y10 = s1 * x0[u].x + c1 * x1[u].x, y11 = s1 * x0[u].y + c1 * x1[u].y;
      const double c2 = cp[u].x, s2 = cp[u].y;
      double z00 = c2 * y00 - s2 * y01, z01 = s2 * y00 + c2 * y01;
      double z10 = c2 * y10 - s2 * y11, z11 = s2 * y10 + c2 * y11;
      if (c8 + 8 * u == kk) { z01 = 0.0; z10 = 0.0; }   // the pair's own element is rotated away
      Sn[nr0 + ncol[u][0]] = z00;
      Sn[nr0 + ncol[u][1]] = z01;
      Sn[nr1 + ncol[u][0]] = z10;
      Sn[nr1 + ncol[u][1]] = z11;
      *(d2*)&Qn[nr0 + 2 * c8 + 16 * u] = c1 * q0[u] - s1 * q1[u];
      *(d2*)&Qn[nr1 + 2 * c8 + 16 * u] = s1 * q0[u] + c1 * q1[u];
    }
    __syncthreads();
  }
  // 63 rounds return every index to its first slot; the result is in buffer 1.
  // Eigenpairs in descending order (the larger values to the pair's lower
  // block): a fixed order keeps values from swapping between blocks from
  // round to round, which otherwise holds the outer iteration at a linear rate.
  // Only the sector's own indices are ordered, over their own positions; a
  // padding index (global index >= n, never rotated) keeps its column, so no
  // eigenvector moves into the padding that k_bjac_finish drops.
  const double* Sf = bj_lds + kSlotMat;
  const double* Qf = bj_lds + 3 * kSlotMat;
  {
    // rank of own entry x among the own entries (descending eigenvalue,
    // ties to the lower entry), counted by the four waves over a quarter of
    // the entries each
    const int x = tid % kSub, part = tid / kSub;
    if (own(x)) {
      const double d = Sf[slot_of(x) * kSP + slot_of(x)];
      int c = 0;
      for (int u = 16 * part; u < 16 * part + 16; u++) {
        if (!own(u)) continue;
        const int pu = slot_of(u);
        const double du = Sf[pu * kSP + pu];
        c += (du > d) || (du == d && u < x);
      }
      if (c) atomicAdd(&cnt[x], c);
    }
  }
  __syncthreads();
  if (tid < kSub) {
    // the rank-th own entry in ascending order: own entries are [0, na) then
    // [32, 32 + nbb)
    const int x = tid;
    int dst = x;
    if (own(x)) {
      const int rk = cnt[x];
      dst = rk < na ? rk : kBJ + (rk - na);
    }
    perm[slot_of(x)] = dst;   // column of Q for the eigenvector in slot slot_of(x)
  }
  __syncthreads();
  for (int e = tid; e < kSub * kSub; e += kSolveThreads) {
    const int i = e % kSub, p = e / kSub;
    Qg[i + kSub * perm[p]] = Qf[p * kSP + i];
  }
}

// tile products of a round on v_mfma_f64_16x16x4 (lane l holds A[l%16][l/16],
// B[l/16][l%16]; result register r holds D[l/16 + 4r][l%16]).  Wave w
// computes rows 16w..16w+15 of the 64 x 64 result; the operand all waves
// share is staged in LDS, the per-wave operand is read from memory.
//   kind 0 (a <= b): X = Qa^T T, R = X Qb for T = G[Ia, Ib]; R written to
//                    G[Ia, Ib] and mirrored into G[Ib, Ia];
//   kind 1: R^T = Qb^T T^T for T = V[rows 64a..64a+63, Ib]: V <- V Qb, the
//           transposed product so that lanes store down V's columns.
constexpr int kTLd = kSub + 1;
__global__ void __launch_bounds__(256) k_bjac_apply(const BJSector* sec, const int4* items, int nitems,
                                                    int r, const int* active, const int* rotated,
                                                    double* Gp, double* Vp, const double* Qp) {
  __shared__ double Bs[kSub * kTLd];
  __shared__ double Xs[kSub * kTLd];
  const int4 it = items[blockIdx.x];
  const int li = it.x;
  const BJSector S = sec[li];
  if (!active[li] || r >= S.nbk - 1) return;
  const int kind = it.w, a = it.y, b = it.z;
  const bool rb = rotated[S.pb + b] != 0;
  const bool ra = kind == 0 && rotated[S.pb + a] != 0;
  if (!ra && !rb) return;   // both factors are the identity
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int l16 = lane & 15, l4 = lane >> 4;
  // global indices of the entries of pairs a and b (their blocks computed once)
  int a0, a1, b0, b1;
  circle_pair(S.nbk, r, a, a0, a1);
  circle_pair(S.nbk, r, b, b0, b1);
  auto ga = [&](int t) { return (t >= kBJ ? a1 : a0) * kBJ + (t & (kBJ - 1)); };
  auto gb = [&](int t) { return (t >= kBJ ? b1 : b0) * kBJ + (t & (kBJ - 1)); };
  const double* qa = Qp + S.qp + (int64_t)a * kSub * kSub;
  const double* qb = Qp + S.qp + (int64_t)b * kSub * kSub;
  double af[16];
  bj_v4d acc[4];
  if (kind == 1) {
    double* V = Vp + S.gp;
#pragma unroll
    for (int s = 0; s < 16; s++) af[s] = qb[(4 * s + l4) + kSub * (16 * w + l16)];   // Qb^T
    double tv[16];   // Bs[k][ii] = V(64a + ii, gj(k)): all loads issued, then the LDS stores
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int e = tid + 256 * u, ii = e % kSub, kk = e / kSub;
      tv[u] = V[(int64_t)(a * kSub + ii) + (int64_t)S.N * gb(kk)];
    }
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int e = tid + 256 * u, ii = e % kSub, kk = e / kSub;
      Bs[kk * kTLd + ii] = tv[u];
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < 4; f++) acc[f] = bj_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 16; s++)
#pragma unroll
      for (int f = 0; f < 4; f++)
        acc[f] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], Bs[(4 * s + l4) * kTLd + 16 * f + l16], acc[f],
                                                      0, 0, 0);
#pragma unroll
    for (int rr = 0; rr < 4; rr++) {
      const int64_t col = (int64_t)S.N * gb(16 * w + l4 + 4 * rr);
#pragma unroll
      for (int f = 0; f < 4; f++) V[col + a * kSub + 16 * f + l16] = acc[f][rr];
    }
    return;
  }
  double* G = Gp + S.gp;
#pragma unroll
  for (int s = 0; s < 16; s++) af[s] = qa[(4 * s + l4) + kSub * (16 * w + l16)];   // Qa^T
  {
    double tv[16], xv[16];   // all loads issued, then the LDS stores
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int e = tid + 256 * u, kk = e % kSub, j = e / kSub;
      tv[u] = G[ga(kk) + (int64_t)S.N * gb(j)];
      xv[u] = qb[kk + kSub * j];
    }
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int e = tid + 256 * u, kk = e % kSub, j = e / kSub;
      Bs[kk * kTLd + j] = tv[u];
      Xs[kk * kTLd + j] = xv[u];
    }
  }
  __syncthreads();
#pragma unroll
  for (int f = 0; f < 4; f++) acc[f] = bj_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 16; s++)
#pragma unroll
    for (int f = 0; f < 4; f++)
      acc[f] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], Bs[(4 * s + l4) * kTLd + 16 * f + l16], acc[f],
                                                    0, 0, 0);
  __syncthreads();   // every wave is done reading T
#pragma unroll
  for (int f = 0; f < 4; f++)
#pragma unroll
    for (int rr = 0; rr < 4; rr++) Bs[(16 * w + l4 + 4 * rr) * kTLd + 16 * f + l16] = acc[f][rr];
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 16; s++) af[s] = Bs[(16 * w + l16) * kTLd + 4 * s + l4];   // X rows of wave w
#pragma unroll
  for (int f = 0; f < 4; f++) acc[f] = bj_v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 16; s++)
#pragma unroll
    for (int f = 0; f < 4; f++)
      acc[f] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[s], Xs[(4 * s + l4) * kTLd + 16 * f + l16], acc[f],
                                                    0, 0, 0);
#pragma unroll
  for (int rr = 0; rr < 4; rr++) {
    const int gi = ga(16 * w + l4 + 4 * rr);
#pragma unroll
    for (int f = 0; f < 4; f++) {
      const int gj = gb(16 * f + l16);
      G[gi + (int64_t)S.N * gj] = acc[f][rr];
      if (a != b) G[gj + (int64_t)S.N * gi] = acc[f][rr];   // G stays symmetric
    }
  }
}

// end of an outer sweep: a sector stays active while its sweep rotated
// anything; the flags start over for the next sweep; *any = some sector active
__global__ void __launch_bounds__(256) k_bjac_sweep_end(int nsec, int* flag, int* act, int* any) {
  __shared__ int s_any;
  if (threadIdx.x == 0) s_any = 0;
  __syncthreads();
  int a = 0;
  for (int li = threadIdx.x; li < nsec; li += blockDim.x) {
    const int f = flag[li];
    act[li] = f;
    flag[li] = 0;
    a |= f;
  }
  if (a) atomicOr(&s_any, 1);
  __syncthreads();
  if (threadIdx.x == 0) *any = s_any;
}

__global__ void k_bjac_finish(const BJSector* sec, int nsec, const int64_t* goff,
                              const int64_t* eoff, const double* Gp, const double* Vp,
                              double* Vout, double* evals) {
  const int li = blockIdx.y;
  if (li >= nsec) return;
  const BJSector S = sec[li];
  const int64_t nn = (int64_t)S.n * S.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nn;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % S.n), j = (int)(e / S.n);
    Vout[goff[S.q] + e] = Vp[S.gp + i + (int64_t)S.N * j];
    if (i == j) evals[eoff[S.q] + i] = Gp[S.gp + i + (int64_t)S.N * i];
  }
}

// host driver of the block Jacobi over the sectors `big` (sizes nq)
static int block_jacobi(const std::vector<int>& big, const std::vector<int>& nq, const int64_t* d_goff,
                        const int64_t* d_eoff, const double* G, double* Vev, double* ev,
                        hipStream_t s, int max_sweeps, double floor_rel) {
  PhaseTimer _phase("block_jacobi");
  const int ns = (int)big.size();
  std::vector<BJSector> sec(ns);
  std::vector<int2> pitems;
  std::vector<int4> titems;
  int64_t gp = 0, qp = 0;
  int maxr = 0;
  for (int li = 0; li < ns; li++) {
    const int q = big[li], n = nq[q];
    const int N = (n + kSub - 1) / kSub * kSub;
    const int nbk = N / kBJ, npair = nbk / 2;
    sec[li] = BJSector{n, N, nbk, q, gp, qp, 0.0, 0.0, (int)pitems.size(), 0};
    gp += (int64_t)N * N;
    qp += (int64_t)npair * kSub * kSub;
    maxr = std::max(maxr, nbk - 1);
    for (int k = 0; k < npair; k++) pitems.push_back(int2{li, k});
    for (int a = 0; a < npair; a++)
      for (int b = a; b < npair; b++) titems.push_back(int4{li, a, b, 0});   // a <= b, mirrored
    for (int rc = 0; rc < N / kSub; rc++)
      for (int b = 0; b < npair; b++) titems.push_back(int4{li, rc, b, 1});
  }
  Scratch sc;
  BJSector* d_sec = sc.get<BJSector>(ns);
  int2* d_pit = sc.get<int2>(pitems.size());
  int4* d_tit = sc.get<int4>(titems.size());
  double* Gp = sc.get<double>(gp);
  double* Vp = sc.get<double>(gp);
  double* Qp = sc.get<double>(qp);
  int* act = sc.get<int>(ns);
  int* flag = sc.get<int>(ns);
  int* rot = sc.get<int>(pitems.size());
  int* anyd = sc.get<int>(1);
  TTG_TRY(sc.ok("svd: block Jacobi"));
  std::vector<int> hact(ns, 1);
  TTG_TRY(h2d_vec(d_sec, sec));
  TTG_TRY(h2d_vec(d_pit, pitems));
  TTG_TRY(h2d_vec(d_tit, titems));
  TTG_TRY(h2d_vec(act, hact));
  TTG_HIP(hipFuncSetAttribute((const void*)k_bjac_solve, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kSolveLds));
  // largest diagonal per sector and the convergence floor, on the device
  if (floor_rel > 0)
    hipLaunchKernelGGL(k_bjac_floor, dim3(ns), dim3(256), 0, s, d_sec, ns, d_goff, G, floor_rel);
  hipLaunchKernelGGL(k_bjac_init, dim3(64, ns), dim3(256), 0, s, d_sec, ns, d_goff, G, Gp, Vp);
  // Outer sweeps.  The active flags move on the device (k_bjac_sweep_end),
  // so the host reads the "any sector active" word only from the sweep at
  // which the previous call of the same kind (LR-reduced or plain Gram
  // matrices) converged, less two: a sweep after convergence finds every
  // sector inactive and its launches return at once.
  TTG_HIP(hipMemsetAsync(flag, 0, sizeof(int) * ns, s));
  static int prev_sweeps[2] = {0, 0};
  int& prev = prev_sweeps[floor_rel > 0 ? 1 : 0];
  const int first_check = std::max(0, prev - 2);
  int sweep = 0;
  for (; sweep < max_sweeps; sweep++) {
    for (int r = 0; r < maxr; r++) {
      hipLaunchKernelGGL(k_bjac_solve, dim3((unsigned)pitems.size()), dim3(kSolveThreads), kSolveLds, s, d_sec, d_pit,
                         (int)pitems.size(), r, act, flag, rot, Gp, Qp);
      hipLaunchKernelGGL(k_bjac_apply, dim3((unsigned)titems.size()), dim3(256), 0, s, d_sec, d_tit,
                         (int)titems.size(), r, act, rot, Gp, Vp, Qp);
    }
    hipLaunchKernelGGL(k_bjac_sweep_end, dim3(1), dim3(256), 0, s, ns, flag, act, anyd);
    if (sweep < first_check && sweep + 1 < max_sweeps) continue;
    int any = 0;
    const D2HPiece pc{&any, anyd, sizeof(int)};
    TTG_TRY(d2h_wait(&pc, 1));
    if (!any) break;
  }
  prev = std::min(sweep, 16);
  hipLaunchKernelGGL(k_bjac_finish, dim3(64, ns), dim3(256), 0, s, d_sec, ns, d_goff, d_eoff, Gp, Vp,
                     Vev, ev);
  TTG_HIP(hipGetLastError());
  return 0;   // the scratch goes back to the pool in stream order
}

// 4c. LR preconditioning of the large sectors (cutoff > 0).  A sector's
// Gram matrix of a DMRG state is numerically rank-deficient (a few hundred
// of its rows carry weight), and the Jacobi iteration on G spends most of
// its sweeps decoupling that null space from the rest.  A pivoted Cholesky
// factorisation P^T G P = R^T R stops at the numerical rank r (largest
// remaining diagonal below 1e-15 of the largest diagonal) and gives
// G' = R R^T (r x r, the same nonzero eigenvalues, no null space).  Jacobi
// on G' with the relative convergence test alone (no absolute floor: G' is
// positive definite) converges in about a third of the sweeps, at size r,
// and G's eigenvectors follow as V = R^T W Lambda^(-1/2), orthonormal to the
// relative accuracy of W.  The n - r null-space columns are left zero with
// eigenvalue 0: a zero singular value is never kept when cutoff > 0
// (qtensor_op.cpp:1294-1300).
// R is stored by original column index (row k of R at k n + j), so neither
// G' nor V needs the permutation.
//   k_pchol_init   d = diag(G), no column pivoted yet, dmax per sector
//   k_pchol_panel  one workgroup per sector: nb pivot steps (largest
//                  remaining diagonal, ties to the lower index), each row of R
//                  from the trailing matrix minus the panel's earlier rows
//                  (left-looking inside the panel, rows held in LDS)
//   k_pchol_update trailing matrix G <- G - R_panel^T R_panel (rank nb)
constexpr double kPcTol = 1e-15;
constexpr int kPcThreads = 1024;
struct PcSector {
  int n, q;
  int64_t goff;   // G (trailing matrix, overwritten) and R, n x n each
  int64_t doff;   // diagonal / pivot flags, n each
};

__global__ void __launch_bounds__(kPcThreads) k_pchol_init(const PcSector* ps, const double* G, double* d,
                                                           int* piv, int* rank, double* dmax) {
  __shared__ double red[kPcThreads / 64];
  const PcSector P = ps[blockIdx.x];
  double mx = 0.0;
  for (int j = threadIdx.x; j < P.n; j += blockDim.x) {
    const double v = G[P.goff + j + (int64_t)P.n * j];
    d[P.doff + j] = v;
    piv[P.doff + j] = 0;
    mx = fmax(mx, v);
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = fmax(m, red[w]);
    dmax[blockIdx.x] = m;
    rank[blockIdx.x] = -1;
  }
}

// dynamic LDS: nb panel rows (nb x n), d (n), pivot flags (n ints)
__global__ void __launch_bounds__(kPcThreads) k_pchol_panel(const PcSector* ps, int k0, int nb,
                                                            const double* A, double* R, double* dg,
                                                            int* pivg, int* rank, const double* dmax) {
  extern __shared__ __attribute__((aligned(16))) double pc_lds[];
  __shared__ double rv[kPcThreads / 64];
  __shared__ int ri[kPcThreads / 64];
  __shared__ int s_p, s_stop;
  __shared__ double s_rkk;
  const PcSector P = ps[blockIdx.x];
  const int n = P.n, tid = threadIdx.x;
  if (k0 >= n || rank[blockIdx.x] >= 0) return;   // finished sector
  double* Rp = pc_lds;                 // panel row s at s * n
  double* d = pc_lds + (int64_t)nb * n;
  int* piv = (int*)(d + n);
  for (int j = tid; j < n; j += blockDim.x) {
    d[j] = dg[P.doff + j];
    piv[j] = pivg[P.doff + j];
  }
  __syncthreads();
  const double tol = kPcTol * dmax[blockIdx.x];
  const int kend = min(k0 + nb, n);
  int k = k0;
  for (; k < kend; k++) {
    double bv = -1.0;
    int bi = n;
    for (int j = tid; j < n; j += blockDim.x)
      if (!piv[j] && (d[j] > bv || (d[j] == bv && j < bi))) { bv = d[j]; bi = j; }
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o);
      const int oi = __shfl_xor(bi, o);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if ((tid & 63) == 0) { rv[tid >> 6] = bv; ri[tid >> 6] = bi; }
    __syncthreads();
    if (tid < 64) {   // the waves' candidates, reduced by wave 0 (same order rule)
      constexpr int kW = kPcThreads / 64;
      double v = tid < kW ? rv[tid] : -1.0;
      int i = tid < kW ? ri[tid] : n;
      for (int o = kW / 2; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o);
        const int oi = __shfl_xor(i, o);
        if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
      }
      if (tid == 0) {
        s_stop = !(v > tol) || i >= n;
        s_p = i;
        s_rkk = s_stop ? 0.0 : sqrt(v);
        if (!s_stop) piv[i] = 1;
      }
    }
    __syncthreads();
    if (s_stop) break;
    const int p = s_p, ks = k - k0;
    const double rkk = s_rkk;
    const double irkk = 1.0 / rkk;
    for (int j = tid; j < n; j += blockDim.x) {
      double v;
      if (j == p) {
        v = rkk;
      } else if (piv[j]) {
        v = 0.0;
      } else {
        v = A[P.goff + j + (int64_t)n * p];   // trailing matrix, column p (symmetric)
#pragma unroll 8
        for (int t = 0; t < ks; t++) v -= Rp[t * n + p] * Rp[t * n + j];
        v *= irkk;
        d[j] -= v * v;
      }
      Rp[ks * n + j] = v;
      R[P.goff + (int64_t)k * n + j] = v;
    }
    __syncthreads();
  }
  for (int j = tid; j < n; j += blockDim.x) {
    dg[P.doff + j] = d[j];
    pivg[P.doff + j] = piv[j];
  }
  if (tid == 0 && k < kend) rank[blockIdx.x] = k;   // stopped inside this panel
}

// G[i, j] -= sum_{k in panel rows} R[k, i] R[k, j], 64 x 64 tiles, 256 threads
// (launched with 256).  The panel rows' two 64-column pieces are staged in LDS
// with one batch of loads (a loop over k with a global load per step would
// wait one memory latency per k); the sums run over k in ascending order.
constexpr int kPcNbMax = 32;
__global__ void __launch_bounds__(256) k_pchol_update(const PcSector* ps, int k0, int nb, const int* rank,
                                                      double* A, const double* R) {
  __shared__ double Ri[kPcNbMax][64];
  __shared__ double Rj[kPcNbMax][64];
  const PcSector P = ps[blockIdx.y];
  const int n = P.n;
  const int nt = (n + 63) / 64;
  if ((int)blockIdx.x >= nt * nt || k0 >= n) return;
  const int rk = rank[blockIdx.y];
  const int k1 = rk >= 0 ? min(rk, min(k0 + nb, n)) : min(k0 + nb, n);
  if (k1 <= k0) return;   // uniform over the workgroup
  const int nk = k1 - k0;   // <= nb <= kPcNbMax (the host caps nb)
  const int i0 = (blockIdx.x % nt) * 64, j0 = (blockIdx.x / nt) * 64;
  {
    constexpr int kPer = kPcNbMax * 64 / 256;   // entries per thread (256 threads)
    double vi[kPer], vj[kPer];
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int e = threadIdx.x + 256 * u, kk = e >> 6, c = e & 63;
      const double* rk_ = R + P.goff + (int64_t)(k0 + kk) * n;
      vi[u] = (kk < nk && i0 + c < n) ? rk_[i0 + c] : 0.0;
      vj[u] = (kk < nk && j0 + c < n) ? rk_[j0 + c] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPer; u++) {
      const int e = threadIdx.x + 256 * u, kk = e >> 6, c = e & 63;
      Ri[kk][c] = vi[u];
      Rj[kk][c] = vj[u];
    }
  }
  __syncthreads();
  const int il = threadIdx.x & 63, jl = (threadIdx.x >> 6) * 16;
  const int i = i0 + il, jb = j0 + jl;
  if (i >= n) return;
  double acc[16];
#pragma unroll
  for (int u = 0; u < 16; u++) acc[u] = 0.0;
  for (int kk = 0; kk < nk; kk++) {
    const double ri_ = Ri[kk][il];
#pragma unroll
    for (int u = 0; u < 16; u++) acc[u] += ri_ * Rj[kk][jl + u];
  }
#pragma unroll
  for (int u = 0; u < 16; u++)
    if (jb + u < n) A[P.goff + i + (int64_t)n * (jb + u)] -= acc[u];
}

// V columns c < r: (R^T W)_c / sqrt(lambda_c), eigenvalue lambda_c; columns
// c >= r: zero, eigenvalue 0
__global__ void k_lr_finish(const PcSector* ps, const int* rk, const int64_t* eoff2, const double* ev2,
                            const int64_t* eoff, double* Vev, double* ev) {
  const PcSector P = ps[blockIdx.y];
  const int n = P.n, r = rk[blockIdx.y];
  const int c = blockIdx.x;
  if (c >= n) return;
  const double lam = c < r ? ev2[eoff2[blockIdx.y] + c] : 0.0;
  const double sc = (c < r && lam > 0.0) ? 1.0 / sqrt(lam) : 0.0;
  double* col = Vev + P.goff + (int64_t)n * c;
  for (int i = threadIdx.x; i < n; i += blockDim.x) col[i] = c < r ? col[i] * sc : 0.0;
  if (threadIdx.x == 0) ev[eoff[P.q] + c] = lam;
}

// 5a. per sector: eigen order (descending), top-k singular values in the
// reference's sector / descending order, selected eigenvectors packed
// ordering key of a value: NaN compares as -inf, so that rank counts over
// any input form a permutation
__device__ inline double rank_key(double v) { return v == v ? v : -INFINITY; }

// every sector's Gram diagonal finite (a Gram matrix's entries are bounded by
// its diagonal, and a NaN input reaches the diagonal); *bad = 1 otherwise
__global__ void __launch_bounds__(256) k_svd_finite(int nQ, const int* nq, const int64_t* goff,
                                                    const double* G, int* bad) {
  const int q = blockIdx.x;
  if (q >= nQ) return;
  const int n = nq[q];
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    if (!isfinite(G[goff[q] + i + (int64_t)n * i])) *bad = 1;
}

// 5a. per sector: the rank of each eigenvalue (descending, ties by index; a
// NaN ranks as -inf, so the ranks are a permutation for any input), the kept
// singular values in rank order and the eigenpair index of each kept rank.
// One thread per eigenvalue (grid: sectors x 256-value groups); the sector's
// eigenvalues pass through LDS in chunks, so every comparison reads one
// broadcast LDS value.
constexpr int kRankChunk = 2048;
__global__ void __launch_bounds__(256) k_svd_select(const int* nq, const int* kq,
                                                    const int64_t* eoff, const double* evals,
                                                    const int64_t* soff, double* svals,
                                                    int* ssec, int* selidx) {
  __shared__ double chunk[kRankChunk];
  const int q = blockIdx.x;
  const int n = nq[q], k = kq[q];
  if ((int)blockIdx.y * 256 >= n) return;   // the whole workgroup: before any barrier
  const int i = blockIdx.y * 256 + threadIdx.x;
  const double* ev = evals + eoff[q];
  const double v = i < n ? rank_key(ev[i]) : 0.0;
  int r = 0;
  for (int c0 = 0; c0 < n; c0 += kRankChunk) {
    const int cn = min(kRankChunk, n - c0);
    __syncthreads();
#pragma unroll 8
    for (int e = threadIdx.x; e < cn; e += 256) chunk[e] = rank_key(ev[c0 + e]);
    __syncthreads();
    for (int j = 0; j < cn; j++) {
      const double u = chunk[j];
      r += (u > v) || (u == v && c0 + j < i);
    }
  }
  if (i < n && r < k) {
    svals[soff[q] + r] = sqrt(fmax(v, 0.0));
    ssec[soff[q] + r] = q;
    selidx[soff[q] + r] = i;
  }
}

// the selected eigenvectors packed in rank order, Esel_q(:, r) = V_q(:, selidx_q[r]):
// workgroups over (sector, columns), consecutive threads on consecutive rows
__global__ void __launch_bounds__(256) k_svd_gather_sel(const int* nq, const int* kq,
                                                        const int64_t* goff, const double* V,
                                                        const int64_t* soff, const int* selidx,
                                                        const int64_t* selo, double* Esel) {
  const int q = blockIdx.x;
  const int n = nq[q], k = kq[q];
  for (int r = blockIdx.y; r < k; r += gridDim.y) {
    const double* src = V + goff[q] + (int64_t)n * selidx[soff[q] + r];
    double* dst = Esel + selo[q] + (int64_t)n * r;
    for (int row = threadIdx.x; row < n; row += blockDim.x) dst[row] = src[row];
  }
}

// 5b. truncation (qtensor_op.cpp:1280-1301): nm and cumsum are summed
// sequentially in the reference's order; the global order is a stable
// descending sort (the reference's std::sort is unstable on ties).
// 5b. global descending order of all singular values (ties by position,
// a stable sort as the reference's pairCompare sort over (sector, desc) order).
// Each sector's values are already non-increasing (k_svd_select), so the rank
// of value v at position j of sector q is j, plus the values >= v of the
// sectors before q (they come first on ties), plus the values > v of the
// sectors after it: one binary search per other sector instead of a pass
// over all values.
__device__ inline int count_ge(const double* a, int n, double v) {   // a non-increasing
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] >= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ inline int count_gt(const double* a, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] > v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__global__ void __launch_bounds__(256) k_svd_rank(int T, int nQ, const int64_t* __restrict__ soff,
                                                  const double* __restrict__ svals, int* __restrict__ order,
                                                  double* __restrict__ sorted) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T) return;
  const double v = svals[i];
  int lo = 0, hi = nQ;   // sector q of i: the last q with soff[q] <= i
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (soff[mid] <= i) lo = mid;
    else hi = mid;
  }
  const int q = lo;
  int r = i - (int)soff[q];
  for (int o = 0; o < nQ; o++) {
    if (o == q) continue;
    const double* a = svals + soff[o];
    const int n = (int)(soff[o + 1] - soff[o]);
    r += o < q ? count_ge(a, n, v) : count_gt(a, n, v);
  }
  order[r] = i;
  sorted[r] = v;
}

// 5c. the reference's keep loop (qtensor_op.cpp:1294-1300), sequential in its
// summation order: nm over (sector, descending) order, cumsum over the sorted
// values
constexpr int kTrChunk = 2048;   // values staged in LDS per pass of the serial loops
constexpr int kTrPer = kTrChunk / 256;   // staged values per thread (launched with 256 threads)
// The serial loops add precomputed terms: every thread forms its values'
// square s*s (for nm) and s*s/nm (for cumsum), each rounded as the
// reference's expressions round them, and one thread adds them in order, so
// the sums are the reference's and the serial chain holds no multiply or
// divide.
__global__ void __launch_bounds__(256) k_svd_truncate(int T, int nQ, const double* svals, const int* ssec,
                                                      const int* order, const double* sorted, double cutoff,
                                                      int64_t K, int* newdim, double* out) {
  __shared__ double val[kTrChunk];
  __shared__ double term[kTrChunk];
  __shared__ double s_nm, s_cum;
  __shared__ long long s_acc;
  __shared__ int s_done;
  const int tid = threadIdx.x;
  // nm: the sum of squares in the (sector, descending) order
  double nm = 0.0;
  for (int c0 = 0; c0 < T; c0 += kTrChunk) {
    const int cn = min(kTrChunk, T - c0);
    {   // a thread's loads are all issued before its stores
      double v[kTrPer];
#pragma unroll
      for (int u = 0; u < kTrPer; u++) {
        const int e = tid + 256 * u;
        v[u] = e < cn ? svals[c0 + e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kTrPer; u++) {
        const int e = tid + 256 * u;
        if (e < cn) term[e] = v[u] * v[u];
      }
    }
    __syncthreads();
    if (tid == 0) {
      // unrolled so that the LDS reads run ahead of the (ordered) adds
#pragma unroll 8
      for (int e = 0; e < cn; e++) nm += term[e];
    }
    __syncthreads();
  }
  if (tid == 0) { s_nm = nm; s_cum = 0.0; s_acc = 0; s_done = 0; }
  __syncthreads();
  nm = s_nm;
  // the keep loop over the sorted values, in order, chunk by chunk
  for (int c0 = 0; c0 < T; c0 += kTrChunk) {
    if (s_done) break;   // read after a barrier: the same value in every thread
    const int cn = min(kTrChunk, T - c0);
    {
      double v[kTrPer];
#pragma unroll
      for (int u = 0; u < kTrPer; u++) {
        const int e = tid + 256 * u;
        v[u] = e < cn ? sorted[c0 + e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kTrPer; u++) {
        const int e = tid + 256 * u;
        if (e < cn) {
          val[e] = v[u];
          term[e] = v[u] * v[u] / nm;
        }
      }
    }
    __syncthreads();
    if (tid == 0) {
      double cumsum = s_cum;
      long long acc = s_acc;
      int done = 0;
#pragma unroll 8
      for (int e = 0; e < cn; e++) {
        cumsum += term[e];
        if (val[e] == 0. && cutoff != 0.) { done = 1; break; }
        acc += 1;
        if ((1 - cumsum < cutoff && cutoff != 0.) || (acc >= K && K != 0)) { done = 1; break; }
      }
      s_cum = cumsum;
      s_acc = acc;
      s_done = done;
    }
    __syncthreads();
  }
  // kept values per sector: the first acc entries of the sorted order
  const long long acc = s_acc;
  for (long long i = tid; i < acc; i += blockDim.x) {
    const int oi = order[i];
    const int q = (oi >= 0 && oi < T) ? ssec[oi] : -1;
    if (q >= 0 && q < nQ) atomicAdd(&newdim[q], 1);   // unreachable otherwise: order and ssec are complete
  }
  if (tid == 0) {
    out[0] = fmax(1 - s_cum, 0.0);
    out[1] = (double)acc;
  }
}

// 6a. block lists of U (left..., mid) and V (mid, right...) and S (mid)
__global__ void __launch_bounds__(1024) k_svd_emit(MetaView A, int nleft, SvdTab t, int nQ,
                                                   const int* newdim, const int* midnum,
                                                   int* uqi, int64_t* uoff, int* vqi,
                                                   int64_t* voff) {
  __shared__ int64_t lds[32];
  const int nuL = (int)t.hdr[0], nuR = (int)t.hdr[1];
  const int rU = nleft + 1, rV = A.rank - nleft + 1;
  // U: entries of uL whose sector keeps values, in (sector, left) order
  int64_t cb = 0, co = 0;
  for (int base = 0; base < nuL; base += 1024) {
    int i = base + threadIdx.x;
    int q = -1;
    int64_t f = 0, sz = 0;
    if (i < nuL) {
      int lo = 0, hi = nQ;
      while (lo < hi) { int m = (lo + hi) >> 1; if (t.qL0[m] <= i) lo = m + 1; else hi = m; }
      q = lo - 1;
      f = newdim[q] > 0;
      sz = f ? t.lsz[i] * newdim[q] : 0;
    }
    int64_t tb, ts;
    int64_t pb = block_excl_scan<1024>(f, lds, &tb);
    int64_t ps = block_excl_scan<1024>(sz, lds, &ts);
    if (f) {
      uint64_t key = t.uL[i] & 0xffffffffull;
      for (int j = nleft - 1; j >= 0; j--) { uqi[(cb + pb) * rU + j] = (int)(key % A.nsec[j]); key /= A.nsec[j]; }
      uqi[(cb + pb) * rU + nleft] = midnum[q];
      uoff[cb + pb] = co + ps;
    }
    cb += tb;
    co += ts;
  }
  if (threadIdx.x == 0) uoff[cb] = co;
  __syncthreads();
  cb = 0;
  co = 0;
  for (int base = 0; base < nuR; base += 1024) {
    int i = base + threadIdx.x;
    int64_t f = 0, sz = 0;
    int q = -1;
    if (i < nuR) {
      const uint64_t qc = t.uR[i] >> 32;
      int lo = 0, hi = nQ;
      while (lo < hi) { int m = (lo + hi) >> 1; if (t.qc[m] < qc) lo = m + 1; else hi = m; }
      q = lo;
      f = newdim[q] > 0;
      sz = f ? t.rsz[i] * newdim[q] : 0;
    }
    int64_t tb, ts;
    int64_t pb = block_excl_scan<1024>(f, lds, &tb);
    int64_t ps = block_excl_scan<1024>(sz, lds, &ts);
    if (f) {
      uint64_t key = t.uR[i] & 0xffffffffull;
      vqi[(cb + pb) * rV] = midnum[q];
      for (int j = A.rank - 1; j >= nleft; j--) {
        vqi[(cb + pb) * rV + 1 + (j - nleft)] = (int)(key % A.nsec[j]);
        key /= A.nsec[j];
      }
      voff[cb + pb] = co + ps;
    }
    cb += tb;
    co += ts;
  }
  if (threadIdx.x == 0) voff[cb] = co;
}

// 6b. slice sector matrices into blocks.  U_q (rows x k, ld rows) ->
// U blocks (lsize x k);  W_q (k x cols, ld k) -> V blocks (k x rsize).
// Workgroups over (sector, columns): a column of a block is contiguous on
// both sides.
__global__ void __launch_bounds__(256) k_svd_scatter(SvdTab t, int nQ, const int* newdim,
                                                     const int* ublk0, const int* vblk0,
                                                     const int64_t* uoffs, const int64_t* voffs,
                                                     const double* Umat, const int64_t* umo,
                                                     const double* Wmat, const int64_t* wmo,
                                                     double* Ud, double* Vd) {
  const int q = blockIdx.x;
  const int k = newdim[q];
  if (k == 0) return;
  const int64_t rows = t.rows[q];
  int ub = ublk0[q];
  for (int i = t.qL0[q]; i < t.qL1[q]; i++, ub++) {
    const int ls = (int)t.lsz[i];
    const int64_t lo = t.loff[i];
    for (int c = blockIdx.y; c < k; c += gridDim.y) {
      double* dst = Ud + uoffs[ub] + (int64_t)ls * c;
      const double* src = Umat + umo[q] + lo + rows * c;
      for (int r = threadIdx.x; r < ls; r += blockDim.x) dst[r] = src[r];
    }
  }
  int vb = vblk0[q];
  for (int i = t.qR0[q]; i < t.qR1[q]; i++, vb++) {
    const int rs = (int)t.rsz[i];
    const int64_t co = t.roff[i];
    for (int c = blockIdx.y; c < rs; c += gridDim.y) {
      double* dst = Vd + voffs[vb] + (int64_t)k * c;
      const double* src = Wmat + wmo[q] + (int64_t)k * (co + c);
      for (int r = threadIdx.x; r < k; r += blockDim.x) dst[r] = src[r];
    }
  }
}

// the unabsorbed factor without a product: W_q = E^T (k x cols) for
// MoveFromRight, by 32 x 32 tiles through LDS (reads along E's columns,
// writes along W's columns); tiles over (sector, tile index)
constexpr int kTT = 32;
__global__ void __launch_bounds__(256) k_svd_transpose(int nQ, const int* nq, const int* newdim,
                                                       const int64_t* selo, const double* E,
                                                       const int64_t* wmo, double* W) {
  __shared__ double tile[kTT][kTT + 1];
  const int q = blockIdx.x;
  const int n = nq[q], k = newdim[q];
  const int ti = (n + kTT - 1) / kTT, tc = (k + kTT - 1) / kTT;
  const int x = threadIdx.x % kTT, y0 = threadIdx.x / kTT;   // 32 x 8 threads
  for (int tt = blockIdx.y; tt < ti * tc; tt += gridDim.y) {
    const int i0 = (tt % ti) * kTT, c0 = (tt / ti) * kTT;
    __syncthreads();
    for (int y = y0; y < kTT; y += 8)   // E(i0 + x, c0 + y)
      if (i0 + x < n && c0 + y < k) tile[y][x] = E[selo[q] + (i0 + x) + (int64_t)n * (c0 + y)];
    __syncthreads();
    for (int y = y0; y < kTT; y += 8)   // W(c0 + x, i0 + y) = E(i0 + y, c0 + x)
      if (c0 + x < k && i0 + y < n) W[wmo[q] + (c0 + x) + (int64_t)k * (i0 + y)] = tile[x][y];
  }
}

// U_q = the first k columns of Esel_q (rows x kq, ld rows) for MoveFromLeft
__global__ void __launch_bounds__(256) k_svd_copy_sel(int nQ, const int64_t* rowsq, const int* newdim,
                                                      const int64_t* selo, const double* E,
                                                      const int64_t* umo, double* U) {
  const int q = blockIdx.x;
  const int64_t tot = rowsq[q] * newdim[q];
  for (int64_t e = (int64_t)blockIdx.y * blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.y * blockDim.x)
    U[umo[q] + e] = E[selo[q] + e];
}

__global__ void k_svd_svals(int nQ, const int* newdim, const int64_t* soff, const double* svals,
                            const int64_t* sout_off, double* S) {
  const int q = blockIdx.x;
  for (int i = threadIdx.x; i < newdim[q]; i += blockDim.x) S[sout_off[q] + i] = svals[soff[q] + i];
}

// grouped GEMM over sectors with host-built tasks
struct HostGemm {
  std::vector<GemmTask> tasks;
  std::vector<CBlock> cblk;
};

static int run_host_gemm(HostGemm& hg, const double* A, const double* B, double* C) {
  if (hg.cblk.empty()) return 0;
  ttg_plan_s p;
  p.prec = TTG_FP64;
  p.cfg = TILE_64x64;
  const int TK = tk_of(p.prec);
  for (CBlock& b : hg.cblk) {   // k-tile offsets of each block's walk
    int kt = 0;
    for (int t = b.t0; t < b.t1; t++) {
      GemmTask& g = hg.tasks[t];
      g.kt0 = kt;
      kt += (g.K / g.ek0) * ((g.ek0 + TK - 1) / TK);
    }
    b.nkt = kt;
  }
  p.ntasks = (int64_t)hg.tasks.size();
  const GemmTask& g0 = hg.tasks[0];
  p.a_mcontig = g0.sam0 == 1;
  p.b_kcontig = g0.sbk0 == 1;
  p.d_tasks = dalloc<GemmTask>(hg.tasks.size());
  p.d_cblk = dalloc<CBlock>(hg.cblk.size());
  TTG_REQUIRE(p.d_tasks && p.d_cblk, (std::string("svd: ") + ttg_last_error()).c_str());
  hipStream_t s = stream();
  TTG_TRY(h2d_vec(p.d_tasks, hg.tasks));
  TTG_TRY(h2d_vec(p.d_cblk, hg.cblk));
  TTG_TRY(build_tiles(p, hg.cblk));
  TTG_TRY(gemm_launch(p, A, B, C, true, s));
  return 0;   // p's destructor returns the device arrays to the pool (stream-ordered reuse)
}

static GemmTask gtask(int64_t a_off, int64_t b_off, int em0, int64_t sam0, int ek0, int64_t sak0,
                      int64_t sbk0, int en0, int64_t sbn0, int K) {
  GemmTask g{};
  g.a_off = a_off; g.b_off = b_off;
  g.em0 = em0; g.sam0 = sam0; g.sam1 = 0;
  g.ek0 = ek0; g.sak0 = sak0; g.sak1 = 0; g.sbk0 = sbk0; g.sbk1 = 0;
  g.en0 = en0; g.sbn0 = sbn0; g.sbn1 = 0;
  g.K = K;
  g.nk1 = K / ek0;
  return g;
}

// host arrays uploaded as one copy into one scratch allocation: add() the
// arrays, upload(), then at<T>(offset) gives each array's device address
class Pack {
 public:
  template <class T>
  size_t add(const std::vector<T>& v) {
    const size_t off = (bytes_.size() + 15) & ~size_t(15);
    bytes_.resize(off + sizeof(T) * std::max<size_t>(v.size(), 1));
    if (!v.empty()) memcpy(bytes_.data() + off, v.data(), sizeof(T) * v.size());
    return off;
  }
  int upload(Scratch& sc, const char* what) {
    d_ = sc.get<char>(std::max<size_t>(bytes_.size(), 1));
    TTG_TRY(sc.ok(what));
    return h2d(d_, bytes_.data(), bytes_.size());
  }
  template <class T>
  T* at(size_t off) const {
    return (T*)(d_ + off);
  }

 private:
  std::vector<char> bytes_;
  char* d_ = nullptr;
};

// eigenpairs of the large sectors `big` through the LR step (see 4c)
static int lr_eigen(const std::vector<int>& big, const std::vector<int>& nq, const std::vector<int64_t>& goff,
                    const std::vector<int64_t>& eoff, const int64_t* d_eoff, int64_t gtotal, double* G,
                    double* Vev, double* ev, hipStream_t s) {
  PhaseTimer _phase("lr_eigen");
  const int ns = (int)big.size();
  std::vector<PcSector> ps(ns);
  int nmax = 0;
  int64_t dtot = 0;
  for (int li = 0; li < ns; li++) {
    const int q = big[li];
    ps[li] = PcSector{nq[q], q, goff[q], dtot};
    dtot += nq[q];
    nmax = std::max(nmax, nq[q]);
  }
  // panel rows: nb rows of n doubles, d and the flags in the 160 KiB of LDS
  const int nb = std::max(1, std::min(kPcNbMax, (int)(163840 / (8 * (int64_t)nmax)) - 2));
  const size_t lds = sizeof(double) * ((size_t)nb * nmax + nmax) + sizeof(int) * nmax;
  TTG_REQUIRE(lds <= 163840, "svd: sector too large for the pivoted Cholesky panel");
  Scratch sc;
  PcSector* d_ps = sc.get<PcSector>(ns);
  double* R = sc.get<double>(gtotal);
  double* dg = sc.get<double>(dtot);
  int* pg = sc.get<int>(dtot);
  int* rank = sc.get<int>(ns);
  double* dmax = sc.get<double>(ns);
  TTG_TRY(sc.ok("svd: LR step"));
  TTG_TRY(h2d_vec(d_ps, ps));
  hipLaunchKernelGGL(k_pchol_init, dim3(ns), dim3(kPcThreads), 0, s, d_ps, G, dg, pg, rank, dmax);
  TTG_HIP(hipFuncSetAttribute((const void*)k_pchol_panel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  const int nt = (nmax + 63) / 64;
  // Panels are launched up to one more than the previous call needed, then
  // the ranks are read back; only if a sector is still running are the rest
  // launched (a finished sector's launches exit at once, but each launch
  // still costs a few microseconds).  The panels run in the same order
  // either way.
  static int prev_panels = 0;
  const int all_panels = (nmax + nb - 1) / nb;
  int target = prev_panels > 0 ? std::min(all_panels, prev_panels + 1) : all_panels;
  std::vector<int> rk(ns);
  int k0 = 0;
  for (;;) {
    for (; k0 < target * nb && k0 < nmax; k0 += nb) {
      hipLaunchKernelGGL(k_pchol_panel, dim3(ns), dim3(kPcThreads), lds, s, d_ps, k0, nb, G, R, dg, pg, rank,
                         dmax);
      hipLaunchKernelGGL(k_pchol_update, dim3(nt * nt, ns), dim3(256), 0, s, d_ps, k0, nb, rank, G, R);
    }
    const D2HPiece pc{rk.data(), rank, sizeof(int) * ns};
    TTG_TRY(d2h_wait(&pc, 1));
    bool running = false;   // a sector with neither a rank nor all its columns done
    for (int li = 0; li < ns; li++) running |= rk[li] < 0 && k0 < ps[li].n;
    if (!running || k0 >= nmax) break;
    target = all_panels;
  }
  {
    int need = 0;   // panels this call needed: up to each sector's stopping panel
    for (int li = 0; li < ns; li++)
      need = std::max(need, rk[li] < 0 ? (ps[li].n + nb - 1) / nb : rk[li] / nb + 1);
    prev_panels = std::min(need, all_panels);
  }
  std::vector<int> r2(ns);
  std::vector<int64_t> goff2(ns + 1, 0), poff2(ns + 1, 0), eoff2(ns + 1, 0);
  int64_t lds_need = 1;
  bool any_small = false;
  std::vector<int> big2;
  for (int li = 0; li < ns; li++) {
    const int r = rk[li] < 0 ? ps[li].n : rk[li];
    r2[li] = r;
    goff2[li + 1] = goff2[li] + (int64_t)r * r;
    const int64_t np = r + (r & 1);
    poff2[li + 1] = poff2[li] + np * (np + 1) / 2;
    eoff2[li + 1] = eoff2[li] + r;
    if (r > kJacBlockMin) big2.push_back(li);
    else if (r > 0) { any_small = true; lds_need = std::max(lds_need, jac_layout(r).total); }
  }
  double* G2 = sc.get<double>(std::max<int64_t>(1, goff2[ns]));
  double* W2 = sc.get<double>(std::max<int64_t>(1, goff2[ns]));
  double* gs2 = sc.get<double>(std::max<int64_t>(1, poff2[ns]));
  double* ev2 = sc.get<double>(std::max<int64_t>(1, eoff2[ns]));
  TTG_TRY(sc.ok("svd: LR step"));
  Pack pk;
  const size_t o_r2 = pk.add(r2), o_goff2 = pk.add(goff2), o_poff2 = pk.add(poff2), o_eoff2 = pk.add(eoff2);
  TTG_TRY(pk.upload(sc, "svd: LR step"));
  int* d_r2 = pk.at<int>(o_r2);
  int64_t* d_goff2 = pk.at<int64_t>(o_goff2);
  int64_t* d_poff2 = pk.at<int64_t>(o_poff2);
  int64_t* d_eoff2 = pk.at<int64_t>(o_eoff2);
  // G' = R R^T: A(k, j) = R[k n + j], B(j, k') = R[k' n + j]
  {
    HostGemm hg;
    for (int li = 0; li < ns; li++) {
      const int n = ps[li].n, r = r2[li];
      if (r == 0) continue;
      hg.tasks.push_back(gtask(ps[li].goff, ps[li].goff, r, n, n, 1, 1, r, n, n));
      const int t = (int)hg.tasks.size() - 1;
      hg.cblk.push_back(CBlock{goff2[li], r, r, t, t + 1});
    }
    TTG_TRY(run_host_gemm(hg, R, R, G2));
  }
  // eigenpairs of G' (relative convergence test only)
  if (any_small) {
    const size_t lb = sizeof(double) * (size_t)std::min<int64_t>(lds_need, kLdsDoubles);
    TTG_HIP(hipFuncSetAttribute((const void*)k_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * kLdsDoubles)));
    hipLaunchKernelGGL(k_jacobi, dim3(ns), dim3(256), lb, s, d_r2, d_goff2, G2, gs2, d_poff2, W2, ev2,
                       d_eoff2, 40, kJacBlockMin, 0.0);
    TTG_HIP(hipGetLastError());
  }
  if (!big2.empty()) TTG_TRY(block_jacobi(big2, r2, d_goff2, d_eoff2, G2, W2, ev2, s, 40, 0.0));
  // V = R^T W Lambda^(-1/2): A(j, k) = R[k n + j], B(k, c) = W[k + r c]
  {
    HostGemm hv;
    for (int li = 0; li < ns; li++) {
      const int n = ps[li].n, r = r2[li];
      if (r == 0) continue;
      hv.tasks.push_back(gtask(ps[li].goff, goff2[li], n, 1, r, n, 1, r, r, r));
      const int t = (int)hv.tasks.size() - 1;
      hv.cblk.push_back(CBlock{ps[li].goff, n, r, t, t + 1});
    }
    TTG_TRY(run_host_gemm(hv, R, W2, Vev));
  }
  hipLaunchKernelGGL(k_lr_finish, dim3(nmax, ns), dim3(256), 0, s, d_ps, d_r2, d_eoff2, ev2, d_eoff, Vev, ev);
  TTG_HIP(hipGetLastError());
  return 0;
}

}  // namespace ttg

using namespace ttg;

extern "C" int ttg_svd(ttg_tensor A, int nleft, int direction, double cutoff, int64_t K,
                       ttg_tensor* Uout, ttg_tensor* Sout, ttg_tensor* Vout, double* trunc_err,
                       int64_t* kept) {
  PhaseTimer _phase("ttg_svd");
  TTG_REQUIRE(A && Uout && Sout && Vout, "svd: null argument");
  const TMeta& m = *A->m;
  TTG_REQUIRE(nleft >= 1 && nleft < m.rank, "svd: need at least one left and one right index");
  TTG_REQUIRE(direction == TTG_MOVE_FROM_LEFT || direction == TTG_MOVE_FROM_RIGHT, "svd: direction");
  double rl = 1, rr = 1;
  for (int j = 0; j < m.rank; j++) (j < nleft ? rl : rr) *= m.nsec[j];
  TTG_REQUIRE(rl < 4294967295.0 && rr < 4294967295.0, "svd: sector-tuple key exceeds 32 bits");
  hipStream_t s = stream();
  const int nb = m.nblocks;
  TTG_REQUIRE(nb > 0, "svd: tensor without blocks");
  // ---- 1. bookkeeping
  Scratch sc;   // every buffer below goes back to the pool on return
  uint64_t *kl = sc.get<uint64_t>(nb), *kr = sc.get<uint64_t>(nb);
  uint64_t *sl = sc.get<uint64_t>(nb), *sr = sc.get<uint64_t>(nb);
  int *ids = sc.get<int>(nb), *sid = sc.get<int>(nb);
  SvdTab t{};
  t.uL = sc.get<uint64_t>(nb); t.uR = sc.get<uint64_t>(nb);
  t.lsz = sc.get<int64_t>(nb); t.loff = sc.get<int64_t>(nb);
  t.rsz = sc.get<int64_t>(nb); t.roff = sc.get<int64_t>(nb);
  // the header and the per-sector tables the host reads, in one allocation
  // (read back with one copy): hdr[4], rows, cols (int64), qc, qL0, qL1,
  // qR0, qR1 (32-bit), nb entries each
  const size_t nbp = ((size_t)nb + 31) & ~size_t(31);
  const size_t tab_bytes = 32 + nbp * (2 * sizeof(int64_t) + 5 * sizeof(int));
  char* tabs = sc.get<char>(tab_bytes);
  TTG_TRY(sc.ok("svd"));
  t.hdr = (int64_t*)tabs;
  t.rows = (int64_t*)(tabs + 32);
  t.cols = t.rows + nbp;
  t.qc = (uint32_t*)(t.cols + nbp);
  t.qL0 = (int*)(t.qc + nbp);
  t.qL1 = t.qL0 + nbp;
  t.qR0 = t.qL1 + nbp;
  t.qR1 = t.qR0 + nbp;
  const unsigned g1 = (nb + 255) / 256;
  hipLaunchKernelGGL(k_svd_keys, dim3(g1), dim3(256), 0, s, m.view(), nleft, kl, kr, ids);
  hipLaunchKernelGGL(k_sort_keys, dim3(g1), dim3(256), 0, s, kl, ids, nb, sl, sid);
  hipLaunchKernelGGL(k_sort_keys, dim3(g1), dim3(256), 0, s, kr, ids, nb, sr, sid);
  hipLaunchKernelGGL(k_svd_unique, dim3(1), dim3(1024), 0, s, m.view(), nleft, sl, sr, nb, t);
  std::vector<char> htab(tab_bytes);
  {
    const D2HPiece pc{htab.data(), tabs, tab_bytes};
    TTG_TRY(d2h_wait(&pc, 1));
  }
  int64_t hdr[4];
  memcpy(hdr, htab.data(), sizeof(hdr));
  const int nuL = (int)hdr[0], nuR = (int)hdr[1], nQ = (int)hdr[2];
  TTG_REQUIRE(nQ >= 0 && nQ <= nb, "svd: bad sector count");
  auto tab = [&](const void* dptr, auto* dst, size_t elem) {
    memcpy(dst, htab.data() + ((const char*)dptr - tabs), elem * nQ);
  };
  std::vector<int64_t> rows(nQ), cols(nQ);
  std::vector<uint32_t> qcs(nQ);
  std::vector<int> qL0(nQ), qL1(nQ), qR0(nQ), qR1(nQ);
  tab(t.rows, rows.data(), sizeof(int64_t));
  tab(t.cols, cols.data(), sizeof(int64_t));
  tab(t.qc, qcs.data(), sizeof(uint32_t));
  tab(t.qL0, qL0.data(), sizeof(int));
  tab(t.qL1, qL1.data(), sizeof(int));
  tab(t.qR0, qR0.data(), sizeof(int));
  tab(t.qR1, qR1.data(), sizeof(int));
  // ---- host layout of the per-sector work arrays
  const bool fromLeft = direction == TTG_MOVE_FROM_LEFT;
  std::vector<int64_t> moff(nQ + 1, 0), goff(nQ + 1, 0), poff(nQ + 1, 0), eoff(nQ + 1, 0);
  std::vector<int> nq(nQ), kq(nQ);
  for (int q = 0; q < nQ; q++) {
    const int64_t n = fromLeft ? rows[q] : cols[q];
    TTG_REQUIRE(n <= 8192, "svd: sector larger than 8192");
    nq[q] = (int)n;
    int kd = (int)std::min(rows[q], cols[q]);
    kq[q] = K > 0 ? (int)std::min<int64_t>(K, kd) : kd;   // rank-K SVD per sector
    moff[q + 1] = moff[q] + rows[q] * cols[q];
    goff[q + 1] = goff[q] + n * n;
    const int64_t np = n + (n & 1);
    poff[q + 1] = poff[q] + np * (np + 1) / 2;
    eoff[q + 1] = eoff[q] + n;
  }
  std::vector<int64_t> soff(nQ + 1, 0), selo(nQ + 1, 0);
  for (int q = 0; q < nQ; q++) {
    soff[q + 1] = soff[q] + kq[q];
    selo[q + 1] = selo[q] + (int64_t)nq[q] * kq[q];
  }
  double* mat = sc.get<double>(moff[nQ]);
  double* G = sc.get<double>(goff[nQ]);
  double* Vev = sc.get<double>(goff[nQ]);
  double* gs = sc.get<double>(poff[nQ]);
  double* ev = sc.get<double>(eoff[nQ]);
  double* Esel = sc.get<double>(selo[nQ]);
  int* selidx = sc.get<int>(soff[nQ]);
  double* svals = sc.get<double>(soff[nQ]);
  int* ssec = sc.get<int>(soff[nQ]);
  int* order = sc.get<int>(soff[nQ]);
  double* ssorted = sc.get<double>(soff[nQ]);
  int* newdim = sc.get<int>(nQ);
  double* tout = sc.get<double>(2);
  TTG_TRY(sc.ok("svd"));
  Pack pk;
  const size_t o_moff = pk.add(moff), o_goff = pk.add(goff), o_poff = pk.add(poff), o_eoff = pk.add(eoff),
               o_soff = pk.add(soff), o_selo = pk.add(selo), o_nq = pk.add(nq), o_kq = pk.add(kq);
  TTG_TRY(pk.upload(sc, "svd"));
  int64_t* d_moff = pk.at<int64_t>(o_moff);
  int64_t* d_goff = pk.at<int64_t>(o_goff);
  int64_t* d_poff = pk.at<int64_t>(o_poff);
  int64_t* d_eoff = pk.at<int64_t>(o_eoff);
  int64_t* d_soff = pk.at<int64_t>(o_soff);
  int64_t* d_selo = pk.at<int64_t>(o_selo);
  int* d_nq = pk.at<int>(o_nq);
  int* d_kq = pk.at<int>(o_kq);
  // ---- 2. gather
  fill_f64(mat, moff[nQ], 0.0);
  hipLaunchKernelGGL(k_svd_gather, dim3(std::min(nb, 4096)), dim3(256), 0, s, m.view(), nleft,
                     A->d_data, t, nuL, nuR, nQ, d_moff, mat);
  // ---- 3. Gram on MFMA
  {
    HostGemm hg;
    for (int q = 0; q < nQ; q++) {
      const int n = nq[q];
      const int64_t r = rows[q], c = cols[q];
      if (fromLeft)   // G = A A^T : A(m=i, k=j) = mat[i + r j]; B(k=j, n=i') = mat[i' + r j]
        hg.tasks.push_back(gtask(moff[q], moff[q], (int)r, 1, (int)c, r, r, (int)r, 1, (int)c));
      else            // G = A^T A : A(m=j, k=i) = mat[i + r j]; B(k=i, n=j') = mat[i + r j']
        hg.tasks.push_back(gtask(moff[q], moff[q], (int)c, r, (int)r, 1, 1, (int)c, r, (int)r));
      hg.cblk.push_back(CBlock{goff[q], n, n, q, q + 1});
    }
    TTG_TRY(run_host_gemm(hg, mat, mat, G));
  }
  // a sector whose Gram matrix overflowed (entries of A above ~1e154) or that
  // holds a NaN has no meaningful eigenpairs: fail before the eigensolver
  {
    int* bad = sc.get<int>(1);
    TTG_TRY(sc.ok("svd"));
    int hbad = 0;
    TTG_TRY(h2d(bad, &hbad, sizeof(int)));
    hipLaunchKernelGGL(k_svd_finite, dim3(nQ), dim3(256), 0, s, nQ, d_nq, d_goff, G, bad);
    const D2HPiece pc{&hbad, bad, sizeof(int)};
    TTG_TRY(d2h_wait(&pc, 1));
    TTG_REQUIRE(!hbad, "svd: a sector's Gram matrix is not finite (tensor entries above ~1e154, or NaN)");
  }
  // ---- 4. eigensolver
  std::vector<int> big;
  int64_t lds_need = 1;
  int nsmall = 0;
  for (int q = 0; q < nQ; q++) {
    if (nq[q] > kJacBlockMin) { big.push_back(q); continue; }
    lds_need = std::max(lds_need, jac_layout(nq[q]).total);
    nsmall = std::max(nsmall, nq[q]);
  }
  // The small sectors' one-workgroup Jacobi runs on the side stream, beside
  // the large sectors' solver on the main stream: the two read and write
  // disjoint sector ranges of G, the packed triangles, Vev and ev.
  bool forked = false;
  if (nsmall > 0) {
    const size_t lds_bytes = sizeof(double) * (size_t)std::min<int64_t>(lds_need, kLdsDoubles);
    TTG_HIP(hipFuncSetAttribute((const void*)k_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(double) * kLdsDoubles)));
    // small sectors: fewer waves per workgroup (cheaper barriers per round)
    const int jac_threads = nsmall <= 96 ? 256 : kJacThreads;
    hipStream_t js = s;
    if (!big.empty()) {
      TTG_TRY(side_fork(&js));
      forked = true;
    }
    hipLaunchKernelGGL(k_jacobi, dim3(nQ), dim3(jac_threads), lds_bytes, js, d_nq, d_goff, G, gs,
                       d_poff, Vev, ev, d_eoff, 40, kJacBlockMin, kJacFloor);
    TTG_HIP(hipGetLastError());
  }
  // large sectors: the LR step when the truncation drops zero singular
  // values (cutoff > 0), else block Jacobi on G
  int brc = 0;
  if (!big.empty())
    brc = cutoff > 0 ? lr_eigen(big, nq, goff, eoff, d_eoff, goff[nQ], G, Vev, ev, s)
                     : block_jacobi(big, nq, d_goff, d_eoff, G, Vev, ev, s, 40, kJacFloor);
  if (forked) {
    // join before anything else, the error path included: the scratch the
    // side stream reads goes back to the pool when this call returns
    const int jrc = side_join();
    if (brc || jrc) (void)hipStreamSynchronize(s);
    TTG_TRY(jrc);
  }
  TTG_TRY(brc);
  // ---- 5. selection and truncation
  {
    int nmax = 1, kmax = 1;
    for (int q = 0; q < nQ; q++) { nmax = std::max(nmax, nq[q]); kmax = std::max(kmax, kq[q]); }
    hipLaunchKernelGGL(k_svd_select, dim3(nQ, (nmax + 255) / 256), dim3(256), 0, s, d_nq, d_kq, d_eoff, ev,
                       d_soff, svals, ssec, selidx);
    hipLaunchKernelGGL(k_svd_gather_sel, dim3(nQ, std::min(kmax, 64)), dim3(256), 0, s, d_nq, d_kq, d_goff,
                       Vev, d_soff, selidx, d_selo, Esel);
  }
  TTG_HIP(hipMemsetAsync(newdim, 0, sizeof(int) * nQ, s));
  {
    const int T = (int)soff[nQ];
    if (T > 0)
      hipLaunchKernelGGL(k_svd_rank, dim3((T + 255) / 256), dim3(256), 0, s, T, nQ, d_soff, svals, order,
                         ssorted);
    hipLaunchKernelGGL(k_svd_truncate, dim3(1), dim3(256), 0, s, T, nQ, svals, ssec, order, ssorted,
                       cutoff, K, newdim, tout);
  }
  std::vector<int> nd(nQ);
  double th[2];
  {
    const D2HPiece pc[2] = {{nd.data(), newdim, sizeof(int) * nQ}, {th, tout, sizeof(th)}};
    TTG_TRY(d2h_wait(pc, 2));
  }
  if (trunc_err) *trunc_err = th[0];
  if (kept) *kept = (int64_t)th[1];
  // ---- 6. outputs: index tables (left..., mid) (mid, right...) (mid)
  std::vector<int> midnum(nQ, -1);
  std::vector<int> mq, md;
  int nm = 0;
  for (int q = 0; q < nQ; q++)
    if (nd[q] > 0) {
      midnum[q] = nm++;
      mq.push_back((int)(qcs[q] >> 16) - 32768);
      mq.push_back((int)(qcs[q] & 0xffff) - 32768);
      md.push_back(nd[q]);
    }
  TTG_REQUIRE(nm > 0, "svd: no singular value kept");
  std::vector<std::pair<const TMeta*, int>> lsrc, rsrc;
  for (int j = 0; j < nleft; j++) lsrc.push_back({&m, j});
  for (int j = nleft; j < m.rank; j++) rsrc.push_back({&m, j});
  MetaPtr mid;
  int arrOut = 1, arrIn = 0;
  TTG_TRY(meta_shell(1, &arrOut, &nm, mq.data(), md.data(), &mid));
  MetaPtr midIn;
  TTG_TRY(meta_shell(1, &arrIn, &nm, mq.data(), md.data(), &midIn));
  auto concat = [](const std::vector<std::pair<const TMeta*, int>>& a) { return a; };
  std::vector<std::pair<const TMeta*, int>> usrc = concat(lsrc), vsrc{{midIn.get(), 0}};
  usrc.push_back({mid.get(), 0});
  vsrc.insert(vsrc.end(), rsrc.begin(), rsrc.end());
  MetaPtr Um = meta_from_indices(usrc), Vm = meta_from_indices(vsrc);
  TTG_REQUIRE(Um && Vm, "svd: meta allocation failed");
  int nUb = 0, nVb = 0;
  int64_t Utot = 0, Vtot = 0, Stot = 0;
  std::vector<int> ublk0(nQ, 0), vblk0(nQ, 0);
  std::vector<int64_t> sout(nQ + 1, 0);
  for (int q = 0; q < nQ; q++) {
    ublk0[q] = nUb;
    vblk0[q] = nVb;
    sout[q + 1] = sout[q] + nd[q];
    if (nd[q] == 0) continue;
    nUb += qL1[q] - qL0[q];
    nVb += qR1[q] - qR0[q];
    Utot += rows[q] * nd[q];
    Vtot += cols[q] * nd[q];
    Stot += nd[q];
  }
  TTG_TRY(meta_alloc_blocks(*Um, nUb, (size_t)Utot));
  TTG_TRY(meta_alloc_blocks(*Vm, nVb, (size_t)Vtot));
  TTG_TRY(meta_alloc_blocks(*mid, nm, (size_t)Stot));
  {
    std::vector<int> sqi(nm);
    std::iota(sqi.begin(), sqi.end(), 0);
    std::vector<int64_t> soffv(nm + 1, 0);
    for (int q = 0, k = 0; q < nQ; q++)
      if (nd[q] > 0) { soffv[k + 1] = soffv[k] + nd[q]; k++; }
    TTG_TRY(h2d_vec(mid->d_qi, sqi));
    TTG_TRY(h2d_vec(mid->d_off, soffv));
  }
  Pack pko;
  const size_t o_mid = pko.add(midnum), o_ub0 = pko.add(ublk0), o_vb0 = pko.add(vblk0), o_sout = pko.add(sout);
  TTG_TRY(pko.upload(sc, "svd"));
  int* d_mid = pko.at<int>(o_mid);
  int* d_ub0 = pko.at<int>(o_ub0);
  int* d_vb0 = pko.at<int>(o_vb0);
  int64_t* d_sout = pko.at<int64_t>(o_sout);
  hipLaunchKernelGGL(k_svd_emit, dim3(1), dim3(1024), 0, s, m.view(), nleft, t, nQ, newdim, d_mid,
                     Um->d_qi, Um->d_off, Vm->d_qi, Vm->d_off);
  ttg_tensor U, V, Sv;
  TTG_TRY(tensor_from_meta(Um, false, &U));
  std::unique_ptr<ttg_tensor_s> Uown(U);
  TTG_TRY(tensor_from_meta(Vm, false, &V));
  std::unique_ptr<ttg_tensor_s> Vown(V);
  TTG_TRY(tensor_from_meta(mid, false, &Sv));
  std::unique_ptr<ttg_tensor_s> Sown(Sv);
  hipLaunchKernelGGL(k_svd_svals, dim3(nQ), dim3(64), 0, s, nQ, newdim, d_soff, svals, d_sout, Sv->d_data);
  // absorbed factor: MoveFromLeft  W = E^T A (k x cols), U = E (rows x k)
  //                  MoveFromRight U = A E (rows x k),   W = E^T (k x cols)
  std::vector<int64_t> umo(nQ + 1, 0), wmo(nQ + 1, 0);
  for (int q = 0; q < nQ; q++) {
    umo[q + 1] = umo[q] + rows[q] * nd[q];
    wmo[q + 1] = wmo[q] + (int64_t)nd[q] * cols[q];
  }
  double* Umat = sc.get<double>(umo[nQ]);
  double* Wmat = sc.get<double>(wmo[nQ]);
  TTG_TRY(sc.ok("svd"));
  Pack pkm;
  const size_t o_umo = pkm.add(umo), o_wmo = pkm.add(wmo);
  TTG_TRY(pkm.upload(sc, "svd"));
  int64_t* d_umo = pkm.at<int64_t>(o_umo);
  int64_t* d_wmo = pkm.at<int64_t>(o_wmo);
  {
    HostGemm hg;
    for (int q = 0; q < nQ; q++) {
      if (nd[q] == 0) continue;
      const int64_t r = rows[q], c = cols[q], n = nq[q];
      const int k = nd[q];
      if (fromLeft)   // W(i, j) = sum_a E(a, i) A(a, j): A-op E^T (m=i,k=a) = E[a + n i]
        hg.tasks.push_back(gtask(selo[q], moff[q], k, n, (int)r, 1, 1, (int)c, r, (int)r));
      else            // U(i, j) = sum_a A(i, a) E(a, j)
        hg.tasks.push_back(gtask(moff[q], selo[q], (int)r, 1, (int)c, r, 1, k, n, (int)c));
      const int tix = (int)hg.tasks.size() - 1;
      hg.cblk.push_back(fromLeft ? CBlock{wmo[q], k, (int)c, tix, tix + 1}
                                 : CBlock{umo[q], (int)r, k, tix, tix + 1});
    }
    TTG_TRY(fromLeft ? run_host_gemm(hg, Esel, mat, Wmat) : run_host_gemm(hg, mat, Esel, Umat));
  }
  if (fromLeft) {   // U_q = E restricted to k columns: Esel already (n = rows) x k, ld rows
    hipLaunchKernelGGL(k_svd_copy_sel, dim3(nQ, 32), dim3(256), 0, s, nQ, t.rows, newdim, d_selo, Esel,
                       d_umo, Umat);
  } else {
    hipLaunchKernelGGL(k_svd_transpose, dim3(nQ, 64), dim3(256), 0, s, nQ, d_nq, newdim, d_selo, Esel,
                       d_wmo, Wmat);
  }
  hipLaunchKernelGGL(k_svd_scatter, dim3(nQ, 64), dim3(256), 0, s, t, nQ, newdim, d_ub0, d_vb0, Um->d_off,
                     Vm->d_off, Umat, d_umo, Wmat, d_wmo, U->d_data, V->d_data);
  TTG_HIP(hipGetLastError());   // the outputs are written in stream order; the scratch returns to the pool
  *Uout = Uown.release();
  *Vout = Vown.release();
  *Sout = Sown.release();
  return 0;
}
