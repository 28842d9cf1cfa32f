// plan.hip -- device QN-sector bookkeeping for C = A * B (qtensor.cpp:523-734).
//
// The reference builds, on the host, per mid-QN q the ordered sets of A's
// free-index position tuples (left), B's free tuples (right) and B's
// contracted tuples (mid), creates one result block per (q, right, left) in
// (ascending q, lexicographic right, lexicographic left) order, and runs one
// block GEMM per mid tuple whose A and B blocks both exist.  Here every set
// is a sorted array of 64-bit keys built on the device:
//   (qcode << 32 | tuple)   with qcode order = QN_t_2 order (util/QN.h:100)
//   tuple = mixed-radix number of the positions, first position most
//           significant, so integer order = std::set<uint_vec> order.
// The planning kernels produce the result block list, its packed offsets, the
// task list in the reference loop order and the grouped-GEMM tile list.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "ttg_device.h"
#include "ttg_internal.h"
#include "ttg_plan.h"

namespace ttg {

__device__ inline uint32_t qcode(int q0, int q1) {
  return ((uint32_t)(q0 + 32768) << 16) | (uint32_t)(q1 + 32768);
}

struct KeyArrays {
  uint64_t *kUL, *kSA, *kUR, *kUM, *kSB;     // raw keys
  int *vSA, *vSB;                            // block ids
  uint64_t *sUL, *sSA, *sUR, *sUM, *sSB;     // sorted
  int *svSA, *svSB;
};

// ---- K1: per-block keys ------------------------------------------------
// Block 0 also clears the plan header that the later planning kernels fill
// (in place of a separate memset launch before this one).
static_assert(sizeof(PlanHdr) % sizeof(int64_t) == 0, "plan header is whole 8-byte words");
__global__ void k_plan_keys(MetaView A, MetaView B, PairSpec ps, KeyArrays k, PlanHdr* hdr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0)
    for (int w = threadIdx.x; w < (int)(sizeof(PlanHdr) / sizeof(int64_t)); w += blockDim.x)
      ((int64_t*)hdr)[w] = 0;
  if (i < A.nblocks) {
    const int* qi = A.qi + (size_t)i * A.rank;
    int q0 = 0, q1 = 0;
    uint64_t left = 0, mid = 0;
    for (int j = 0; j < ps.nfa; j++) {
      int p = ps.fa[j], s = qi[p];
      const int* e = A.idx + 3 * (A.idx_off[p] + s);
      int sg = A.arrow[p] == 0 ? 1 : -1;          // Inward +, Outward - (:617-621)
      q0 += sg * e[0];
      q1 += sg * e[1];
      left = left * A.nsec[p] + s;
    }
    for (int t = 0; t < ps.np; t++) {
      int p = ps.pa[t];
      mid = mid * A.nsec[p] + qi[p];
    }
    k.kUL[i] = ((uint64_t)qcode(q0, q1) << 32) | left;
    k.kSA[i] = (left << 32) | mid;
    k.vSA[i] = i;
  }
  if (i < B.nblocks) {
    const int* qi = B.qi + (size_t)i * B.rank;
    int q0 = 0, q1 = 0;
    uint64_t right = 0, mid = 0;
    for (int j = 0; j < ps.nfb; j++) {
      int p = ps.fb[j], s = qi[p];
      const int* e = B.idx + 3 * (B.idx_off[p] + s);
      int sg = B.arrow[p] == 0 ? -1 : 1;          // Inward -, Outward + (:629-633)
      q0 += sg * e[0];
      q1 += sg * e[1];
      right = right * B.nsec[p] + s;
    }
    for (int t = 0; t < ps.np; t++) {
      int p = ps.pb[t];
      mid = mid * B.nsec[p] + qi[p];
    }
    uint64_t qc = (uint64_t)qcode(q0, q1) << 32;
    k.kUR[i] = qc | right;
    k.kUM[i] = qc | mid;
    k.kSB[i] = (mid << 32) | right;
    k.vSB[i] = i;
  }
}

// ---- K2: stable rank sort (five arrays, grid.y selects) -----------------
struct SortJob {
  const uint64_t* in;
  const int* vin;
  uint64_t* out;
  int* vout;
  int n;
};
struct SortJobs { SortJob j[5]; };

__global__ void __launch_bounds__(256) k_rank_sort(SortJobs jobs) {
  const SortJob& J = jobs.j[blockIdx.y];
  __shared__ uint64_t tile[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if ((int)(blockIdx.x * 256) >= J.n) return;
  const uint64_t key = i < J.n ? J.in[i] : 0;
  int rank = 0;
  for (int base = 0; base < J.n; base += 256) {
    __syncthreads();
    if (base + (int)threadIdx.x < J.n) tile[threadIdx.x] = J.in[base + threadIdx.x];
    __syncthreads();
    const int lim = min(256, J.n - base);
    for (int t = 0; t < lim; t++) {
      uint64_t o = tile[t];
      rank += (o < key) || (o == key && base + t < i);
    }
  }
  if (i < J.n) {
    J.out[rank] = key;
    if (J.vout) J.vout[rank] = J.vin ? J.vin[i] : i;
  }
}

// ---- K3: unique sets, per-q ranges, result-block bases -------------------
struct QTab {
  int *qL0, *qL1, *qR0, *qR1, *qM0, *qM1;
  int64_t* qbase;   // nQ + 1
};

__device__ int compact_unique(const uint64_t* s, int n, uint64_t* u, int64_t* lds) {
  int64_t carry = 0;
  for (int base = 0; base < n; base += 1024) {
    int i = base + threadIdx.x;
    int f = 0;
    if (i < n) f = (i == 0) || (s[i] != s[i - 1]);
    int64_t tot;
    int64_t pos = block_excl_scan<1024>((int64_t)f, lds, &tot);
    if (f) u[carry + pos] = s[i];
    carry += tot;
  }
  __syncthreads();
  return (int)carry;
}

__global__ void __launch_bounds__(1024) k_plan_unique(KeyArrays k, int nA, int nB, uint64_t* uL,
                                                      uint64_t* uR, uint64_t* uM, QTab q,
                                                      PlanHdr* hdr) {
  __shared__ int64_t lds[32];
  const int nuL = compact_unique(k.sUL, nA, uL, lds);
  const int nuR = compact_unique(k.sUR, nB, uR, lds);
  const int nuM = compact_unique(k.sUM, nB, uM, lds);
  // q groups of uL (ascending q; A's mid-QN set, :622-623)
  int64_t carry = 0;
  for (int base = 0; base < nuL; base += 1024) {
    int i = base + threadIdx.x;
    int f = 0;
    if (i < nuL) f = (i == 0) || ((uL[i] >> 32) != (uL[i - 1] >> 32));
    int64_t tot;
    int64_t pos = block_excl_scan<1024>((int64_t)f, lds, &tot);
    if (f) q.qL0[carry + pos] = i;
    carry += tot;
  }
  __syncthreads();
  const int nQ = (int)carry;
  for (int t = threadIdx.x; t < nQ; t += 1024) {
    int l0 = q.qL0[t];
    int l1 = (t + 1 < nQ) ? q.qL0[t + 1] : nuL;
    q.qL1[t] = l1;
    uint64_t qc = uL[l0] >> 32;
    q.qR0[t] = lower_bound_u64(uR, nuR, qc << 32);
    q.qR1[t] = lower_bound_u64(uR, nuR, (qc + 1) << 32);
    q.qM0[t] = lower_bound_u64(uM, nuM, qc << 32);
    q.qM1[t] = lower_bound_u64(uM, nuM, (qc + 1) << 32);
  }
  __syncthreads();
  carry = 0;
  int64_t ncand = 0;   // every result block of group t tries each of its nM mids
  for (int base = 0; base < nQ; base += 1024) {
    int t = base + threadIdx.x;
    int64_t cnt = 0, cand = 0;
    if (t < nQ) {
      cnt = (int64_t)(q.qL1[t] - q.qL0[t]) * (q.qR1[t] - q.qR0[t]);
      cand = cnt * (q.qM1[t] - q.qM0[t]);
    }
    int64_t tot, tot2;
    int64_t pos = block_excl_scan<1024>(cnt, lds, &tot);
    (void)block_excl_scan<1024>(cand, lds, &tot2);
    if (t < nQ) q.qbase[t] = carry + pos;
    carry += tot;
    ncand += tot2;
  }
  if (threadIdx.x == 0) {
    q.qbase[nQ] = carry;
    hdr->nuL = nuL;
    hdr->nuR = nuR;
    hdr->nuM = nuM;
    hdr->nQ = nQ;
    hdr->nblocksC = carry;
    hdr->ncand = ncand;
  }
}

// ---- K4: per result block: tuple, dims, task enumeration ----------------
struct BlkArrays {
  int64_t* size;      // nblocksC
  int64_t* ntask;     // nblocksC
  double* flops;      // nblocksC
  const int64_t* task_off;   // nblocksC + 1 (phase 1)
  int* c_qi;          // nblocksC * rankC
  const int64_t* c_off;      // nblocksC + 1 (phase 1)
};
struct EmitArrays {
  GemmTask* tasks;
  CBlock* cblk;
  int* tuples;
};

__device__ inline int find_sorted(const uint64_t* a, int n, uint64_t key) {
  int p = lower_bound_u64(a, n, key);
  return (p < n && a[p] == key) ? p : -1;
}

// fold (extent, stride) components: drop extent-1, merge contiguous runs;
// returns the component count (components beyond 2 are reported, not kept)
__device__ inline int fold2(int n, const int* ext, const int64_t* st, int& e0, int64_t& s0,
                            int64_t& s1) {
  int cnt = 0;
  int64_t ce = 1, cs = 0;
  e0 = 1; s0 = 0; s1 = 0;
  int ncomp = 0;
  int64_t ce0 = 0;
  for (int j = 0; j < n; j++) {
    if (ext[j] == 1) continue;
    if (cnt > 0 && st[j] == ce * cs) {
      ce *= ext[j];
      continue;
    }
    if (cnt > 0) {  // close previous
      if (ncomp == 0) { ce0 = ce; s0 = cs; }
      else if (ncomp == 1) { s1 = cs; }
      ncomp++;
    }
    cnt = 1;
    ce = ext[j];
    cs = st[j];
  }
  if (cnt > 0) {
    if (ncomp == 0) { ce0 = ce; s0 = cs; }
    else if (ncomp == 1) { s1 = cs; }
    ncomp++;
  }
  e0 = ncomp >= 1 ? (int)ce0 : 1;
  return ncomp;
}

// K fold with two operands sharing the contraction index order
__device__ inline int foldk(int n, const int* ext, const int64_t* sa, const int64_t* sb, int& e0,
                            int64_t& a0, int64_t& a1, int64_t& b0, int64_t& b1) {
  int ncomp = 0, open = 0;
  int64_t ce = 1, ca = 0, cb = 0, ce0 = 1;
  a0 = a1 = b0 = b1 = 0;
  for (int j = 0; j < n; j++) {
    if (ext[j] == 1) continue;
    if (open && sa[j] == ce * ca && sb[j] == ce * cb) {
      ce *= ext[j];
      continue;
    }
    if (open) {
      if (ncomp == 0) { ce0 = ce; a0 = ca; b0 = cb; }
      else if (ncomp == 1) { a1 = ca; b1 = cb; }
      ncomp++;
    }
    open = 1;
    ce = ext[j]; ca = sa[j]; cb = sb[j];
  }
  if (open) {
    if (ncomp == 0) { ce0 = ce; a0 = ca; b0 = cb; }
    else if (ncomp == 1) { a1 = ca; b1 = cb; }
    ncomp++;
  }
  e0 = ncomp >= 1 ? (int)ce0 : 1;
  return ncomp;
}

// One wave per result block.  The block's tasks are its mid tuples in
// ascending order (qtensor.cpp:696-722); the lanes take 64 mids at a time,
// and each task's position in the list and the k-tile offset of its walk
// are exclusive prefix sums over the lanes in mid order, so the task list is
// the one a walk over the mids in order gives.
__device__ inline int64_t wave_sum_i64(int64_t v) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}
__device__ inline double wave_sum_f64(double v) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}
__device__ inline int64_t wave_max_i64(int64_t v) {
  for (int d = 32; d >= 1; d >>= 1) {
    const int64_t o = __shfl_xor(v, d);
    v = o > v ? o : v;
  }
  return v;
}

__global__ void __launch_bounds__(256) k_plan_blocks(MetaView A, MetaView B, PairSpec ps,
                                                     KeyArrays k, const uint64_t* uL,
                                                     const uint64_t* uR, const uint64_t* uM,
                                                     QTab q, int nQ, int nblocksC, int phase,
                                                     BlkArrays ba, EmitArrays ea, PlanHdr* hdr,
                                                     int nA, int nB, int TK) {
  const int lane = threadIdx.x & 63;
  const int b = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
  if (b >= nblocksC) return;   // the whole wave
  // q group containing b: last t with qbase[t] <= b
  int lo = 0, hi = nQ;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (q.qbase[mid] <= b) lo = mid + 1;
    else hi = mid;
  }
  const int t = lo - 1;
  const int64_t local = b - q.qbase[t];
  const int nL = q.qL1[t] - q.qL0[t];
  const int ri = (int)(local / nL), li = (int)(local % nL);
  const uint64_t left = uL[q.qL0[t] + li] & 0xffffffffull;
  const uint64_t right = uR[q.qR0[t] + ri] & 0xffffffffull;
  // decode tuples (first position most significant)
  int lq[MAXR], rq[MAXR];
  {
    uint64_t x = left;
    for (int j = ps.nfa - 1; j >= 0; j--) { int p = ps.fa[j]; lq[j] = (int)(x % A.nsec[p]); x /= A.nsec[p]; }
    x = right;
    for (int j = ps.nfb - 1; j >= 0; j--) { int p = ps.fb[j]; rq[j] = (int)(x % B.nsec[p]); x /= B.nsec[p]; }
  }
  const int rankC = ps.nfa + ps.nfb;
  int64_t M = 1, N = 1;
  for (int j = 0; j < ps.nfa; j++) M *= A.idx[3 * (A.idx_off[ps.fa[j]] + lq[j]) + 2];
  for (int j = 0; j < ps.nfb; j++) N *= B.idx[3 * (B.idx_off[ps.fb[j]] + rq[j]) + 2];
  if (phase == 0 && lane == 0) {
    for (int j = 0; j < ps.nfa; j++) ba.c_qi[(size_t)b * rankC + j] = lq[j];
    for (int j = 0; j < ps.nfb; j++) ba.c_qi[(size_t)b * rankC + ps.nfa + j] = rq[j];
    ba.size[b] = M * N;
  }
  const int64_t tw = phase == 1 ? ba.task_off[b] : 0;
  int64_t nt = 0;   // tasks of the block so far (uniform)
  int64_t kt = 0;   // k-tiles of the block's walk so far (uniform)
  // per-lane partial sums and maxima, reduced over the wave at the end
  double fl = 0.0;
  int64_t sumK = 0;
  int64_t votes[4] = {0, 0, 0, 0};
  int64_t ktmax = 0, opmax = 0;
  int64_t bad = 0;
  for (int m0 = q.qM0[t]; m0 < q.qM1[t]; m0 += 64) {
    const int m = m0 + lane;
    bool valid = false;
    GemmTask g{};
    int64_t K = 1, ktl = 0;
    int ablk = 0, bblk = 0;
    if (m < q.qM1[t]) {
      const uint64_t mid = uM[m] & 0xffffffffull;
      const int ia = find_sorted(k.sSA, nA, (left << 32) | mid);
      const int ib = ia < 0 ? -1 : find_sorted(k.sSB, nB, (mid << 32) | right);
      if (ib >= 0) {
        valid = true;
        ablk = k.svSA[ia];
        bblk = k.svSB[ib];
        // strides of the stored (unpermuted) blocks
        int dA[MAXR], dB[MAXR];
        int64_t sA[MAXR], sB[MAXR];
        int64_t acc = 1;
        for (int j = 0; j < A.rank; j++) {
          dA[j] = A.idx[3 * (A.idx_off[j] + A.qi[(size_t)ablk * A.rank + j]) + 2];
          sA[j] = acc;
          acc *= dA[j];
        }
        acc = 1;
        for (int j = 0; j < B.rank; j++) {
          dB[j] = B.idx[3 * (B.idx_off[j] + B.qi[(size_t)bblk * B.rank + j]) + 2];
          sB[j] = acc;
          acc *= dB[j];
        }
        int em[MAXR], ek[MAXR], en[MAXR];
        int64_t smA[MAXR], skA[MAXR], skB[MAXR], snB[MAXR];
        for (int j = 0; j < ps.nfa; j++) { em[j] = dA[ps.fa[j]]; smA[j] = sA[ps.fa[j]]; }
        for (int p = 0; p < ps.np; p++) {
          ek[p] = dB[ps.pb[p]];
          skA[p] = sA[ps.pa[p]];
          skB[p] = sB[ps.pb[p]];
          K *= ek[p];
        }
        for (int j = 0; j < ps.nfb; j++) { en[j] = dB[ps.fb[j]]; snB[j] = sB[ps.fb[j]]; }
        g.a_off = A.off[ablk];
        g.b_off = B.off[bblk];
        g.K = (int)K;
        int cm = fold2(ps.nfa, em, smA, g.em0, g.sam0, g.sam1);
        int cn = fold2(ps.nfb, en, snB, g.en0, g.sbn0, g.sbn1);
        int ck = foldk(ps.np, ek, skA, skB, g.ek0, g.sak0, g.sak1, g.sbk0, g.sbk1);
        {  // reversed contraction order: taken when it folds better, or when it
           // puts the longer component first (the GEMM walks k in chunks of the
           // first component, so a short first component wastes chunk slots)
          int ekr[MAXR];
          int64_t skAr[MAXR], skBr[MAXR];
          for (int p = 0; p < ps.np; p++) {
            ekr[p] = ek[ps.np - 1 - p];
            skAr[p] = skA[ps.np - 1 - p];
            skBr[p] = skB[ps.np - 1 - p];
          }
          int e0r;
          int64_t a0r, a1r, b0r, b1r;
          const int ckr = foldk(ps.np, ekr, skAr, skBr, e0r, a0r, a1r, b0r, b1r);
          if ((ck > 2 && ckr <= 2) || (ck == 2 && ckr == 2 && e0r > g.ek0)) {
            ck = ckr;
            g.ek0 = e0r; g.sak0 = a0r; g.sak1 = a1r; g.sbk0 = b0r; g.sbk1 = b1r;
          }
        }
        if (cm > 2 || cn > 2 || ck > 2) bad = 1;
        ktl = (int64_t)(K / g.ek0) * (int64_t)((g.ek0 + TK - 1) / TK);
        if (phase == 0) {
          votes[0] += (g.sam0 == 1) || M == 1;
          votes[1] += (g.sak0 == 1) || K == 1;
          votes[2] += (g.sbk0 == 1) || K == 1;
          votes[3] += (g.sbn0 == 1) || N == 1;
        }
        ktmax = K > ktmax ? K : ktmax;
        opmax = M * K > opmax ? M * K : opmax;
        opmax = K * N > opmax ? K * N : opmax;
        // K walk in TK-wide chunks of the first k component (gemm.hip)
        sumK += (int64_t)(K / g.ek0) * (int64_t)((g.ek0 + TK - 1) / TK * TK);
        fl += 2.0 * (double)M * (double)N * (double)K;
      }
    }
    // position and k-tile offset: exclusive prefix sums over the lanes
    const uint64_t vm = __ballot(valid);
    const int pos = __popcll(vm & ((1ull << lane) - 1ull));
    int64_t incl = valid ? ktl : 0;
    for (int d = 1; d < 64; d <<= 1) {
      const int64_t y = __shfl_up(incl, d);
      if (lane >= d) incl += y;
    }
    if (valid && phase == 1) {
      g.kt0 = (int)(kt + incl - ktl);
      g.nk1 = (int)(K / g.ek0);
      ea.tasks[tw + nt + pos] = g;
      if (ea.tuples) {
        int* tu = ea.tuples + 6 * (tw + nt + pos);
        tu[0] = b; tu[1] = ablk; tu[2] = bblk; tu[3] = (int)M; tu[4] = (int)N; tu[5] = (int)K;
      }
    }
    nt += __popcll(vm);
    kt += __shfl(incl, 63);
  }
  if (phase == 0) {
    fl = wave_sum_f64(fl);
    sumK = wave_sum_i64(sumK);
    for (int v = 0; v < 4; v++) votes[v] = wave_sum_i64(votes[v]);
    ktmax = wave_max_i64(ktmax);
    opmax = wave_max_i64(opmax);
    bad = wave_max_i64(bad);
    if (lane != 0) return;
    ba.ntask[b] = nt;
    ba.flops[b] = fl;
    for (int c = 0; c < TILE_NCFG; c++) {
      const int TMc = cfg_tm(c), TNc = cfg_tn(c);
      int64_t tm = (M + TMc - 1) / TMc, tn = (N + TNc - 1) / TNc;
      atomicAdd((unsigned long long*)&hdr->tiles[c], (unsigned long long)(tm * tn));
      atomicAdd(&hdr->padded[c], (double)(tm * TMc) * (double)(tn * TNc) * (double)sumK);
    }
    atomicMax((unsigned long long*)&hdr->kmax, (unsigned long long)sumK);
    if (nt) {
      atomicMax((unsigned long long*)&hdr->nmax, (unsigned long long)N);
      atomicMax((unsigned long long*)&hdr->ktmax, (unsigned long long)ktmax);
      atomicMax((unsigned long long*)&hdr->opmax, (unsigned long long)opmax);
    }
    for (int v = 0; v < 4; v++)
      if (votes[v]) atomicAdd((unsigned long long*)&hdr->votes[v], (unsigned long long)votes[v]);
    if (bad) atomicAdd((unsigned long long*)&hdr->err, 1ull);
  } else if (lane == 0) {
    CBlock cb;
    cb.c_off = ba.c_off[b];
    cb.M = (int)M;
    cb.N = (int)N;
    cb.t0 = (int)tw;
    cb.t1 = (int)(tw + nt);
    cb.nkt = (int)kt;
    cb.pad = 0;
    ea.cblk[b] = cb;
  }
}

// ---- K5: scans over result blocks ---------------------------------------
__global__ void __launch_bounds__(1024) k_plan_scan(int n, const int64_t* size, const int64_t* ntask,
                                                    const double* flops, int64_t* c_off,
                                                    int64_t* task_off, PlanHdr* hdr) {
  __shared__ int64_t lds[32];
  __shared__ double red[16];
  int64_t cs = 0, ct = 0;
  double fsum = 0.0;
  for (int base = 0; base < n; base += 1024) {
    int i = base + threadIdx.x;
    int64_t s = i < n ? size[i] : 0, t = i < n ? ntask[i] : 0;
    double f = i < n ? flops[i] : 0.0;
    int64_t ts, tt;
    int64_t ps = block_excl_scan<1024>(s, lds, &ts);
    int64_t pt = block_excl_scan<1024>(t, lds, &tt);
    if (i < n) {
      c_off[i] = cs + ps;
      task_off[i] = ct + pt;
    }
    fsum += block_sum<1024>(f, red);
    cs += ts;
    ct += tt;
  }
  if (threadIdx.x == 0) {
    c_off[n] = cs;
    task_off[n] = ct;
    hdr->totalC = cs;
    hdr->ntasks = ct;
    hdr->flops = fsum;
  }
}

}  // namespace ttg

ttg_plan_s::~ttg_plan_s() {
  ttg::dfree(d_tasks);
  ttg::dfree(d_cblk);
  ttg::dfree(d_tiles);   // d_groups lies in the same allocation
  ttg::dfree(d_part);
  ttg::dfree(d_tuples);
}

namespace ttg {

// Tile-shape autotuner (analytic).  For tile shape c a workgroup's k-tile
// costs, per wave (one wave per SIMD):
//   mfma  = fragments x MFMA cycles per k-tile (FP64 f64 16x16x4: 64 cycles,
//           4 per 16 k; FP32 f32 16x16x4: 32; split-BF16: three 16-cycle
//           16x16x32 per 32 k),
//   issue = 4 cycles x the staging / address / LDS instructions of the tile's
//           load slots (about 6 per FP64/FP32 element, 8 per split-BF16
//           element including the hi/lo split) plus fragment loads and loop,
//   lat   = global-load latency (kLatency) shared by the occ workgroups a CU
//           keeps resident (occupancy of the actual kernel instantiation).
// A CU retires one workgroup k-tile per max(mfma, issue, lat / occ) cycles;
// each work item adds its first load and its epilogue.  The launch takes
// max(total work over the CUs, the longest item after split-K run at the
// single-workgroup rate max(mfma, issue, lat)).
// Split-K (build_tiles) cuts the longest walk into at most kMaxSplit pieces.
constexpr int kMinPiece = 8;   // fewest k-tiles one split piece owns
constexpr double kLatency = 2000.0;

static int choose_cfg(const PlanHdr& h, int prec, int ncu, bool am, bool bk) {
  const int TK = tk_of(prec);
  int best = 0;
  double bestc = 1e300;
  for (int c = 0; c < TILE_NCFG; c++) {
    const double TM = cfg_tm(c), TN = cfg_tn(c);
    const double wm = TN == 16 ? 4 : 2, wn = 4 / wm;
    const double fm = TM / (16 * wm), fn = TN / (16 * wn);
    double mfma, issue;
    if (prec == TTG_BF16X3) {
      mfma = fm * fn * 3 * 16;
      const double slots = std::ceil(TM * TK / 8 / 256.0) + std::ceil(TN * TK / 8 / 256.0);
      issue = 4.0 * (slots * 8 * 8 + 4 * (fm + fn) + 24);
    } else {
      mfma = fm * fn * (TK / 4) * (prec == TTG_FP64 ? 64.0 : 32.0);
      const double slots = (TM + TN) * TK / 256.0;
      issue = 4.0 * (6 * slots + 2 * (fm + fn) + 24);
    }
    const double occ = gemm_occupancy(prec, c, am, bk);
    const double per_ktile = std::max(std::max(mfma, issue), kLatency / occ);
    const double single = std::max(std::max(mfma, issue), kLatency);
    const double ktiles = h.padded[c] / (TM * TN * (double)TK);
    const double item = kLatency + TM * TN / 256.0 * 16.0;   // first load + epilogue stores
    const double work = (ktiles * per_ktile + (double)h.tiles[c] * item / occ) / ncu;
    const double kwalk = (double)h.kmax / TK;
    const double longest = std::max(kwalk / kMaxSplit, std::min(kwalk, (double)kMinPiece));
    const double critical = longest * single + item;
    const double cost = std::max(work, critical);
    if (cost < bestc) {
      bestc = cost;
      best = c;
    }
  }
  return best;
}

// Work list of a plan.  Every output block is cut into TM x TN tiles; a tile
// whose k walk is longer than `piece` k-tiles is split into S <= kMaxSplit
// consecutive k ranges, each written to its own partial slot and summed in
// slot order by the reduction pass (deterministic).  `piece` is the whole
// walk over 6 x the CU count, so the longest item is short against the
// launch while the partial tiles (written, then read by the reduction) stay
// few: at 8 x the CU count the m=2048 operator's FP64 stage 3 spent ~30 us of
// 116 us on them, at 6 x it runs in 88 us (whole matvec 0.253 ms against
// 0.257-0.261 ms at 4 x and 5 x and 0.281 ms at 8 x; the m=8192 operator is
// unchanged within 2%).  Items keep output-block order.
int build_tiles(ttg_plan_s& p, const std::vector<CBlock>& cb) {
  const int TM = cfg_tm(p.cfg), TN = cfg_tn(p.cfg);
  const int64_t G = 3 * (int64_t)rt().num_cu;
  int64_t work = 0;
  for (const CBlock& b : cb)
    work += (int64_t)((b.M + TM - 1) / TM) * ((b.N + TN - 1) / TN) * std::max(b.nkt, 1);
  const int64_t piece = std::max<int64_t>(kMinPiece, (work + 2 * G - 1) / (2 * G));
  std::vector<Tile> tiles;
  std::vector<SplitGroup> groups;
  int64_t nslots = 0;
  for (size_t c = 0; c < cb.size(); c++) {
    const CBlock& b = cb[c];
    if (p.cfg == TILE_SKINNY) {   // row chunks, whole walk, no split
      for (int m0 = 0; m0 < b.M; m0 += TM) tiles.push_back(Tile{(int)c, m0, 0, 0, b.nkt, -1});
      continue;
    }
    const int S = (int)std::min<int64_t>(kMaxSplit, std::max<int64_t>(1, (b.nkt + piece - 1) / piece));
    for (int m0 = 0; m0 < b.M; m0 += TM)
      for (int n0 = 0; n0 < b.N; n0 += TN) {
        if (S == 1) {
          tiles.push_back(Tile{(int)c, m0, n0, 0, b.nkt, -1});
          continue;
        }
        groups.push_back(SplitGroup{(int)c, m0, n0, (int)nslots, S, 0});
        for (int sp = 0; sp < S; sp++)
          tiles.push_back(Tile{(int)c, m0, n0, (int)((int64_t)b.nkt * sp / S),
                               (int)((int64_t)b.nkt * (sp + 1) / S), (int)(nslots + sp)});
        nslots += S;
      }
  }
  TTG_REQUIRE(tiles.size() < (size_t)INT32_MAX && nslots < INT32_MAX, "plan: work list too long");
  // Items stay in output-block creation order (q -> right -> left): the
  // tiles of one block share A rows, and consecutive blocks share B blocks,
  // so consecutive items (one XCD's run, see xcd_first_item) reuse its L2.
  // A longest-walk-first order balanced the tail no better and ran the
  // m=2048 and m=8192 operators 1-4% slower.
  p.ntiles = (int)tiles.size();
  p.ngroups = (int)groups.size();
  p.nslots = nslots;
  // tiles and reduction groups share one allocation and one upload
  const size_t tbytes = sizeof(Tile) * tiles.size(), goff = (tbytes + 255) & ~size_t(255);
  const size_t gbytes = sizeof(SplitGroup) * groups.size();
  p.d_tiles = (Tile*)dmalloc(goff + std::max<size_t>(gbytes, sizeof(SplitGroup)));
  p.d_groups = p.d_tiles ? (SplitGroup*)((char*)p.d_tiles + goff) : nullptr;
  const size_t acc = p.prec == TTG_FP64 ? sizeof(double) : sizeof(float);
  p.d_part = nslots ? dmalloc((size_t)nslots * TM * TN * acc) : nullptr;
  TTG_REQUIRE(p.d_tiles && (p.d_part || !nslots), "plan: work list allocation failed");
  const H2DPiece pc[2] = {{tiles.data(), 0, tbytes}, {groups.data(), goff, gbytes}};
  TTG_TRY(h2d_pieces(p.d_tiles, pc, 2));
  return 0;
}

// replace a plan's work list by one for tile shape cfg (an MFMA shape)
int rebuild_tiles(ttg_plan_s& p, int cfg) {
  TTG_REQUIRE(cfg >= 0 && (cfg < TILE_NCFG || cfg == TILE_SKINNY), "plan: bad tile shape");
  if ((int)p.hcb.size() != p.nblocksC) {
    p.hcb.resize(p.nblocksC);
    const D2HPiece pc{p.hcb.data(), p.d_cblk, sizeof(CBlock) * p.nblocksC};
    TTG_TRY(d2h_wait(&pc, 1));
  }
  dfree(p.d_tiles); dfree(p.d_part);   // d_groups lies in d_tiles' allocation
  p.d_tiles = nullptr; p.d_groups = nullptr; p.d_part = nullptr;
  p.ngroups = 0;
  p.nslots = 0;
  p.cfg = cfg;
  p.tiles_pending = false;
  return build_tiles(p, p.hcb);
}

// k-tile offsets of every task of a block's k walk, and the block's k-tile
// count, for k-tile depth TK (one thread per output block)
__global__ void __launch_bounds__(256) k_plan_ktiles(int nblocks, CBlock* cblk, GemmTask* tasks, int TK) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  const CBlock cb = cblk[b];
  int kt = 0;
  for (int t = cb.t0; t < cb.t1; t++) {
    const GemmTask& g = tasks[t];
    tasks[t].kt0 = kt;
    kt += (g.K / g.ek0) * ((g.ek0 + TK - 1) / TK);
  }
  cblk[b].nkt = kt;
}

int plan_clone(const ttg_plan_s& src, int prec, ttg_plan* out) {
  TTG_REQUIRE(prec == TTG_FP64 || prec == TTG_FP32 || prec == TTG_BF16X3, "unknown precision");
  std::unique_ptr<ttg_plan_s> p(new ttg_plan_s());
  p->A = src.A;
  p->B = src.B;
  p->C = src.C;
  p->ps = src.ps;
  p->prec = prec;
  p->nblocksC = src.nblocksC;
  p->ntasks = src.ntasks;
  p->flops = src.flops;
  p->cfg = src.cfg;   // packed_in / split_out start false: the new operator sets them
  p->a_mcontig = src.a_mcontig;
  p->b_kcontig = src.b_kcontig;
  const int64_t nt = std::max<int64_t>(1, p->ntasks);
  const int64_t nb = std::max<int64_t>(1, p->nblocksC);
  p->d_tasks = dalloc<GemmTask>(nt);
  p->d_cblk = dalloc<CBlock>(nb);
  p->d_tuples = dalloc<int>(6 * nt);
  TTG_REQUIRE(p->d_tasks && p->d_cblk && p->d_tuples, (std::string("plan: ") + ttg_last_error()).c_str());
  hipStream_t s = stream();
  TTG_HIP(hipMemcpyAsync(p->d_tasks, src.d_tasks, sizeof(GemmTask) * nt, hipMemcpyDeviceToDevice, s));
  TTG_HIP(hipMemcpyAsync(p->d_cblk, src.d_cblk, sizeof(CBlock) * nb, hipMemcpyDeviceToDevice, s));
  TTG_HIP(hipMemcpyAsync(p->d_tuples, src.d_tuples, sizeof(int) * 6 * nt, hipMemcpyDeviceToDevice, s));
  if (tk_of(prec) != tk_of(src.prec) && p->nblocksC > 0)
    hipLaunchKernelGGL(k_plan_ktiles, dim3((p->nblocksC + 255) / 256), dim3(256), 0, s, p->nblocksC,
                       p->d_cblk, p->d_tasks, tk_of(prec));
  if (tk_of(prec) == tk_of(src.prec) && (int)src.hcb.size() == p->nblocksC) {
    p->hcb = src.hcb;   // the same k-tile counts
  } else {
    p->hcb.resize(p->nblocksC);
    const D2HPiece pc{p->hcb.data(), p->d_cblk, sizeof(CBlock) * (size_t)p->nblocksC};
    TTG_TRY(d2h_wait(&pc, 1));
  }
  TTG_TRY(build_tiles(*p, p->hcb));
  *out = p.release();
  return 0;
}

// pinned host copy of a plan header, allocated once for the process: plans
// are built one at a time, and every read of it follows a stream sync
static PlanHdr* pinned_hdr() {
  static PlanHdr* h = nullptr;
  if (!h && hipHostMalloc((void**)&h, sizeof(PlanHdr)) != hipSuccess) {
    (void)hipGetLastError();
    h = nullptr;
  }
  return h;
}

int plan_build(const MetaPtr& A, const MetaPtr& B, int np, const int* pa, const int* pb, int prec,
               int force_cfg, ttg_plan* out, bool defer_tiles) {
  PhaseTimer _phase("plan_build");
  TTG_TRY(ensure_init());
  TTG_REQUIRE(np >= 0 && np <= A->rank && np <= B->rank, "plan: bad pair count");
  PairSpec ps{};
  ps.np = np;
  std::vector<bool> usedA(A->rank, false), usedB(B->rank, false);
  for (int p = 0; p < np; p++) {
    TTG_REQUIRE(pa[p] >= 0 && pa[p] < A->rank && pb[p] >= 0 && pb[p] < B->rank,
                "plan: pair position out of range");
    TTG_REQUIRE(!usedA[pa[p]] && !usedB[pb[p]], "plan: index paired twice");
    usedA[pa[p]] = usedB[pb[p]] = true;
    ps.pa[p] = pa[p];
    ps.pb[p] = pb[p];
    // the pair must be the same index seen from both sides (tag match)
    TTG_REQUIRE(A->nsec[pa[p]] == B->nsec[pb[p]], "plan: paired indices differ in sectors");
    for (int s = 0; s < A->nsec[pa[p]]; s++)
      TTG_REQUIRE(A->q0(pa[p], s) == B->q0(pb[p], s) && A->q1(pa[p], s) == B->q1(pb[p], s) &&
                      A->dim(pa[p], s) == B->dim(pb[p], s),
                  "plan: paired indices differ in sectors");
  }
  for (int j = 0; j < A->rank; j++) if (!usedA[j]) ps.fa[ps.nfa++] = j;
  for (int j = 0; j < B->rank; j++) if (!usedB[j]) ps.fb[ps.nfb++] = j;
  TTG_REQUIRE(ps.nfa + ps.nfb <= MAXR, "plan: result rank exceeds 8");
  auto radix_ok = [](const TMeta& m, const int* pos, int n) {
    double r = 1;
    for (int j = 0; j < n; j++) r *= m.nsec[pos[j]];
    return r < 4294967295.0;
  };
  TTG_REQUIRE(radix_ok(*A, ps.fa, ps.nfa) && radix_ok(*B, ps.fb, ps.nfb) && radix_ok(*A, ps.pa, np),
              "plan: sector-tuple key exceeds 32 bits");

  const int nA = A->nblocks, nB = B->nblocks;
  hipStream_t s = stream();
  // scratch (returned to the pool on every exit path)
  Scratch sc;
  KeyArrays k{};
  const int nmax = std::max(1, std::max(nA, nB));
  k.kUL = sc.get<uint64_t>(nmax); k.kSA = sc.get<uint64_t>(nmax); k.kUR = sc.get<uint64_t>(nmax);
  k.kUM = sc.get<uint64_t>(nmax); k.kSB = sc.get<uint64_t>(nmax);
  k.sUL = sc.get<uint64_t>(nmax); k.sSA = sc.get<uint64_t>(nmax); k.sUR = sc.get<uint64_t>(nmax);
  k.sUM = sc.get<uint64_t>(nmax); k.sSB = sc.get<uint64_t>(nmax);
  k.vSA = sc.get<int>(nmax); k.vSB = sc.get<int>(nmax); k.svSA = sc.get<int>(nmax); k.svSB = sc.get<int>(nmax);
  uint64_t* uL = sc.get<uint64_t>(nmax);
  uint64_t* uR = sc.get<uint64_t>(nmax);
  uint64_t* uM = sc.get<uint64_t>(nmax);
  QTab q{};
  q.qL0 = sc.get<int>(nmax); q.qL1 = sc.get<int>(nmax); q.qR0 = sc.get<int>(nmax);
  q.qR1 = sc.get<int>(nmax); q.qM0 = sc.get<int>(nmax); q.qM1 = sc.get<int>(nmax);
  q.qbase = sc.get<int64_t>(nmax + 1);
  PlanHdr* d_hdr = sc.get<PlanHdr>(1);
  TTG_TRY(sc.ok("plan"));
  PlanHdr* h_hdr = pinned_hdr();
  TTG_REQUIRE(h_hdr, "plan: pinned host header allocation failed");

  hipLaunchKernelGGL(k_plan_keys, dim3((nmax + 255) / 256), dim3(256), 0, s, A->view(), B->view(),
                     ps, k, d_hdr);
  SortJobs jobs;
  jobs.j[0] = {k.kUL, nullptr, k.sUL, nullptr, nA};
  jobs.j[1] = {k.kSA, k.vSA, k.sSA, k.svSA, nA};
  jobs.j[2] = {k.kUR, nullptr, k.sUR, nullptr, nB};
  jobs.j[3] = {k.kUM, nullptr, k.sUM, nullptr, nB};
  jobs.j[4] = {k.kSB, k.vSB, k.sSB, k.svSB, nB};
  hipLaunchKernelGGL(k_rank_sort, dim3((nmax + 255) / 256, 5), dim3(256), 0, s, jobs);
  hipLaunchKernelGGL(k_plan_unique, dim3(1), dim3(1024), 0, s, k, nA, nB, uL, uR, uM, q, d_hdr);
  TTG_HIP(hipMemcpyAsync(h_hdr, d_hdr, sizeof(PlanHdr), hipMemcpyDeviceToHost, s));
  TTG_HIP(hipStreamSynchronize(s));
  const int64_t nblocksC = h_hdr->nblocksC;
  const int nQ = (int)h_hdr->nQ;
  const int64_t ncand = h_hdr->ncand;
  TTG_REQUIRE(nblocksC < (1 << 30), "plan: too many result blocks");
  TTG_REQUIRE(ncand < (1 << 30), "plan: too many block pairs");

  // result structure
  std::vector<std::pair<const TMeta*, int>> cidx;
  for (int j = 0; j < ps.nfa; j++) cidx.push_back({A.get(), ps.fa[j]});
  for (int j = 0; j < ps.nfb; j++) cidx.push_back({B.get(), ps.fb[j]});
  MetaPtr C = meta_from_indices(cidx);
  if (!C) return 1;
  TTG_TRY(meta_alloc_blocks(*C, (int)nblocksC, 0));
  BlkArrays ba{};
  ba.size = sc.get<int64_t>(nblocksC + 1);
  ba.ntask = sc.get<int64_t>(nblocksC + 1);
  ba.flops = sc.get<double>(nblocksC + 1);
  int64_t* task_off = sc.get<int64_t>(nblocksC + 1);
  TTG_TRY(sc.ok("plan"));
  ba.c_qi = C->d_qi;
  EmitArrays ea{};
  const unsigned gb = (unsigned)std::max<int64_t>(1, (nblocksC + 3) / 4);   // one wave per result block
  // the task list is emitted right after the count pass, into arrays sized
  // by the candidate count (an upper bound of the task count), so one
  // read-back (header and block table) serves both passes
  std::unique_ptr<ttg_plan_s> p(new ttg_plan_s());
  const int64_t ncap = std::max<int64_t>(1, ncand);
  p->d_tasks = dalloc<GemmTask>(ncap);
  p->d_cblk = dalloc<CBlock>(std::max<int64_t>(1, nblocksC));
  p->d_tuples = dalloc<int>(6 * ncap);
  if (!p->d_tasks || !p->d_cblk || !p->d_tuples) {
    set_error(std::string("plan: ") + ttg_last_error());
    return 1;
  }
  hipLaunchKernelGGL(k_plan_blocks, dim3(gb), dim3(256), 0, s, A->view(), B->view(), ps, k, uL, uR,
                     uM, q, nQ, (int)nblocksC, 0, ba, ea, d_hdr, nA, nB, tk_of(prec));
  hipLaunchKernelGGL(k_plan_scan, dim3(1), dim3(1024), 0, s, (int)nblocksC, ba.size, ba.ntask,
                     ba.flops, C->d_off, task_off, d_hdr);
  ba.task_off = task_off;
  ba.c_off = C->d_off;
  ea.tasks = p->d_tasks;
  ea.cblk = p->d_cblk;
  ea.tuples = p->d_tuples;
  hipLaunchKernelGGL(k_plan_blocks, dim3(gb), dim3(256), 0, s, A->view(), B->view(), ps, k, uL, uR,
                     uM, q, nQ, (int)nblocksC, 1, ba, ea, d_hdr, nA, nB, tk_of(prec));
  p->hcb.resize(nblocksC);
  {
    const D2HPiece pc[2] = {{h_hdr, d_hdr, sizeof(PlanHdr)},
                            {p->hcb.data(), p->d_cblk, sizeof(CBlock) * (size_t)nblocksC}};
    TTG_TRY(d2h_wait(pc, 2));
  }
  TTG_REQUIRE(!h_hdr->err, "plan: operand layout needs more than two strides per GEMM dimension");
  TTG_REQUIRE(h_hdr->ntasks <= ncand, "plan: task count above its bound");
  C->total = (size_t)h_hdr->totalC;
  // the GEMM's 32-bit byte offsets per block
  TTG_REQUIRE(h_hdr->opmax < (int64_t(1) << 28), "plan: an operand block exceeds 2^28 elements");

  p->A = A;
  p->B = B;
  p->C = C;
  p->ps = ps;
  p->prec = prec;
  p->nblocksC = (int)nblocksC;
  p->ntasks = h_hdr->ntasks;
  p->flops = h_hdr->flops;
  p->a_mcontig = h_hdr->votes[0] >= h_hdr->votes[1];
  p->b_kcontig = h_hdr->votes[2] >= h_hdr->votes[3];
  if (force_cfg < 0) {   // TTG_TILE_CFG=<n> pins every plan's tile shape (tuning runs)
    const char* e = getenv("TTG_TILE_CFG");
    if (e && *e) force_cfg = atoi(e);
  }
  const bool skinny_ok = h_hdr->nmax <= kSkinnyN && h_hdr->ktmax <= kSkinnyK;
  if (force_cfg == TILE_SKINNY && skinny_ok) p->cfg = TILE_SKINNY;
  else if (force_cfg >= 0 && force_cfg < TILE_NCFG) p->cfg = force_cfg;
  else p->cfg = skinny_ok ? TILE_SKINNY
                          : choose_cfg(*h_hdr, prec, rt().num_cu, p->a_mcontig, p->b_kcontig);
  if (defer_tiles) p->tiles_pending = true;
  else TTG_TRY(build_tiles(*p, p->hcb));
  *out = p.release();
  return 0;
}

}  // namespace ttg

using namespace ttg;

extern "C" {

int ttg_contract_plan(ttg_tensor A, ttg_tensor B, int npairs, const int* pa, const int* pb, int prec,
                      ttg_plan* out) {
  TTG_REQUIRE(A && B && out, "contract_plan: null argument");
  TTG_REQUIRE(prec == TTG_FP64 || prec == TTG_FP32 || prec == TTG_BF16X3, "unknown precision");
  return plan_build(A->m, B->m, npairs, pa, pb, prec, -1, out);
}

int ttg_plan_info(ttg_plan p, int* nblocks, int64_t* ntasks, double* flops) {
  TTG_REQUIRE(p, "null plan");
  if (nblocks) *nblocks = p->nblocksC;
  if (ntasks) *ntasks = p->ntasks;
  if (flops) *flops = p->flops;
  return 0;
}

int ttg_plan_tasks_download(ttg_plan p, int* tasks6) {
  TTG_REQUIRE(p, "null plan");
  if (p->ntasks)
    TTG_HIP(hipMemcpyAsync(tasks6, p->d_tuples, sizeof(int) * 6 * (size_t)p->ntasks,
                           hipMemcpyDeviceToHost, stream()));
  TTG_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int ttg_plan_new_output(ttg_plan p, ttg_tensor* C) {
  TTG_REQUIRE(p, "null plan");
  return tensor_from_meta(p->C, false, C);
}

int ttg_contract_exec(ttg_plan p, ttg_tensor A, ttg_tensor B, ttg_tensor C) {
  TTG_REQUIRE(p && A && B && C, "contract_exec: null argument");
  // the plan's task offsets are only valid for the block lists it was built on
  TTG_REQUIRE(same_meta(A->m, p->A) && same_meta(B->m, p->B) && same_meta(C->m, p->C),
              "contract_exec: tensors do not have the block lists the plan was built on");
  if (p->prec == TTG_FP64) {
    TTG_TRY(gemm_launch(*p, A->d_data, B->d_data, C->d_data, true, stream()));
  } else {
    const float *a32, *b32;
    TTG_TRY(tensor_f32(A, &a32));
    TTG_TRY(tensor_f32(B, &b32));
    TTG_TRY(gemm_launch(*p, a32, b32, C->d_data, true, stream()));
  }
  mark_written(C);
  return 0;
}

int ttg_plan_free(ttg_plan p) {
  delete p;
  return 0;
}

int ttg_contract(ttg_tensor A, ttg_tensor B, int npairs, const int* pa, const int* pb, int prec,
                 ttg_tensor* C) {
  PhaseTimer _phase("ttg_contract");
  ttg_plan p = nullptr;
  TTG_TRY(ttg_contract_plan(A, B, npairs, pa, pb, prec, &p));
  int r = ttg_plan_new_output(p, C);
  if (r == 0) r = ttg_contract_exec(p, A, B, *C);
  delete p;
  return r;
}

}  // extern "C"
