// tt_oracle.cpp -- FP64 CPU restatement of the tensor-tools DMRG hot path.
// TEST INFRASTRUCTURE ONLY (the oracle): never linked into the product.
// Reference file:line citations are to ClarkResearchGroup/tensor-tools.
#include "tt_oracle.hpp"

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>

#ifdef _OPENMP
#include <omp.h>
#include <dlfcn.h>
#endif

namespace orc {

static void die(const char* msg) {
  fprintf(stderr, "oracle: %s\n", msg);
  throw std::runtime_error(msg);
}

std::string qstr(const QN& q) {
  return std::to_string(q.q0) + "|" + std::to_string(q.q1) + "|";
}

// ---------------------------------------------------------------- Index
// qtensor_index.h:35-54 -- addQN appends, then the list is re-sorted by QN
// (std::sort with qnCompare) and the tag rebuilt.
void Index::addQN(QN q, unsigned d) {
  qn.push_back({q, d});
  setTag();
}

void Index::setTag() {
  if (qn.size() > 1)
    std::sort(qn.begin(), qn.end(),
              [](const std::pair<QN, unsigned>& a, const std::pair<QN, unsigned>& b) {
                return a.first < b.first;
              });
  tagbase = name + " " + std::to_string((int)type) + " ";
  for (auto& p : qn)
    tagbase += "qn" + qstr(p.first) + "qdim" + std::to_string(p.second) + " ";
}

// qtensor_index.h:66 -- `_arrow+" "+...` is pointer arithmetic on the
// literal " ": Inward(0) gives " ", Outward(1) gives "".
std::string Index::tag() const {
  return std::string(arrow == Inward ? " " : "") + std::to_string(level) + " " + tagbase;
}
std::string Index::tagNoArrow() const { return std::to_string(level) + " " + tagbase; }

// ---------------------------------------------------------------- QTensor
static std::string key_of(const std::vector<QN>& qns) {
  std::string s;
  for (auto& q : qns) s += qstr(q) + " ";
  return s;
}

void QTensor::clearBlock() {
  rank = (unsigned)idx_set.size();
  block.clear();
  block_index_qn.clear();
  block_index_qd.clear();
  block_index_qi.clear();
  block_id_by_qn_str.clear();
  initted = false;
}

void QTensor::reset(const std::vector<Index>& idx) {
  clearBlock();
  idx_set = idx;
  rank = (unsigned)idx.size();
}

size_t QTensor::total_size() const {
  size_t t = 0;
  for (auto& b : block) t += b.size();
  return t;
}

// qtensor.cpp:302-345 -- enumerate QN-position tuples with index 0 fastest,
// keep those with zero total flux (Inward +, Outward -).
void QTensor::initBlock() {
  if (rank > 0 && !initted) {
    std::vector<unsigned> stride(1, 1);
    for (unsigned i = 0; i < rank; i++) stride.push_back(idx_set[i].size() * stride[i]);
    unsigned n = stride.back();
    for (unsigned i = 0; i < n; i++) {
      std::vector<unsigned> pos(rank);
      std::vector<int> qds(rank);
      std::vector<QN> qns(rank);
      size_t size = 1;
      QN total;
      for (unsigned j = 0; j < rank; j++) {
        pos[j] = (i / stride[j]) % idx_set[j].size();
        qns[j] = idx_set[j].qnAt(pos[j]);
        qds[j] = (int)idx_set[j].qdim(pos[j]);
        size *= (size_t)qds[j];
        if (idx_set[j].arrow == Inward) total += qns[j]; else total -= qns[j];
      }
      if (total == QN()) {
        block.emplace_back(size, 0.0);
        block_id_by_qn_str[key_of(qns)] = (unsigned)block.size() - 1;
        block_index_qn.push_back(std::move(qns));
        block_index_qd.push_back(std::move(qds));
        block_index_qi.push_back(std::move(pos));
      }
    }
  }
  initted = true;
}

void QTensor::setZero() { for (auto& b : block) std::fill(b.begin(), b.end(), 0.0); }
void QTensor::setOne() { for (auto& b : block) std::fill(b.begin(), b.end(), 1.0); }

static inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// uniform [0,1) from the top 53 bits; the product uses the same counter RNG
// for the Davidson re-randomisation path.
static inline double u01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

void QTensor::setRandom(uint64_t seed) {
  uint64_t s = seed;
  for (auto& b : block)
    for (auto& v : b) v = u01(s);
}

QTensor& QTensor::operator*=(double c) {
  for (auto& b : block)
    for (auto& v : b) v *= c;
  return *this;
}

// qtensor.cpp:540-572 -- pair indices whose dagged tag matches.
void pair_indices(const QTensor& A, const QTensor& B,
                  std::vector<std::pair<unsigned, unsigned>>& idx_pair,
                  std::vector<unsigned>& A_perm, std::vector<unsigned>& B_perm) {
  std::unordered_map<std::string, int> marker;
  idx_pair.clear(); A_perm.clear(); B_perm.clear();
  for (unsigned i = 0; i < A.rank; i++) marker[A.idx_set[i].tag()] = (int)i;
  for (unsigned i = 0; i < B.rank; i++) {
    Index tp = B.idx_set[i];
    tp.dag();
    auto it = marker.find(tp.tag());
    if (it == marker.end()) {
      marker[B.idx_set[i].tag()] = (int)i;
    } else {
      idx_pair.push_back({(unsigned)it->second, i});
      marker[tp.tag()] = -1;
      marker[B.idx_set[i].tag()] = -1;
    }
  }
  for (unsigned i = 0; i < A.rank; i++)
    if (marker[A.idx_set[i].tag()] != -1) A_perm.push_back(i);
  for (auto& p : idx_pair) { A_perm.push_back(p.first); B_perm.push_back(p.second); }
  for (unsigned i = 0; i < B.rank; i++)
    if (marker[B.idx_set[i].tag()] != -1) B_perm.push_back(i);
}

std::vector<int> g_last_tasks;   // (c, a, b, M, N, K) of the last operator*

// Host BLAS for the packed block products (the reference's path ends in
// MKL dgemm under CTF): cblas_dgemm from the MKL runtime library when it can
// be loaded (libmkl_rt.so, sequential inside each call, since the caller runs
// one block per OpenMP thread), else the column-axpy loop below.  The
// library in use is reported by orc_blas_name().
namespace {
typedef void (*cblas_dgemm_t)(int, int, int, int, int, int, double, const double*, int, const double*, int,
                              double, double*, int);
struct HostBlas {
  cblas_dgemm_t dgemm = nullptr;
  const char* name = "none (column-axpy loop)";
  HostBlas() {
    const char* off = getenv("ORACLE_NO_BLAS");
    if (off && *off && *off != '0') return;
    for (const char* lib : {"libmkl_rt.so", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"}) {
      void* h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
      if (!h) continue;
      auto f = (cblas_dgemm_t)dlsym(h, "cblas_dgemm");
      auto seq = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
      if (!f) continue;
      if (seq) seq(1);   // MKL_THREADING_SEQUENTIAL
      dgemm = f;
      name = "MKL (libmkl_rt, cblas_dgemm, sequential per block)";
      return;
    }
  }
};
HostBlas& host_blas() {
  static HostBlas b;
  return b;
}
// C (M x N) += A (M x K) * B (K x N), all column-major; false: no BLAS
bool host_dgemm(size_t M, size_t N, size_t K, const double* A, const double* B, double* C) {
  HostBlas& hb = host_blas();
  if (!hb.dgemm || M == 0 || N == 0 || K == 0 || M > (1u << 30) || N > (1u << 30) || K > (1u << 30))
    return false;
  hb.dgemm(102 /* CblasColMajor */, 111 /* CblasNoTrans */, 111, (int)M, (int)N, (int)K, 1.0, A, (int)M, B,
           (int)K, 1.0, C, (int)M);
  return true;
}
}  // namespace

const char* blas_name() { return host_blas().name; }

namespace {
struct Task { unsigned a, b; };
// One matched block pair C += A*B.  Operands keep their stored layout
// (qtensor.cpp:719-721 hands the un-permuted blocks to CTF's einsum).
void block_gemm(double* C, const std::vector<double>& Ab, const std::vector<int>& dA,
                const std::vector<double>& Bb, const std::vector<int>& dB,
                const std::vector<unsigned>& A_perm, const std::vector<unsigned>& B_perm,
                unsigned nfreeA, unsigned npair, unsigned nfreeB) {
  auto strides = [](const std::vector<int>& d) {
    std::vector<size_t> s(d.size());
    size_t acc = 1;
    for (size_t i = 0; i < d.size(); i++) { s[i] = acc; acc *= (size_t)d[i]; }
    return s;
  };
  auto sA = strides(dA), sB = strides(dB);
  // offsets of every m (A-free multi-index, first fastest), k, n
  auto offsets = [](const std::vector<int>& ext, const std::vector<size_t>& st) {
    size_t n = 1;
    for (int e : ext) n *= (size_t)e;
    std::vector<size_t> off(n);
    std::vector<int> ctr(ext.size(), 0);
    for (size_t i = 0; i < n; i++) {
      size_t o = 0;
      for (size_t j = 0; j < ext.size(); j++) o += (size_t)ctr[j] * st[j];
      off[i] = o;
      for (size_t j = 0; j < ext.size(); j++) {
        if (++ctr[j] < ext[j]) break;
        ctr[j] = 0;
      }
    }
    return off;
  };
  std::vector<int> em, ek, en;
  std::vector<size_t> sAm, sAk, sBk, sBn;
  for (unsigned j = 0; j < nfreeA; j++) { em.push_back(dA[A_perm[j]]); sAm.push_back(sA[A_perm[j]]); }
  for (unsigned p = 0; p < npair; p++) {
    ek.push_back(dA[A_perm[nfreeA + p]]);
    sAk.push_back(sA[A_perm[nfreeA + p]]);
    sBk.push_back(sB[B_perm[p]]);
  }
  for (unsigned j = 0; j < nfreeB; j++) { en.push_back(dB[B_perm[npair + j]]); sBn.push_back(sB[B_perm[npair + j]]); }
  auto oAm = offsets(em, sAm), oAk = offsets(ek, sAk), oBk = offsets(ek, sBk), oBn = offsets(en, sBn);
  const size_t M = oAm.size(), K = oAk.size(), N = oBn.size();
  // pack A (M x K) and B (K x N) column-major, then column-axpy GEMM
  std::vector<double> Ap(M * K), Bp(K * N);
  for (size_t k = 0; k < K; k++)
    for (size_t m = 0; m < M; m++) Ap[m + M * k] = Ab[oAm[m] + oAk[k]];
  for (size_t n = 0; n < N; n++)
    for (size_t k = 0; k < K; k++) Bp[k + K * n] = Bb[oBk[k] + oBn[n]];
  if (host_dgemm(M, N, K, Ap.data(), Bp.data(), C)) return;
  for (size_t n = 0; n < N; n++) {
    double* c = C + M * n;
    for (size_t k = 0; k < K; k++) {
      const double b = Bp[k + K * n];
      const double* a = Ap.data() + M * k;
      for (size_t m = 0; m < M; m++) c[m] += a[m] * b;
    }
  }
}
}  // namespace

// qtensor.cpp:523-734
QTensor QTensor::operator*(const QTensor& other) const {
  if (!(initted || other.initted)) die("operator*: both operands uninitialised");
  if (initted && !other.initted) return *this;   // :526-529
  if (other.initted && !initted) return other;   // :530-533
  std::vector<std::pair<unsigned, unsigned>> idx_pair;
  std::vector<unsigned> A_perm, B_perm;
  pair_indices(*this, other, idx_pair, A_perm, B_perm);
  std::vector<Index> A_idx(rank), B_idx(other.rank);
  for (unsigned i = 0; i < rank; i++) A_idx[i] = idx_set[A_perm[i]];
  for (unsigned i = 0; i < other.rank; i++) B_idx[i] = other.idx_set[B_perm[i]];
  std::unordered_map<std::string, unsigned> A_id, B_id;
  for (size_t i = 0; i < block.size(); i++) {
    std::string s;
    for (unsigned j = 0; j < rank; j++) s += qstr(block_index_qn[i][A_perm[j]]) + " ";
    A_id[s] = (unsigned)i;
  }
  for (size_t i = 0; i < other.block.size(); i++) {
    std::string s;
    for (unsigned j = 0; j < other.rank; j++) s += qstr(other.block_index_qn[i][B_perm[j]]) + " ";
    B_id[s] = (unsigned)i;
  }
  const unsigned num_rep = (unsigned)idx_pair.size();
  std::vector<Index> res_idx;
  for (unsigned i = 0; i < rank - num_rep; i++) res_idx.push_back(A_idx[i]);
  for (unsigned i = num_rep; i < other.rank; i++) res_idx.push_back(B_idx[i]);
  QTensor res(res_idx);
  res.initted = true;
  // :609-646 -- per mid-QN sets of left / right / paired position tuples
  std::set<QN> mid_QN_set;
  std::map<QN, std::set<std::vector<unsigned>>> left_qi, mid_qi_m, right_qi;
  for (size_t i = 0; i < block.size(); i++) {
    QN q;
    std::vector<unsigned> l;
    for (unsigned j = 0; j < rank - num_rep; j++) {
      l.push_back(block_index_qi[i][A_perm[j]]);
      if (A_idx[j].arrow == Inward) q += block_index_qn[i][A_perm[j]];
      else q -= block_index_qn[i][A_perm[j]];
    }
    mid_QN_set.insert(q);
    left_qi[q].insert(l);
  }
  for (size_t i = 0; i < other.block.size(); i++) {
    QN q;
    std::vector<unsigned> r, m;
    for (unsigned j = num_rep; j < other.rank; j++) {
      r.push_back(other.block_index_qi[i][B_perm[j]]);
      if (B_idx[j].arrow == Inward) q -= other.block_index_qn[i][B_perm[j]];
      else q += other.block_index_qn[i][B_perm[j]];
    }
    for (unsigned j = 0; j < num_rep; j++) m.push_back(other.block_index_qi[i][B_perm[j]]);
    right_qi[q].insert(r);
    mid_qi_m[q].insert(m);
  }
  // :652-726 -- result blocks in q -> right -> left order; tasks in mid order
  std::vector<std::vector<Task>> tasks;
  for (auto& q : mid_QN_set) {
    const auto& Ls = left_qi[q];
    const auto& Rs = right_qi[q];
    const auto& Ms = mid_qi_m[q];
    for (auto& r : Rs) {
      for (auto& l : Ls) {
        std::vector<unsigned> bqi;
        std::vector<int> bqd;
        std::vector<QN> bqn;
        size_t bsize = 1;
        std::vector<QN> Aq(rank), Bq(other.rank);
        for (size_t i = 0; i < l.size(); i++) {
          bqi.push_back(l[i]);
          bqd.push_back((int)A_idx[i].qdim(l[i]));
          bqn.push_back(A_idx[i].qnAt(l[i]));
          bsize *= (size_t)bqd.back();
          Aq[i] = bqn.back();
        }
        for (size_t i = 0; i < r.size(); i++) {
          bqi.push_back(r[i]);
          bqd.push_back((int)B_idx[num_rep + i].qdim(r[i]));
          bqn.push_back(B_idx[num_rep + i].qnAt(r[i]));
          bsize *= (size_t)bqd.back();
          Bq[num_rep + i] = bqn.back();
        }
        res.block.emplace_back(bsize, 0.0);
        res.block_id_by_qn_str[key_of(bqn)] = (unsigned)res.block.size() - 1;
        res.block_index_qn.push_back(bqn);
        res.block_index_qd.push_back(bqd);
        res.block_index_qi.push_back(bqi);
        tasks.emplace_back();
        for (auto& m : Ms) {
          std::vector<QN> Ac(Aq), Bc(Bq);
          for (unsigned i = 0; i < num_rep; i++) {
            Ac[i + rank - num_rep] = A_idx[i + rank - num_rep].qnAt(m[i]);
            Bc[i] = B_idx[i].qnAt(m[i]);
          }
          auto ait = A_id.find(key_of(Ac));
          auto bit = B_id.find(key_of(Bc));
          if (ait != A_id.end() && bit != B_id.end())
            tasks.back().push_back({ait->second, bit->second});
        }
      }
    }
  }
  const unsigned nfA = rank - num_rep, nfB = other.rank - num_rep;
  const long nres = (long)res.block.size();
  g_last_tasks.clear();
  for (long c = 0; c < nres; c++) {
    long M = 1, N = 1;
    for (unsigned j = 0; j < nfA; j++) M *= res.block_index_qd[c][j];
    for (unsigned j = 0; j < nfB; j++) N *= res.block_index_qd[c][nfA + j];
    for (auto& t : tasks[c]) {
      long K = 1;
      for (unsigned p = 0; p < num_rep; p++) K *= block_index_qd[t.a][A_perm[nfA + p]];
      g_last_tasks.insert(g_last_tasks.end(), {(int)c, (int)t.a, (int)t.b, (int)M, (int)N, (int)K});
    }
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (long c = 0; c < nres; c++) {
    for (auto& t : tasks[c])
      block_gemm(res.block[c].data(), block[t.a], block_index_qd[t.a], other.block[t.b],
                 other.block_index_qd[t.b], A_perm, B_perm, nfA, num_rep, nfB);
  }
  return res;
}

// qtensor.cpp:1095-1119 -- full contraction, matched by key string
double QTensor::contract(const QTensor& A) const {
  double res = 0;
  for (auto& kv : block_id_by_qn_str) {
    auto it = A.block_id_by_qn_str.find(kv.first);
    if (it == A.block_id_by_qn_str.end()) continue;
    const auto& x = block[kv.second];
    const auto& y = A.block[it->second];
    double s = 0;
    for (size_t i = 0; i < x.size(); i++) s += x[i] * y[i];
    res += s;
  }
  return res;
}

// qtensor.cpp:1127-1156 -- this += c*A; blocks missing in this are appended
// in A's key-map iteration order.
void QTensor::add(const QTensor& A, double c) {
  for (auto& kv : A.block_id_by_qn_str) {
    unsigned aid = kv.second;
    auto it = block_id_by_qn_str.find(kv.first);
    if (it != block_id_by_qn_str.end()) {
      auto& x = block[it->second];
      const auto& y = A.block[aid];
      if (x.size() != y.size()) die("add: block size mismatch");
      for (size_t i = 0; i < x.size(); i++) x[i] += c * y[i];
    } else {
      block_index_qn.push_back(A.block_index_qn[aid]);
      block_index_qd.push_back(A.block_index_qd[aid]);
      block_index_qi.push_back(A.block_index_qi[aid]);
      std::vector<double> nb(A.block[aid].size());
      for (size_t i = 0; i < nb.size(); i++) nb[i] = c * A.block[aid][i];
      block.push_back(std::move(nb));
      block_id_by_qn_str[kv.first] = (unsigned)block.size() - 1;
    }
  }
}

// qtensor.cpp:1161-1179
double QTensor::inner_product(const QTensor& A) const { return contract(A); }

// qtensor.cpp:1606-1616
double QTensor::norm() const {
  double r = 0;
  for (auto& b : block)
    for (double v : b) r += v * v;
  return std::sqrt(r);
}

// qtensor.cpp:1638-1647
double QTensor::normalize() {
  double n = norm();
  *this *= (1.0 / n);
  return n;
}

// ---------------------------------------------------------------- BigQTensor
void BigQTensor::addMid(const QTensor* m) { mid = mid * (*m); }

QTensor BigQTensor::product(const QTensor& v) const {
  QTensor res;
  if (L) res = (*L) * v;
  if (res.rank > 0) res = res * mid;
  else res = v * mid;
  if (R) res = res * (*R);
  res.prime(-1);
  return res;
}

double BigQTensor::expec(const QTensor& v) const {
  QTensor r = product(v);
  return r.contract(v);
}

// big_qtensor.cpp:96-118 (size_ is cached, the mid factor re-applied per call)
size_t BigQTensor::size() const {
  if (size_ == 0) {
    size_ = 1;
    if (L) for (auto& ix : L->idx_set) if (ix.level > 0) size_ *= ix.size();
    if (R) for (auto& ix : R->idx_set) if (ix.level > 0) size_ *= ix.size();
  }
  for (auto& ix : mid.idx_set) if (ix.type == Site) size_ *= ix.size();
  return size_;
}

// ---------------------------------------------------------------- DIAG
// Cyclic Jacobi eigensolver (textbook; Golub & Van Loan 8.5) for the small
// symmetric Krylov matrix.  Output matches dsyev(jobz=V) semantics: values
// ascending, A overwritten by column eigenvectors.
void DIAG(int n, double* A, double* evals) {
  std::vector<double> a(A, A + n * n), v(n * n, 0.0);
  for (int i = 0; i < n; i++) v[i + n * i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += a[i + n * j] * a[i + n * j];
        if (i != j) off += a[i + n * j] * a[i + n * j];
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = a[p + n * q];
        if (apq == 0.0) continue;
        double app = a[p + n * p], aqq = a[q + n * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {  // columns p,q
          double akp = a[k + n * p], akq = a[k + n * q];
          a[k + n * p] = c * akp - s * akq;
          a[k + n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {  // rows p,q
          double apk = a[p + n * k], aqk = a[q + n * k];
          a[p + n * k] = c * apk - s * aqk;
          a[q + n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = v[k + n * p], vkq = v[k + n * q];
          v[k + n * p] = c * vkp - s * vkq;
          v[k + n * q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return a[x + n * x] < a[y + n * y]; });
  for (int j = 0; j < n; j++) {
    evals[j] = a[ord[j] + n * ord[j]];
    for (int i = 0; i < n; i++) A[i + n * j] = v[i + n * ord[j]];
  }
}

// ---------------------------------------------------------------- Davidson
// tensor_davidson.cpp:192-347 (ITensor-style, no preconditioner)
double tensor_davidsonIT(const BigQTensor& A, QTensor& x, int m, int max_restart,
                         double tol, char mode, uint64_t* rng, int* n_matvec) {
  (void)m; (void)mode;
  int nmv = 0;
  double x_norm = x.norm();
  x *= 1.0 / x_norm;
  int maxsize = (int)A.size();
  int actual_maxiter = std::min(max_restart, maxsize - 1);
  std::vector<QTensor> v(actual_maxiter + 2), va(actual_maxiter + 2);
  const int S = actual_maxiter + 2;
  std::vector<double> M(S * S, 0.0), Mnext(S * S, 0.0), evals(S, 0.0);
  double last_lambda = 1000.;
  const double Approx0 = 1e-12;
  v[0] = x;
  va[0] = A.product(v[0]); nmv++;
  double initEn = va[0].inner_product(v[0]);
  double lambda = 0.;
  for (int ii = 0; ii < actual_maxiter + 1; ii++) {
    int ni = ii + 1;
    QTensor& q = v[ni];
    if (ii == 0) {
      lambda = initEn;
      M[0] = lambda;
      Mnext[0] = lambda;
      q = va[0];
      q.add(v[0], -lambda);
    } else {
      for (int mi = 0; mi < ni; mi++)
        for (int mj = 0; mj <= mi; mj++) {
          Mnext[mj + mi * (ni + 1)] = M[mj + mi * ni];
          Mnext[mi + mj * (ni + 1)] = M[mi + mj * ni];
        }
      DIAG(ni, M.data(), evals.data());
      lambda = evals[0];
      x = v[0]; x *= M[0];
      q = va[0]; q *= M[0];
      for (int k = 1; k <= ii; k++) {
        x.add(v[k], M[k]);
        q.add(va[k], M[k]);
      }
      q.add(x, -lambda);
      if (M[0] < 0.) { x *= -1.; q *= -1.; }
    }
    double qnorm = q.norm();
    bool converged = (qnorm < tol && std::fabs(lambda - last_lambda) < tol) ||
                     qnorm < std::max(Approx0, tol * 1e-3);
    last_lambda = lambda;
    if ((qnorm < 1e-20) || (converged && ii >= 1) || (ii == actual_maxiter)) break;
    const int Npass = 1;
    std::vector<double> Vq(ni);
    int pass = 1, tot_pass = 0;
    bool give_up = false;
    while (pass <= Npass) {
      ++tot_pass;
      for (int k = 0; k < ni; k++) Vq[k] = q.contract(v[k]);
      for (int k = 0; k < ni; k++) q.add(v[k], -Vq[k]);
      double qnrm = q.norm();
      if (qnrm < 1e-10) {
        q = v[ni - 1];
        q.setRandom(splitmix64(*rng));
        qnrm = q.norm();
        --pass;
        if (ni >= maxsize) { give_up = true; break; }
        if (tot_pass > Npass * 3) { give_up = true; break; }
      }
      q *= 1.0 / qnrm;
      ++pass;
    }
    if (give_up) break;
    va[ni] = A.product(v[ni]); nmv++;
    std::swap(Mnext, M);
    int l = ni;
    for (int k = 0; k < ni + 1; k++) {
      M[l + k * (ni + 1)] = va[l].inner_product(v[k]);
      M[k + l * (ni + 1)] = M[l + k * (ni + 1)];
    }
  }
  x.normalize();
  if (n_matvec) *n_matvec = nmv;
  return lambda;
}

// ---------------------------------------------------------------- dense SVD
// One-sided Jacobi (Hestenes 1958; Demmel & Veselic 1992) on the columns of
// a tall matrix; a wide matrix is handled through its transpose.  Columns of
// zero norm are completed to an orthonormal set (Gram-Schmidt on unit
// vectors) so U (or V) is always an isometry, as LAPACK's is.
static void complete_orthonormal(int m, int k, std::vector<double>& Q, const std::vector<bool>& ok) {
  // Q is m x k column-major; columns with ok[j]==false are replaced.
  int e = 0;
  for (int j = 0; j < k; j++) {
    if (ok[j]) continue;
    for (; e < m; e++) {
      std::vector<double> v(m, 0.0);
      v[e] = 1.0;
      for (int pass = 0; pass < 2; pass++)
        for (int c = 0; c < k; c++) {
          if (c == j || (!ok[c] && c > j)) continue;
          double d = 0;
          for (int i = 0; i < m; i++) d += Q[i + (size_t)m * c] * v[i];
          for (int i = 0; i < m; i++) v[i] -= d * Q[i + (size_t)m * c];
        }
      double nv = 0;
      for (int i = 0; i < m; i++) nv += v[i] * v[i];
      nv = std::sqrt(nv);
      if (nv > 1e-8) {
        for (int i = 0; i < m; i++) Q[i + (size_t)m * j] = v[i] / nv;
        e++;
        break;
      }
    }
  }
}

static void jacobi_tall(int m, int n, std::vector<double>& W, std::vector<double>& Vm,
                        std::vector<double>& sig) {
  // W: m x n (m >= n) overwritten by U*S; Vm: n x n
  Vm.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) Vm[i + (size_t)n * i] = 1.0;
  for (int sweep = 0; sweep < 80; sweep++) {
    bool rotated = false;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double al = 0, be = 0, ga = 0;
        const double* wp = &W[(size_t)m * p];
        const double* wq = &W[(size_t)m * q];
        for (int i = 0; i < m; i++) { al += wp[i] * wp[i]; be += wq[i] * wq[i]; ga += wp[i] * wq[i]; }
        if (ga == 0.0 || std::fabs(ga) <= 1e-15 * std::sqrt(al * be)) continue;
        rotated = true;
        double zeta = (be - al) / (2.0 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        double* Wp = &W[(size_t)m * p];
        double* Wq = &W[(size_t)m * q];
        for (int i = 0; i < m; i++) {
          double a = Wp[i], b = Wq[i];
          Wp[i] = c * a - s * b;
          Wq[i] = s * a + c * b;
        }
        double* Vp = &Vm[(size_t)n * p];
        double* Vq = &Vm[(size_t)n * q];
        for (int i = 0; i < n; i++) {
          double a = Vp[i], b = Vq[i];
          Vp[i] = c * a - s * b;
          Vq[i] = s * a + c * b;
        }
      }
    if (!rotated) break;
  }
  sig.resize(n);
  for (int j = 0; j < n; j++) {
    double s = 0;
    for (int i = 0; i < m; i++) s += W[i + (size_t)m * j] * W[i + (size_t)m * j];
    sig[j] = std::sqrt(s);
  }
}

void dense_svd(int rows, int cols, const double* A, std::vector<double>& U,
               std::vector<double>& S, std::vector<double>& Vt) {
  const bool tall = rows >= cols;
  const int m = tall ? rows : cols, n = tall ? cols : rows;
  std::vector<double> W((size_t)m * n), Vm, sig;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++)
      W[i + (size_t)m * j] = tall ? A[i + (size_t)rows * j] : A[j + (size_t)rows * i];
  jacobi_tall(m, n, W, Vm, sig);
  std::vector<int> ord(n);
  for (int i = 0; i < n; i++) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return sig[a] > sig[b]; });
  const double smax = n > 0 ? sig[ord[0]] : 0.0;
  std::vector<double> Qm((size_t)m * n), Qn((size_t)n * n);
  std::vector<bool> ok(n, true);
  S.resize(n);
  for (int j = 0; j < n; j++) {
    int o = ord[j];
    S[j] = sig[o];
    for (int i = 0; i < n; i++) Qn[i + (size_t)n * j] = Vm[i + (size_t)n * o];
    if (sig[o] > 0 && sig[o] > 1e-300 && !(smax > 0 && sig[o] < smax * 1e-300)) {
      for (int i = 0; i < m; i++) Qm[i + (size_t)m * j] = W[i + (size_t)m * o] / sig[o];
    } else {
      ok[j] = false;
      S[j] = 0.0;
    }
  }
  complete_orthonormal(m, n, Qm, ok);
  const int k = n;
  if (tall) {  // A = Qm S Qn^T
    U = Qm;
    Vt.assign((size_t)k * cols, 0.0);
    for (int j = 0; j < cols; j++)
      for (int i = 0; i < k; i++) Vt[i + (size_t)k * j] = Qn[j + (size_t)n * i];
  } else {  // A^T = Qm S Qn^T  ->  A = Qn S Qm^T
    U = Qn;  // rows x k
    Vt.assign((size_t)k * cols, 0.0);
    for (int j = 0; j < cols; j++)
      for (int i = 0; i < k; i++) Vt[i + (size_t)k * j] = Qm[j + (size_t)m * i];
  }
}

// ---------------------------------------------------------------- svd
static void find_index_permutation(const std::vector<Index>& from, const std::vector<Index>& to,
                                   std::vector<unsigned>& perm) {
  perm.clear();
  for (size_t i = 0; i < to.size(); i++)
    for (size_t j = 0; j < from.size(); j++)
      if (to[i] == from[j]) { perm.push_back((unsigned)j); break; }
  if (perm.size() != to.size()) die("svd: index permutation not found");
}

// qtensor_op.cpp:1097-1453
double svd(const QTensor& A, const std::vector<Index>& left, const std::vector<Index>& right,
           QTensor& U, QTensor& V, QTensor& S, int direction, double cutoff, unsigned K,
           unsigned* kept_out) {
  S.clearBlock();
  std::vector<Index> new_idx(left);
  new_idx.insert(new_idx.end(), right.begin(), right.end());
  std::vector<unsigned> perm;
  find_index_permutation(A.idx_set, new_idx, perm);
  std::unordered_map<std::string, unsigned> A_id;
  for (size_t i = 0; i < A.block.size(); i++) {
    std::string s;
    for (unsigned j = 0; j < A.rank; j++) s += qstr(A.block_index_qn[i][perm[j]]) + " ";
    A_id[s] = (unsigned)i;
  }
  std::set<QN> mid_Q_set;
  std::map<QN, std::set<std::vector<unsigned>>> l_qi, r_qi;
  for (size_t i = 0; i < A.block.size(); i++) {
    QN q;
    std::vector<unsigned> l, r;
    for (size_t j = 0; j < left.size(); j++) {
      if (left[j].arrow == Inward) q += A.block_index_qn[i][perm[j]];
      else q -= A.block_index_qn[i][perm[j]];
      l.push_back(A.block_index_qi[i][perm[j]]);
    }
    for (size_t j = 0; j < right.size(); j++) r.push_back(A.block_index_qi[i][perm[j + left.size()]]);
    mid_Q_set.insert(q);
    l_qi[q].insert(l);
    r_qi[q].insert(r);
  }
  std::vector<QN> mid_Q(mid_Q_set.begin(), mid_Q_set.end());
  const size_t nq = mid_Q.size();
  std::vector<std::vector<std::string>> lstr(nq), rstr(nq);
  std::vector<std::vector<int>> lsz(nq), rsz(nq);
  std::vector<unsigned> mid_QDim(nq);
  for (size_t ii = 0; ii < nq; ii++) {
    unsigned row = 0, col = 0;
    for (auto& l : l_qi[mid_Q[ii]]) {
      unsigned s1 = 1; std::string st;
      for (size_t j = 0; j < l.size(); j++) { s1 *= left[j].qdim(l[j]); st += qstr(left[j].qnAt(l[j])) + " "; }
      lstr[ii].push_back(st); lsz[ii].push_back((int)s1); row += s1;
    }
    for (auto& r : r_qi[mid_Q[ii]]) {
      unsigned s1 = 1; std::string st;
      for (size_t j = 0; j < r.size(); j++) { s1 *= right[j].qdim(r[j]); st += qstr(right[j].qnAt(r[j])) + " "; }
      rstr[ii].push_back(st); rsz[ii].push_back((int)s1); col += s1;
    }
    mid_QDim[ii] = std::min(row, col);
  }
  std::vector<std::vector<double>> mU(nq), mS(nq), mV(nq);
  std::vector<int> rowsv(nq), colsv(nq), kv(nq);
  std::vector<std::pair<double, unsigned>> S_and_block;
  double nm = 0.0;
  // A's blocks hold the un-permuted layout; gather with the permutation so a
  // sector matrix has rows = left multi-index, cols = right multi-index.
  std::vector<size_t> strideA;
  for (size_t ii = 0; ii < nq; ii++) {
    int row = 0, col = 0;
    for (int s : lsz[ii]) row += s;
    for (int s : rsz[ii]) col += s;
    rowsv[ii] = row; colsv[ii] = col;
    std::vector<double> mA((size_t)row * col, 0.0);
    int c_row = 0;
    for (size_t i = 0; i < lstr[ii].size(); i++) {
      int c_col = 0;
      for (size_t j = 0; j < rstr[ii].size(); j++) {
        auto it = A_id.find(lstr[ii][i] + rstr[ii][j]);
        if (it != A_id.end()) {
          const auto& blk = A.block[it->second];
          const auto& d = A.block_index_qd[it->second];
          std::vector<size_t> st(A.rank);
          size_t acc = 1;
          for (unsigned t = 0; t < A.rank; t++) { st[t] = acc; acc *= (size_t)d[t]; }
          // iterate permuted multi-index, left part first-fastest then right
          std::vector<int> ext(A.rank);
          for (unsigned t = 0; t < A.rank; t++) ext[t] = d[perm[t]];
          std::vector<int> ctr(A.rank, 0);
          for (size_t e = 0; e < blk.size(); e++) {
            size_t src = 0;
            for (unsigned t = 0; t < A.rank; t++) src += (size_t)ctr[t] * st[perm[t]];
            size_t lin_l = 0, mul = 1;
            for (size_t t = 0; t < left.size(); t++) { lin_l += ctr[t] * mul; mul *= ext[t]; }
            size_t lin_r = 0; mul = 1;
            for (size_t t = left.size(); t < A.rank; t++) { lin_r += ctr[t] * mul; mul *= ext[t]; }
            mA[(c_row + lin_l) + (size_t)row * (c_col + lin_r)] = blk[src];
            for (unsigned t = 0; t < A.rank; t++) { if (++ctr[t] < ext[t]) break; ctr[t] = 0; }
          }
        }
        c_col += rsz[ii][j];
      }
      c_row += lsz[ii][i];
    }
    std::vector<double>& Us = mU[ii];
    std::vector<double>& Ss = mS[ii];
    std::vector<double>& Vs = mV[ii];
    int k;
    if (row == 1 && col == 2) {  // :1247-1267 brute-force workaround
      k = 1;
      Us.assign(1, 1.0);
      double d = std::sqrt(mA[0] * mA[0] + mA[1] * mA[1]);
      Ss.assign(1, d);
      Vs = mA;
      if (d != 0.) { Vs[0] /= d; Vs[1] /= d; }
      else { Vs[0] = Vs[1] = 1.0 / std::sqrt(2.0); }
    } else {
      std::vector<double> Uf, Sf, Vf;
      dense_svd(row, col, mA.data(), Uf, Sf, Vf);
      int kf = std::min(row, col);
      k = (K != 0) ? std::min<int>((int)K, kf) : kf;  // rank-K SVD per sector
      Us.assign(Uf.begin(), Uf.begin() + (size_t)row * k);
      Ss.assign(Sf.begin(), Sf.begin() + k);
      Vs.assign((size_t)k * col, 0.0);
      for (int j = 0; j < col; j++)
        for (int i = 0; i < k; i++) Vs[i + (size_t)k * j] = Vf[i + (size_t)kf * j];
    }
    kv[ii] = k;
    if (direction == MoveFromLeft) {
      for (int j = 0; j < col; j++)
        for (int i = 0; i < k; i++) Vs[i + (size_t)k * j] *= Ss[i];
    } else {
      for (int j = 0; j < k; j++)
        for (int i = 0; i < row; i++) Us[i + (size_t)row * j] *= Ss[j];
    }
    for (int j = 0; j < k; j++) {
      S_and_block.push_back({Ss[j], (unsigned)ii});
      nm += Ss[j] * Ss[j];
    }
  }
  // :1289-1301 global truncation
  std::sort(S_and_block.begin(), S_and_block.end(),
            [](const std::pair<double, unsigned>& a, const std::pair<double, unsigned>& b) {
              return a.first > b.first;
            });
  double cumsum = 0;
  unsigned long accBD = 0;
  std::vector<unsigned> new_QDim(nq, 0);
  for (size_t i = 0; i < S_and_block.size(); i++) {
    cumsum += S_and_block[i].first * S_and_block[i].first / nm;
    if (S_and_block[i].first == 0. && cutoff != 0.) break;
    new_QDim[S_and_block[i].second] += 1;
    accBD += 1;
    if ((1 - cumsum < cutoff && cutoff != 0.) || (accBD >= K && K != 0)) break;
  }
  double trunc = std::max(1 - cumsum, 0.);
  Index midCut(Outward, "svdmid", Link, 0);
  for (size_t ii = 0; ii < nq; ii++)
    if (new_QDim[ii] > 0) midCut.addQN(mid_Q[ii], new_QDim[ii]);
  std::vector<Index> Uidx(left);
  Uidx.push_back(midCut);
  std::vector<Index> Vidx;
  Index md = midCut; md.dag();
  Vidx.push_back(md);
  Vidx.insert(Vidx.end(), right.begin(), right.end());
  U.reset(Uidx); U.initted = true;
  V.reset(Vidx); V.initted = true;
  S.reset(std::vector<Index>{midCut}); S.initted = true;
  unsigned mid_num = 0;
  for (size_t ii = 0; ii < nq; ii++) {
    QN q = mid_Q[ii];
    int ndim = (int)new_QDim[ii];
    if (ndim == 0) continue;
    int row = rowsv[ii], col = colsv[ii], k = kv[ii];
    int c_row = 0;
    for (auto& l : l_qi[q]) {
      std::vector<unsigned> bqi(l);
      std::vector<int> bqd;
      std::vector<QN> bqn;
      size_t lsize = 1;
      for (size_t j = 0; j < l.size(); j++) {
        bqd.push_back((int)left[j].qdim(l[j]));
        bqn.push_back(left[j].qnAt(l[j]));
        lsize *= (size_t)bqd.back();
      }
      bqi.push_back(mid_num); bqd.push_back(ndim); bqn.push_back(q);
      std::vector<double> blk(lsize * ndim);
      for (int c = 0; c < ndim; c++)
        for (size_t r = 0; r < lsize; r++) blk[r + lsize * c] = mU[ii][(c_row + r) + (size_t)row * c];
      U.block.push_back(std::move(blk));
      U.block_id_by_qn_str[key_of(bqn)] = (unsigned)U.block.size() - 1;
      U.block_index_qn.push_back(bqn); U.block_index_qd.push_back(bqd); U.block_index_qi.push_back(bqi);
      c_row += (int)lsize;
    }
    int c_col = 0;
    for (auto& r : r_qi[q]) {
      std::vector<unsigned> bqi{mid_num};
      std::vector<int> bqd{ndim};
      std::vector<QN> bqn{q};
      size_t rsize = 1;
      for (size_t j = 0; j < r.size(); j++) {
        bqi.push_back(r[j]);
        bqd.push_back((int)right[j].qdim(r[j]));
        bqn.push_back(right[j].qnAt(r[j]));
        rsize *= (size_t)right[j].qdim(r[j]);
      }
      std::vector<double> blk((size_t)ndim * rsize);
      for (size_t c = 0; c < rsize; c++)
        for (int rr = 0; rr < ndim; rr++) blk[rr + (size_t)ndim * c] = mV[ii][rr + (size_t)k * (c_col + c)];
      V.block.push_back(std::move(blk));
      V.block_id_by_qn_str[key_of(bqn)] = (unsigned)V.block.size() - 1;
      V.block_index_qn.push_back(bqn); V.block_index_qd.push_back(bqd); V.block_index_qi.push_back(bqi);
      c_col += (int)rsize;
    }
    (void)col;
    S.block.emplace_back(mS[ii].begin(), mS[ii].begin() + ndim);
    S.block_index_qd.push_back({ndim});
    S.block_index_qn.push_back({q});
    S.block_index_qi.push_back({mid_num});
    mid_num++;
  }
  if (kept_out) *kept_out = (unsigned)accBD;
  return trunc;
}

double svd_bond(const QTensor& combined, QTensor& A_left, QTensor& A_right, const Index& mid,
                QTensor& S, int direction, double cutoff, unsigned long K, unsigned* kept) {
  QTensor U, V;
  std::vector<Index> left, right;
  Index dag_mid(mid); dag_mid.dag();
  for (auto& ix : A_left.idx_set) if (ix.tag() != mid.tag()) left.push_back(ix);
  for (auto& ix : A_right.idx_set) if (ix.tag() != dag_mid.tag()) right.push_back(ix);
  double tr = svd(combined, left, right, U, V, S, direction, cutoff, (unsigned)K, kept);
  U.idx_set.back().rename(mid.name);
  U.idx_set.back().retype(mid.type);
  V.idx_set[0].rename(dag_mid.name);
  V.idx_set[0].retype(dag_mid.type);
  if (U.idx_set.back().level != mid.level) U.idx_set.back().prime((int)mid.level - (int)U.idx_set.back().level);
  if (V.idx_set[0].level != dag_mid.level) V.idx_set[0].prime((int)dag_mid.level - (int)V.idx_set[0].level);
  A_right = V;
  A_left = U;
  return tr;
}

// ---------------------------------------------------------------- qMPS
static std::string link_name(unsigned id, unsigned i) { return "ID" + std::to_string(id) + "Link" + std::to_string(i); }

void QTT::allocateMPS(const unsigned* product_state) {
  std::vector<std::set<QN>> Link_QN(length + 1);
  Link_QN[0].insert(QN());
  Link_QN[length].insert(totalQ);
  for (unsigned j = 1; j < length; j++)
    for (auto& q : Link_QN[j - 1])
      for (auto& p : phy_qn) Link_QN[j].insert(q + p);
  for (unsigned j = length - 1; j > 0; j--) {
    std::vector<QN> rem;
    for (auto& q : Link_QN[j]) {
      bool valid = false;
      for (auto& p : phy_qn)
        if (Link_QN[j + 1].count(q + p)) { valid = true; break; }
      if (!valid) rem.push_back(q);
    }
    for (auto& q : rem) Link_QN[j].erase(q);
  }
  A.clear();
  for (unsigned i = 0; i < length; i++) {
    QTensor t({Index(Inward, link_name(id, i), Link, 0), Index(Inward, "Site" + std::to_string(i), Site, 0),
               Index(Outward, link_name(id, i + 1), Link, 0)});
    for (auto& q : Link_QN[i]) t.idx_set[0].addQN(q, 1);
    for (auto& q : phy_qn) t.idx_set[1].addQN(q, 1);
    for (auto& q : Link_QN[i + 1]) t.idx_set[2].addQN(q, 1);
    t.initBlock();
    A.push_back(std::move(t));
  }
  if (product_state) {
    QN last;
    for (unsigned i = 0; i < length; i++) {
      std::string s = qstr(last) + " " + qstr(phy_qn[product_state[i]]) + " ";
      last += phy_qn[product_state[i]];
      s += qstr(last) + " ";
      auto it = A[i].block_id_by_qn_str.find(s);
      if (it == A[i].block_id_by_qn_str.end()) die("product state block missing");
      std::fill(A[i].block[it->second].begin(), A[i].block[it->second].end(), 1.0);
    }
    center = (int)length - 1;
  }
}

static void rename_to(Index& ix, const Index& mid) {
  ix.rename(mid.name);
  ix.prime((int)mid.level - (int)ix.level);
}

void QTT::rc() {
  QTensor U, V, S;
  for (unsigned i = length - 1; i > 0; i--) {
    std::vector<Index> left, right;
    Index mid;
    std::string LN = link_name(id, i);
    for (auto& ix : A[i].idx_set) {
      if (ix.name == LN) { left.push_back(ix); mid = ix; } else right.push_back(ix);
    }
    svd(A[i], left, right, U, V, S, MoveFromRight, 0, 0, nullptr);
    bond_dims[i] = (unsigned)S.total_size();
    A[i] = V;
    rename_to(A[i].idx_set[0], mid);
    A[i - 1] = A[i - 1] * U;
    rename_to(A[i - 1].idx_set.back(), mid);
  }
  center = 0;
}

void QTT::lc() {
  QTensor U, V, S;
  for (unsigned i = 0; i + 1 < length; i++) {
    std::vector<Index> left, right;
    Index mid;
    std::string LN = link_name(id, i + 1);
    for (auto& ix : A[i].idx_set) {
      if (ix.name == LN) { right.push_back(ix); mid = ix; } else left.push_back(ix);
    }
    svd(A[i], left, right, U, V, S, MoveFromLeft, 0, 0, nullptr);
    bond_dims[i + 1] = (unsigned)S.total_size();
    A[i] = U;
    rename_to(A[i].idx_set.back(), mid);
    A[i + 1] = V * A[i + 1];
    rename_to(A[i + 1].idx_set[0], mid);
  }
  center = (int)length - 1;
}

void QTT::normalize() {
  if (center == -1) die("normalize without canonical centre is not restated");
  double nm = A[center].norm();
  A[center] *= 1.0 / nm;
}

double QTT::position(int site) {
  if (center == -1) rc();
  QTensor U, V, S;
  bool moved = false;
  while (center != site) {
    std::vector<Index> left, right;
    Index mid;
    moved = true;
    if (center > site) {
      std::string LN = link_name(id, center);
      for (auto& ix : A[center].idx_set) {
        if (ix.name == LN) { left.push_back(ix); mid = ix; } else right.push_back(ix);
      }
      svd(A[center], left, right, U, V, S, MoveFromRight, 0, 0, nullptr);
      bond_dims[center] = (unsigned)S.total_size();
      A[center] = V;
      rename_to(A[center].idx_set[0], mid);
      A[center - 1] = A[center - 1] * U;
      rename_to(A[center - 1].idx_set.back(), mid);
      --center;
    } else {
      std::string LN = link_name(id, center + 1);
      for (auto& ix : A[center].idx_set) {
        if (ix.name == LN) { right.push_back(ix); mid = ix; } else left.push_back(ix);
      }
      svd(A[center], left, right, U, V, S, MoveFromLeft, 0, 0, nullptr);
      bond_dims[center + 1] = (unsigned)S.total_size();
      A[center] = U;
      rename_to(A[center].idx_set.back(), mid);
      A[center + 1] = V * A[center + 1];
      rename_to(A[center + 1].idx_set[0], mid);
      ++center;
    }
  }
  if (!moved) return 0.;
  double e = 0;
  for (auto& b : S.block)
    for (double s : b) if (s > 1e-24) e -= s * s * std::log(s * s);
  return e;
}

// ---------------------------------------------------------------- sites
// models/sites/spinhalf.cpp:23-46 (Dn=0, Up=1)
double spinhalf_op(unsigned bra, const std::string& op, unsigned ket) {
  const unsigned Dn = 0, Up = 1;
  if (op == "Id" || op == "I") return ket == bra ? 1 : 0;
  if (op == "Sz") {
    if (ket == Dn && bra == Dn) return -0.5;
    if (ket == Up && bra == Up) return 0.5;
    return 0;
  }
  if (op == "Sx") return (ket != bra) ? 0.5 : 0;
  if (op == "S+" || op == "Sp") return (ket == Dn && bra == Up) ? 1 : 0;
  if (op == "S-" || op == "Sm") return (ket == Up && bra == Dn) ? 1 : 0;
  die("unsupported spin-1/2 operator");
  return 0;
}
// spinhalf.cpp:95-106
QN spinhalf_div(const std::string& op) {
  if (op == "Sz" || op == "Id") return QN();
  if (op == "S+") return QN(2, 0);
  if (op == "S-") return QN(-2, 0);
  die("div: unknown op");
  return QN();
}

static std::vector<QN> spin_phy_qn() { return {QN(-1, 0), QN(1, 0)}; }

// Heisenberg.cpp:136-180 -- write val*<b|op|k> into the block (Qi,k,b,Qo)
// at offset s0*r+s3*c (s3*c on site 0); CTF write overwrites.
static void add_operator(QTT& H, unsigned site, unsigned r, unsigned c, const std::string& op,
                         double val, QN Qi, QN Qo) {
  bool added = false;
  QTensor& W = H.A[site];
  for (unsigned k = 0; k < H.phy_dim; k++)
    for (unsigned b = 0; b < H.phy_dim; b++) {
      std::string s = qstr(Qi) + " " + qstr(H.phy_qn[k]) + " " + qstr(H.phy_qn[b]) + " " + qstr(Qo) + " ";
      auto it = W.block_id_by_qn_str.find(s);
      if (it == W.block_id_by_qn_str.end()) continue;
      added = true;
      unsigned bp = it->second;
      size_t s0 = 1;
      size_t s1 = s0 * W.idx_set[0].qdim(W.block_index_qi[bp][0]);
      size_t s2 = s1 * W.idx_set[1].qdim(W.block_index_qi[bp][1]);
      size_t s3 = s2 * W.idx_set[2].qdim(W.block_index_qi[bp][2]);
      double v = val * spinhalf_op(b, op, k);
      size_t off = (site == 0) ? s3 * c : s0 * r + s3 * c;
      if (off >= W.block[bp].size()) die("add_operator: offset out of block");
      W.block[bp][off] = v;
    }
  if (!added) die("add_operator: no block");
}

static void mpo_shell(QTT& H, unsigned N, unsigned id) {
  H.id = id; H.length = N; H.phy_dim = 2; H.N = 2;
  H.phy_qn = spin_phy_qn();
  H.bond_dims.assign(N + 1, 1);
  H.center = -1;
  H.A.clear();
}

// Heisenberg.cpp:333-402
void heisenberg_mpo_nn(unsigned N, unsigned id, QTT& H) {
  mpo_shell(H, N, id);
  for (unsigned i = 0; i < N; i++) {
    QTensor t({Index(Inward, link_name(id, i), Link, 0), Index(Outward, "Site" + std::to_string(i), Site, 0),
               Index(Inward, "Site" + std::to_string(i), Site, 1), Index(Outward, link_name(id, i + 1), Link, 0)});
    if (i == 0) t.idx_set[0].addQN(QN(), 1);
    else { t.idx_set[0].addQN(QN(), 3); t.idx_set[0].addQN(QN(-2, 0), 1); t.idx_set[0].addQN(QN(2, 0), 1); }
    for (auto& q : H.phy_qn) t.idx_set[1].addQN(q, 1);
    for (auto& q : H.phy_qn) t.idx_set[2].addQN(q, 1);
    if (i == N - 1) t.idx_set[3].addQN(QN(), 1);
    else { t.idx_set[3].addQN(QN(), 3); t.idx_set[3].addQN(QN(-2, 0), 1); t.idx_set[3].addQN(QN(2, 0), 1); }
    t.initBlock();
    H.A.push_back(std::move(t));
  }
  for (unsigned site = 0; site < N; site++) {
    if (site > 0) add_operator(H, site, 0, 0, "Id", 1.0, QN(), QN());
    if (site < N - 1) add_operator(H, site, 1, 1, "Id", 1.0, QN(), QN());
    if (site > 0) add_operator(H, site, 2, 0, "Sz", 1.0, QN(), QN());
    if (site < N - 1) add_operator(H, site, 1, 2, "Sz", 1.0, QN(), QN());
    if (site > 0) add_operator(H, site, 0, 0, "S+", 1.0, QN(-2, 0), QN());
    if (site > 0) add_operator(H, site, 0, 0, "S-", 1.0, QN(2, 0), QN());
    if (site < N - 1) add_operator(H, site, 1, 0, "S-", 0.5, QN(), QN(-2, 0));
    if (site < N - 1) add_operator(H, site, 1, 0, "S+", 0.5, QN(), QN(2, 0));
  }
}

// ---------------------------------------------------------------- AutoMPO
// models/hams/AutoMPO.{h,cpp}: SiteTerm order (site, then op), HTerm keys
// ordered by LessNoCoef, duplicates summed.
namespace {
struct SiteTerm {
  std::string op; int i = -1;
  bool operator==(const SiteTerm& o) const { return op == o.op && i == o.i; }
  bool operator!=(const SiteTerm& o) const { return !(*this == o); }
  bool operator<(const SiteTerm& o) const { if (i != o.i) return i < o.i; return op < o.op; }
};
struct HTerm { double coef = 1; std::vector<SiteTerm> ops;
  const SiteTerm& first() const { return ops.front(); }
  const SiteTerm& last() const { return ops.back(); } };
struct LessNoCoef {
  bool operator()(const HTerm& a, const HTerm& b) const {
    if (a.ops.size() != b.ops.size()) return a.ops.size() < b.ops.size();
    for (size_t j = 0; j < a.ops.size(); j++)
      if (a.ops[j] != b.ops[j]) return a.ops[j] < b.ops[j];
    return false;
  }
};
struct SiteQN { SiteTerm st; QN q; };
// AutoMPO.cpp:190-214 -- keep ops in site order
void hterm_add(HTerm& t, const std::string& op, int i) {
  auto it = t.ops.begin();
  while (it != t.ops.end() && it->i <= i) ++it;
  t.ops.insert(it, SiteTerm{op, i});
}
unsigned basisToIndex(QN qni, std::map<QN, unsigned>& qnToMaxIdx,
                      std::unordered_map<unsigned, unsigned>& rcToIdx, unsigned i) {
  if (qnToMaxIdx.count(qni) == 0) qnToMaxIdx[qni] = 1;
  if (rcToIdx.count(i) == 0) rcToIdx[i] = qnToMaxIdx[qni]++;
  return rcToIdx[i] - 1;
}
}  // namespace

std::vector<Term> heisenberg_terms(const std::vector<std::pair<int, int>>& bonds,
                                   const std::vector<double>& J) {
  std::vector<Term> t;
  for (size_t b = 0; b < bonds.size(); b++) {
    double j = J.size() == bonds.size() ? J[b] : J.at(0);
    t.push_back({j, "Sz", bonds[b].first, "Sz", bonds[b].second});
    t.push_back({0.5 * j, "S+", bonds[b].first, "S-", bonds[b].second});
    t.push_back({0.5 * j, "S-", bonds[b].first, "S+", bonds[b].second});
  }
  return t;
}

// Heisenberg.cpp:416-635
void autompo_mpo(unsigned N, unsigned id, const std::vector<Term>& in_terms, QTT& H) {
  std::set<HTerm, LessNoCoef> terms;
  for (auto& tm : in_terms) {   // AutoMPO::add (AutoMPO.cpp:369-397)
    HTerm h;
    h.coef = 1.0;
    hterm_add(h, tm.op1, tm.i1);
    hterm_add(h, tm.op2, tm.i2);
    h.coef *= tm.coef;
    if (std::fabs(h.coef) == 0.0) continue;
    auto it = terms.find(h);
    if (it == terms.end()) terms.insert(h);
    else { HTerm nt = h; nt.coef += it->coef; terms.erase(it); terms.insert(nt); }
  }
  mpo_shell(H, N, id);
  SiteTerm IL{"IL", 0}, HL{"HL", 0};
  std::vector<std::vector<SiteQN>> basis(N + 1);
  for (unsigned n = 0; n < N; n++) basis[n].push_back({IL, QN()});
  for (unsigned n = 1; n <= N; n++) basis[n].push_back({HL, QN()});
  for (auto& ht : terms) {
    for (int n = ht.first().i + 1; n <= ht.last().i + 1; ++n) {
      auto& bn = basis.at(n);
      bool has = false;
      for (auto& sq : bn) if (sq.st == ht.first()) { has = true; break; }
      if (!has) bn.push_back({ht.first(), -spinhalf_div(ht.first().op)});
    }
  }
  const QN Zero;
  auto qn_comp = [&Zero](const SiteQN& a, const SiteQN& b) {
    if (a.q == Zero && b.q != Zero) return true;
    if (b.q == Zero && a.q != Zero) return false;
    return a.q < b.q;
  };
  for (auto& bn : basis) std::sort(bn.begin(), bn.end(), qn_comp);
  std::vector<std::vector<std::pair<QN, unsigned>>> links(N + 1);
  for (unsigned n = 0; n <= N; n++) {
    auto& bn = basis[n];
    QN currq = bn.front().q;
    unsigned currm = 0;
    for (auto& sq : bn) {
      if (sq.q == currq) ++currm;
      else { links[n].push_back({currq, currm}); currq = sq.q; currm = 1; }
    }
    links[n].push_back({currq, currm});
  }
  std::vector<std::vector<HTerm>> ht_by_n(N + 1);
  for (auto& ht : terms)
    for (auto& st : ht.ops) ht_by_n.at(st.i + 1).push_back(ht);
  std::vector<std::map<QN, unsigned>> qnToMaxIdx(N + 1);
  std::vector<std::unordered_map<unsigned, unsigned>> rcToIdx(N + 1);
  for (unsigned i = 0; i < N; i++) {
    QTensor t({Index(Inward, link_name(id, i), Link, 0), Index(Outward, "Site" + std::to_string(i), Site, 0),
               Index(Inward, "Site" + std::to_string(i), Site, 1), Index(Outward, link_name(id, i + 1), Link, 0)});
    for (auto& rq : links[i]) t.idx_set[0].addQN(rq.first, rq.second);
    if (i != N - 1) for (auto& cq : links[i + 1]) t.idx_set[3].addQN(cq.first, cq.second);
    else t.idx_set[3].addQN(QN(), 1);
    for (auto& q : H.phy_qn) { t.idx_set[1].addQN(q, 1); t.idx_set[2].addQN(q, 1); }
    t.initBlock();
    H.A.push_back(std::move(t));
    auto& bn1 = basis[i];
    auto& bn = basis[i + 1];
    unsigned c_max = (i == N - 1) ? 1 : (unsigned)bn.size();
    for (unsigned r_ = 0; r_ < bn1.size(); r_++) {
      for (unsigned c_ = 0; c_ < c_max; c_++) {
        const SiteTerm& rst = bn1[r_].st;
        const SiteTerm& cst = bn[c_].st;
        QN qr = -bn1[r_].q, qc = -bn[c_].q;
        unsigned r = (i == 0) ? 1 : r_;
        unsigned c = c_;
        if (cst.i == (int)i && rst == IL) {
          const std::string& op = cst.op;
          if (op != "HL" && op != "IL") {
            unsigned thisR = basisToIndex(-qr, qnToMaxIdx[i], rcToIdx[i], r);
            unsigned thisC = basisToIndex(qc, qnToMaxIdx[i + 1], rcToIdx[i + 1], c);
            if (i == 0) thisR = 1;
            add_operator(H, i, thisR, thisC, op, 1.0, qr, qc);
          }
        }
        if (cst == rst) {
          unsigned thisR = basisToIndex(-qr, qnToMaxIdx[i], rcToIdx[i], r);
          unsigned thisC = basisToIndex(qc, qnToMaxIdx[i + 1], rcToIdx[i + 1], c);
          add_operator(H, i, thisR, thisC, "Id", 1.0, qr, qc);
        }
        if (cst == HL) {
          for (auto& ht : ht_by_n[i + 1]) {
            if (rst == ht.first() && ht.last().i == (int)i) {
              unsigned thisR = basisToIndex(-qr, qnToMaxIdx[i], rcToIdx[i], r);
              unsigned thisC = basisToIndex(qc, qnToMaxIdx[i + 1], rcToIdx[i + 1], c);
              add_operator(H, i, thisR, thisC, ht.last().op, ht.coef, qr, qc);
            }
          }
        }
        if (rst == IL && cst == HL) {
          for (auto& ht : ht_by_n[i + 1]) {
            if (ht.first().i == ht.last().i) {
              unsigned thisR = basisToIndex(-qr, qnToMaxIdx[i], rcToIdx[i], r);
              unsigned thisC = basisToIndex(qc, qnToMaxIdx[i + 1], rcToIdx[i + 1], c);
              add_operator(H, i, thisR, thisC, ht.first().op, ht.coef, qr, qc);
            }
          }
        }
      }
    }
  }
}

// models/lattice/square.h:25-108 (1-based site numbers converted to 0-based)
std::vector<std::pair<int, int>> square_lattice(int Nx, int Ny, bool yperiodic,
                                                std::vector<int>* type, bool nnn) {
  yperiodic = yperiodic && (Ny > 2);
  std::vector<std::pair<int, int>> b;
  auto add = [&](int s1, int s2, int t) { b.push_back({s1 - 1, s2 - 1}); if (type) type->push_back(t); };
  int N = Nx * Ny;
  for (int n = 1; n <= N; ++n) {
    int x = (n - 1) / Ny + 1, y = (n - 1) % Ny + 1;
    if (x < Nx) add(n, n + Ny, 1);
    if (Ny > 1) {
      if (y < Ny) add(n, n + 1, 1);
      if (yperiodic && y == 1) add(n, n + Ny - 1, 1);
    }
    if (nnn && x < Nx && Ny > 1) {
      if (y < Ny) add(n, n + Ny + 1, 2);
      if (y > 1) add(n, n + Ny - 1, 2);
      if (yperiodic && y == Ny) add(n, n + 1, 2);
      if (yperiodic && y == 1) add(n, n + 2 * Ny - 1, 2);
    }
  }
  return b;
}

// ---------------------------------------------------------------- DMRG
// dmrg.cpp:56-84
void buildEnv(QTT& psi, QTT& H, std::vector<QTensor>& TR, std::vector<QTensor>& TL) {
  unsigned L = psi.length;
  {
    QTensor le({psi.A[0].idx_set[0], H.A[0].idx_set[0], psi.A[0].idx_set[0]});
    le.idx_set[0].prime(); le.idx_set[0].dag();
    le.initBlock(); le.setOne(); le.dag();
    TL[0] = le;
  }
  {
    QTensor re({psi.A[L - 1].idx_set.back(), H.A[L - 1].idx_set.back(), psi.A[L - 1].idx_set.back()});
    re.idx_set[0].prime(); re.idx_set[0].dag();
    re.initBlock(); re.setOne(); re.dag();
    TR[L - 1] = re;
  }
  for (unsigned i = L - 1; i > 0; i--) {
    QTensor t1 = psi.A[i]; t1.dag(); t1.prime();
    QTensor t4 = t1 * TR[i];
    QTensor t5 = t4 * H.A[i];
    TR[i - 1] = t5 * psi.A[i];
  }
}

// dmrg.cpp:491-509
void updateEnv(QTT& psi, QTT& H, std::vector<QTensor>& TR, std::vector<QTensor>& TL,
               unsigned site, int direction) {
  QTensor t1 = psi.A[site]; t1.dag(); t1.prime();
  if (direction == MoveFromLeft) {
    QTensor t4 = t1 * TL[site];
    QTensor t5 = t4 * H.A[site];
    TL[site + 1] = t5 * psi.A[site];
  } else {
    QTensor t4 = t1 * TR[site];
    QTensor t5 = t4 * H.A[site];
    TR[site - 1] = t5 * psi.A[site];
  }
}

// dmrg.cpp:300-372
SiteStats updateSite(QTT& psi, QTT& H, std::vector<QTensor>& TR, std::vector<QTensor>& TL,
                     unsigned site, int direction, int max_bd, double cutoff, int max_restart,
                     double tol, uint64_t* rng) {
  SiteStats st{};
  unsigned a = (direction == MoveFromLeft) ? site : site - 1;
  BigQTensor A;
  A.setLeft(&TL[a]);
  A.addMid(&H.A[a]);
  A.addMid(&H.A[a + 1]);
  A.setRight(&TR[a + 1]);
  QTensor x = psi.A[a] * psi.A[a + 1];
  st.energy = tensor_davidsonIT(A, x, 3, max_restart, tol, 'S', rng, &st.n_matvec);
  QTensor S;
  Index mid;
  std::string LN = link_name(psi.id, a + 1);
  for (auto& ix : psi.A[a].idx_set) if (ix.name == LN) { mid = ix; break; }
  st.trunc_err = svd_bond(x, psi.A[a], psi.A[a + 1], mid, S,
                          direction, cutoff, (unsigned long)max_bd, &st.kept);
  psi.bond_dims[a + 1] = (unsigned)S.total_size();
  psi.center = (direction == MoveFromLeft) ? (int)site + 1 : (int)site - 1;
  return st;
}

// dmrg.cpp:760-813
double dmrg(QTT& psi, QTT& H, int num_sweeps, const std::vector<int>& max_bd,
            const std::vector<double>& cutoff, const std::vector<int>& max_restart,
            double tol, DmrgLog* log) {
  int L = (int)H.length;
  uint64_t rng = 0x5eed;
  psi.position(0);
  psi.normalize();
  std::vector<QTensor> TR(L), TL(L);
  buildEnv(psi, H, TR, TL);
  double Energy = 0;
  {
    BigQTensor A;
    A.setLeft(&TL[0]); A.addMid(&H.A[0]); A.setRight(&TR[0]);
    Energy = A.expec(psi.A[0]);
    if (log) log->initial_energy = Energy;
  }
  for (int sweep = 0; sweep < num_sweeps; sweep++) {
    int lbd = sweep % (int)max_bd.size();
    int lc = sweep % (int)cutoff.size();
    int lmr = sweep % (int)max_restart.size();
    for (int l = 2 * sweep; l < 2 * (sweep + 1); l++) {
      int direction = (l % 2 == 0) ? MoveFromLeft : MoveFromRight;
      for (int i = 0; i < L - 1; i++) {
        int site = (direction == MoveFromLeft) ? i : L - 1 - i;
        SiteStats s = updateSite(psi, H, TR, TL, (unsigned)site, direction, max_bd[lbd],
                                 cutoff[lc], max_restart[lmr], tol, &rng);
        Energy = s.energy;
        if (log) {
          log->bond_energy.push_back(s.energy);
          log->trunc_err.push_back(s.trunc_err);
          log->kept.push_back(s.kept);
        }
        updateEnv(psi, H, TR, TL, (unsigned)site, direction);
      }
      if (log) log->halfsweep_energy.push_back(Energy);
    }
    psi.normalize();
  }
  psi.position(0);
  return Energy;
}

}  // namespace orc
