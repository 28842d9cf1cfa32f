// oracle_capi.cpp -- C entry points of the oracle for the Python tests.
// TEST INFRASTRUCTURE ONLY (the oracle); see tt_oracle.hpp.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <sstream>
#include <stdexcept>

#include "tt_oracle.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

namespace {
std::vector<std::string> split_names(const char* names, int rank) {
  std::vector<std::string> out;
  std::string s(names ? names : "");
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, '\n')) out.push_back(item);
  while ((int)out.size() < rank) out.push_back("");
  return out;
}

std::vector<Index> make_indices(int rank, const int* arrows, const int* levels, const int* types,
                                const char* names, const int* nsec, const int* qn, const int* qdim) {
  auto nm = split_names(names, rank);
  std::vector<Index> idx;
  int off = 0;
  for (int i = 0; i < rank; i++) {
    Index ix((Arrow)arrows[i], nm[i], (IType)types[i], (unsigned)levels[i]);
    for (int s = 0; s < nsec[i]; s++) ix.qn.push_back({QN(qn[2 * (off + s)], qn[2 * (off + s) + 1]), (unsigned)qdim[off + s]});
    ix.setTag();
    for (int s = 0; s < nsec[i]; s++)
      if (ix.qn[s].first != QN(qn[2 * (off + s)], qn[2 * (off + s) + 1]))
        throw std::runtime_error("index sectors must be given sorted by QN");
    off += nsec[i];
    idx.push_back(ix);
  }
  return idx;
}

const char* g_err = "";
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err; }

void* orc_tensor_import(int rank, const int* arrows, const int* levels, const int* types,
                        const char* names, const int* nsec, const int* qn, const int* qdim,
                        int nblocks, const int* block_qi, const double* data) {
  try {
    QTensor* t = new QTensor(make_indices(rank, arrows, levels, types, names, nsec, qn, qdim));
    t->initted = true;
    size_t off = 0;
    for (int b = 0; b < nblocks; b++) {
      std::vector<unsigned> qi(rank);
      std::vector<int> qd(rank);
      std::vector<QN> qs(rank);
      size_t sz = 1;
      std::string key;
      for (int j = 0; j < rank; j++) {
        qi[j] = (unsigned)block_qi[b * rank + j];
        qd[j] = (int)t->idx_set[j].qdim(qi[j]);
        qs[j] = t->idx_set[j].qnAt(qi[j]);
        sz *= (size_t)qd[j];
        key += qstr(qs[j]) + " ";
      }
      t->block.emplace_back(data ? data + off : nullptr, data ? data + off + sz : nullptr);
      if (!data) t->block.back().assign(sz, 0.0);
      off += sz;
      t->block_index_qi.push_back(qi);
      t->block_index_qd.push_back(qd);
      t->block_index_qn.push_back(qs);
      t->block_id_by_qn_str[key] = (unsigned)b;
    }
    return t;
  } catch (std::exception& e) {
    g_err = "import failed";
    return nullptr;
  }
}

void* orc_init_block(int rank, const int* arrows, const int* levels, const int* types,
                     const char* names, const int* nsec, const int* qn, const int* qdim) {
  QTensor* t = new QTensor(make_indices(rank, arrows, levels, types, names, nsec, qn, qdim));
  t->initBlock();
  return t;
}

// out: [rank, nsec_total, nblocks, names_len]; *total = number of doubles
void orc_tensor_info(void* h, int* out, long* total) {
  QTensor* t = (QTensor*)h;
  int ns = 0, nl = 0;
  for (auto& ix : t->idx_set) { ns += (int)ix.size(); nl += (int)ix.name.size() + 1; }
  out[0] = (int)t->rank; out[1] = ns; out[2] = (int)t->block.size(); out[3] = nl;
  *total = (long)t->total_size();
}

void orc_tensor_export(void* h, int* arrows, int* levels, int* types, char* names, int* nsec,
                       int* qn, int* qdim, int* block_qi, long* block_off, double* data) {
  QTensor* t = (QTensor*)h;
  std::string nm;
  int off = 0;
  for (unsigned i = 0; i < t->rank; i++) {
    auto& ix = t->idx_set[i];
    arrows[i] = (int)ix.arrow; levels[i] = (int)ix.level; types[i] = (int)ix.type;
    nm += ix.name + "\n";
    nsec[i] = (int)ix.size();
    for (unsigned s = 0; s < ix.size(); s++) {
      qn[2 * (off + s)] = ix.qn[s].first.q0;
      qn[2 * (off + s) + 1] = ix.qn[s].first.q1;
      qdim[off + s] = (int)ix.qn[s].second;
    }
    off += (int)ix.size();
  }
  memcpy(names, nm.c_str(), nm.size());
  long o = 0;
  for (size_t b = 0; b < t->block.size(); b++) {
    for (unsigned j = 0; j < t->rank; j++) block_qi[b * t->rank + j] = (int)t->block_index_qi[b][j];
    block_off[b] = o;
    if (data) memcpy(data + o, t->block[b].data(), t->block[b].size() * sizeof(double));
    o += (long)t->block[b].size();
  }
  block_off[t->block.size()] = o;
}

void orc_tensor_free(void* h) { delete (QTensor*)h; }
void* orc_tensor_copy(void* h) { return new QTensor(*(QTensor*)h); }

void* orc_contract(void* A, void* B) {
  try {
    return new QTensor((*(QTensor*)A) * (*(QTensor*)B));
  } catch (std::exception&) { g_err = "contract failed"; return nullptr; }
}

// index pairing of operator*: writes npairs, pairs (posA,posB)
int orc_pairs(void* A, void* B, int* pa, int* pb) {
  std::vector<std::pair<unsigned, unsigned>> pr;
  std::vector<unsigned> ap, bp;
  pair_indices(*(QTensor*)A, *(QTensor*)B, pr, ap, bp);
  for (size_t i = 0; i < pr.size(); i++) { pa[i] = (int)pr[i].first; pb[i] = (int)pr[i].second; }
  return (int)pr.size();
}

int orc_last_tasks(int* out, int cap) {
  int n = (int)g_last_tasks.size() / 6;
  for (int i = 0; i < std::min(cap, n) * 6; i++) out[i] = g_last_tasks[i];
  return n;
}

double orc_inner(void* A, void* B) { return ((QTensor*)A)->inner_product(*(QTensor*)B); }
double orc_norm(void* A) { return ((QTensor*)A)->norm(); }
void orc_add(void* A, void* B, double c) { ((QTensor*)A)->add(*(QTensor*)B, c); }
void orc_scale(void* A, double c) { *((QTensor*)A) *= c; }
void orc_prime(void* A, int inc) { ((QTensor*)A)->prime(inc); }
void orc_dag(void* A) { ((QTensor*)A)->dag(); }

// ---- effective Hamiltonian (big_qtensor) -------------------------------
void* orc_big_new(void* L, void* W1, void* W2, void* R) {
  BigQTensor* b = new BigQTensor();
  b->setLeft((QTensor*)L);
  if (W1) b->addMid((QTensor*)W1);
  if (W2) b->addMid((QTensor*)W2);
  b->setRight((QTensor*)R);
  return b;
}
void orc_big_free(void* b) { delete (BigQTensor*)b; }
void* orc_big_mid(void* b) { return new QTensor(((BigQTensor*)b)->mid); }
void* orc_big_product(void* b, void* x) {
  try { return new QTensor(((BigQTensor*)b)->product(*(QTensor*)x)); }
  catch (std::exception&) { g_err = "product failed"; return nullptr; }
}
double orc_big_expec(void* b, void* x) { return ((BigQTensor*)b)->expec(*(QTensor*)x); }
long orc_big_size(void* b) { ((BigQTensor*)b)->size_ = 0; return (long)((BigQTensor*)b)->size(); }

double orc_davidson(void* b, void* x, int max_restart, double tol, int* nmv) {
  uint64_t rng = 0x5eed;
  ((BigQTensor*)b)->size_ = 0;
  return tensor_davidsonIT(*(BigQTensor*)b, *(QTensor*)x, 3, max_restart, tol, 'S', &rng, nmv);
}

// seconds for n products (CPU baseline timing)
double orc_heff_time(void* b, void* x, int n) {
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < n; i++) {
    QTensor y = ((BigQTensor*)b)->product(*(QTensor*)x);
    (void)y;
  }
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

// host BLAS behind the oracle's block products (name, for the CPU baseline record)
const char* orc_blas_name() { return blas_name(); }

int orc_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---- env step: ((psi^dag' * T) * W) * psi  (dmrg.cpp:491-509) ------------
void* orc_env_step(void* psi, void* T, void* W) {
  try {
    QTensor t1 = *(QTensor*)psi;
    t1.dag(); t1.prime();
    QTensor t4 = t1 * (*(QTensor*)T);
    QTensor t5 = t4 * (*(QTensor*)W);
    return new QTensor(t5 * (*(QTensor*)psi));
  } catch (std::exception&) { g_err = "env step failed"; return nullptr; }
}

// ---- SVD ------------------------------------------------------------------
// svd_bond(combined, A_left, A_right, mid, ...) with mid = the index of
// A_left named mid_name (dmrg.cpp:318-327).  A_left / A_right are replaced.
double orc_svd_bond(void* combined, void* Al, void* Ar, const char* mid_name, int dir,
                    double cutoff, long K, int* kept, void** S_out) {
  try {
    QTensor* L = (QTensor*)Al;
    Index mid;
    bool found = false;
    for (auto& ix : L->idx_set) if (ix.name == mid_name) { mid = ix; found = true; break; }
    if (!found) { g_err = "mid index not found"; return -1; }
    QTensor S;
    unsigned kp = 0;
    double tr = svd_bond(*(QTensor*)combined, *L, *(QTensor*)Ar, mid, S, dir, cutoff, (unsigned long)K, &kp);
    if (kept) *kept = (int)kp;
    if (S_out) {
      // S as a rank-1 tensor over midCut with vector blocks
      QTensor* s = new QTensor(S);
      *S_out = s;
    }
    return tr;
  } catch (std::exception&) { g_err = "svd failed"; return -1; }
}

void orc_dense_svd(int rows, int cols, const double* A, double* U, double* S, double* Vt) {
  std::vector<double> u, s, v;
  dense_svd(rows, cols, A, u, s, v);
  memcpy(U, u.data(), u.size() * sizeof(double));
  memcpy(S, s.data(), s.size() * sizeof(double));
  memcpy(Vt, v.data(), v.size() * sizeof(double));
}

void orc_diag(int n, double* A, double* evals) { DIAG(n, A, evals); }

// ---- models -------------------------------------------------------------
void* orc_mpo_heisenberg_nn(int N, int id) {
  QTT* H = new QTT();
  heisenberg_mpo_nn((unsigned)N, (unsigned)id, *H);
  return H;
}

static const char* op_name(int c) { return c == 0 ? "Sz" : (c == 1 ? "S+" : "S-"); }

void* orc_mpo_autompo(int N, int id, int nterms, const double* coef, const int* op1, const int* i1,
                      const int* op2, const int* i2) {
  try {
    std::vector<Term> t;
    for (int k = 0; k < nterms; k++) t.push_back({coef[k], op_name(op1[k]), i1[k], op_name(op2[k]), i2[k]});
    QTT* H = new QTT();
    autompo_mpo((unsigned)N, (unsigned)id, t, *H);
    return H;
  } catch (std::exception&) { g_err = "autompo failed"; return nullptr; }
}

// Heisenberg terms over a lattice: kind 0 = square NN, 1 = square NN + NNN
// (J2 on type-2 bonds).  Returns the number of terms written (3 per bond).
int orc_heisenberg_terms(int kind, int Nx, int Ny, int yper, double J1, double J2, int cap,
                         double* coef, int* op1, int* i1, int* op2, int* i2) {
  std::vector<int> type;
  auto b = square_lattice(Nx, Ny, yper != 0, &type, kind == 1);
  std::vector<double> J;
  for (int t : type) J.push_back(t == 1 ? J1 : J2);
  auto terms = heisenberg_terms(b, J);
  int n = 0;
  for (auto& tm : terms) {
    if (n >= cap) break;
    coef[n] = tm.coef;
    op1[n] = tm.op1 == "Sz" ? 0 : (tm.op1 == "S+" ? 1 : 2);
    op2[n] = tm.op2 == "Sz" ? 0 : (tm.op2 == "S+" ? 1 : 2);
    i1[n] = tm.i1; i2[n] = tm.i2;
    n++;
  }
  return (int)terms.size();
}

void* orc_mps_product(int N, int id, const int* state) {
  QTT* p = new QTT();
  p->id = (unsigned)id; p->length = (unsigned)N; p->phy_dim = 2; p->N = 1;
  p->bond_dims.assign(N + 1, 1);
  p->phy_qn = {QN(-1, 0), QN(1, 0)};
  std::vector<unsigned> st(N);
  QN tot;
  for (int i = 0; i < N; i++) { st[i] = (unsigned)state[i]; tot += p->phy_qn[st[i]]; }
  p->totalQ = tot;
  p->allocateMPS(st.data());
  return p;
}

void* orc_qtt_tensor(void* q, int i) { return new QTensor(((QTT*)q)->A.at(i)); }
void orc_qtt_set_tensor(void* q, int i, void* t) { ((QTT*)q)->A.at(i) = *(QTensor*)t; }
int orc_qtt_length(void* q) { return (int)((QTT*)q)->length; }
int orc_qtt_center(void* q) { return ((QTT*)q)->center; }
void orc_qtt_free(void* q) { delete (QTT*)q; }
double orc_qtt_position(void* q, int site) { return ((QTT*)q)->position(site); }
void orc_qtt_normalize(void* q) { ((QTT*)q)->normalize(); }
void orc_qtt_bond_dims(void* q, int* out) {
  QTT* p = (QTT*)q;
  for (size_t i = 0; i < p->bond_dims.size(); i++) out[i] = (int)p->bond_dims[i];
}

// DMRG (dmrg.cpp:760-813).  halfsweep_E has 2*nsweeps entries; returns E.
// bond_E, trunc, kept (optional): per bond update, 2 (L-1) per sweep
double orc_dmrg(void* psi, void* H, int nsweeps, int nmb, const int* maxbd, int nc, const double* cut,
                int nmr, const int* mr, double tol, double* halfsweep_E, double* init_E,
                double* seconds, double* bond_E, double* trunc, int* kept) {
  try {
    std::vector<int> m(maxbd, maxbd + nmb), r(mr, mr + nmr);
    std::vector<double> c(cut, cut + nc);
    DmrgLog log;
    auto t0 = std::chrono::steady_clock::now();
    double E = dmrg(*(QTT*)psi, *(QTT*)H, nsweeps, m, c, r, tol, &log);
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    for (size_t i = 0; i < log.halfsweep_energy.size(); i++) halfsweep_E[i] = log.halfsweep_energy[i];
    for (size_t i = 0; i < log.bond_energy.size(); i++) {
      if (bond_E) bond_E[i] = log.bond_energy[i];
      if (trunc) trunc[i] = log.trunc_err[i];
      if (kept) kept[i] = (int)log.kept[i];
    }
    if (init_E) *init_E = log.initial_energy;
    return E;
  } catch (std::exception&) { g_err = "dmrg failed"; return 0; }
}

// -1 marks the tensor train as not canonical (position() then runs rc())
void orc_qtt_set_center(void* q, int c) { ((QTT*)q)->center = c; }

// CPU timing sample of a DMRG sweep: canonicalise to site `start`,
// normalise, build the right environments and the left ones up to `start`
// (out[0] seconds), then nbonds bond updates of a left-to-right half-sweep
// from `start`, each updateSite + updateEnv as in dmrg() (out[1] seconds in
// total, out[2] the last bond energy).
int orc_dmrg_bond_sample(void* psi_, void* H_, int max_bd, double cutoff, int max_restart,
                         double tol, int start, int nbonds, double* out) {
  try {
    QTT& psi = *(QTT*)psi_;
    QTT& H = *(QTT*)H_;
    const int L = (int)H.length;
    if (start < 0 || start >= L - 1) { g_err = "bond sample: bad start site"; return 1; }
    uint64_t rng = 0x5eed;
    auto t0 = std::chrono::steady_clock::now();
    psi.position(start);
    psi.normalize();
    std::vector<QTensor> TR(L), TL(L);
    buildEnv(psi, H, TR, TL);
    for (int i = 0; i < start; i++) updateEnv(psi, H, TR, TL, (unsigned)i, MoveFromLeft);
    auto t1 = std::chrono::steady_clock::now();
    double E = 0;
    for (int i = start; i < start + nbonds && i < L - 1; i++) {
      SiteStats s = updateSite(psi, H, TR, TL, (unsigned)i, MoveFromLeft, max_bd, cutoff,
                               max_restart, tol, &rng);
      E = s.energy;
      updateEnv(psi, H, TR, TL, (unsigned)i, MoveFromLeft);
    }
    auto t2 = std::chrono::steady_clock::now();
    out[0] = std::chrono::duration<double>(t1 - t0).count();
    out[1] = std::chrono::duration<double>(t2 - t1).count();
    out[2] = E;
    return 0;
  } catch (std::exception&) { g_err = "dmrg bond sample failed"; return 1; }
}

}  // extern "C"
