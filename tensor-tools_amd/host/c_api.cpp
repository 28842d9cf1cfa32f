// c_api.cpp -- extern "C" entry points of the C++ host layer (qtensor,
// big_qtensor, tensor_davidsonIT, svd_bond, qMPS/qMPO, dmrg) for bindings
// that cannot use C++ directly (the Python binding, bench.py).  Tensors
// cross this boundary in the flat format of include/ttgpu.h plus index names,
// prime levels and types.
#include <chrono>
#include <cstring>
#include <limits>
#include <memory>
#include <sstream>

#include "tt_gpu.hpp"

using namespace ttgpu;

namespace {
std::vector<qtensor_index> make_indices(int rank, const int* arrows, const int* levels,
                                        const int* types, const char* names, const int* nsec,
                                        const int* qn, const int* qdim) {
  std::vector<std::string> nm;
  std::stringstream ss(names ? names : "");
  std::string item;
  while (std::getline(ss, item, '\n')) nm.push_back(item);
  while ((int)nm.size() < rank) nm.push_back("");
  std::vector<qtensor_index> idx;
  int off = 0;
  for (int i = 0; i < rank; i++) {
    qtensor_index ix((arrow_type)arrows[i], nm[i], (index_type)types[i], (unsigned)levels[i]);
    for (int s = 0; s < nsec[i]; s++)
      ix.addQN(QN_t(qn[2 * (off + s)], qn[2 * (off + s) + 1]), (unsigned)qdim[off + s]);
    off += nsec[i];
    idx.push_back(ix);
  }
  return idx;
}
using Q = qtensor<double>;
std::string g_err;
}  // namespace

// every entry point converts a ttgpu::Error (or any other exception) into an
// error return: nullptr / NaN / -1, with the message in tth_last_error();
// nothing thrown below crosses the C boundary
#define TTH_GUARD(err, ...)                          \
  try {                                              \
    g_err.clear();                                   \
    __VA_ARGS__                                      \
  } catch (std::exception & e) {                     \
    g_err = e.what();                                \
    return err;                                      \
  } catch (...) {                                    \
    g_err = "unknown exception in the host layer";   \
    return err;                                      \
  }
static const double kNaN = std::numeric_limits<double>::quiet_NaN();

extern "C" {

const char* tth_last_error() { return g_err.c_str(); }

int tth_init(int device) { return ttg_init(device); }
void tth_set_precision(int prec) { set_heff_precision(prec); }
void tth_set_env_precision(int prec) { set_env_precision(prec); }

void* tth_tensor_import(int rank, const int* arrows, const int* levels, const int* types,
                        const char* names, const int* nsec, const int* qn, const int* qdim,
                        int nblocks, const int* block_qi, const double* data) {
  TTH_GUARD(nullptr,
    check(ttg_init(0), "init");
    std::unique_ptr<Q> t(new Q(make_indices(rank, arrows, levels, types, names, nsec, qn, qdim)));
    std::vector<std::vector<unsigned>> qi(nblocks, std::vector<unsigned>(rank));
    size_t tot = 0;
    for (int b = 0; b < nblocks; b++) {
      size_t sz = 1;
      for (int j = 0; j < rank; j++) {
        qi[b][j] = (unsigned)block_qi[b * rank + j];
        sz *= t->idx_set[j].qdim(qi[b][j]);
      }
      tot += sz;
    }
    std::vector<double> d(data, data + tot);
    t->from_host(qi, d);
    return t.release();
  )
}

void* tth_init_block(int rank, const int* arrows, const int* levels, const int* types,
                     const char* names, const int* nsec, const int* qn, const int* qdim) {
  TTH_GUARD(nullptr,
    check(ttg_init(0), "init");
    std::unique_ptr<Q> t(new Q(make_indices(rank, arrows, levels, types, names, nsec, qn, qdim)));
    t->initBlock();
    return t.release();
  )
}

// out: [rank, nsec_total, nblocks, names_len]
int tth_tensor_info(void* h, int* out, long* total) {
  TTH_GUARD(-1,
  Q* t = (Q*)h;
  int ns = 0, nl = 0;
  for (auto& ix : t->idx_set) { ns += (int)ix.size(); nl += (int)ix.name().size() + 1; }
  out[0] = (int)t->rank; out[1] = ns; out[2] = t->nblocks(); out[3] = nl;
  *total = (long)t->total_size();
  return 0;
  )
}

int tth_tensor_export(void* h, int* arrows, int* levels, int* types, char* names, int* nsec,
                      int* qn, int* qdim, int* block_qi, long* block_off, double* data) {
  TTH_GUARD(-1,
  Q* t = (Q*)h;
  std::string nm;
  int off = 0;
  for (unsigned i = 0; i < t->rank; i++) {
    auto& ix = t->idx_set[i];
    arrows[i] = (int)ix.arrow(); levels[i] = (int)ix.level(); types[i] = (int)ix.type();
    nm += ix.name() + "\n";
    nsec[i] = (int)ix.size();
    for (unsigned s = 0; s < ix.size(); s++) {
      qn[2 * (off + s)] = ix.qn(s).q0;
      qn[2 * (off + s) + 1] = ix.qn(s).q1;
      qdim[off + s] = (int)ix.qdim(s);
    }
    off += (int)ix.size();
  }
  memcpy(names, nm.c_str(), nm.size());
  int nb = t->nblocks();
  std::vector<int> qi((size_t)nb * t->rank + 1);
  std::vector<int64_t> bo(nb + 1);
  check(ttg_tensor_meta_download(t->dev.get(), qi.data(), bo.data()), "meta");
  for (size_t i = 0; i < (size_t)nb * t->rank; i++) block_qi[i] = qi[i];
  for (int b = 0; b <= nb; b++) block_off[b] = (long)bo[b];
  if (data) {
    auto d = t->to_host();
    memcpy(data, d.data(), d.size() * sizeof(double));
  }
  return 0;
  )
}

void tth_tensor_free(void* h) { delete (Q*)h; }
void* tth_tensor_copy(void* h) { TTH_GUARD(nullptr, return new Q(*(Q*)h);) }
void* tth_contract(void* A, void* B) { TTH_GUARD(nullptr, return new Q((*(Q*)A) * (*(Q*)B));) }
double tth_inner(void* A, void* B) { TTH_GUARD(kNaN, return ((Q*)A)->inner_product(*(Q*)B);) }
double tth_norm(void* A) { TTH_GUARD(kNaN, return ((Q*)A)->norm();) }
int tth_add(void* A, void* B, double c) { TTH_GUARD(-1, ((Q*)A)->add(*(Q*)B, c); return 0;) }
int tth_scale(void* A, double c) { TTH_GUARD(-1, *((Q*)A) *= c; return 0;) }
void tth_prime(void* A, int inc) { ((Q*)A)->prime(inc); }
int tth_dag(void* A) { TTH_GUARD(-1, ((Q*)A)->dag(); return 0;) }

// ---- big_qtensor ----------------------------------------------------------
void* tth_big_new(void* L, void* W1, void* W2, void* R) {
  TTH_GUARD(nullptr,
    std::unique_ptr<big_qtensor<double>> b(new big_qtensor<double>());
    b->setLeft((Q*)L);
    if (W1) b->addMid((Q*)W1);
    if (W2) b->addMid((Q*)W2);
    b->setRight((Q*)R);
    return b.release();
  )
}
void tth_big_free(void* b) { delete (big_qtensor<double>*)b; }
void* tth_big_product(void* b, void* x) { TTH_GUARD(nullptr, return new Q(((big_qtensor<double>*)b)->product(*(Q*)x));) }
double tth_big_expec(void* b, void* x) { TTH_GUARD(kNaN, return ((big_qtensor<double>*)b)->expec(*(Q*)x);) }
long tth_big_size(void* b) { TTH_GUARD(-1, return (long)((big_qtensor<double>*)b)->size();) }
// the device operator (plans it for x's indices first)
void* tth_big_heff(void* b, void* x) {
  TTH_GUARD(nullptr,
    auto* bq = (big_qtensor<double>*)b;
    Q y;
    bq->product_into(*(Q*)x, y);
    return bq->op();
  )
}
// device time per product (ms) over n graph replays, x widened to the
// operator's block list first; ms3: the three stage kernels timed apart
int tth_big_time(void* b, void* x, int n, double* ms, double* ms3) {
  TTH_GUARD(-1,
    auto* bq = (big_qtensor<double>*)b;
    Q w = bq->widen(*(Q*)x);
    Q y;
    bq->product_into(w, y);
    check(ttg_heff_time(bq->op(), w.dev.get(), y.dev.get(), n, ms), "heff_time");
    if (ms3) check(ttg_heff_stage_time(bq->op(), w.dev.get(), n, ms3), "heff_stage_time");
    return 0;
  )
}
// flops per product, per-stage flops / task counts / tile shape / tile count
int tth_big_info(void* b, void* x, double* flops, double* sflops3, long* ntasks3, int* cfg3,
                 int* ntiles3) {
  TTH_GUARD(-1,
    auto* bq = (big_qtensor<double>*)b;
    Q y;
    bq->product_into(*(Q*)x, y);
    int64_t nt[3];
    check(ttg_heff_info(bq->op(), flops, nt, sflops3), "heff_info");
    for (int k = 0; k < 3; k++) ntasks3[k] = (long)nt[k];
    check(ttg_heff_tiles(bq->op(), cfg3, ntiles3), "heff_tiles");
    return 0;
  )
}
double tth_davidson(void* b, void* x, int max_restart, double tol, int* nmv, double* resid) {
  TTH_GUARD(kNaN,
    DavidsonStats st;
    double e = tensor_davidsonIT(*(big_qtensor<double>*)b, *(Q*)x, 3, max_restart, tol, 'S', &st);
    if (nmv) *nmv = st.n_matvec;
    if (resid) *resid = st.residual;
    return e;
  )
}

// ---- svd ------------------------------------------------------------------
double tth_svd_bond(void* combined, void* Al, void* Ar, const char* mid_name, int dir, double cutoff,
                    long K, int* kept, void** S_out) {
  TTH_GUARD(-1,
    Q* L = (Q*)Al;
    qtensor_index mid;
    bool found = false;
    for (auto& ix : L->idx_set) if (ix.name() == mid_name) { mid = ix; found = true; break; }
    if (!found) { g_err = "svd_bond: mid index not found"; return -1; }
    Q S;
    unsigned kp = 0;
    double tr = svd_bond(*(Q*)combined, *L, *(Q*)Ar, mid, S, dir, cutoff, (long unsigned)K, &kp);
    if (kept) *kept = (int)kp;
    if (S_out) *S_out = new Q(S);
    return tr;
  )
}

// ---- models / MPS ------------------------------------------------------------
void* tth_mpo_heisenberg_nn(int N, int id) {
  TTH_GUARD(nullptr,
    check(ttg_init(0), "init");
    return new qMPO<double>(heisenberg_mpo_nn((unsigned)N, (unsigned)id));
  )
}
static const char* opn(int c) { return c == 0 ? "Sz" : (c == 1 ? "S+" : "S-"); }
void* tth_mpo_autompo(int N, int id, int nt, const double* coef, const int* o1, const int* i1,
                      const int* o2, const int* i2) {
  TTH_GUARD(nullptr,
    check(ttg_init(0), "init");
    std::vector<SpinTerm> t;
    for (int k = 0; k < nt; k++) t.push_back({coef[k], opn(o1[k]), i1[k], opn(o2[k]), i2[k]});
    return new qMPO<double>(autompo_mpo((unsigned)N, (unsigned)id, t));
  )
}
int tth_heisenberg_terms(int kind, int Nx, int Ny, int yper, double J1, double J2, int cap,
                         double* coef, int* o1, int* i1, int* o2, int* i2) {
  TTH_GUARD(-1,
  std::vector<int> type;
  auto b = square_lattice(Nx, Ny, yper != 0, &type, kind == 1);
  std::vector<double> J;
  for (int x : type) J.push_back(x == 1 ? J1 : J2);
  auto t = heisenberg_terms(b, J);
  int n = 0;
  for (auto& x : t) {
    if (n >= cap) break;
    coef[n] = x.coef;
    o1[n] = x.op1 == "Sz" ? 0 : (x.op1 == "S+" ? 1 : 2);
    o2[n] = x.op2 == "Sz" ? 0 : (x.op2 == "S+" ? 1 : 2);
    i1[n] = x.i1; i2[n] = x.i2;
    n++;
  }
  return (int)t.size();
  )
}
void* tth_mps_product(int N, int id, const int* state) {
  TTH_GUARD(nullptr,
    check(ttg_init(0), "init");
    std::vector<unsigned> st(state, state + N);
    return new qMPS<double>(product_state_mps((unsigned)id, st));
  )
}
void* tth_mpo_tensor(void* q, int i) { TTH_GUARD(nullptr, return new Q(((qMPO<double>*)q)->A.at(i));) }
int tth_mpo_length(void* q) { return (int)((qMPO<double>*)q)->length; }
void* tth_qtt_tensor(void* q, int i) { TTH_GUARD(nullptr, return new Q(((qMPS<double>*)q)->A.at(i));) }
int tth_qtt_set_tensor(void* q, int i, void* t) { TTH_GUARD(-1, ((qMPS<double>*)q)->A.at(i) = *(Q*)t; return 0;) }
int tth_qtt_length(void* q) { return (int)((qMPS<double>*)q)->length; }
int tth_qtt_center(void* q) { return ((qMPS<double>*)q)->center; }
// -1 marks the tensor train as not canonical (the next position() runs rc())
void tth_qtt_set_center(void* q, int c) { ((qMPS<double>*)q)->center = c; }
// empty trains, filled by load (checkpoint.cpp)
void* tth_mps_new() { TTH_GUARD(nullptr, return new qMPS<double>();) }
void* tth_mpo_new() { TTH_GUARD(nullptr, return new qMPO<double>();) }
void tth_mps_free(void* q) { delete (qMPS<double>*)q; }
void tth_mpo_free(void* q) { delete (qMPO<double>*)q; }
double tth_mps_position(void* q, int site) { TTH_GUARD(kNaN, return ((qMPS<double>*)q)->position(site);) }
int tth_mps_normalize(void* q) { TTH_GUARD(-1, ((qMPS<double>*)q)->normalize(); return 0;) }
void tth_qtt_bond_dims(void* q, int* out) {
  auto* p = (qMPS<double>*)q;
  for (size_t i = 0; i < p->bond_dims.size(); i++) out[i] = (int)p->bond_dims[i];
}

// mps/observables.cpp:64-105, 415-450
double tth_psiHphi(void* psi, void* H, void* phi) {
  TTH_GUARD(kNaN, return psiHphi(*(qMPS<double>*)psi, *(qMPO<double>*)H, *(qMPS<double>*)phi);)
}
double tth_psiphi(void* psi, void* phi) {
  TTH_GUARD(kNaN, return psiphi(*(qMPS<double>*)psi, *(qMPS<double>*)phi);)
}

void tth_set_dmrg_progress(int every) { set_dmrg_progress(every); }

// stats: [t_davidson, t_svd, t_env, n_matvec, seconds, largest truncation
// error, sum over the last half-sweep, largest device-pool live bytes after a
// bond]; bond_E, trunc, kept, env (optional): per bond update in sweep order,
// `per` bonds per half-sweep (env: elements of the environment it produced)
static void copy_log(const DmrgLog& log, double secs, size_t per, double* halfsweep_E,
                     double* halfsweep_s, double* init_E, double* stats, double* bond_E, double* trunc,
                     int* kept, double* env) {
  for (size_t i = 0; i < log.halfsweep_energy.size(); i++) {
    halfsweep_E[i] = log.halfsweep_energy[i];
    if (halfsweep_s) halfsweep_s[i] = log.halfsweep_seconds[i];
  }
  if (init_E) *init_E = log.initial_energy;
  for (size_t i = 0; i < log.bond_energy.size(); i++) {
    if (bond_E) bond_E[i] = log.bond_energy[i];
    if (trunc) trunc[i] = log.trunc_err[i];
    if (kept) kept[i] = (int)log.kept[i];
    if (env) env[i] = (double)log.env_elems[i];
  }
  if (stats) {
    stats[0] = log.t_dav; stats[1] = log.t_svd; stats[2] = log.t_env;
    stats[3] = (double)log.n_matvec;
    stats[4] = secs;
    double tmax = 0, tlast = 0;
    for (double t : log.trunc_err) tmax = std::max(tmax, t);
    const size_t nb = log.trunc_err.size();
    for (size_t i = (nb > per ? nb - per : 0); i < nb; i++) tlast += log.trunc_err[i];
    stats[5] = tmax;
    stats[6] = tlast;
    stats[7] = (double)log.pool_live_max;
  }
}

static double secs_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// dmrg.cpp:702-755 (schedules cycled per sweep)
double tth_dmrg(void* psi, void* H, int nsweeps, int nmb, const int* maxbd, int nc, const double* cut,
                int nmr, const int* mr, double tol, double* halfsweep_E, double* halfsweep_s,
                double* init_E, double* stats, double* bond_E, double* trunc, int* kept, double* env) {
  TTH_GUARD(kNaN,
    std::vector<int> m(maxbd, maxbd + nmb), r(mr, mr + nmr);
    std::vector<double> c(cut, cut + nc);
    DmrgLog log;
    auto t0 = std::chrono::steady_clock::now();
    auto* q = (qMPS<double>*)psi;
    double E = dmrg(*q, *(qMPO<double>*)H, nsweeps, m, c, r, tol, &log);
    copy_log(log, secs_since(t0), q->length - 1, halfsweep_E, halfsweep_s, init_E, stats, bond_E, trunc,
             kept, env);
    return E;
  )
}

// dmrg.cpp:760-813 (bonds of sites [start, stop])
double tth_dmrg_range(void* psi, void* H, int start, int stop, int nsweeps, int nmb, const int* maxbd,
                      int nc, const double* cut, int nmr, const int* mr, double tol, double* halfsweep_E,
                      double* halfsweep_s, double* init_E, double* stats, double* bond_E, double* trunc,
                      int* kept, double* env) {
  TTH_GUARD(kNaN,
    if (start < 0 || stop <= start) fail("dmrg: bad site range");
    std::vector<int> m(maxbd, maxbd + nmb), r(mr, mr + nmr);
    std::vector<double> c(cut, cut + nc);
    DmrgLog log;
    auto t0 = std::chrono::steady_clock::now();
    double E = dmrg(*(qMPS<double>*)psi, *(qMPO<double>*)H, (unsigned)start, (unsigned)stop, nsweeps, m, c,
                    r, tol, &log);
    copy_log(log, secs_since(t0), (size_t)(stop - start), halfsweep_E, halfsweep_s, init_E, stats, bond_E,
             trunc, kept, env);
    return E;
  )
}

// dmrg.cpp:653-697 (half-sweeps start_sweep .. num_half-1; the resume form)
double tth_dmrg_resume(void* psi, void* H, int num_half, int max_bd, double cutoff, int max_restart,
                       int start_sweep, double tol, double* halfsweep_E, double* halfsweep_s,
                       double* init_E, double* stats, double* bond_E, double* trunc, int* kept,
                       double* env) {
  TTH_GUARD(kNaN,
    DmrgLog log;
    auto t0 = std::chrono::steady_clock::now();
    auto* q = (qMPS<double>*)psi;
    double E = dmrg(*q, *(qMPO<double>*)H, num_half, max_bd, cutoff, 'S', 3, max_restart, start_sweep,
                    tol, &log);
    copy_log(log, secs_since(t0), q->length - 1, halfsweep_E, halfsweep_s, init_E, stats, bond_E, trunc,
             kept, env);
    return E;
  )
}

void tth_set_env_checkpoint_stride(int k) { set_env_checkpoint_stride(k < 1 ? 1 : k); }

}  // extern "C"
