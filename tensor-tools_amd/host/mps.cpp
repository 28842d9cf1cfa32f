// mps.cpp -- qMPS/qMPO containers, spin-1/2 Heisenberg MPO builders, SVD
// wrappers and the two-site DMRG driver of the reference, over device blocks.
//   mps/qtt.cpp:60-78, 205-311, 821-994; models/hams/Heisenberg.cpp:136-180,
//   333-635; models/hams/AutoMPO.{h,cpp}; models/lattice/square.h;
//   qtensor/qtensor_op.cpp:1097-1544; algos/dmrg/dmrg.cpp:56-84, 300-372,
//   491-509, 760-813.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <unordered_map>

#include "tt_gpu.hpp"

namespace ttgpu {

// ---------------------------------------------------------------- host blocks
// Block enumeration on the host for the input builders only (MPO / product
// state coefficients are written by key before upload).  Same order as
// qtensor::initBlock (qtensor.cpp:302-345).
namespace {
struct HostBlocks {
  std::vector<std::vector<unsigned>> qi;
  std::vector<size_t> off;
  std::unordered_map<std::string, unsigned> by_key;
  std::vector<double> data;
};
std::string key_of(const std::vector<QN_t>& q) {
  std::string s;
  for (auto& x : q) s += to_string(x) + " ";
  return s;
}
HostBlocks host_init_blocks(const std::vector<qtensor_index>& idx) {
  HostBlocks hb;
  unsigned rank = (unsigned)idx.size();
  std::vector<unsigned> stride(1, 1);
  for (unsigned i = 0; i < rank; i++) stride.push_back(idx[i].size() * stride[i]);
  size_t total = 0;
  for (unsigned i = 0; i < stride.back(); i++) {
    std::vector<unsigned> pos(rank);
    std::vector<QN_t> qns(rank);
    size_t size = 1;
    QN_t tot;
    for (unsigned j = 0; j < rank; j++) {
      pos[j] = (i / stride[j]) % idx[j].size();
      qns[j] = idx[j].qn(pos[j]);
      size *= idx[j].qdim(pos[j]);
      if (idx[j].arrow() == Inward) tot += qns[j]; else tot -= qns[j];
    }
    if (tot == QN_t()) {
      hb.by_key[key_of(qns)] = (unsigned)hb.qi.size();
      hb.qi.push_back(pos);
      hb.off.push_back(total);
      total += size;
    }
  }
  hb.off.push_back(total);
  hb.data.assign(total, 0.0);
  return hb;
}
std::string link_name(unsigned id, unsigned i) { return "ID" + std::to_string(id) + "Link" + std::to_string(i); }
}  // namespace

// ---------------------------------------------------------------- sites
// models/sites/spinhalf.cpp:23-46 (Dn = 0, Up = 1)
static double spinhalf_op(unsigned bra, const std::string& op, unsigned ket) {
  if (op == "Id" || op == "I") return ket == bra ? 1 : 0;
  if (op == "Sz") return ket != bra ? 0 : (ket == 0 ? -0.5 : 0.5);
  if (op == "S+") return (ket == 0 && bra == 1) ? 1 : 0;
  if (op == "S-") return (ket == 1 && bra == 0) ? 1 : 0;
  fail("spin-1/2 operator lookup");
}
static QN_t spinhalf_div(const std::string& op) {   // spinhalf.cpp:95-106
  if (op == "S+") return QN_t(2, 0);
  if (op == "S-") return QN_t(-2, 0);
  return QN_t();
}

// ---------------------------------------------------------------- MPO builders
namespace {
struct MpoBuild {
  unsigned N, id;
  std::vector<QN_t> phy_qn{QN_t(-1, 0), QN_t(1, 0)};
  std::vector<std::vector<qtensor_index>> idx;
  std::vector<HostBlocks> hb;
  // Heisenberg.cpp:136-180: write val*<b|op|k> into block (Qi, k, b, Qo) at
  // offset s0*r + s3*c (s3*c on site 0); a write overwrites
  void add_operator(unsigned site, unsigned r, unsigned c, const std::string& op, double val,
                    QN_t Qi, QN_t Qo) {
    bool added = false;
    for (unsigned k = 0; k < 2; k++)
      for (unsigned b = 0; b < 2; b++) {
        std::string s = to_string(Qi) + " " + to_string(phy_qn[k]) + " " + to_string(phy_qn[b]) +
                        " " + to_string(Qo) + " ";
        auto it = hb[site].by_key.find(s);
        if (it == hb[site].by_key.end()) continue;
        added = true;
        unsigned bp = it->second;
        const auto& q = hb[site].qi[bp];
        size_t s1 = idx[site][0].qdim(q[0]);
        size_t s2 = s1 * idx[site][1].qdim(q[1]);
        size_t s3 = s2 * idx[site][2].qdim(q[2]);
        size_t off = (site == 0) ? s3 * c : r + s3 * c;
        if (off >= hb[site].off[bp + 1] - hb[site].off[bp]) fail("MPO coefficient outside its block");
        hb[site].data[hb[site].off[bp] + off] = val * spinhalf_op(b, op, k);
      }
    if (!added) fail("MPO operator has no block");
  }
  qMPO<double> finish() {
    qMPO<double> H;
    H._id = id; H.length = N; H.phy_dim = 2; H.phy_qn = phy_qn;
    H.bond_dims.assign(N + 1, 1);
    for (unsigned i = 0; i < N; i++) {
      qtensor<double> t(idx[i]);
      t.from_host(hb[i].qi, hb[i].data);
      H.A.push_back(std::move(t));
    }
    H.tensors_allocated = true;
    return H;
  }
  std::vector<qtensor_index> site_indices(unsigned i) {
    return {qtensor_index(Inward, link_name(id, i), Link, 0),
            qtensor_index(Outward, "Site" + std::to_string(i), Site, 0),
            qtensor_index(Inward, "Site" + std::to_string(i), Site, 1),
            qtensor_index(Outward, link_name(id, i + 1), Link, 0)};
  }
};
}  // namespace

// Heisenberg.cpp:333-402
qMPO<double> heisenberg_mpo_nn(unsigned N, unsigned id) {
  MpoBuild b{N, id};
  for (unsigned i = 0; i < N; i++) {
    auto ix = b.site_indices(i);
    if (i == 0) ix[0].addQN(QN_t(), 1);
    else { ix[0].addQN(QN_t(), 3); ix[0].addQN(QN_t(-2, 0), 1); ix[0].addQN(QN_t(2, 0), 1); }
    for (auto& q : b.phy_qn) { ix[1].addQN(q, 1); ix[2].addQN(q, 1); }
    if (i == N - 1) ix[3].addQN(QN_t(), 1);
    else { ix[3].addQN(QN_t(), 3); ix[3].addQN(QN_t(-2, 0), 1); ix[3].addQN(QN_t(2, 0), 1); }
    b.hb.push_back(host_init_blocks(ix));
    b.idx.push_back(ix);
  }
  for (unsigned s = 0; s < N; s++) {
    if (s > 0) b.add_operator(s, 0, 0, "Id", 1.0, QN_t(), QN_t());
    if (s < N - 1) b.add_operator(s, 1, 1, "Id", 1.0, QN_t(), QN_t());
    if (s > 0) b.add_operator(s, 2, 0, "Sz", 1.0, QN_t(), QN_t());
    if (s < N - 1) b.add_operator(s, 1, 2, "Sz", 1.0, QN_t(), QN_t());
    if (s > 0) b.add_operator(s, 0, 0, "S+", 1.0, QN_t(-2, 0), QN_t());
    if (s > 0) b.add_operator(s, 0, 0, "S-", 1.0, QN_t(2, 0), QN_t());
    if (s < N - 1) b.add_operator(s, 1, 0, "S-", 0.5, QN_t(), QN_t(-2, 0));
    if (s < N - 1) b.add_operator(s, 1, 0, "S+", 0.5, QN_t(), QN_t(2, 0));
  }
  return b.finish();
}

// AutoMPO.{h,cpp}: SiteTerm order (site, then op); HTerm set ordered by
// LessNoCoef with duplicates summed.  Heisenberg.cpp:416-635: link bases
// (IL, HL and every first operator of a term crossing the link), the
// zero-QN-first std::sort of each basis, sector grouping, basisToIndex
// numbering and the start / continue / end operator placements.
namespace {
struct SiteTerm {
  std::string op; int i = -1;
  bool operator==(const SiteTerm& o) const { return op == o.op && i == o.i; }
  bool operator!=(const SiteTerm& o) const { return !(*this == o); }
  bool operator<(const SiteTerm& o) const { return i != o.i ? i < o.i : op < o.op; }
};
struct HTerm {
  double coef = 1;
  std::vector<SiteTerm> ops;
  const SiteTerm& first() const { return ops.front(); }
  const SiteTerm& last() const { return ops.back(); }
};
struct LessNoCoef {
  bool operator()(const HTerm& a, const HTerm& b) const {
    if (a.ops.size() != b.ops.size()) return a.ops.size() < b.ops.size();
    for (size_t j = 0; j < a.ops.size(); j++)
      if (a.ops[j] != b.ops[j]) return a.ops[j] < b.ops[j];
    return false;
  }
};
struct SiteQN { SiteTerm st; QN_t q; };
unsigned basis_to_index(QN_t q, std::map<QN_t, unsigned>& maxidx,
                        std::unordered_map<unsigned, unsigned>& rc, unsigned i) {
  if (!maxidx.count(q)) maxidx[q] = 1;
  if (!rc.count(i)) rc[i] = maxidx[q]++;
  return rc[i] - 1;
}
}  // namespace

qMPO<double> autompo_mpo(unsigned N, unsigned id, const std::vector<SpinTerm>& in_terms) {
  std::set<HTerm, LessNoCoef> terms;
  for (auto& tm : in_terms) {
    HTerm h;
    for (auto& op : {std::make_pair(tm.op1, tm.i1), std::make_pair(tm.op2, tm.i2)}) {
      auto it = h.ops.begin();
      while (it != h.ops.end() && it->i <= op.second) ++it;
      h.ops.insert(it, SiteTerm{op.first, op.second});
    }
    h.coef = tm.coef;
    if (h.coef == 0.0) continue;
    auto it = terms.find(h);
    if (it == terms.end()) terms.insert(h);
    else { HTerm nt = h; nt.coef += it->coef; terms.erase(it); terms.insert(nt); }
  }
  MpoBuild b{N, id};
  const SiteTerm IL{"IL", 0}, HL{"HL", 0};
  std::vector<std::vector<SiteQN>> basis(N + 1);
  for (unsigned n = 0; n < N; n++) basis[n].push_back({IL, QN_t()});
  for (unsigned n = 1; n <= N; n++) basis[n].push_back({HL, QN_t()});
  for (auto& ht : terms)
    for (int n = ht.first().i + 1; n <= ht.last().i + 1; ++n) {
      auto& bn = basis.at(n);
      bool has = false;
      for (auto& sq : bn) if (sq.st == ht.first()) { has = true; break; }
      if (!has) bn.push_back({ht.first(), -spinhalf_div(ht.first().op)});
    }
  const QN_t Zero;
  auto qn_comp = [&Zero](const SiteQN& a, const SiteQN& c) {
    if (a.q == Zero && c.q != Zero) return true;
    if (c.q == Zero && a.q != Zero) return false;
    return a.q < c.q;
  };
  for (auto& bn : basis) std::sort(bn.begin(), bn.end(), qn_comp);
  std::vector<std::vector<std::pair<QN_t, unsigned>>> links(N + 1);
  for (unsigned n = 0; n <= N; n++) {
    QN_t cq = basis[n].front().q;
    unsigned cm = 0;
    for (auto& sq : basis[n]) {
      if (sq.q == cq) ++cm;
      else { links[n].push_back({cq, cm}); cq = sq.q; cm = 1; }
    }
    links[n].push_back({cq, cm});
  }
  std::vector<std::vector<HTerm>> ht_by_n(N + 1);
  for (auto& ht : terms)
    for (auto& st : ht.ops) ht_by_n.at(st.i + 1).push_back(ht);
  std::vector<std::map<QN_t, unsigned>> maxidx(N + 1);
  std::vector<std::unordered_map<unsigned, unsigned>> rcidx(N + 1);
  for (unsigned i = 0; i < N; i++) {
    auto ix = b.site_indices(i);
    for (auto& rq : links[i]) ix[0].addQN(rq.first, rq.second);
    if (i != N - 1) for (auto& cq : links[i + 1]) ix[3].addQN(cq.first, cq.second);
    else ix[3].addQN(QN_t(), 1);
    for (auto& q : b.phy_qn) { ix[1].addQN(q, 1); ix[2].addQN(q, 1); }
    b.hb.push_back(host_init_blocks(ix));
    b.idx.push_back(ix);
    auto& bn1 = basis[i];
    auto& bn = basis[i + 1];
    unsigned c_max = (i == N - 1) ? 1 : (unsigned)bn.size();
    for (unsigned r_ = 0; r_ < bn1.size(); r_++)
      for (unsigned c_ = 0; c_ < c_max; c_++) {
        const SiteTerm& rst = bn1[r_].st;
        const SiteTerm& cst = bn[c_].st;
        QN_t qr = -bn1[r_].q, qc = -bn[c_].q;
        unsigned r = (i == 0) ? 1 : r_, c = c_;
        auto R_ = [&]() { return basis_to_index(-qr, maxidx[i], rcidx[i], r); };
        auto C_ = [&]() { return basis_to_index(qc, maxidx[i + 1], rcidx[i + 1], c); };
        if (cst.i == (int)i && rst == IL && cst.op != "HL" && cst.op != "IL") {
          unsigned tr = R_(), tc = C_();
          if (i == 0) tr = 1;
          b.add_operator(i, tr, tc, cst.op, 1.0, qr, qc);
        }
        if (cst == rst) {
          unsigned tr = R_(), tc = C_();
          b.add_operator(i, tr, tc, "Id", 1.0, qr, qc);
        }
        if (cst == HL)
          for (auto& ht : ht_by_n[i + 1])
            if (rst == ht.first() && ht.last().i == (int)i) {
              unsigned tr = R_(), tc = C_();
              b.add_operator(i, tr, tc, ht.last().op, ht.coef, qr, qc);
            }
        if (rst == IL && cst == HL)
          for (auto& ht : ht_by_n[i + 1])
            if (ht.first().i == ht.last().i) {
              unsigned tr = R_(), tc = C_();
              b.add_operator(i, tr, tc, ht.first().op, ht.coef, qr, qc);
            }
      }
  }
  return b.finish();
}

// models/lattice/square.h:25-108 (0-based sites)
std::vector<std::pair<int, int>> square_lattice(int Nx, int Ny, bool yperiodic,
                                                std::vector<int>* type, bool nnn) {
  yperiodic = yperiodic && (Ny > 2);
  std::vector<std::pair<int, int>> out;
  auto add = [&](int a, int c, int t) { out.push_back({a - 1, c - 1}); if (type) type->push_back(t); };
  for (int n = 1; n <= Nx * Ny; ++n) {
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
  return out;
}

std::vector<SpinTerm> heisenberg_terms(const std::vector<std::pair<int, int>>& bonds,
                                       const std::vector<double>& J) {
  std::vector<SpinTerm> t;
  for (size_t k = 0; k < bonds.size(); k++) {
    double j = J.size() == bonds.size() ? J[k] : J.at(0);
    t.push_back({j, "Sz", bonds[k].first, "Sz", bonds[k].second});
    t.push_back({0.5 * j, "S+", bonds[k].first, "S-", bonds[k].second});
    t.push_back({0.5 * j, "S-", bonds[k].first, "S+", bonds[k].second});
  }
  return t;
}

// ---------------------------------------------------------------- MPS
// qtt.cpp:205-311 -- link QN sets (forward reachable, backward pruned),
// all legal blocks, the product-state path set to 1
qMPS<double> product_state_mps(unsigned id, const std::vector<unsigned>& state) {
  qMPS<double> p;
  const unsigned L = (unsigned)state.size();
  p._id = id; p.length = L; p.phy_dim = 2; p.phy_qn = {QN_t(-1, 0), QN_t(1, 0)};
  p.bond_dims.assign(L + 1, 1);
  for (unsigned s : state) p.totalQ += p.phy_qn.at(s);
  std::vector<std::set<QN_t>> LQ(L + 1);
  LQ[0].insert(QN_t());
  LQ[L].insert(p.totalQ);
  for (unsigned j = 1; j < L; j++)
    for (auto& q : LQ[j - 1])
      for (auto& ph : p.phy_qn) LQ[j].insert(q + ph);
  for (unsigned j = L - 1; j > 0; j--) {
    std::vector<QN_t> rem;
    for (auto& q : LQ[j]) {
      bool ok = false;
      for (auto& ph : p.phy_qn) if (LQ[j + 1].count(q + ph)) { ok = true; break; }
      if (!ok) rem.push_back(q);
    }
    for (auto& q : rem) LQ[j].erase(q);
  }
  QN_t last;
  for (unsigned i = 0; i < L; i++) {
    std::vector<qtensor_index> ix{qtensor_index(Inward, link_name(id, i), Link, 0),
                                  qtensor_index(Inward, "Site" + std::to_string(i), Site, 0),
                                  qtensor_index(Outward, link_name(id, i + 1), Link, 0)};
    for (auto& q : LQ[i]) ix[0].addQN(q, 1);
    for (auto& q : p.phy_qn) ix[1].addQN(q, 1);
    for (auto& q : LQ[i + 1]) ix[2].addQN(q, 1);
    HostBlocks hb = host_init_blocks(ix);
    std::string key = to_string(last) + " " + to_string(p.phy_qn[state[i]]) + " ";
    last += p.phy_qn[state[i]];
    key += to_string(last) + " ";
    auto it = hb.by_key.find(key);
    if (it == hb.by_key.end()) fail("product-state block missing");
    hb.data[hb.off[it->second]] = 1.0;
    qtensor<double> t(ix);
    t.from_host(hb.qi, hb.data);
    p.A.push_back(std::move(t));
  }
  p.center = (int)L - 1;
  p.tensors_allocated = true;
  return p;
}

// ---------------------------------------------------------------- SVD
// qtensor_op.cpp:1097-1453 on the device: the left indices must be A's
// leading indices in order (true for every caller in the DMRG flow).
double svd(qtensor<double>& A, std::vector<qtensor_index>& left, std::vector<qtensor_index>& right,
           qtensor<double>& U, qtensor<double>& V, qtensor<double>& S, int direction, double cutoff,
           unsigned K, unsigned* kept) {
  if (left.size() + right.size() != A.rank) fail("svd: index count");
  for (size_t j = 0; j < left.size(); j++)
    if (!(A.idx_set[j] == left[j])) fail("svd: left indices must lead A's index list");
  for (size_t j = 0; j < right.size(); j++)
    if (!(A.idx_set[left.size() + j] == right[j])) fail("svd: right indices must follow the left ones");
  ttg_tensor u, s, v;
  double trunc = 0;
  int64_t kp = 0;
  check(ttg_svd(A.dev.get(), (int)left.size(), direction, cutoff, (int64_t)K, &u, &s, &v, &trunc, &kp),
        "svd");
  // midCut: Outward, sectors of the kept singular values (qtensor_op.cpp:1303-1306)
  int nsec = 0;
  check(ttg_tensor_index_download(s, 0, &nsec, nullptr, nullptr), "svd mid");
  std::vector<int> qn(2 * std::max(nsec, 1)), qd(std::max(nsec, 1));
  check(ttg_tensor_index_download(s, 0, &nsec, qn.data(), qd.data()), "svd mid");
  qtensor_index midCut(Outward, "svdmid", Link, 0);
  for (int k = 0; k < nsec; k++) midCut.addQN(QN_t(qn[2 * k], qn[2 * k + 1]), (unsigned)qd[k]);
  std::vector<qtensor_index> uidx(left);
  uidx.push_back(midCut);
  qtensor_index md = midCut;
  md.dag();
  std::vector<qtensor_index> vidx{md};
  vidx.insert(vidx.end(), right.begin(), right.end());
  U = qtensor<double>(uidx); U.dev = adopt(u); U._initted = true;
  V = qtensor<double>(vidx); V.dev = adopt(v); V._initted = true;
  S = qtensor<double>(std::vector<qtensor_index>{midCut}); S.dev = adopt(s); S._initted = true;
  if (kept) *kept = (unsigned)kp;
  return trunc;
}

// qtensor_op.cpp:1506-1544
double svd_bond(qtensor<double>& combined, qtensor<double>& A_left, qtensor<double>& A_right,
                qtensor_index& mid, qtensor<double>& S, int direction, double cutoff,
                long unsigned K, unsigned* kept) {
  qtensor<double> U, V;
  std::vector<qtensor_index> left, right;
  qtensor_index dag_mid(mid);
  dag_mid.dag();
  for (auto& ix : A_left.idx_set) if (ix.tag() != mid.tag()) left.push_back(ix);
  for (auto& ix : A_right.idx_set) if (ix.tag() != dag_mid.tag()) right.push_back(ix);
  double tr = svd(combined, left, right, U, V, S, direction, cutoff, (unsigned)K, kept);
  U.idx_set.back().rename(mid.name());
  U.idx_set.back().retype(mid.type());
  V.idx_set[0].rename(dag_mid.name());
  V.idx_set[0].retype(dag_mid.type());
  if (U.idx_set.back().level() != mid.level())
    U.idx_set.back().prime((int)mid.level() - (int)U.idx_set.back().level());
  if (V.idx_set[0].level() != dag_mid.level())
    V.idx_set[0].prime((int)dag_mid.level() - (int)V.idx_set[0].level());
  A_right = std::move(V);
  A_left = std::move(U);
  return tr;
}

// qtensor.h:166-173: -sum s^2 log s^2 over s > cutoff
double calcEntropy(qtensor<double>& S, double cutoff) {
  std::vector<double> s = S.to_host();
  double e = 0;
  for (double v : s) if (v > cutoff) e -= v * v * std::log(v * v);
  return e;
}

// progress lines on stderr every g_progress_every sites or bonds (0: silent)
static int g_progress_every = 0;
void set_dmrg_progress(int every) { g_progress_every = every; }
// environments kept between half-sweeps: every g_env_stride-th site (and the
// range ends); the others are recomputed from the nearest kept one when a
// bond needs them (1: keep every environment, one full set)
static int g_env_stride = 1;
void set_env_checkpoint_stride(int k) {
  if (k < 1) fail("env checkpoint stride must be >= 1");
  g_env_stride = k;
}

// ---------------------------------------------------------------- canonical form
static void rename_like(qtensor_index& ix, const qtensor_index& mid) {
  ix.rename(mid.name());
  ix.prime((int)mid.level() - (int)ix.level());
}

template <unsigned N>
void qTensorTrain<double, N>::rc() {   // qtt.cpp:821-850
  qtensor<double> U, V, S;
  for (unsigned i = length - 1; i > 0; i--) {
    std::vector<qtensor_index> left, right;
    qtensor_index mid;
    std::string LN = link_name(_id, i);
    for (auto& ix : A[i].idx_set) {
      if (ix.name() == LN) { left.push_back(ix); mid = ix; } else right.push_back(ix);
    }
    svd(A[i], left, right, U, V, S, MoveFromRight);
    bond_dims[i] = (unsigned)S.total_size();
    A[i] = V;
    rename_like(A[i].idx_set[0], mid);
    A[i - 1] = A[i - 1] * U;
    rename_like(A[i - 1].idx_set.back(), mid);
    if (g_progress_every > 0 && (i % g_progress_every == 0 || i == 1)) {
      ttg_sync();
      fprintf(stderr, "[dmrg] right-canonical form: site %u/%u, bond dim %u\n", i, length - 1, bond_dims[i]);
      fflush(stderr);
    }
  }
  center = 0;
}

template <unsigned N>
void qTensorTrain<double, N>::lc() {   // qtt.cpp:857-886
  qtensor<double> U, V, S;
  for (unsigned i = 0; i + 1 < length; i++) {
    std::vector<qtensor_index> left, right;
    qtensor_index mid;
    std::string LN = link_name(_id, i + 1);
    for (auto& ix : A[i].idx_set) {
      if (ix.name() == LN) { right.push_back(ix); mid = ix; } else left.push_back(ix);
    }
    svd(A[i], left, right, U, V, S, MoveFromLeft);
    bond_dims[i + 1] = (unsigned)S.total_size();
    A[i] = U;
    rename_like(A[i].idx_set.back(), mid);
    A[i + 1] = V * A[i + 1];
    rename_like(A[i + 1].idx_set[0], mid);
  }
  center = (int)length - 1;
}

template <unsigned N>
void qTensorTrain<double, N>::normalize() {   // qtt.cpp:899-909
  if (center == -1) fail("normalize without a canonical centre is not supported");
  double nm = A[center].norm();
  A[center] *= 1.0 / nm;
}

template <unsigned N>
double qTensorTrain<double, N>::position(int site) {   // qtt.cpp:932-994
  if (center == -1) rc();
  qtensor<double> U, V, S;
  bool moved = false;
  while (center != site) {
    std::vector<qtensor_index> left, right;
    qtensor_index mid;
    moved = true;
    if (center > site) {
      std::string LN = link_name(_id, center);
      for (auto& ix : A[center].idx_set) {
        if (ix.name() == LN) { left.push_back(ix); mid = ix; } else right.push_back(ix);
      }
      svd(A[center], left, right, U, V, S, MoveFromRight);
      bond_dims[center] = (unsigned)S.total_size();
      A[center] = V;
      rename_like(A[center].idx_set[0], mid);
      A[center - 1] = A[center - 1] * U;
      rename_like(A[center - 1].idx_set.back(), mid);
      --center;
    } else {
      std::string LN = link_name(_id, center + 1);
      for (auto& ix : A[center].idx_set) {
        if (ix.name() == LN) { right.push_back(ix); mid = ix; } else left.push_back(ix);
      }
      svd(A[center], left, right, U, V, S, MoveFromLeft);
      bond_dims[center + 1] = (unsigned)S.total_size();
      A[center] = U;
      rename_like(A[center].idx_set.back(), mid);
      A[center + 1] = V * A[center + 1];
      rename_like(A[center + 1].idx_set[0], mid);
      ++center;
    }
  }
  if (!moved) return 0.;
  return calcEntropy(S);
}

template class qTensorTrain<double, 1>;
template class qTensorTrain<double, 2>;

// ---------------------------------------------------------------- DMRG
// One environment step ((psi^dag' * env) * W) * psi (dmrg.cpp:494-508, 76-83).
// The index pairs of the three operator* calls are found by tag, as the
// reference does; when they are the reference layouts' pairs (every DMRG
// environment), the step runs as one ttg_env_update call, else as three
// contractions.
static qtensor<double> env_step(const qtensor<double>& psi, const qtensor<double>& env,
                                const qtensor<double>& W, int dir) {
  qtensor<double> t1 = psi;
  t1.conj(); t1.dag(); t1.prime();
  IndexPairing p1 = pair_indices(t1.idx_set, env.idx_set);
  IndexPairing p2 = pair_indices(p1.result, W.idx_set);
  IndexPairing p3 = pair_indices(p2.result, psi.idx_set);
  int want[16];
  check(ttg_env_pairs(dir, want), "env pairs");
  bool std_layout = psi._initted && env._initted && W._initted;
  const IndexPairing* ps[3] = {&p1, &p2, &p3};
  int k = 0;
  for (int s = 0; s < 3 && std_layout; s++) {
    const int n = want[k++];
    std_layout = (int)ps[s]->pa.size() == n;
    for (int i = 0; i < n && std_layout; i++, k += 2)
      std_layout = ps[s]->pa[i] == want[k] && ps[s]->pb[i] == want[k + 1];
  }
  if (!std_layout) return ((t1 * env) * W) * psi;
  qtensor<double> res(p3.result);
  ttg_tensor out = nullptr;
  check(ttg_env_update(psi.dev.get(), env.dev.get(), W.dev.get(), dir, env_precision(), &out),
        "env_update");
  res.dev = adopt(out);
  res._initted = true;
  return res;
}

// dmrg.cpp:56-84
void buildEnv(qMPS<double>& psi, qMPO<double>& H, std::vector<qtensor<double>>& TR,
              std::vector<qtensor<double>>& TL) {
  const unsigned L = psi.length;
  {
    qtensor<double> le({psi.A[0].idx_set[0], H.A[0].idx_set[0], psi.A[0].idx_set[0]});
    le.idx_set[0].primeLink(); le.idx_set[0].dag();
    le.initBlock(); le.setOne(); le.dag();
    TL[0] = std::move(le);
  }
  {
    qtensor<double> re({psi.A[L - 1].idx_set.back(), H.A[L - 1].idx_set.back(), psi.A[L - 1].idx_set.back()});
    re.idx_set[0].primeLink(); re.idx_set[0].dag();
    re.initBlock(); re.setOne(); re.dag();
    TR[L - 1] = std::move(re);
  }
  for (unsigned i = L - 1; i > 0; i--) {
    TR[i - 1] = env_step(psi.A[i], TR[i], H.A[i], MoveFromRight);
    // with a checkpoint stride, only every k-th environment and the end stay
    if (g_env_stride > 1 && i < L - 1 && i % (unsigned)g_env_stride != 0) TR[i] = qtensor<double>();
  }
}

// The mirror of buildEnv: the open-boundary ends TL[0] and TR[L-1], then
// TL[i+1] for every site from the left (the environments a sweep that
// starts at the right end reads).
static void buildEnvLeft(qMPS<double>& psi, qMPO<double>& H, std::vector<qtensor<double>>& TR,
                         std::vector<qtensor<double>>& TL) {
  const unsigned L = psi.length;
  {
    qtensor<double> le({psi.A[0].idx_set[0], H.A[0].idx_set[0], psi.A[0].idx_set[0]});
    le.idx_set[0].primeLink(); le.idx_set[0].dag();
    le.initBlock(); le.setOne(); le.dag();
    TL[0] = std::move(le);
  }
  {
    qtensor<double> re({psi.A[L - 1].idx_set.back(), H.A[L - 1].idx_set.back(), psi.A[L - 1].idx_set.back()});
    re.idx_set[0].primeLink(); re.idx_set[0].dag();
    re.initBlock(); re.setOne(); re.dag();
    TR[L - 1] = std::move(re);
  }
  for (unsigned i = 0; i + 1 < L; i++) {
    TL[i + 1] = env_step(psi.A[i], TL[i], H.A[i], MoveFromLeft);
    if (g_env_stride > 1 && i > 0 && i % (unsigned)g_env_stride != 0) TL[i] = qtensor<double>();
  }
}

// dmrg.cpp:89-134: TL[start] holds sites 0..start-1 and TR[stop] sites
// stop+1..L-1, each contracted into one end tensor; TR[i-1] for
// stop >= i > start as in the full build
void buildEnv(qMPS<double>& psi, qMPO<double>& H, std::vector<qtensor<double>>& TR,
              std::vector<qtensor<double>>& TL, unsigned start, unsigned stop) {
  const unsigned L = psi.length;
  if (!(start < stop && stop < L) || TR.size() < L || TL.size() < L) fail("buildEnv: bad site range");
  {
    qtensor<double> le({psi.A[0].idx_set[0], H.A[0].idx_set[0], psi.A[0].idx_set[0]});
    le.idx_set[0].primeLink(); le.idx_set[0].dag();
    le.initBlock(); le.setOne(); le.dag();
    TL[start] = std::move(le);
  }
  for (unsigned i = 0; i < start; i++) TL[start] = env_step(psi.A[i], TL[start], H.A[i], MoveFromLeft);
  {
    qtensor<double> re({psi.A[L - 1].idx_set.back(), H.A[L - 1].idx_set.back(), psi.A[L - 1].idx_set.back()});
    re.idx_set[0].primeLink(); re.idx_set[0].dag();
    re.initBlock(); re.setOne(); re.dag();
    TR[stop] = std::move(re);
  }
  for (unsigned i = L - 1; i > stop; i--) TR[stop] = env_step(psi.A[i], TR[stop], H.A[i], MoveFromRight);
  for (unsigned i = stop; i > start; i--) {
    TR[i - 1] = env_step(psi.A[i], TR[i], H.A[i], MoveFromRight);
    if (g_env_stride > 1 && i < stop && i % (unsigned)g_env_stride != 0) TR[i] = qtensor<double>();
  }
}

// dmrg.cpp:491-509
void updateEnv(qMPS<double>& psi, qMPO<double>& H, std::vector<qtensor<double>>& TR,
               std::vector<qtensor<double>>& TL, const unsigned& site, int& direction) {
  if (direction == MoveFromLeft) TL[site + 1] = env_step(psi.A[site], TL[site], H.A[site], MoveFromLeft);
  else TR[site - 1] = env_step(psi.A[site], TR[site], H.A[site], MoveFromRight);
}

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// dmrg.cpp:300-372
void updateSite(qMPS<double>& psi, qMPO<double>& H, std::vector<qtensor<double>>& TR,
                std::vector<qtensor<double>>& TL, const unsigned& site, double& energy,
                int& direction, int max_bd, double cutoff, char mode, int search_space_size,
                int max_restart, double tol, SiteLog* log) {
  const unsigned a = (direction == MoveFromLeft) ? site : site - 1;
  double t0 = now_s();
  big_qtensor<double> A;
  A.setLeft(&TL[a]);
  A.addMid(&H.A[a]);
  A.addMid(&H.A[a + 1]);
  A.setRight(&TR[a + 1]);
  qtensor<double> x = psi.A[a] * psi.A[a + 1];
  DavidsonStats st;
  energy = tensor_davidsonIT(A, x, search_space_size, max_restart, tol, mode, &st);
  ttg_sync();
  double t1 = now_s();
  qtensor<double> S;
  qtensor_index mid;
  std::string LN = link_name(psi._id, a + 1);
  for (auto& ix : psi.A[a].idx_set) if (ix.name() == LN) { mid = ix; break; }
  unsigned kept = 0;
  double tr = svd_bond(x, psi.A[a], psi.A[a + 1], mid, S, direction, cutoff, (long unsigned)max_bd, &kept);
  psi.bond_dims[a + 1] = (unsigned)S.total_size();
  psi.center = (direction == MoveFromLeft) ? (int)site + 1 : (int)site - 1;
  ttg_sync();
  double t2 = now_s();
  if (log) { log->energy = energy; log->trunc_err = tr; log->kept = kept; log->n_matvec = st.n_matvec;
             log->t_dav = t1 - t0; log->t_svd = t2 - t1; }
}

// dmrg.cpp:760-813
// ---------------------------------------------------------------- observables
// mps/observables.cpp:64-105.  Left to right: acc <- ((acc * psi_i^+') * W_i) * phi_i,
// closed by all-ones end tensors on the open boundary links.
double psiHphi(qMPS<double>& psi, qMPO<double>& H, qMPS<double>& phi) {
  if (psi.length != phi.length || psi.length != H.length || psi.length == 0)
    fail("psiHphi: chain lengths differ");
  qtensor<double> acc;
  double res = 0;
  for (unsigned i = 0; i < psi.length; i++) {
    qtensor<double> t1 = psi.A[i];
    t1.dag();
    t1.prime();
    qtensor<double>& t2 = H.A[i];
    qtensor<double>& t3 = phi.A[i];
    if (i == 0) {
      qtensor<double> le({t1.idx_set[0], t2.idx_set[0], t3.idx_set[0]});
      le.initBlock();
      le.setOne();
      le.dag();
      acc = ((le * t1) * t2) * t3;
    } else {
      acc = ((acc * t1) * t2) * t3;
    }
    if (i + 1 == psi.length) {
      qtensor<double> re({t1.idx_set.back(), t2.idx_set.back(), t3.idx_set.back()});
      re.initBlock();
      re.setOne();
      re.dag();
      res = acc.contract(re);
    }
  }
  return res;
}

// mps/observables.cpp:415-450 (links of psi primed, sites shared)
double psiphi(qMPS<double>& psi, qMPS<double>& phi) {
  if (psi.length != phi.length || psi.length == 0) fail("psiphi: chain lengths differ");
  qtensor<double> acc;
  double res = 0;
  for (unsigned i = 0; i < psi.length; i++) {
    qtensor<double> t1 = psi.A[i];
    t1.dag();
    t1.primeLink();
    qtensor<double>& t2 = phi.A[i];
    if (i == 0) {
      qtensor<double> le({t1.idx_set[0], t2.idx_set[0]});
      le.initBlock();
      le.setOne();
      le.dag();
      acc = (le * t1) * t2;
    } else {
      acc = (t1 * acc) * t2;
    }
    if (i + 1 == psi.length) {
      qtensor<double> re({t1.idx_set.back(), t2.idx_set.back()});
      re.initBlock();
      re.setOne();
      re.dag();
      res = acc.contract(re);
    }
  }
  return res;
}

// Initial energy <psi|H|psi> at the centre site c (dmrg.cpp:717-724).
static double initial_energy(qMPS<double>& psi, qMPO<double>& H, std::vector<qtensor<double>>& TR,
                             std::vector<qtensor<double>>& TL, unsigned c, DmrgLog* log) {
  big_qtensor<double> A;
  A.setLeft(&TL[c]);
  A.addMid(&H.A[c]);
  A.setRight(&TR[c]);
  const double E = A.expec(psi.A[c]);
  if (log) log->initial_energy = E;
  if (g_progress_every > 0) {
    fprintf(stderr, "[dmrg] canonical form and environments ready, E0=%.12f\n", E);
    fflush(stderr);
  }
  return E;
}

static size_t env_elems(const qtensor<double>& t) { return t.dev ? t.total_size() : 0; }

// Make sure the environments bond (a, a+1) reads exist.  With a checkpoint
// stride above 1 the environments between kept sites were released after
// their last use of the previous half-sweep; they are recomputed here from
// the nearest kept one on the far side of the bond, from site tensors that
// have not changed since (bit-identical to the released ones).  The
// intermediate ones stay: the next bonds of the half-sweep read them.
static void ensure_env(qMPS<double>& psi, qMPO<double>& H, std::vector<qtensor<double>>& TR,
                       std::vector<qtensor<double>>& TL, unsigned a, unsigned lo, unsigned hi,
                       int direction) {
  if (direction == MoveFromLeft) {
    if (TR[a + 1].dev) return;
    unsigned j = a + 2;
    while (j <= hi && !TR[j].dev) j++;
    if (j > hi) fail("dmrg: no kept right environment to rebuild from");
    for (unsigned t = j; t > a + 1; t--) TR[t - 1] = env_step(psi.A[t], TR[t], H.A[t], MoveFromRight);
  } else {
    if (TL[a].dev) return;
    if (a == 0) fail("dmrg: left end environment missing");
    int j = (int)a - 1;
    while (j >= (int)lo && !TL[j].dev) j--;
    if (j < (int)lo) fail("dmrg: no kept left environment to rebuild from");
    for (unsigned t = (unsigned)j; t < a; t++) TL[t + 1] = env_step(psi.A[t], TL[t], H.A[t], MoveFromLeft);
  }
}

// One half-sweep over the bonds of [lo, hi] (dmrg.cpp:731-740, 785-792):
// MoveFromLeft visits sites lo .. hi-1, MoveFromRight sites hi .. lo+1.
static void half_sweep(qMPS<double>& psi, qMPO<double>& H, std::vector<qtensor<double>>& TR,
                       std::vector<qtensor<double>>& TL, int direction, unsigned lo, unsigned hi,
                       int max_bd, double cutoff, int max_restart, double tol, double& Energy,
                       int sweep_label, DmrgLog* log) {
  const int search_space_size = 3;
  const double t0 = now_s();
  const unsigned nb = hi - lo;
  const unsigned k = (unsigned)g_env_stride;
  for (unsigned i = 0; i < nb; i++) {
    const unsigned site = (direction == MoveFromLeft) ? lo + i : hi - i;
    const unsigned a = (direction == MoveFromLeft) ? site : site - 1;
    const double tr0 = now_s();
    ensure_env(psi, H, TR, TL, a, lo, hi, direction);
    const double t_rebuild = now_s() - tr0;
    SiteLog sl{};
    updateSite(psi, H, TR, TL, site, Energy, direction, max_bd, cutoff, 'S', search_space_size,
               max_restart, tol, &sl);
    const double te = now_s();
    updateEnv(psi, H, TR, TL, site, direction);
    // The environment this bond read on the far side is stale: the next
    // half-sweep recomputes it before any bond reads it.  Releasing it
    // keeps one environment set on the device instead of two (results are
    // unchanged; the range's ends TL[lo] and TR[hi] are kept).  With a
    // checkpoint stride k > 1, the environment this bond read on the near
    // side (TL[site] / TR[site]) is kept for the next half-sweep only at
    // sites that are multiples of k.
    if (direction == MoveFromLeft) {
      if (site + 1 < hi) TR[site + 1] = qtensor<double>();
      if (site == lo) TR[lo] = qtensor<double>();
      if (k > 1 && site > lo && site % k != 0) TL[site] = qtensor<double>();
    } else {
      if (site - 1 > lo) TL[site - 1] = qtensor<double>();
      if (site == hi) TL[hi] = qtensor<double>();
      if (k > 1 && site < hi && site % k != 0) TR[site] = qtensor<double>();
    }
    ttg_sync();
    if (g_progress_every > 0 && (i % g_progress_every == 0 || i + 1 == nb)) {
      size_t in_use = 0, peak = 0, cached = 0;
      ttg_mem_info(&in_use, &peak, &cached);
      fprintf(stderr, "[dmrg] sweep %d half %d bond %u/%u E=%.12f kept=%u %.2fs (device pool %.1f GB live, "
              "%.1f GB cached)\n", sweep_label, direction == MoveFromLeft ? 0 : 1, i + 1, nb, sl.energy,
              sl.kept, now_s() - t0, in_use / 1e9, cached / 1e9);
      fflush(stderr);
    }
    if (log) {
      log->bond_energy.push_back(sl.energy);
      log->trunc_err.push_back(sl.trunc_err);
      log->kept.push_back(sl.kept);
      log->t_dav += sl.t_dav;
      log->t_svd += sl.t_svd;
      log->t_env += now_s() - te + t_rebuild;
      log->n_matvec += sl.n_matvec;
      log->env_elems.push_back(env_elems(direction == MoveFromLeft ? TL[site + 1] : TR[site - 1]));
      size_t in_use = 0, peak = 0, cached = 0;
      ttg_mem_info(&in_use, &peak, &cached);
      log->pool_live_max = std::max(log->pool_live_max, in_use);
    }
  }
  ttg_sync();
  if (log) {
    log->halfsweep_energy.push_back(Energy);
    log->halfsweep_seconds.push_back(now_s() - t0);
  }
}

// dmrg.cpp:702-755: schedules cycled per sweep (max_bd[sweep % size], ...)
double dmrg(qMPS<double>& psi, qMPO<double>& H, int num_sweeps, const std::vector<int>& max_bd,
            const std::vector<double>& cutoff, const std::vector<int>& max_restart, double tol,
            DmrgLog* log) {
  return dmrg(psi, H, 0, H.length - 1, num_sweeps, max_bd, cutoff, max_restart, tol, log);
}

// dmrg.cpp:760-813: sweeps over the bonds of sites [start, stop] with the
// environments outside the range contracted once (buildEnv(start, stop))
double dmrg(qMPS<double>& psi, qMPO<double>& H, unsigned start, unsigned stop, int num_sweeps,
            const std::vector<int>& max_bd, const std::vector<double>& cutoff,
            const std::vector<int>& max_restart, double tol, DmrgLog* log) {
  const unsigned L = H.length;
  if (psi.length != L || !(start < stop && stop < L)) fail("dmrg: bad site range");
  if (max_bd.empty() || cutoff.empty() || max_restart.empty()) fail("dmrg: empty schedule");
  psi.position((int)start);
  psi.normalize();
  std::vector<qtensor<double>> TR(L), TL(L);
  if (start == 0 && stop == L - 1) buildEnv(psi, H, TR, TL);
  else buildEnv(psi, H, TR, TL, start, stop);
  double Energy = initial_energy(psi, H, TR, TL, start, log);
  for (int sweep = 0; sweep < num_sweeps; sweep++) {
    const int lbd = sweep % (int)max_bd.size();
    const int lc = sweep % (int)cutoff.size();
    const int lmr = sweep % (int)max_restart.size();
    for (int l = 2 * sweep; l < 2 * (sweep + 1); l++)
      half_sweep(psi, H, TR, TL, (l % 2 == 0) ? MoveFromLeft : MoveFromRight, start, stop, max_bd[lbd],
                 cutoff[lc], max_restart[lmr], tol, Energy, sweep, log);
    psi.normalize();
  }
  psi.position(0);
  return Energy;
}

// dmrg.cpp:653-697: half-sweeps start_sweep .. num_sweeps-1 (even ones move
// from the left), one bond-dimension / cutoff / restart setting; a run
// saved after half-sweep k resumes with start_sweep = k + 1
double dmrg(qMPS<double>& psi, qMPO<double>& H, int num_sweeps, int max_bd, double cutoff, char mode,
            int search_space_size, int max_restart, int start_sweep, double tol, DmrgLog* log) {
  (void)mode; (void)search_space_size;
  const unsigned L = H.length;
  if (psi.length != L || L < 2) fail("dmrg: chain lengths differ");
  if (start_sweep < 0) fail("dmrg: negative start_sweep");
  const bool from_right = start_sweep % 2 == 1;
  psi.position(from_right ? (int)L - 1 : 0);
  psi.normalize();
  std::vector<qtensor<double>> TR(L), TL(L);
  // an even start reads TL[0] and the right environments, which buildEnv
  // makes; an odd start moves from the right first and reads the left
  // environments, built here from site 0 (the reference calls buildEnv for
  // both and leaves TL[1..L-1] empty on an odd start)
  if (!from_right) buildEnv(psi, H, TR, TL);
  else buildEnvLeft(psi, H, TR, TL);
  double Energy = initial_energy(psi, H, TR, TL, from_right ? L - 1 : 0, log);
  for (int l = start_sweep; l < num_sweeps; l++)
    half_sweep(psi, H, TR, TL, (l % 2 == 0) ? MoveFromLeft : MoveFromRight, 0, L - 1, max_bd, cutoff,
               max_restart, tol, Energy, l / 2, log);
  psi.position(0);
  psi.normalize();
  return Energy;
}

}  // namespace ttgpu
