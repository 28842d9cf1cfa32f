// qtensor.cpp -- host side of qtensor<double> / big_qtensor<double> over the C-ABI.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <unordered_map>

#include "tt_gpu.hpp"

namespace ttgpu {

void fail(const char* what) {
  throw Error(std::string(what) + " failed: " + ttg_last_error());
}

std::string to_string(const QN_t& q) {
  return std::to_string(q.q0) + "|" + std::to_string(q.q1) + "|";
}

// ---------------------------------------------------------------- index
qtensor_index::qtensor_index(arrow_type a, std::string name, index_type t, unsigned level)
    : _arrow(a), _name(std::move(name)), _type(t), _level(level) {
  setTag();
}

// qtensor_index.h:35-42: append, re-sort by QN, rebuild the tag
void qtensor_index::addQN(QN_t q, unsigned dim) {
  _qn.push_back({q, dim});
  setTag();
}

void qtensor_index::setTag() {
  if (_qn.size() > 1)
    std::sort(_qn.begin(), _qn.end(),
              [](const std::pair<QN_t, unsigned>& a, const std::pair<QN_t, unsigned>& b) {
                return a.first < b.first;
              });
  _tag = _name + " " + std::to_string((int)_type) + " ";
  for (auto& p : _qn) _tag += "qn" + to_string(p.first) + "qdim" + std::to_string(p.second) + " ";
}

// `_arrow+" "` is char-pointer arithmetic in the reference: " " for
// Inward, "" for Outward (qtensor_index.h:66).
std::string qtensor_index::tag() const {
  return std::string(_arrow == Inward ? " " : "") + std::to_string(_level) + " " + _tag;
}
std::string qtensor_index::tagNoArrow() const { return std::to_string(_level) + " " + _tag; }

// qtensor.cpp:536-607
IndexPairing pair_indices(const std::vector<qtensor_index>& A, const std::vector<qtensor_index>& B) {
  std::unordered_map<std::string, int> marker;
  IndexPairing r;
  for (size_t i = 0; i < A.size(); i++) marker[A[i].tag()] = (int)i;
  for (size_t i = 0; i < B.size(); i++) {
    qtensor_index tp = B[i];
    tp.dag();
    auto it = marker.find(tp.tag());
    if (it == marker.end()) {
      marker[B[i].tag()] = (int)i;
    } else {
      r.pa.push_back(it->second);
      r.pb.push_back((int)i);
      marker[tp.tag()] = -1;
      marker[B[i].tag()] = -1;
    }
  }
  for (size_t i = 0; i < A.size(); i++)
    if (marker[A[i].tag()] != -1) r.result.push_back(A[i]);
  for (size_t i = 0; i < B.size(); i++)
    if (marker[B[i].tag()] != -1) r.result.push_back(B[i]);
  return r;
}

// ---------------------------------------------------------------- tensor
DevHandle adopt(ttg_tensor t) { return DevHandle(t, [](ttg_tensor p) { ttg_tensor_free(p); }); }

namespace {
struct IdxArrays {
  std::vector<int> arrows, nsec, qn, qdim;
};
IdxArrays idx_arrays(const std::vector<qtensor_index>& idx) {
  IdxArrays a;
  for (auto& ix : idx) {
    a.arrows.push_back((int)ix.arrow());
    a.nsec.push_back((int)ix.size());
    for (unsigned s = 0; s < ix.size(); s++) {
      a.qn.push_back(ix.qn(s).q0);
      a.qn.push_back(ix.qn(s).q1);
      a.qdim.push_back((int)ix.qdim(s));
    }
  }
  if (a.qn.empty()) { a.qn.push_back(0); a.qdim.push_back(0); }
  return a;
}
}  // namespace

qtensor<double>::qtensor(const qtensor& o) : rank(o.rank), idx_set(o.idx_set), _initted(o._initted) {
  if (o.dev) {
    ttg_tensor t;
    check(ttg_tensor_copy(o.dev.get(), &t), "tensor copy");
    dev = adopt(t);
  }
}

qtensor<double>& qtensor<double>::operator=(const qtensor& o) {
  if (this != &o) {
    qtensor tmp(o);
    *this = std::move(tmp);
  }
  return *this;
}

void qtensor<double>::initBlock() {
  if (rank > 0 && !_initted) {
    IdxArrays a = idx_arrays(idx_set);
    ttg_tensor t;
    check(ttg_tensor_init_blocks((int)rank, a.arrows.data(), a.nsec.data(), a.qn.data(),
                                 a.qdim.data(), &t),
          "initBlock");
    dev = adopt(t);
  }
  _initted = true;
}

void qtensor<double>::setZero() { if (dev) check(ttg_tensor_fill(dev.get(), 0.0), "setZero"); }
void qtensor<double>::setOne() { if (dev) check(ttg_tensor_fill(dev.get(), 1.0), "setOne"); }
void qtensor<double>::setRandom(uint64_t seed) {
  if (dev) check(ttg_tensor_set_random(dev.get(), seed), "setRandom");
}

size_t qtensor<double>::total_size() const {
  if (!dev) return 0;
  size_t tot = 0;
  check(ttg_tensor_info(dev.get(), nullptr, nullptr, &tot), "info");
  return tot;
}

int qtensor<double>::nblocks() const {
  if (!dev) return 0;
  int nb = 0;
  check(ttg_tensor_info(dev.get(), nullptr, &nb, nullptr), "info");
  return nb;
}

void qtensor<double>::from_host(const std::vector<std::vector<unsigned>>& block_qi,
                                const std::vector<double>& data) {
  IdxArrays a = idx_arrays(idx_set);
  std::vector<int> qi;
  for (auto& b : block_qi)
    for (unsigned v : b) qi.push_back((int)v);
  if (qi.empty()) qi.push_back(0);
  ttg_tensor t;
  check(ttg_tensor_create((int)rank, a.arrows.data(), a.nsec.data(), a.qn.data(), a.qdim.data(),
                          (int)block_qi.size(), qi.data(), &t),
        "tensor create");
  dev = adopt(t);
  check(ttg_tensor_upload(t, data.data(), data.size()), "upload");
  _initted = true;
}

std::vector<double> qtensor<double>::to_host() const {
  std::vector<double> out(total_size());
  if (!out.empty()) check(ttg_tensor_download(dev.get(), out.data(), out.size()), "download");
  return out;
}

std::vector<std::vector<unsigned>> qtensor<double>::block_list() const {
  int nb = nblocks();
  std::vector<int> qi((size_t)nb * rank + 1);
  std::vector<int64_t> off(nb + 1);
  check(ttg_tensor_meta_download(dev.get(), qi.data(), off.data()), "meta");
  std::vector<std::vector<unsigned>> out(nb, std::vector<unsigned>(rank));
  for (int b = 0; b < nb; b++)
    for (unsigned j = 0; j < rank; j++) out[b][j] = (unsigned)qi[(size_t)b * rank + j];
  return out;
}

// qtensor.cpp:523-734: index pairing on the host, blocks on the device
qtensor<double> qtensor<double>::operator*(const qtensor& B) const {
  if (!(_initted || B._initted)) fail("operator* on two uninitialised tensors");
  if (_initted && !B._initted) return *this;
  if (B._initted && !_initted) return B;
  IndexPairing p = pair_indices(idx_set, B.idx_set);
  qtensor res(p.result);
  ttg_tensor c;
  check(ttg_contract(dev.get(), B.dev.get(), (int)p.pa.size(), p.pa.data(), p.pb.data(), TTG_FP64,
                     &c),
        "contract");
  res.dev = adopt(c);
  res._initted = true;
  return res;
}

void qtensor<double>::dag() {
  for (auto& i : idx_set) i.dag();
  if (dev) check(ttg_tensor_dag(dev.get()), "dag");
}

qtensor<double>& qtensor<double>::operator*=(double c) {
  if (dev) check(ttg_scale(dev.get(), c), "scale");
  return *this;
}

// qtensor.cpp:1095-1119: blocks matched by key, missing ones add nothing
double qtensor<double>::contract(const qtensor& A) const {
  double r = 0;
  int same = 0;
  check(ttg_same_blocks(dev.get(), A.dev.get(), &same), "contract");
  if (same) check(ttg_dot(dev.get(), A.dev.get(), &r), "contract");
  else check(ttg_dot_matched(dev.get(), A.dev.get(), &r), "contract");
  return r;
}
double qtensor<double>::inner_product(const qtensor& A) const { return contract(A); }

void qtensor<double>::add(const qtensor& A, double c) {
  check(ttg_axpy(c, A.dev.get(), dev.get()), "add");
}

double qtensor<double>::norm() const {
  double r = 0;
  check(ttg_nrm2(dev.get(), &r), "norm");
  return r;
}

double qtensor<double>::normalize() {
  double n = norm();
  *this *= 1.0 / n;
  return n;
}

// ---------------------------------------------------------------- big_qtensor
static int g_heff_prec = TTG_FP64;
void set_heff_precision(int prec) { g_heff_prec = prec; }
int heff_precision() { return g_heff_prec; }
static int g_env_prec = TTG_FP64;
void set_env_precision(int prec) { g_env_prec = prec; }
int env_precision() { return g_env_prec; }

big_qtensor<double>::~big_qtensor() { drop(); }

void big_qtensor<double>::drop() {
  if (heff) ttg_heff_free(heff);
  if (heff64) ttg_heff_free(heff64);
  heff = nullptr;
  heff64 = nullptr;
  full = qtensor<double>();
}

void big_qtensor<double>::addMid(qtensor<double>* m) {
  mid = mid * (*m);
  drop();
}

static bool same_tags(const std::vector<qtensor_index>& a, const std::vector<qtensor_index>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); i++)
    if (a[i].tag() != b[i].tag()) return false;
  return true;
}

bool same_blocks(const qtensor<double>& a, const qtensor<double>& b) {
  int s = 0;
  check(ttg_same_blocks(a.dev.get(), b.dev.get(), &s), "same_blocks");
  return s != 0;
}

// One operator per x index set: the three stage pairings are the
// reference's operator* index logic applied to L, x, mid, R in turn, and
// the operand layout is the full legal block list of x's indices.
void big_qtensor<double>::ensure_plan(const qtensor<double>& v) {
  if (heff && plan_prec == g_heff_prec && same_tags(full.idx_set, v.idx_set)) return;
  drop();
  if (!L || !R) fail("big_qtensor::product needs both environments (L*x*mid*R)");
  full = qtensor<double>(v.idx_set);
  full.initBlock();
  IndexPairing p1 = pair_indices(L->idx_set, full.idx_set);
  IndexPairing p2 = pair_indices(p1.result, mid.idx_set);
  IndexPairing p3 = pair_indices(p2.result, R->idx_set);
  std::vector<int> enc;
  for (auto* p : {&p1, &p2, &p3}) {
    enc.push_back((int)p->pa.size());
    for (size_t i = 0; i < p->pa.size(); i++) { enc.push_back(p->pa[i]); enc.push_back(p->pb[i]); }
  }
  check(ttg_heff_create(L->dev.get(), mid.dev.get(), R->dev.get(), full.dev.get(), enc.data(),
                        g_heff_prec, &heff),
        "heff_create");
  if (g_heff_prec != TTG_FP64)   // FP64 twin on the same plans, for the correction products
    check(ttg_heff_create_twin(heff, TTG_FP64, &heff64), "heff_create (FP64 correction)");
  plan_prec = g_heff_prec;
  // TTG_HEFF_LOG set (and not "0"): one stderr line per operator with its
  // per-stage FLOPs, task counts, work shapes and work-item counts
  static const bool log = [] {
    const char* e = std::getenv("TTG_HEFF_LOG");
    return e && *e && *e != '0';
  }();
  if (log) {
    double f = 0, sf[3] = {0, 0, 0};
    int64_t nt[3] = {0, 0, 0};
    int cfg[3] = {0, 0, 0}, ntl[3] = {0, 0, 0}, rank = 0, nb = 0;
    size_t total = 0;
    if (ttg_heff_info(heff, &f, nt, sf) == 0 && ttg_heff_tiles(heff, cfg, ntl) == 0 &&
        ttg_tensor_info(full.dev.get(), &rank, &nb, &total) == 0)
      std::fprintf(stderr, "[heff] x %zu elems %d blocks, GF %.4f %.4f %.4f, tasks %lld %lld %lld, shape %d %d %d, items %d %d %d\n",
                   total, nb, sf[0] * 1e-9, sf[1] * 1e-9, sf[2] * 1e-9,
                   (long long)nt[0], (long long)nt[1], (long long)nt[2], cfg[0], cfg[1], cfg[2], ntl[0], ntl[1],
                   ntl[2]);
  }
}

// x in the operator's layout (the full legal block list of its indices)
qtensor<double> big_qtensor<double>::widen(const qtensor<double>& v) {
  ensure_plan(v);
  if (same_blocks(v, full)) return v;
  qtensor<double> w(full);
  check(ttg_embed(v.dev.get(), w.dev.get()), "embed");
  return w;
}

void big_qtensor<double>::product_into(qtensor<double>& v, qtensor<double>& y) {
  ensure_plan(v);
  if (!y.dev || !same_blocks(y, full)) {
    ttg_tensor t;
    check(ttg_heff_new_output(heff, &t), "heff output");
    y.dev = adopt(t);
    y.idx_set = v.idx_set;
    y.rank = v.rank;
    y._initted = true;
  }
  if (same_blocks(v, full)) {
    check(ttg_heff_apply(heff, v.dev.get(), y.dev.get()), "heff apply");
  } else {
    qtensor<double> w = widen(v);
    check(ttg_heff_apply(heff, w.dev.get(), y.dev.get()), "heff apply");
  }
}

void big_qtensor<double>::product_into_fp64(qtensor<double>& v, qtensor<double>& y) {
  ensure_plan(v);
  if (!heff64) {
    product_into(v, y);
    return;
  }
  if (!y.dev || !same_blocks(y, full)) {
    ttg_tensor t;
    check(ttg_heff_new_output(heff64, &t), "heff output");
    y.dev = adopt(t);
    y.idx_set = v.idx_set;
    y.rank = v.rank;
    y._initted = true;
  }
  qtensor<double> w = same_blocks(v, full) ? v : widen(v);
  check(ttg_heff_apply(heff64, w.dev.get(), y.dev.get()), "heff apply (FP64)");
}

// big_qtensor.cpp:23-40 (prime(-1) of the result is implicit: y takes x's indices)
qtensor<double> big_qtensor<double>::product(qtensor<double>& v) {
  qtensor<double> y;
  product_into(v, y);
  return y;
}

double big_qtensor<double>::expec(qtensor<double>& v) {
  qtensor<double> y = product(v);
  qtensor<double> w = widen(v);
  return y.contract(w);
}

double big_qtensor<double>::flops_per_product() const {
  double f = 0;
  if (heff) ttg_heff_info(heff, &f, nullptr, nullptr);
  return f;
}

// big_qtensor.cpp:96-118 (the mid factor is re-applied on every call, as
// in the reference)
size_t big_qtensor<double>::size() const {
  if (size_ == 0) {
    size_ = 1;
    if (L) for (auto& ix : L->idx_set) if (ix.level() > 0) size_ *= ix.size();
    if (R) for (auto& ix : R->idx_set) if (ix.level() > 0) size_ *= ix.size();
  }
  for (auto& ix : mid.idx_set) if (ix.type() == Site) size_ *= ix.size();
  return size_;
}

}  // namespace ttgpu
