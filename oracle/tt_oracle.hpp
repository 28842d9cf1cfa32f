// tt_oracle.hpp -- FP64 CPU restatement of the tensor-tools DMRG hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle the GPU path is checked
// against; it is never linked into, loaded by, or called from the product
// library.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it.
//
// Each routine restates the reference algorithm from the cited file:line of
// ClarkResearchGroup/tensor-tools (parallel branch).  Blocks are dense,
// column-major (first index fastest), exactly as the reference's CTF blocks.
//
// Pinning (what checks each function; tests/test_oracle.py and
// tests/test_oracle_pin.py):
//   QN / index tags / initBlock .......... brute-force enumeration of legal tuples
//   operator* (contract_tags) ............ dense numpy einsum of densified tensors;
//                                         block and task order against an
//                                         independent statement (tests/ref_order.py)
//   big_qtensor::product ................. explicit dense H_eff (numpy)
//   tensor_davidsonIT ................... lowest eigenvalue of the dense H_eff
//                                         (numpy.linalg.eigh), variational bound
//   jacobi_eigh / DIAG .................. numpy.linalg.eigh
//   svd / svd_bond ...................... numpy.linalg.svd singular values, U.S.V
//                                         reconstruction, orthonormality; keep rule
//                                         against invariants of the sector values
//   position / rc / lc .................. state unchanged (dense), orthonormality
//   dmrg ................................ exact diagonalisation (scipy eigsh)
//   MPO builders ........................ dense MPO-contracted H vs. Kronecker H
#pragma once
#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace orc {

// util/QN.h:79-134 -- two-component abelian quantum number, lexicographic '<'.
struct QN {
  int q0 = 0, q1 = 0;
  QN() = default;
  QN(int a, int b) : q0(a), q1(b) {}
  QN operator+(const QN& o) const { return {q0 + o.q0, q1 + o.q1}; }
  QN operator-(const QN& o) const { return {q0 - o.q0, q1 - o.q1}; }
  QN operator-() const { return {-q0, -q1}; }
  QN& operator+=(const QN& o) { q0 += o.q0; q1 += o.q1; return *this; }
  QN& operator-=(const QN& o) { q0 -= o.q0; q1 -= o.q1; return *this; }
  bool operator==(const QN& o) const { return q0 == o.q0 && q1 == o.q1; }
  bool operator!=(const QN& o) const { return !(*this == o); }
  bool operator<(const QN& o) const {
    return std::make_pair(q0, q1) < std::make_pair(o.q0, o.q1);
  }
};
// QN.h:120-124: "q0|q1|"
std::string qstr(const QN& q);

enum Arrow { Inward = 0, Outward = 1 };
enum IType { Link = 0, Site = 1, AllTypes = 2 };

// qtensor/qtensor_index.h:23-115
struct Index {
  Arrow arrow = Inward;
  std::string name;
  IType type = Link;
  unsigned level = 0;
  std::vector<std::pair<QN, unsigned>> qn;
  std::string tagbase;
  Index() { setTag(); }
  Index(Arrow a, std::string n, IType t = Link, unsigned l = 0)
      : arrow(a), name(std::move(n)), type(t), level(l) { setTag(); }
  void addQN(QN q, unsigned d);
  void setTag();
  std::string tag() const;         // qtensor_index.h:66 (arrow char-pointer quirk)
  std::string tagNoArrow() const;  // qtensor_index.h:67
  unsigned size() const { return (unsigned)qn.size(); }
  QN qnAt(unsigned i) const { return qn.at(i).first; }
  unsigned qdim(unsigned i) const { return qn.at(i).second; }
  void dag() { arrow = (arrow == Inward) ? Outward : Inward; }
  void prime(int inc = 1) { level = (unsigned)((int)level + inc); }
  void rename(const std::string& n) { name = n; setTag(); }
  void retype(IType t) { type = t; setTag(); }
  bool operator==(const Index& o) const {  // qtensor_index.cpp:98-100
    return tagbase == o.tagbase && arrow == o.arrow && level == o.level;
  }
};

// qtensor/qtensor.h:47-164 (double only)
struct QTensor {
  unsigned rank = 0;
  std::vector<Index> idx_set;
  std::vector<std::vector<double>> block;
  std::vector<std::vector<QN>> block_index_qn;
  std::vector<std::vector<int>> block_index_qd;
  std::vector<std::vector<unsigned>> block_index_qi;
  std::unordered_map<std::string, unsigned> block_id_by_qn_str;
  bool initted = false;

  QTensor() = default;
  explicit QTensor(const std::vector<Index>& idx) : rank((unsigned)idx.size()), idx_set(idx) {}
  void initBlock();                    // qtensor.cpp:302-345
  void clearBlock();
  void reset(const std::vector<Index>& idx);
  void setZero();
  void setOne();
  void setRandom(uint64_t seed);       // qtensor.cpp:382-388 (CTF fill_random(0,1))
  QTensor operator*(const QTensor& B) const;   // qtensor.cpp:523-734
  QTensor& operator*=(double c);       // qtensor.cpp:1031-1040
  double contract(const QTensor& A) const;     // qtensor.cpp:1095-1119
  void add(const QTensor& A, double c = 1.0);  // qtensor.cpp:1127-1156
  double inner_product(const QTensor& A) const;  // qtensor.cpp:1161-1179
  double norm() const;                 // qtensor.cpp:1606-1616
  double normalize();                  // qtensor.cpp:1638-1647
  void prime(int inc = 1) { for (auto& i : idx_set) i.prime(inc); }
  void dag() { for (auto& i : idx_set) i.dag(); }
  size_t total_size() const;
};

extern std::vector<int> g_last_tasks;  // (c,a,b,M,N,K) of the last operator*

// index pairing used by operator*: returns pairs (posA, posB) in B order,
// plus A_perm / B_perm (qtensor.cpp:540-572)
void pair_indices(const QTensor& A, const QTensor& B,
                  std::vector<std::pair<unsigned, unsigned>>& pairs,
                  std::vector<unsigned>& A_perm, std::vector<unsigned>& B_perm);

// qtensor/big_qtensor.{h,cpp}
struct BigQTensor {
  const QTensor* L = nullptr;
  const QTensor* R = nullptr;
  QTensor mid;
  mutable size_t size_ = 0;
  void setLeft(const QTensor* l) { L = l; }
  void setRight(const QTensor* r) { R = r; }
  void addMid(const QTensor* m);            // big_qtensor.h:56
  QTensor product(const QTensor& v) const;  // big_qtensor.cpp:23-40
  double expec(const QTensor& v) const;     // big_qtensor.cpp:45-64
  size_t size() const;                      // big_qtensor.cpp:96-118
};

// linalg/lapack_wrapper.cpp:40-54 -- symmetric eigensolver (ascending values,
// eigenvectors in columns, column-major, overwriting A).
void DIAG(int n, double* A, double* evals);

// linalg/tensor_davidson.cpp:192-347
double tensor_davidsonIT(const BigQTensor& A, QTensor& x, int m, int max_restart,
                         double tol, char mode, uint64_t* rng_state, int* n_matvec);

// qtensor/qtensor_op.cpp:1097-1453.  Returns the truncation error
// max(1-cumsum,0) printed at :1301 and the number of kept values.
double svd(const QTensor& A, const std::vector<Index>& left,
           const std::vector<Index>& right, QTensor& U, QTensor& V, QTensor& S,
           int direction, double cutoff, unsigned K, unsigned* kept);
// qtensor_op.cpp:1506-1544
double svd_bond(const QTensor& combined, QTensor& A_left, QTensor& A_right,
                const Index& mid, QTensor& S, int direction, double cutoff,
                unsigned long K, unsigned* kept);

// dense SVD used per sector: one-sided Jacobi (Hestenes), FP64.
// A is rows x cols column-major.  Produces k=min(rows,cols) triplets sorted
// descending: U (rows x k), S (k), Vt (k x cols), all column-major.
void dense_svd(int rows, int cols, const double* A, std::vector<double>& U,
               std::vector<double>& S, std::vector<double>& Vt);

constexpr int MoveFromLeft = 0;
constexpr int MoveFromRight = 1;

// mps/qtt.{h,cpp}
struct QTT {
  unsigned id = 0;
  unsigned length = 0;
  unsigned phy_dim = 0;
  unsigned N = 1;   // 1 = MPS, 2 = MPO
  std::vector<unsigned> bond_dims;
  int center = -1;
  QN totalQ;
  std::vector<QN> phy_qn;
  std::vector<QTensor> A;
  void allocateMPS(const unsigned* product_state);   // qtt.cpp:205-311
  void rc();                       // qtt.cpp:821-850
  void lc();                       // qtt.cpp:857-886
  double position(int site);       // qtt.cpp:932-994
  void normalize();                // qtt.cpp:899-909
};

// spin-1/2 sites, QN = (2Sz, 0) (models/sites/spinhalf.{h,cpp})
double spinhalf_op(unsigned bra, const std::string& op, unsigned ket);
QN spinhalf_div(const std::string& op);

// models/hams/Heisenberg.cpp:333-402 -- hand-built NN Heisenberg qMPO (w=5)
void heisenberg_mpo_nn(unsigned N, unsigned id, QTT& H);

// models/hams/AutoMPO + Heisenberg.cpp:416-635 for spin-1/2 two-site terms
struct Term { double coef; std::string op1; int i1; std::string op2; int i2; };
void autompo_mpo(unsigned N, unsigned id, const std::vector<Term>& terms, QTT& H);
// Heisenberg bonds: J*(Sz Sz + 1/2 S+S- + 1/2 S-S+) over a bond list.
std::vector<Term> heisenberg_terms(const std::vector<std::pair<int, int>>& bonds,
                                   const std::vector<double>& J);

// models/lattice/square.h
std::vector<std::pair<int, int>> square_lattice(int Nx, int Ny, bool yperiodic,
                                                std::vector<int>* type = nullptr,
                                                bool nnn = false);

// algos/dmrg/dmrg.cpp
void buildEnv(QTT& psi, QTT& H, std::vector<QTensor>& TR, std::vector<QTensor>& TL);
void updateEnv(QTT& psi, QTT& H, std::vector<QTensor>& TR, std::vector<QTensor>& TL,
               unsigned site, int direction);
struct SiteStats { double energy; double trunc_err; unsigned kept; int n_matvec; };
SiteStats updateSite(QTT& psi, QTT& H, std::vector<QTensor>& TR,
                     std::vector<QTensor>& TL, unsigned site, int direction,
                     int max_bd, double cutoff, int max_restart, double tol,
                     uint64_t* rng);
struct DmrgLog { std::vector<double> halfsweep_energy; std::vector<double> bond_energy;
                 std::vector<double> trunc_err; std::vector<unsigned> kept; double initial_energy = 0; };
double dmrg(QTT& psi, QTT& H, int num_sweeps, const std::vector<int>& max_bd,
            const std::vector<double>& cutoff, const std::vector<int>& max_restart,
            double tol, DmrgLog* log);

// host BLAS library the block products use ("none ..." for the built-in loop)
const char* blas_name();

}  // namespace orc
