// davidson.cpp -- tensor_davidsonIT (linalg/tensor_davidson.cpp:192-347) with
// device-resident Krylov vectors.  The control flow, convergence tests,
// sign fix and the single Gram-Schmidt pass with re-randomisation are the
// reference's; every vector operation and the FP64 Rayleigh-Ritz
// eigensolver (DIAG, lapack_wrapper.cpp:40-54) run on the GPU.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tt_gpu.hpp"

namespace ttgpu {

static inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double tensor_davidsonIT(big_qtensor<double>& A, qtensor<double>& x_in, int m, int max_restart,
                         double tol, char mode, DavidsonStats* st) {
  (void)m; (void)mode;
  if (!(max_restart >= 1 && tol > 0)) fail("tensor_davidsonIT arguments");
  uint64_t rng = 0x5eed;
  int nmv = 0;
  // all Krylov vectors share the operator's block list
  qtensor<double> x = A.widen(x_in);
  double x_norm = x.norm();
  x *= 1.0 / x_norm;
  const int maxsize = (int)A.size();
  const int actual_maxiter = std::min(max_restart, maxsize - 1);
  const int S = actual_maxiter + 2;
  std::vector<qtensor<double>> v(S), va(S);
  std::vector<double> M(S * S, 0.0), Mnext(S * S, 0.0), evals(S), evecs(S * S);
  double last_lambda = 1000.;
  const double Approx0 = 1e-12;
  v[0] = x;
  A.product_into(v[0], va[0]);
  nmv++;
  const double initEn = va[0].inner_product(v[0]);
  double lambda = 0., qnorm = 0.;
  int ii = 0;
  for (ii = 0; ii < actual_maxiter + 1; ii++) {
    const int ni = ii + 1;
    qtensor<double>& q = v[ni];
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
      // DIAG(ni, M, evals): M overwritten by the column eigenvectors
      check(ttg_small_eigh(ni, M.data(), evals.data(), evecs.data()), "Rayleigh-Ritz");
      for (int i = 0; i < ni * ni; i++) M[i] = evecs[i];
      lambda = evals[0];
      // x = sum_k M[k] v_k ; q = sum_k M[k] va_k - lambda x ; sign fix
      const double sgn = (M[0] < 0.) ? -1.0 : 1.0;
      std::vector<ttg_tensor> xs, qs;
      std::vector<double> cx, cq;
      for (int k = 0; k <= ii; k++) {
        xs.push_back(v[k].dev.get());
        cx.push_back(sgn * M[k]);
      }
      check(ttg_lincomb((int)xs.size(), xs.data(), cx.data(), x.dev.get()), "lincomb x");
      if (!q.dev) q = x;
      for (int k = 0; k <= ii; k++) {
        qs.push_back(va[k].dev.get());
        cq.push_back(sgn * M[k]);
      }
      qs.push_back(x.dev.get());
      cq.push_back(-lambda);
      check(ttg_lincomb((int)qs.size(), qs.data(), cq.data(), q.dev.get()), "lincomb q");
    }
    qnorm = q.norm();
    const bool converged = (qnorm < tol && std::fabs(lambda - last_lambda) < tol) ||
                           qnorm < std::max(Approx0, tol * 1e-3);
    last_lambda = lambda;
    if ((qnorm < 1e-20) || (converged && ii >= 1) || (ii == actual_maxiter)) break;
    // one Gram-Schmidt pass; re-randomise on orthogonalisation failure
    const int Npass = 1;
    std::vector<double> Vq(ni);
    int pass = 1, tot_pass = 0;
    bool give_up = false;
    while (pass <= Npass) {
      ++tot_pass;
      for (int k = 0; k < ni; k++) Vq[k] = q.contract(v[k]);
      {
        std::vector<ttg_tensor> xs{q.dev.get()};
        std::vector<double> c{1.0};
        for (int k = 0; k < ni; k++) { xs.push_back(v[k].dev.get()); c.push_back(-Vq[k]); }
        check(ttg_lincomb((int)xs.size(), xs.data(), c.data(), q.dev.get()), "gram-schmidt");
      }
      double qnrm = q.norm();
      if (qnrm < 1e-10) {
        q = v[ni - 1];
        q.setRandom(splitmix64(rng));
        qnrm = q.norm();
        --pass;
        if (ni >= maxsize || tot_pass > Npass * 3) { give_up = true; break; }
      }
      q *= 1.0 / qnrm;
      ++pass;
    }
    if (give_up) break;
    A.product_into(v[ni], va[ni]);
    nmv++;
    std::swap(Mnext, M);
    const int l = ni;
    for (int k = 0; k < ni + 1; k++) {
      M[l + k * (ni + 1)] = va[l].inner_product(v[k]);
      M[k + l * (ni + 1)] = M[l + k * (ni + 1)];
    }
  }
  x.normalize();
  // Reduced-precision forms: an FP64 residual-correction step.  The FP64
  // product y = H x of the reduced-precision Ritz vector gives the FP64
  // residual r = y - <x,y> x; r (orthonormalised against x) extends the
  // basis to {x, r}, whose 2x2 Rayleigh-Ritz problem is solved in FP64 with
  // a second FP64 product H r.  The corrected vector c0 x + c1 r goes on to
  // the SVD, and its FP64 Rayleigh quotient is the energy: the step removes
  // the component of the reduced-precision error along the residual, and the
  // quotient's error is second order in what remains.
  const double lambda_reduced = lambda;
  bool corrected = false;
  if (heff_precision() != TTG_FP64) {
    qtensor<double> y;
    A.product_into_fp64(x, y);
    nmv++;
    const double a = y.inner_product(x);
    qtensor<double> r = y;
    r.add(x, -a);
    r.add(x, -r.inner_product(x));   // re-orthogonalise against x (rounding)
    const double rn = r.norm();
    lambda = a;
    qnorm = rn;
    if (rn > 1e-14 * std::max(1.0, std::fabs(a))) {
      r *= 1.0 / rn;
      qtensor<double> z;
      A.product_into_fp64(r, z);
      nmv++;
      const double b = z.inner_product(x), d = z.inner_product(r);
      // lowest eigenpair of [[a, b], [b, d]]: lambda = (a+d)/2 - sqrt(((a-d)/2)^2 + b^2),
      // eigenvector (d - lambda, -b) with a non-negative x component
      const double h = 0.5 * (a - d);
      const double lam = 0.5 * (a + d) - std::sqrt(h * h + b * b);
      double c0 = d - lam, c1 = -b;
      const double cn = std::sqrt(c0 * c0 + c1 * c1);
      if (cn > 0) {
        c0 /= cn;
        c1 /= cn;
        x *= c0;
        x.add(r, c1);
        y *= c0;
        y.add(z, c1);
        x.normalize();
        lambda = lam;
        y.add(x, -lambda);
        qnorm = y.norm();
      }
    }
    corrected = true;
  }
  x_in = std::move(x);
  if (st) {
    st->n_matvec = nmv;
    st->iterations = ii;
    st->residual = qnorm;
    st->fp64_correction = corrected;
    st->lambda_reduced = lambda_reduced;
  }
  return lambda;
}

}  // namespace ttgpu
