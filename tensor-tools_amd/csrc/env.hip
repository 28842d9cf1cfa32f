// env.hip -- DMRG environment update (updateEnv, dmrg.cpp:491-509) as one
// library call: env' = ((psi^dag' * env) * W) * psi, three planned grouped
// GEMM stages on the library stream.
//
// Index layouts are the reference's (qtt.cpp:246-253, Heisenberg.cpp:560-567,
// dmrg.cpp:56-84):
//   psi = (l In, s In, r Out)            one MPS site tensor
//   W   = (w_l In, s Out lvl0, s In lvl1, w_r Out)
//   TL  = (l' In lvl1, w_l Out, l Out)   left environment of the site
//   TR  = (r' Out lvl1, w_r In, r In)    right environment of the site
// psi^dag' has every arrow flipped and links and sites primed, so its data is
// psi's (the dag changes arrows only; the prime changes tags only).  The
// pairs of each stage are those qtensor::operator* finds by tag for these
// layouts (B-order pairing, qtensor.cpp:540-555), listed in env_pairs().
// The intermediates live in pool scratch for the call; the result is a new
// FP64 tensor: TL of the next site (MoveFromLeft) or TR of the previous one
// (MoveFromRight).
#include <algorithm>
#include <memory>

#include "ttg_internal.h"
#include "ttg_plan.h"

namespace ttg {

// stage s: n pairs (pa, pb) of (stage-s left operand, stage-s right operand)
struct EnvPairs {
  int n[3];
  int pa[3][2], pb[3][2];
};

static EnvPairs env_pairs(int dir) {
  EnvPairs e{};
  if (dir == TTG_MOVE_FROM_LEFT) {
    // psi^dag'(l', s', r') * TL(l', w_l, l) -> (s', r', w_l, l)
    e.n[0] = 1; e.pa[0][0] = 0; e.pb[0][0] = 0;
    // * W(w_l, s, s', w_r) -> (r', l, s, w_r)
    e.n[1] = 2; e.pa[1][0] = 2; e.pb[1][0] = 0; e.pa[1][1] = 0; e.pb[1][1] = 2;
    // * psi(l, s, r) -> (r', w_r, r)
    e.n[2] = 2; e.pa[2][0] = 1; e.pb[2][0] = 0; e.pa[2][1] = 2; e.pb[2][1] = 1;
  } else {
    // psi^dag'(l', s', r') * TR(r', w_r, r) -> (l', s', w_r, r)
    e.n[0] = 1; e.pa[0][0] = 2; e.pb[0][0] = 0;
    // * W(w_l, s, s', w_r) -> (l', r, w_l, s)
    e.n[1] = 2; e.pa[1][0] = 1; e.pb[1][0] = 2; e.pa[1][1] = 2; e.pb[1][1] = 3;
    // * psi(l, s, r) -> (l', w_l, l)
    e.n[2] = 2; e.pa[2][0] = 3; e.pb[2][0] = 1; e.pa[2][1] = 1; e.pb[2][1] = 2;
  }
  return e;
}

// index j of a and index k of b pair up: same sectors, opposite arrows
static bool pairable(const TMeta& a, int j, const TMeta& b, int k) {
  if (a.nsec[j] != b.nsec[k] || a.arrow[j] == b.arrow[k]) return false;
  for (int s = 0; s < a.nsec[j]; s++)
    if (a.q0(j, s) != b.q0(k, s) || a.q1(j, s) != b.q1(k, s) || a.dim(j, s) != b.dim(k, s))
      return false;
  return true;
}

static int check_pairs(const TMeta& a, const TMeta& b, int n, const int* pa, const int* pb, int stage) {
  for (int p = 0; p < n; p++) {
    if (pairable(a, pa[p], b, pb[p])) continue;
    set_error("env_update: stage " + std::to_string(stage + 1) +
              ": operands do not have the reference's index layout (index " +
              std::to_string(pa[p]) + " of the left operand does not pair with index " +
              std::to_string(pb[p]) + " of the right one)");
    return 1;
  }
  return 0;
}

}  // namespace ttg

using namespace ttg;

extern "C" {

int ttg_env_pairs(int direction, int* pairs3) {
  TTG_REQUIRE(pairs3, "env_pairs: null argument");
  TTG_REQUIRE(direction == TTG_MOVE_FROM_LEFT || direction == TTG_MOVE_FROM_RIGHT,
              "env_pairs: unknown direction");
  const EnvPairs e = env_pairs(direction);
  int* p = pairs3;
  for (int s = 0; s < 3; s++) {
    *p++ = e.n[s];
    for (int i = 0; i < e.n[s]; i++) { *p++ = e.pa[s][i]; *p++ = e.pb[s][i]; }
  }
  return 0;
}

int ttg_env_update(ttg_tensor psi, ttg_tensor env, ttg_tensor W, int direction, int prec,
                   ttg_tensor* out) {
  PhaseTimer _phase("ttg_env_update");
  TTG_REQUIRE(psi && env && W && out, "env_update: null argument");
  TTG_REQUIRE(direction == TTG_MOVE_FROM_LEFT || direction == TTG_MOVE_FROM_RIGHT,
              "env_update: unknown direction");
  TTG_REQUIRE(prec == TTG_FP64 || prec == TTG_FP32 || prec == TTG_BF16X3, "unknown precision");
  TTG_REQUIRE(psi->m->rank == 3 && env->m->rank == 3 && W->m->rank == 4,
              "env_update: psi, env and W must have ranks 3, 3 and 4");
  TTG_TRY(ensure_init());
  const EnvPairs e = env_pairs(direction);
  MetaPtr psid;
  TTG_TRY(meta_dag(*psi->m, &psid));
  TTG_TRY(check_pairs(*psid, *env->m, e.n[0], e.pa[0], e.pb[0], 0));
  std::unique_ptr<ttg_plan_s> p[3];
  {
    ttg_plan q = nullptr;
    TTG_TRY(plan_build(psid, env->m, e.n[0], e.pa[0], e.pb[0], prec, -1, &q));
    p[0].reset(q);
    TTG_TRY(check_pairs(*p[0]->C, *W->m, e.n[1], e.pa[1], e.pb[1], 1));
    TTG_TRY(plan_build(p[0]->C, W->m, e.n[1], e.pa[1], e.pb[1], prec, -1, &q));
    p[1].reset(q);
    TTG_TRY(check_pairs(*p[1]->C, *psi->m, e.n[2], e.pa[2], e.pb[2], 2));
    TTG_TRY(plan_build(p[1]->C, psi->m, e.n[2], e.pa[2], e.pb[2], prec, -1, &q));
    p[2].reset(q);
  }
  const bool red = prec != TTG_FP64;
  const size_t es = red ? sizeof(float) : sizeof(double);
  Scratch sc;
  void* T1 = sc.get<char>(std::max<size_t>(1, p[0]->C->total) * es);
  void* T2 = sc.get<char>(std::max<size_t>(1, p[1]->C->total) * es);
  TTG_TRY(sc.ok("env_update"));
  ttg_tensor res = nullptr;
  TTG_TRY(tensor_from_meta(p[2]->C, false, &res));
  std::unique_ptr<ttg_tensor_s> hold(res);
  const void *a0, *b0, *b1, *b2;
  bool env_shadow = false;
  if (red) {
    // FP32 operands; the environment's shadow is dropped again after the
    // call unless it existed before (environments are the largest tensors)
    env_shadow = env->d_f32 == nullptr;
    const float *fp, *fe, *fw;
    TTG_TRY(tensor_f32(psi, &fp));
    TTG_TRY(tensor_f32(env, &fe));
    TTG_TRY(tensor_f32(W, &fw));
    a0 = fp; b0 = fe; b1 = fw; b2 = fp;
  } else {
    a0 = psi->d_data; b0 = env->d_data; b1 = W->d_data; b2 = psi->d_data;
  }
  // the planner's analytic tile choice: an environment update runs once per
  // bond, so timing candidate shapes (as the H_eff operator does for its
  // repeated applies) would cost more than it saves
  hipStream_t s = stream();
  int rc = gemm_launch(*p[0], a0, b0, T1, !red, s);
  if (!rc) rc = gemm_launch(*p[1], T1, b1, T2, !red, s);
  if (!rc) rc = gemm_launch(*p[2], T2, b2, res->d_data, true, s);
  if (env_shadow) {
    dfree(env->d_f32);   // pool reuse is ordered by the library stream
    env->d_f32 = nullptr;
    env->f32_version = 0;
  }
  TTG_TRY(rc);
  // the plans' device lists and the scratch go back to the pool, ordered by
  // the stream after the launches that read them
  *out = hold.release();
  return 0;
}

}  // extern "C"
