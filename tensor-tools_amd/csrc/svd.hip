// svd.hip -- sector SVD with global truncation (qtensor_op.cpp:1097-1453).
#include "ttg_internal.h"

using namespace ttg;

extern "C" int ttg_svd(ttg_tensor A, int nleft, int direction, double cutoff, int64_t K,
                       ttg_tensor* U, ttg_tensor* S, ttg_tensor* V, double* trunc_err,
                       int64_t* kept) {
  (void)A; (void)nleft; (void)direction; (void)cutoff; (void)K; (void)U; (void)S; (void)V;
  (void)trunc_err; (void)kept;
  set_error("ttg_svd: not built yet");
  return 1;
}
