/* ttgpu.h -- C-ABI of the MI355X (gfx950) DMRG hot-path library libttgpu.so.
 *
 * The reference (ClarkResearchGroup/tensor-tools) is C++ over CTF; every
 * entry point below replaces the CTF/MKL work behind one reference call site,
 * cited per function.  Plain pointers and sizes only; 0 = success, non-zero =
 * failure with a message in ttg_last_error() (the C++ layer in
 * tensor-tools_amd/host turns a failure into an abort, matching the
 * reference's assert-and-abort convention).
 *
 * Tensor format (the reference's qtensor, qtensor/qtensor.h:84-92):
 *   rank indices; index j has nsec[j] sectors, sector s has QN (qn[2s],
 *   qn[2s+1]) and dimension qdim[s], sectors sorted ascending by QN
 *   (qtensor_index.h:47).  arrows[j]: 0 = Inward, 1 = Outward.
 *   Blocks: nblocks tuples of sector positions block_qi[b*rank + j]; data
 *   packed in block order, each block column-major (first index fastest);
 *   block offsets are the prefix sums of the block sizes.
 * Not thread-safe; one device, one stream per library instance.
 */
#ifndef TTGPU_H
#define TTGPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TTG_MAX_RANK 8

typedef struct ttg_tensor_s* ttg_tensor;
typedef struct ttg_plan_s* ttg_plan;
typedef struct ttg_heff_s* ttg_heff;

/* arithmetic of the block contractions */
enum { TTG_FP64 = 0, TTG_FP32 = 1, TTG_BF16X3 = 2 };
/* sweep directions (util/types_and_headers.h:147-148) */
enum { TTG_MOVE_FROM_LEFT = 0, TTG_MOVE_FROM_RIGHT = 1 };

/* ---- runtime ---------------------------------------------------------- */
/* replaces CTF::World construction (project/main.cpp:85): bind one GPU */
int ttg_init(int device);
int ttg_finalize(void);
const char* ttg_last_error(void);
int ttg_sync(void);
/* bytes currently held by the device pool / high-water mark */
int ttg_mem_info(size_t* in_use, size_t* peak, size_t* cached);
/* fault injection for tests (no reference counterpart): the n-th device
 * allocation request from now on is refused as if the pool were full
 * (n >= 1; 0 disables).  Every entry point must then fail with an error
 * before it launches work, and leave the library usable. */
int ttg_pool_fail_after(int64_t n);

/* ---- tensor store (qtensor storage, qtensor.h:84-92) ------------------- */
/* allocate a tensor with the given block list, data zeroed */
int ttg_tensor_create(int rank, const int* arrows, const int* nsec, const int* qn,
                      const int* qdim, int nblocks, const int* block_qi, ttg_tensor* out);
/* qtensor::initBlock (qtensor.cpp:302-345): legal blocks enumerated on the
 * device, index 0 fastest, zero total flux; data zeroed */
int ttg_tensor_init_blocks(int rank, const int* arrows, const int* nsec, const int* qn,
                           const int* qdim, ttg_tensor* out);
int ttg_tensor_upload(ttg_tensor t, const double* host, size_t n);
int ttg_tensor_download(ttg_tensor t, double* host, size_t n);
int ttg_tensor_info(ttg_tensor t, int* rank, int* nblocks, size_t* total);
/* block list and offsets (nblocks+1 entries) of a tensor, for the bit-exact
 * bookkeeping checks */
int ttg_tensor_meta_download(ttg_tensor t, int* block_qi, int64_t* offsets);
/* qtensor::dag (qtensor.cpp:1575-1581): flip every index arrow.  Blocks
 * and data are unchanged (a zero-flux block stays zero-flux). */
int ttg_tensor_dag(ttg_tensor t);
/* *same = 1 when a and b have identical index sectors and block lists
 * (the precondition of the packed-array vector ops below) */
int ttg_same_blocks(ttg_tensor a, ttg_tensor b, int* same);
/* qtensor copy constructor (qtensor.cpp:178-189) */
int ttg_tensor_copy(ttg_tensor src, ttg_tensor* out);
int ttg_tensor_free(ttg_tensor t);
/* setZero/setOne (qtensor.cpp:392-411): fill every element */
int ttg_tensor_fill(ttg_tensor t, double v);
/* setRandom (qtensor.cpp:382-388): uniform [0,1) from a SplitMix64 counter
 * stream seeded with seed (the reference draws from CTF fill_random) */
int ttg_tensor_set_random(ttg_tensor t, uint64_t seed);

/* ---- vector ops on packed blocks (qtensor.cpp:1031-1179, 1606-1647) ----
 * Blocks are matched by sector tuple; x and y must share the block list
 * (as every Davidson vector does), which the call verifies. */
int ttg_dot(ttg_tensor x, ttg_tensor y, double* out);            /* inner_product / contract */
/* qtensor::contract for different block lists (qtensor.cpp:1095-1119): the
 * sum over x's blocks that y also holds, matched by sector tuple; index
 * sectors must agree, arrows may differ (the reference dags A's indices) */
int ttg_dot_matched(ttg_tensor x, ttg_tensor y, double* out);
int ttg_nrm2(ttg_tensor x, double* out);                         /* norm */
int ttg_scale(ttg_tensor x, double a);                           /* operator*= */
int ttg_axpy(double a, ttg_tensor x, ttg_tensor y);              /* y.add(x, a) */

/* ---- block-sparse contraction (qtensor::operator*, qtensor.cpp:523-734) --
 * pa[p], pb[p]: positions in A and B of the p-th contracted index pair, in
 * the reference's pairing order (B order, :540-555).  The result index list
 * is A's free indices (A order) then B's free indices (B order).  The result
 * block list, offsets and the matched (C, A, B, M, N, K) task list are built
 * on the device in the reference's q -> right -> left -> mid order. */
int ttg_contract_plan(ttg_tensor A, ttg_tensor B, int npairs, const int* pa, const int* pb,
                      int prec, ttg_plan* out);
/* nblocks of C, number of matched block GEMMs, sum of 2*M*N*K */
int ttg_plan_info(ttg_plan p, int* nblocks, int64_t* ntasks, double* flops);
/* tasks as 6-tuples (c, a, b, M, N, K) in reference loop order */
int ttg_plan_tasks_download(ttg_plan p, int* tasks6);
/* allocate an output tensor with the plan's result block list */
int ttg_plan_new_output(ttg_plan p, ttg_tensor* C);
/* C = A*B for tensors whose block lists equal those the plan was built on */
int ttg_contract_exec(ttg_plan p, ttg_tensor A, ttg_tensor B, ttg_tensor C);
int ttg_plan_free(ttg_plan p);
/* one-shot plan + output + exec + free */
int ttg_contract(ttg_tensor A, ttg_tensor B, int npairs, const int* pa, const int* pb,
                 int prec, ttg_tensor* C);

/* ---- lazy effective Hamiltonian (big_qtensor::product, big_qtensor.cpp:23-40)
 * y = ((L*x)*W12)*R with the three stages planned once for x's block list.
 * pairs3: for stage s = 0,1,2: [np, pa0, pb0, pa1, pb1, ...] concatenated.
 * prec: TTG_FP64, TTG_FP32 or TTG_BF16X3 (split-BF16, three MFMA products,
 * FP32 accumulation); the output y is always FP64.  Applies are launched
 * as one captured HIP graph per operator. */
int ttg_heff_create(ttg_tensor L, ttg_tensor W12, ttg_tensor R, ttg_tensor x_like,
                    const int* pairs3, int prec, ttg_heff* out);
/* the same operator at another precision, on src's plans (output blocks and
 * task lists copied on the device, no QN bookkeeping) and src's tile shapes,
 * applied by direct stage launches (no graph: it serves a few products);
 * src's L, W12 and R are read by both.  The FP64 twin serves the residual
 * correction of the reduced forms (tensor_davidson.cpp:255-276 computes its
 * products with the one FP64 operator) */
int ttg_heff_create_twin(ttg_heff src, int prec, ttg_heff* out);
int ttg_heff_new_output(ttg_heff h, ttg_tensor* y);
int ttg_heff_apply(ttg_heff h, ttg_tensor x, ttg_tensor y);
/* algorithmic FLOPs of one apply (sum of 2*M*N*K over the 3 stages) and the
 * per-stage task counts */
int ttg_heff_info(ttg_heff h, double* flops, int64_t* ntasks3, double* stage_flops3);
/* device time of n whole applies (x copied in, graph launch, y copied out)
 * measured with HIP events on the library stream (ms per apply) */
int ttg_heff_time(ttg_heff h, ttg_tensor x, ttg_tensor y, int n, double* ms_per_apply);
/* per-stage device time (ms, mean over n runs) of the three grouped-GEMM
 * launches, HIP events on the library stream around each launch */
int ttg_heff_stage_time(ttg_heff h, ttg_tensor x, int n, double* ms3);
/* work shape the planner chose for each stage (0..10: MFMA tiles 64x64,
 * 32x32, 64x32, 32x64, 128x64, 256x16, 128x16, 64x16, 96x96, 96x32, 32x96;
 * 11: the vector-ALU kernel for K <= 32, N <= 16) and its work-item count */
int ttg_heff_tiles(ttg_heff h, int* tile_cfg3, int* ntiles3);
/* the plan of stage 0, 1 or 2 (L*x, *W12, *R), borrowed: valid while h
 * lives and never passed to ttg_plan_free.  For the bit-exact checks of each
 * stage's block list, offsets and task list (ttg_plan_info,
 * ttg_plan_tasks_download, ttg_plan_new_output), the three operator* calls of
 * big_qtensor::product (big_qtensor.cpp:27-39). */
int ttg_heff_stage_plan(ttg_heff h, int stage, ttg_plan* out);
int ttg_heff_free(ttg_heff h);

/* ---- environment update (updateEnv, dmrg.cpp:491-509; buildEnv's steps,
 * dmrg.cpp:76-83) ------------------------------------------------------
 * env' = ((psi^dag' * env) * W) * psi in three planned grouped-GEMM stages,
 * for the reference's index layouts:
 *   psi = (l In, s In, r Out); W = (w_l In, s Out lvl0, s In lvl1, w_r Out);
 *   MoveFromLeft:  env = TL[i]   = (l' In lvl1, w_l Out, l Out) -> TL[i+1]
 *   MoveFromRight: env = TR[i]   = (r' Out lvl1, w_r In, r In)  -> TR[i-1]
 * psi^dag' is psi with every arrow flipped (and primed, a tag change only);
 * its data is psi's.  Operands whose paired indices do not carry equal
 * sectors with opposite arrows are rejected before any launch.  prec selects
 * the arithmetic of the three stages (FP64 output in every form).  *out is a
 * new tensor with the block list operator* gives the last stage. */
int ttg_env_update(ttg_tensor psi, ttg_tensor env, ttg_tensor W, int direction, int prec,
                   ttg_tensor* out);
/* the (psi^dag'*env, *W, *psi) index pairs ttg_env_update uses for a
 * direction, in ttg_heff_create's pairs3 encoding (host only, no device
 * work): the pairs qtensor::operator* finds by tag (qtensor.cpp:540-555) for
 * the layouts above */
int ttg_env_pairs(int direction, int* pairs3);

/* ---- sector SVD with global truncation (qtensor_op.cpp:1097-1453) -------
 * A's first nleft indices (in order) form the rows, the rest the columns
 * (the caller permutes indices so that this holds, as svd_bond's callers in
 * dmrg do).  Per mid-QN sector: FP64 Gram matrix on MFMA, FP64 Jacobi
 * eigensolver, singular values, global descending sort and the reference's
 * truncation rule (cutoff on 1-cumsum of s^2/sum s^2, cap K), then U and V
 * sliced into blocks.  direction MoveFromLeft absorbs S into V, MoveFromRight
 * into U.  Outputs: U (left..., mid Outward), V (mid Inward, right...),
 * S (mid), truncation error max(1-cumsum,0) and kept count. */
int ttg_svd(ttg_tensor A, int nleft, int direction, double cutoff, int64_t K,
            ttg_tensor* U, ttg_tensor* S, ttg_tensor* V, double* trunc_err, int64_t* kept);
/* sector layout of the mid index of the last ttg_svd: nsec, qn pairs, dims */
int ttg_tensor_index_download(ttg_tensor t, int j, int* nsec, int* qn, int* qdim);

/* ---- Davidson support (tensor_davidson.cpp:192-347) --------------------
 * Rayleigh-Ritz step on the device: symmetric eigensolver of the n x n
 * Krylov matrix M (column-major, FP64) returning ascending eigenvalues and
 * column eigenvectors (dsyev semantics, lapack_wrapper.cpp:40-54). */
int ttg_small_eigh(int n, const double* M, double* evals, double* evecs);
/* y = sum_k c[k] * x_k over nx tensors sharing one block list */
int ttg_lincomb(int nx, const ttg_tensor* xs, const double* c, ttg_tensor y);
/* dst = 0, then every block of src copied into the block of dst with the
 * same sector tuple (the block-appending half of qtensor::add,
 * qtensor.cpp:1141-1154): widens a tensor to a larger block list */
int ttg_embed(ttg_tensor src, ttg_tensor dst);
/* ttg_embed without any device allocation (no reference counterpart): for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc.  Its scratch (the
 * sorted destination keys, the block map, the error count) is a fixed 1.75
 * MiB device array of the library, located at ttg_init; src and dst may hold
 * at most 65536 blocks each (larger inputs are refused with an error before
 * any launch; use ttg_embed).  Every call on a device shares that device's
 * array: calls that overlap on one device race, so issue them from one
 * thread (the library's single stream orders them). */
int ttg_embed_noalloc(ttg_tensor src, ttg_tensor dst);

#ifdef __cplusplus
}
#endif
#endif /* TTGPU_H */
