// ttg_internal.h -- shared types of the gfx950 DMRG hot-path library.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/ttgpu.h"

#define MAXR TTG_MAX_RANK

namespace ttg {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
#define TTG_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::ttg::set_error(std::string(#call) + ": " + hipGetErrorString(_e));   \
      return 1;                                                              \
    }                                                                        \
  } while (0)
#define TTG_REQUIRE(cond, msg)        \
  do {                                \
    if (!(cond)) {                    \
      ::ttg::set_error(msg);          \
      return 1;                       \
    }                                 \
  } while (0)
#define TTG_TRY(call)           \
  do {                          \
    int _r = (call);            \
    if (_r != 0) return _r;     \
  } while (0)

// ---------------------------------------------------------------- runtime
struct Runtime {
  int device = -1;
  hipStream_t stream = nullptr;
  // a second stream (default priority, like the first) for independent work
  // that overlaps the main stream's, joined back through events
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  int num_cu = 256;
  bool inited = false;
};
Runtime& rt();
int ensure_init();
// wall time of a host phase, accumulated under `name` when TTG_HOST_PROF is
// set (scope-bound; the name must outlive the program, e.g. a literal)
class PhaseTimer {
 public:
  explicit PhaseTimer(const char* name);
  ~PhaseTimer();
  PhaseTimer(const PhaseTimer&) = delete;
  PhaseTimer& operator=(const PhaseTimer&) = delete;

 private:
  const char* name_;
  bool on_;
  double t0_ = 0;
};
// set-up of the allocation-free entry points on the current device (device
// address of their fixed scratch array; called by ttg_init)
int noalloc_init();
inline hipStream_t stream() { return rt().stream; }
// the side stream, after every piece of work already queued on the main one
int side_fork(hipStream_t* side);
// the main stream waits for everything queued on the side stream so far
int side_join();
// host-to-device copy on the library stream through pinned staging: returns
// once the bytes are staged (the source may be released), the copy itself is
// ordered on the stream
int h2d(void* dst, const void* src, size_t bytes);
template <class T>
int h2d_vec(T* dst, const std::vector<T>& v) {
  return h2d(dst, v.data(), sizeof(T) * v.size());
}
// host-to-device copy of several host ranges into one device buffer, at byte
// offsets `off` (the bytes between pieces are left unspecified), as one copy
// on the library stream; returns once the bytes are staged, like h2d
struct H2DPiece {
  const void* src;
  size_t off;
  size_t bytes;
};
int h2d_pieces(void* dst, const H2DPiece* pieces, int n);
// device-to-host copies of several ranges into pinned staging, one wait for
// the library stream, then into the destinations
struct D2HPiece {
  void* dst;
  const void* src;
  size_t bytes;
};
int d2h_wait(const D2HPiece* pieces, int n);

// caching device pool (size classes; buffers go back to the pool, reuse is
// ordered by the library's single stream)
void* dmalloc(size_t bytes);
void dfree(void* p);
template <class T>
T* dalloc(size_t n) { return (T*)dmalloc(n * sizeof(T) + (n == 0 ? 8 : 0)); }

// Scratch buffers of one library call.  Every buffer goes back to the pool
// when the object leaves scope, on the error returns too, and a refused
// allocation is remembered: the caller checks ok() after a group of
// allocations and fails with the pool's reason before it launches anything
// on a null pointer.
class Scratch {
 public:
  Scratch() = default;
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  ~Scratch() {
    for (void* p : bufs_) dfree(p);
  }
  // Buffers up to kSmall bytes are carved (256-byte aligned) from shared
  // chunks of kChunk bytes, so a planning call's two dozen small tables cost
  // one or two pool allocations instead of one each; larger ones are pool
  // allocations of their own.
  template <class T>
  T* get(size_t n) {
    const size_t bytes = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~size_t(255);
    if (bytes > kSmall) {
      T* p = dalloc<T>(n);
      hold(p);
      return p;
    }
    if (!chunk_ || used_ + bytes > kChunk) {
      chunk_ = (char*)dmalloc(kChunk);
      hold(chunk_);
      used_ = 0;
      if (!chunk_) return nullptr;
    }
    T* p = (T*)(chunk_ + used_);
    used_ += bytes;
    return p;
  }
  // take over a buffer allocated elsewhere (null marks a failure)
  void hold(void* p) {
    if (p) bufs_.push_back(p);
    else failed_ = true;
  }
  bool failed() const { return failed_; }
  // 0, or 1 with "<what>: <the pool's reason>" as the error
  int ok(const char* what) const;

 private:
  static constexpr size_t kSmall = size_t(64) << 10, kChunk = size_t(1) << 20;
  std::vector<void*> bufs_;
  char* chunk_ = nullptr;
  size_t used_ = 0;
  bool failed_ = false;
};

// ---------------------------------------------------------------- tensors
// Kernel-side view of a tensor's block structure.  Index sectors are stored
// as (q0, q1, dim) int triples, all indices concatenated; idx_off[j] is the
// first sector of index j.
struct MetaView {
  int rank;
  int nblocks;
  int arrow[MAXR];
  int nsec[MAXR];
  int idx_off[MAXR];
  const int* idx;         // 3 ints per sector
  const int* qi;          // nblocks * rank
  const int64_t* off;     // nblocks + 1
};

// Immutable block structure shared by all tensors with the same block list.
struct TMeta {
  int rank = 0;
  int nblocks = 0;
  size_t total = 0;
  int arrow[MAXR] = {0};
  int nsec[MAXR] = {0};
  int idx_off[MAXR + 1] = {0};
  std::vector<int> h_idx;          // host copy of the (q0,q1,dim) triples
  int* d_idx = nullptr;
  int* d_qi = nullptr;
  int64_t* d_off = nullptr;
  // lazily built: sector-combination -> block id table (for embedding)
  int* d_lut = nullptr;
  int64_t lut_size = 0;
  ~TMeta();
  MetaView view() const;
  int q0(int j, int s) const { return h_idx[3 * (idx_off[j] + s)]; }
  int q1(int j, int s) const { return h_idx[3 * (idx_off[j] + s) + 1]; }
  int dim(int j, int s) const { return h_idx[3 * (idx_off[j] + s) + 2]; }
  bool same_indices(const TMeta& o) const;
};
using MetaPtr = std::shared_ptr<TMeta>;

// index tables only (no blocks yet)
int meta_shell(int rank, const int* arrows, const int* nsec, const int* qn, const int* qdim,
               MetaPtr* out);
// index tables copied from selected indices of other metas
MetaPtr meta_from_indices(const std::vector<std::pair<const TMeta*, int>>& src);
int meta_alloc_blocks(TMeta& m, int nblocks, size_t total);

}  // namespace ttg

struct ttg_tensor_s {
  ttg::MetaPtr m;
  double* d_data = nullptr;
  float* d_f32 = nullptr;          // FP32 shadow of d_data, built on demand
  uint64_t version = 1;            // bumped on every write of d_data
  uint64_t f32_version = 0;
  ~ttg_tensor_s();
};

namespace ttg {
int tensor_from_meta(const MetaPtr& m, bool zero, ttg_tensor* out);
inline void mark_written(ttg_tensor t) { t->version++; }
int tensor_f32(ttg_tensor t, const float** out);
// flipped-arrow copy of a block structure (qtensor::dag)
int meta_dag(const TMeta& o, MetaPtr* out);
// same index sectors and block list (hence the same packed offsets); equal
// pointers short-cut, otherwise the block lists are compared on the host
bool same_meta(const MetaPtr& a, const MetaPtr& b);
bool same_structure(ttg_tensor a, ttg_tensor b);
int block_scan_sum(const int64_t* in, int64_t* out_excl, int n);

// device reductions shared by the vector ops
int dev_dot(const double* x, const double* y, size_t n, double* d_out);
}  // namespace ttg
