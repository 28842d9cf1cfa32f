// runtime.hip -- device binding, stream, error string, caching pool, pinned
// staging of small host-device copies and the host-side phase timers.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "ttg_internal.h"

namespace ttg {

static std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }

static Runtime g_rt;
Runtime& rt() { return g_rt; }

// ---------------------------------------------------------------- pool
// Size classes: round up to 256 B, then to a multiple of 1/8 of the next
// power of two, so a freed buffer is reused by any request within ~12%.
// The pool holds at most `budget` bytes (live plus cached): the device's free
// memory at ttg_init less a reserve for the HIP runtime's own allocations
// (queues, kernel scratch), which fail with the device full of cached buffers.
// A request that would pass the budget first returns the cache to HIP, and
// is refused (an error, not a device-wide failure) if live buffers alone
// would pass it.
namespace {
struct Pool {
  std::multimap<size_t, void*> free_list;
  std::map<void*, size_t> live;
  size_t in_use = 0, peak = 0, cached = 0;
  size_t budget = SIZE_MAX;
  int64_t fail_at = 0;   // > 0: the fail_at-th request from now is refused (tests)
} g_pool;

// hand every cached buffer back to HIP (after the device is idle: a buffer
// freed by the host may still be read by queued work)
void release_cache() {
  if (g_pool.free_list.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto& kv : g_pool.free_list) (void)hipFree(kv.second);
  g_pool.free_list.clear();
  g_pool.cached = 0;
}

size_t size_class(size_t b) {
  if (b < 256) return 256;
  size_t p = 1;
  while (p < b) p <<= 1;
  size_t step = p >> 3;
  if (step < 256) step = 256;
  return (b + step - 1) / step * step;
}
}  // namespace

void* dmalloc(size_t bytes) {
  if (g_pool.fail_at > 0 && --g_pool.fail_at == 0) {
    set_error("device allocation of " + std::to_string(bytes) + " bytes refused (injected failure)");
    return nullptr;
  }
  size_t sz = size_class(bytes);
  auto it = g_pool.free_list.lower_bound(sz);
  void* p = nullptr;
  if (it != g_pool.free_list.end() && it->first <= sz + sz / 4) {
    p = it->second;
    sz = it->first;
    g_pool.cached -= sz;
    g_pool.free_list.erase(it);
  } else {
    if (g_pool.in_use + g_pool.cached + sz > g_pool.budget) release_cache();
    if (g_pool.in_use + sz > g_pool.budget) {
      set_error("device allocation of " + std::to_string(sz) + " bytes refused: " +
                std::to_string(g_pool.in_use >> 20) + " MiB live, budget " + std::to_string(g_pool.budget >> 20) +
                " MiB (device memory less a reserve for the HIP runtime)");
      return nullptr;
    }
    if (hipMalloc(&p, sz) != hipSuccess) {
      // release cached buffers and retry once.  The failure is handled here,
      // so it is cleared from the runtime's last-error state: the next
      // launch check (hipGetLastError) would otherwise report it.
      (void)hipGetLastError();
      release_cache();
      if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        set_error("device allocation of " + std::to_string(sz) + " bytes failed");
        return nullptr;
      }
    }
  }
  g_pool.live[p] = sz;
  g_pool.in_use += sz;
  if (g_pool.in_use > g_pool.peak) g_pool.peak = g_pool.in_use;
  return p;
}

void dfree(void* p) {
  if (!p) return;
  auto it = g_pool.live.find(p);
  if (it == g_pool.live.end()) return;
  g_pool.in_use -= it->second;
  g_pool.cached += it->second;
  g_pool.free_list.insert({it->second, p});
  g_pool.live.erase(it);
}

int Scratch::ok(const char* what) const {
  if (!failed_) return 0;
  set_error(std::string(what) + ": " + g_error);   // g_error holds the pool's reason
  return 1;
}

int ensure_init() {
  if (g_rt.inited) return 0;
  return ttg_init(0);
}

int side_fork(hipStream_t* side) {
  TTG_TRY(ensure_init());
  if (!g_rt.side) {
    TTG_HIP(hipStreamCreateWithFlags(&g_rt.side, hipStreamNonBlocking));
    TTG_HIP(hipEventCreateWithFlags(&g_rt.fork, hipEventDisableTiming));
    TTG_HIP(hipEventCreateWithFlags(&g_rt.join, hipEventDisableTiming));
  }
  TTG_HIP(hipEventRecord(g_rt.fork, g_rt.stream));
  TTG_HIP(hipStreamWaitEvent(g_rt.side, g_rt.fork, 0));
  *side = g_rt.side;
  return 0;
}

int side_join() {
  if (!g_rt.side) return 0;
  TTG_HIP(hipEventRecord(g_rt.join, g_rt.side));
  TTG_HIP(hipStreamWaitEvent(g_rt.stream, g_rt.join, 0));
  return 0;
}

// ---------------------------------------------------------------- staging
// Small host-to-device copies go through a pinned ring: the bytes are copied
// into the ring and the device copy is queued from there, so the call returns
// at once and the caller's buffer may go out of scope (a copy queued from
// pageable memory makes the host wait for the stream).  Every ring copy is
// queued on the library stream; a copy that would wrap the ring first waits
// for that stream, which has then finished every earlier copy out of the
// ring.  Copies above a quarter of the ring are queued directly.
namespace {
constexpr size_t kRingBytes = size_t(8) << 20;
struct Ring {
  char* p = nullptr;
  size_t off = 0;
} g_ring;
struct Readback {
  char* p = nullptr;
  size_t cap = 0;
} g_rb;
}  // namespace

int h2d(void* dst, const void* src, size_t bytes) {
  if (!bytes) return 0;
  TTG_TRY(ensure_init());
  hipStream_t s = g_rt.stream;
  if (bytes > kRingBytes / 4) {
    TTG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    TTG_HIP(hipStreamSynchronize(s));   // the caller's buffer may be released on return
    return 0;
  }
  if (!g_ring.p) TTG_HIP(hipHostMalloc((void**)&g_ring.p, kRingBytes));
  size_t off = (g_ring.off + 255) & ~size_t(255);
  if (off + bytes > kRingBytes) {
    TTG_HIP(hipStreamSynchronize(s));
    off = 0;
  }
  memcpy(g_ring.p + off, src, bytes);
  TTG_HIP(hipMemcpyAsync(dst, g_ring.p + off, bytes, hipMemcpyHostToDevice, s));
  g_ring.off = off + bytes;
  return 0;
}

int h2d_pieces(void* dst, const H2DPiece* pieces, int n) {
  size_t total = 0;
  for (int i = 0; i < n; i++) total = std::max(total, pieces[i].off + pieces[i].bytes);
  if (!total) return 0;
  TTG_TRY(ensure_init());
  hipStream_t s = g_rt.stream;
  if (total > kRingBytes / 4) {   // large: one direct copy per piece
    for (int i = 0; i < n; i++) TTG_TRY(h2d((char*)dst + pieces[i].off, pieces[i].src, pieces[i].bytes));
    return 0;
  }
  if (!g_ring.p) TTG_HIP(hipHostMalloc((void**)&g_ring.p, kRingBytes));
  size_t off = (g_ring.off + 255) & ~size_t(255);
  if (off + total > kRingBytes) {
    TTG_HIP(hipStreamSynchronize(s));
    off = 0;
  }
  for (int i = 0; i < n; i++)
    if (pieces[i].bytes) memcpy(g_ring.p + off + pieces[i].off, pieces[i].src, pieces[i].bytes);
  TTG_HIP(hipMemcpyAsync(dst, g_ring.p + off, total, hipMemcpyHostToDevice, s));
  g_ring.off = off + total;
  return 0;
}

int d2h_wait(const D2HPiece* pieces, int n) {
  TTG_TRY(ensure_init());
  hipStream_t s = g_rt.stream;
  size_t total = 0;
  for (int i = 0; i < n; i++) total += (pieces[i].bytes + 255) & ~size_t(255);
  if (total > g_rb.cap) {
    if (g_rb.p) TTG_HIP(hipHostFree(g_rb.p));
    g_rb.p = nullptr;
    g_rb.cap = 0;
    const size_t cap = std::max<size_t>(total, size_t(1) << 20);
    TTG_HIP(hipHostMalloc((void**)&g_rb.p, cap));
    g_rb.cap = cap;
  }
  size_t off = 0;
  for (int i = 0; i < n; i++) {
    if (pieces[i].bytes)
      TTG_HIP(hipMemcpyAsync(g_rb.p + off, pieces[i].src, pieces[i].bytes, hipMemcpyDeviceToHost, s));
    off += (pieces[i].bytes + 255) & ~size_t(255);
  }
  TTG_HIP(hipStreamSynchronize(s));
  off = 0;
  for (int i = 0; i < n; i++) {
    if (pieces[i].bytes) memcpy(pieces[i].dst, g_rb.p + off, pieces[i].bytes);
    off += (pieces[i].bytes + 255) & ~size_t(255);
  }
  return 0;
}

// ---------------------------------------------------------------- phase timers
// Wall-clock seconds and counts per named host phase, when the environment
// variable TTG_HOST_PROF is set; printed to stderr at process exit.  A phase
// covers the host code of a library call, its synchronisations included.
namespace {
struct PhaseAcc {
  std::map<std::string, std::pair<double, long>> acc;
  bool on = false;
  PhaseAcc() {
    const char* e = getenv("TTG_HOST_PROF");
    on = e && *e && *e != '0';
  }
  ~PhaseAcc() {
    if (!on || acc.empty()) return;
    fprintf(stderr, "[ttg host phases] seconds  calls  phase\n");
    for (auto& kv : acc)
      fprintf(stderr, "[ttg host phases] %8.4f %6ld  %s\n", kv.second.first, kv.second.second, kv.first.c_str());
  }
};
PhaseAcc& phases() {
  static PhaseAcc p;
  return p;
}
}  // namespace

PhaseTimer::PhaseTimer(const char* name) : name_(name), on_(phases().on) {
  if (on_) t0_ = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
PhaseTimer::~PhaseTimer() {
  if (!on_) return;
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
  auto& a = phases().acc[name_];
  a.first += t - t0_;
  a.second += 1;
}

}  // namespace ttg

using namespace ttg;

extern "C" {

int ttg_init(int device) {
  if (g_rt.inited) {
    TTG_REQUIRE(device == g_rt.device, "ttg_init: already bound to another device");
    return 0;
  }
  int n = 0;
  TTG_HIP(hipGetDeviceCount(&n));
  TTG_REQUIRE(device >= 0 && device < n, "ttg_init: no such device");
  TTG_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  TTG_HIP(hipGetDeviceProperties(&prop, device));
  g_rt.num_cu = prop.multiProcessorCount;
  TTG_HIP(hipStreamCreateWithFlags(&g_rt.stream, hipStreamNonBlocking));
  size_t mfree = 0, mtotal = 0;
  TTG_HIP(hipMemGetInfo(&mfree, &mtotal));
  const size_t reserve = std::max<size_t>(mtotal / 8, (size_t)8 << 30);
  g_pool.budget = mfree > reserve ? mfree - reserve : mfree / 2;
  g_rt.device = device;
  g_rt.inited = true;
  TTG_TRY(noalloc_init());
  return 0;
}

int ttg_finalize(void) {
  if (!g_rt.inited) return 0;
  TTG_HIP(hipStreamSynchronize(g_rt.stream));
  release_cache();
  return 0;
}

const char* ttg_last_error(void) { return g_error.c_str(); }

int ttg_sync(void) {
  TTG_TRY(ensure_init());
  TTG_HIP(hipStreamSynchronize(g_rt.stream));
  return 0;
}

int ttg_pool_fail_after(int64_t n) {
  TTG_REQUIRE(n >= 0, "pool_fail_after: n must be >= 0");
  g_pool.fail_at = n;
  return 0;
}

int ttg_mem_info(size_t* in_use, size_t* peak, size_t* cached) {
  if (in_use) *in_use = g_pool.in_use;
  if (peak) *peak = g_pool.peak;
  if (cached) *cached = g_pool.cached;
  return 0;
}

}  // extern "C"
