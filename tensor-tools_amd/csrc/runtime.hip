// runtime.hip -- device binding, stream, error string and caching pool.
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
namespace {
struct Pool {
  std::multimap<size_t, void*> free_list;
  std::map<void*, size_t> live;
  size_t in_use = 0, peak = 0, cached = 0;
} g_pool;

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
  size_t sz = size_class(bytes);
  auto it = g_pool.free_list.lower_bound(sz);
  void* p = nullptr;
  if (it != g_pool.free_list.end() && it->first <= sz + sz / 4) {
    p = it->second;
    sz = it->first;
    g_pool.cached -= sz;
    g_pool.free_list.erase(it);
  } else {
    if (hipMalloc(&p, sz) != hipSuccess) {
      // release cached buffers and retry once.  The failure is handled here,
      // so it is cleared from the runtime's last-error state: the next
      // launch check (hipGetLastError) would otherwise report it.
      (void)hipGetLastError();
      (void)hipDeviceSynchronize();
      for (auto& kv : g_pool.free_list) (void)hipFree(kv.second);
      g_pool.free_list.clear();
      g_pool.cached = 0;
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

int ensure_init() {
  if (g_rt.inited) return 0;
  return ttg_init(0);
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
  g_rt.device = device;
  g_rt.inited = true;
  return 0;
}

int ttg_finalize(void) {
  if (!g_rt.inited) return 0;
  TTG_HIP(hipStreamSynchronize(g_rt.stream));
  for (auto& kv : g_pool.free_list) (void)hipFree(kv.second);
  g_pool.free_list.clear();
  g_pool.cached = 0;
  return 0;
}

const char* ttg_last_error(void) { return g_error.c_str(); }

int ttg_sync(void) {
  TTG_TRY(ensure_init());
  TTG_HIP(hipStreamSynchronize(g_rt.stream));
  return 0;
}

int ttg_mem_info(size_t* in_use, size_t* peak, size_t* cached) {
  if (in_use) *in_use = g_pool.in_use;
  if (peak) *peak = g_pool.peak;
  if (cached) *cached = g_pool.cached;
  return 0;
}

}  // extern "C"
