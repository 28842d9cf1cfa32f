// davidson.hip -- device pieces of tensor_davidsonIT (tensor_davidson.cpp:192-347)
// that are not BLAS-1: the FP64 Rayleigh-Ritz eigensolver of the small
// Krylov matrix (DIAG/dsyev, lapack_wrapper.cpp:40-54) and block-list
// embedding (qtensor::add appends missing blocks, qtensor.cpp:1127-1156).
#include <algorithm>

#include "ttg_device.h"
#include "ttg_internal.h"

namespace ttg {

// Cyclic Jacobi (Golub & Van Loan 8.5) on an n x n symmetric matrix, one
// thread (n <= 16).  Values ascending, vectors in columns (column-major).
__global__ void k_small_eigh(int n, const double* __restrict__ Min, double* __restrict__ evals,
                             double* __restrict__ evecs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a[16 * 16], v[16 * 16];
  for (int i = 0; i < n * n; i++) a[i] = Min[i];
  for (int i = 0; i < n * n; i++) v[i] = 0.0;
  for (int i = 0; i < n; i++) v[i + n * i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n * n; i++) tot += a[i] * a[i];
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++)
        if (i != j) off += a[i + n * j] * a[i + n * j];
    if (off == 0.0 || off <= 1e-32 * tot) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = a[p + n * q];
        if (apq == 0.0) continue;
        double theta = (a[q + n * q] - a[p + n * p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {
          double x = a[k + n * p], y = a[k + n * q];
          a[k + n * p] = c * x - s * y;
          a[k + n * q] = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          double x = a[p + n * k], y = a[q + n * k];
          a[p + n * k] = c * x - s * y;
          a[q + n * k] = s * x + c * y;
        }
        for (int k = 0; k < n; k++) {
          double x = v[k + n * p], y = v[k + n * q];
          v[k + n * p] = c * x - s * y;
          v[k + n * q] = s * x + c * y;
        }
      }
  }
  // stable selection sort by eigenvalue
  int ord[16];
  for (int i = 0; i < n; i++) ord[i] = i;
  for (int i = 1; i < n; i++) {
    int o = ord[i];
    int j = i - 1;
    while (j >= 0 && a[ord[j] + n * ord[j]] > a[o + n * o]) {
      ord[j + 1] = ord[j];
      j--;
    }
    ord[j + 1] = o;
  }
  for (int j = 0; j < n; j++) {
    evals[j] = a[ord[j] + n * ord[j]];
    for (int i = 0; i < n; i++) evecs[i + n * j] = v[i + n * ord[j]];
  }
}

// dst (zeroed) receives every block of src at the same sector tuple
__global__ void k_embed_map(MetaView src, const uint64_t* dkeys, const int* dids, int nd,
                            const int64_t* doff, int* map, int* err) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= src.nblocks) return;
  uint64_t k = 0;
  for (int j = 0; j < src.rank; j++) k = k * src.nsec[j] + src.qi[(size_t)b * src.rank + j];
  int p = lower_bound_u64(dkeys, nd, k);
  if (p >= nd || dkeys[p] != k) {
    atomicAdd(err, 1);
    map[b] = -1;
    return;
  }
  map[b] = dids[p];
}

__global__ void k_embed_copy(int nb, const int* map, const int64_t* soff, const int64_t* doff,
                             const double* __restrict__ s, double* __restrict__ d) {
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const int db = map[b];
    if (db < 0) continue;
    const int64_t n = soff[b + 1] - soff[b];
#pragma unroll 8
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) d[doff[db] + i] = s[soff[b] + i];
  }
}

// sum over the blocks of x that y also holds (map: y block of each x block)
__global__ void __launch_bounds__(256) k_dot_mapped(int nb, const int* map, const int64_t* xoff,
                                                    const int64_t* yoff, const double* __restrict__ x,
                                                    const double* __restrict__ y, double* part) {
  double s = 0.0;
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    const int yb = map[b];
    if (yb < 0) continue;
    const int64_t n = xoff[b + 1] - xoff[b];
#pragma unroll 8
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += x[xoff[b] + i] * y[yoff[yb] + i];
  }
  __shared__ double red[256];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(512) k_sum_parts(const double* part, int n, double* out) {
  __shared__ double red[512];
  double s = threadIdx.x < n ? part[threadIdx.x] : 0.0;
  s = block_sum<512>(s, red);
  if (threadIdx.x == 0) *out = s;
}

__global__ void k_fill_i32(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- embed without allocation
// Device scratch of ttg_embed_noalloc: the sorted destination keys and ids,
// the source-to-destination block map and an error counter, carved from one
// file-scope array.  Each device holds its own copy; host code reaches it
// only through the address hipGetSymbolAddress returns for that device.
constexpr int kNoAllocMaxBlocks = 65536;
constexpr size_t kNoAllocBytes = (size_t)kNoAllocMaxBlocks * (8 + 8 + 4 + 4) +   // keys, skeys, ids, sids
                                 (size_t)kNoAllocMaxBlocks * 4 + 256;               // map, err
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
struct NoAllocDev {
  unsigned char* scratch = nullptr;   // device address of `scratch` on this device
  int* flag = nullptr;                // pinned host word for the error count
};
static NoAllocDev g_noalloc[kMaxDevices];

// Set-up, called by ttg_init with the device current: the device address of
// the scratch array (each device holds its own copy) and a pinned host word
// for reading the error count back.
int noalloc_init() {
  int dev = -1;
  TTG_HIP(hipGetDevice(&dev));
  TTG_REQUIRE(dev >= 0 && dev < kMaxDevices, "embed_noalloc: device id out of range");
  NoAllocDev& d = g_noalloc[dev];
  if (!d.scratch) {
    void* p = nullptr;
    TTG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TTG_REQUIRE(p, "embed_noalloc: no device address for the scratch array");
    d.scratch = (unsigned char*)p;
  }
  if (!d.flag) TTG_HIP(hipHostMalloc((void**)&d.flag, sizeof(int), hipHostMallocDefault));
  return 0;
}

}  // namespace ttg

using namespace ttg;

extern "C" {

int ttg_dot_matched(ttg_tensor x, ttg_tensor y, double* out) {
  PhaseTimer _phase("ttg_dot_matched");
  TTG_REQUIRE(x && y && out, "dot_matched: null argument");
  const TMeta& a = *x->m;
  const TMeta& b = *y->m;
  bool same = a.rank == b.rank && a.h_idx == b.h_idx;
  for (int j = 0; same && j < a.rank; j++) same = a.nsec[j] == b.nsec[j];
  TTG_REQUIRE(same, "dot_matched: index sectors differ");
  hipStream_t s = stream();
  const int nx = a.nblocks, ny = b.nblocks;
  const int grid = std::max(1, std::min(nx, 512));
  Scratch sc;
  uint64_t* keys = sc.get<uint64_t>(std::max(ny, 1));
  uint64_t* skeys = sc.get<uint64_t>(std::max(ny, 1));
  int* ids = sc.get<int>(std::max(ny, 1));
  int* sids = sc.get<int>(std::max(ny, 1));
  int* map = sc.get<int>(std::max(nx, 1));
  int* err = sc.get<int>(1);
  double* part = sc.get<double>(grid + 1);
  TTG_TRY(sc.ok("dot_matched"));
  double r = 0.0;
  if (nx && ny) {
    hipLaunchKernelGGL(k_tuple_keys, dim3((ny + 255) / 256), dim3(256), 0, s, b.view(), keys, ids);
    hipLaunchKernelGGL(k_sort_keys, dim3((ny + 255) / 256), dim3(256), 0, s, keys, ids, ny, skeys, sids);
    hipLaunchKernelGGL(k_embed_map, dim3((nx + 255) / 256), dim3(256), 0, s, a.view(), skeys, sids,
                       ny, b.d_off, map, err);   // blocks y lacks map to -1 (they add zero)
    hipLaunchKernelGGL(k_dot_mapped, dim3(grid), dim3(256), 0, s, nx, map, a.d_off, b.d_off,
                       x->d_data, y->d_data, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(512), 0, s, part, grid, part + grid);
    const D2HPiece pc{&r, part + grid, sizeof(double)};
    TTG_TRY(d2h_wait(&pc, 1));
  }
  *out = r;
  return 0;
}

int ttg_small_eigh(int n, const double* M, double* evals, double* evecs) {
  PhaseTimer _phase("ttg_small_eigh");
  TTG_TRY(ensure_init());
  TTG_REQUIRE(n >= 1 && n <= 16, "small_eigh: 1 <= n <= 16");
  Scratch sc;
  double* d = sc.get<double>(n * n * 2 + n);
  TTG_TRY(sc.ok("small_eigh"));
  TTG_TRY(h2d(d, M, sizeof(double) * n * n));
  hipLaunchKernelGGL(k_small_eigh, dim3(1), dim3(64), 0, stream(), n, d, d + 2 * n * n, d + n * n);
  const D2HPiece pc[2] = {{evecs, d + n * n, sizeof(double) * n * n}, {evals, d + 2 * n * n, sizeof(double) * n}};
  return d2h_wait(pc, 2);
}

int ttg_embed(ttg_tensor src, ttg_tensor dst) {
  PhaseTimer _phase("ttg_embed");
  TTG_REQUIRE(src && dst, "embed: null tensor");
  TTG_REQUIRE(src->m->same_indices(*dst->m), "embed: index sectors differ");
  hipStream_t s = stream();
  const int ns = src->m->nblocks, nd = dst->m->nblocks;
  Scratch sc;
  uint64_t* keys = sc.get<uint64_t>(std::max(nd, 1));
  uint64_t* skeys = sc.get<uint64_t>(std::max(nd, 1));
  int* ids = sc.get<int>(std::max(nd, 1));
  int* sids = sc.get<int>(std::max(nd, 1));
  int* map = sc.get<int>(std::max(ns, 1));
  int* err = sc.get<int>(1);
  TTG_TRY(sc.ok("embed"));
  int herr = 0;
  TTG_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
  fill_f64(dst->d_data, dst->m->total, 0.0);
  if (nd) {
    hipLaunchKernelGGL(k_tuple_keys, dim3((nd + 255) / 256), dim3(256), 0, s, dst->m->view(), keys, ids);
    hipLaunchKernelGGL(k_sort_keys, dim3((nd + 255) / 256), dim3(256), 0, s, keys, ids, nd, skeys, sids);
  }
  if (ns) {
    hipLaunchKernelGGL(k_embed_map, dim3((ns + 255) / 256), dim3(256), 0, s, src->m->view(), skeys,
                       sids, nd, dst->m->d_off, map, err);
    hipLaunchKernelGGL(k_embed_copy, dim3(std::min(ns, 2048)), dim3(256), 0, s, ns, map,
                       src->m->d_off, dst->m->d_off, src->d_data, dst->d_data);
  }
  {
    const D2HPiece pc{&herr, err, sizeof(int)};
    TTG_TRY(d2h_wait(&pc, 1));
  }
  mark_written(dst);
  TTG_REQUIRE(herr == 0, "embed: source block missing from destination block list");
  return 0;
}

int ttg_embed_noalloc(ttg_tensor src, ttg_tensor dst) {
  TTG_REQUIRE(src && dst, "embed_noalloc: null tensor");
  TTG_REQUIRE(src->m->same_indices(*dst->m), "embed_noalloc: index sectors differ");
  const int ns = src->m->nblocks, nd = dst->m->nblocks;
  TTG_REQUIRE(ns <= kNoAllocMaxBlocks && nd <= kNoAllocMaxBlocks,
              "embed_noalloc: more than 65536 blocks in the source or the destination "
              "(the limit of the library's fixed scratch; use ttg_embed)");
  TTG_TRY(ensure_init());
  int dev = -1;
  TTG_HIP(hipGetDevice(&dev));
  TTG_REQUIRE(dev >= 0 && dev < kMaxDevices && g_noalloc[dev].scratch && g_noalloc[dev].flag,
              "embed_noalloc: library not initialised on this device (ttg_init)");
  uint64_t* keys = (uint64_t*)g_noalloc[dev].scratch;
  uint64_t* skeys = keys + kNoAllocMaxBlocks;
  int* ids = (int*)(skeys + kNoAllocMaxBlocks);
  int* sids = ids + kNoAllocMaxBlocks;
  int* map = sids + kNoAllocMaxBlocks;
  int* err = map + kNoAllocMaxBlocks;
  hipStream_t s = stream();
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, err, 1, 0);
  fill_f64(dst->d_data, dst->m->total, 0.0);
  if (nd) {
    hipLaunchKernelGGL(k_tuple_keys, dim3((nd + 255) / 256), dim3(256), 0, s, dst->m->view(), keys, ids);
    hipLaunchKernelGGL(k_sort_keys, dim3((nd + 255) / 256), dim3(256), 0, s, keys, ids, nd, skeys, sids);
  }
  if (ns) {
    hipLaunchKernelGGL(k_embed_map, dim3((ns + 255) / 256), dim3(256), 0, s, src->m->view(), skeys,
                       sids, nd, dst->m->d_off, map, err);
    hipLaunchKernelGGL(k_embed_copy, dim3(std::min(ns, 2048)), dim3(256), 0, s, ns, map,
                       src->m->d_off, dst->m->d_off, src->d_data, dst->d_data);
  }
  TTG_HIP(hipGetLastError());
  int* flag = g_noalloc[dev].flag;
  TTG_HIP(hipMemcpyAsync(flag, err, sizeof(int), hipMemcpyDeviceToHost, s));
  TTG_HIP(hipStreamSynchronize(s));
  mark_written(dst);
  TTG_REQUIRE(*flag == 0, "embed_noalloc: source block missing from destination block list");
  return 0;
}

}  // extern "C"
