// ttg_device.h -- device-side helpers (64-wide wavefront scans/reductions).
#pragma once
#include "ttg_internal.h"

namespace ttg {

void fill_f64(double* d, size_t n, double v);

// Sum over a workgroup of NT threads (NT a multiple of 64).  lds >= NT/64.
template <int NT>
__device__ inline double block_sum(double v, double* lds) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double t = 0.0;
  if (w == 0) {
    t = lane < NT / 64 ? lds[lane] : 0.0;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d, 64);
    if (lane == 0) lds[0] = t;
  }
  __syncthreads();
  double r = lds[0];
  __syncthreads();
  return r;
}

// Exclusive prefix sum over a workgroup of NT threads; lds >= NT/64 entries.
template <int NT>
__device__ inline int64_t block_excl_scan(int64_t v, int64_t* lds, int64_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int64_t y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  if (w == 0) {
    int64_t t = lane < NT / 64 ? lds[lane] : 0;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      int64_t y = __shfl_up(t, d, 64);
      if (lane >= d) t += y;
    }
    if (lane < NT / 64) lds[lane] = t;
  }
  __syncthreads();
  int64_t base = w > 0 ? lds[w - 1] : 0;
  *total = lds[NT / 64 - 1];
  __syncthreads();
  return base + x - v;
}

// lower_bound over a sorted uint64 array
__device__ inline int lower_bound_u64(const uint64_t* a, int n, uint64_t key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// full-tuple key of every block of a structure (index 0 most significant)
static __global__ void k_tuple_keys(MetaView m, uint64_t* keys, int* ids) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= m.nblocks) return;
  uint64_t k = 0;
  for (int j = 0; j < m.rank; j++) k = k * m.nsec[j] + m.qi[(size_t)b * m.rank + j];
  keys[b] = k;
  ids[b] = b;
}

static __global__ void __launch_bounds__(256) k_sort_keys(const uint64_t* in, const int* vin, int n,
                                                   uint64_t* out, int* vout) {
  __shared__ uint64_t tile[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const uint64_t key = i < n ? in[i] : 0;
  int rank = 0;
  for (int base = 0; base < n; base += 256) {
    __syncthreads();
    if (base + (int)threadIdx.x < n) tile[threadIdx.x] = in[base + threadIdx.x];
    __syncthreads();
    const int lim = min(256, n - base);
    for (int t = 0; t < lim; t++) rank += (tile[t] < key) || (tile[t] == key && base + t < i);
  }
  if (i < n) {
    out[rank] = key;
    vout[rank] = vin[i];
  }
}

}  // namespace ttg
