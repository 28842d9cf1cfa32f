// tensor.hip -- device tensor store and the packed-block vector operations.
#include <algorithm>
#include <cstring>

#include "ttg_internal.h"
#include "ttg_device.h"

namespace ttg {

// ---------------------------------------------------------------- meta
TMeta::~TMeta() {
  dfree(d_idx);
  dfree(d_qi);
  dfree(d_off);
  dfree(d_lut);
}

MetaView TMeta::view() const {
  MetaView v;
  v.rank = rank;
  v.nblocks = nblocks;
  for (int j = 0; j < MAXR; j++) {
    v.arrow[j] = arrow[j];
    v.nsec[j] = nsec[j];
    v.idx_off[j] = idx_off[j];
  }
  v.idx = d_idx;
  v.qi = d_qi;
  v.off = d_off;
  return v;
}

bool TMeta::same_indices(const TMeta& o) const {
  if (rank != o.rank) return false;
  for (int j = 0; j < rank; j++)
    if (arrow[j] != o.arrow[j] || nsec[j] != o.nsec[j]) return false;
  return h_idx == o.h_idx;
}

static int upload_idx(TMeta& m) {
  m.d_idx = dalloc<int>(m.h_idx.size());
  TTG_REQUIRE(m.d_idx, "meta: device allocation failed");
  return h2d_vec(m.d_idx, m.h_idx);
}

int meta_shell(int rank, const int* arrows, const int* nsec, const int* qn, const int* qdim,
               MetaPtr* out) {
  TTG_TRY(ensure_init());
  TTG_REQUIRE(rank >= 0 && rank <= MAXR, "tensor rank out of range (max 8)");
  auto m = std::make_shared<TMeta>();
  m->rank = rank;
  int off = 0;
  for (int j = 0; j < rank; j++) {
    TTG_REQUIRE(arrows[j] == 0 || arrows[j] == 1, "arrow must be 0 (Inward) or 1 (Outward)");
    TTG_REQUIRE(nsec[j] >= 1, "every index needs at least one sector");
    m->arrow[j] = arrows[j];
    m->nsec[j] = nsec[j];
    m->idx_off[j] = off;
    for (int s = 0; s < nsec[j]; s++) {
      int a = qn[2 * (off + s)], b = qn[2 * (off + s) + 1], d = qdim[off + s];
      TTG_REQUIRE(d >= 1, "sector dimension must be >= 1");
      TTG_REQUIRE(a > -32768 && a < 32767 && b > -32768 && b < 32767, "QN component out of range");
      if (s > 0) {
        int pa = qn[2 * (off + s - 1)], pb = qn[2 * (off + s - 1) + 1];
        TTG_REQUIRE(std::make_pair(pa, pb) < std::make_pair(a, b),
                    "index sectors must be sorted by QN and distinct");
      }
      m->h_idx.push_back(a);
      m->h_idx.push_back(b);
      m->h_idx.push_back(d);
    }
    off += nsec[j];
  }
  m->idx_off[rank] = off;
  TTG_TRY(upload_idx(*m));
  *out = m;
  return 0;
}

MetaPtr meta_from_indices(const std::vector<std::pair<const TMeta*, int>>& src) {
  auto m = std::make_shared<TMeta>();
  m->rank = (int)src.size();
  int off = 0;
  for (int j = 0; j < m->rank; j++) {
    const TMeta* s = src[j].first;
    int k = src[j].second;
    m->arrow[j] = s->arrow[k];
    m->nsec[j] = s->nsec[k];
    m->idx_off[j] = off;
    for (int t = 0; t < s->nsec[k]; t++) {
      m->h_idx.push_back(s->q0(k, t));
      m->h_idx.push_back(s->q1(k, t));
      m->h_idx.push_back(s->dim(k, t));
    }
    off += s->nsec[k];
  }
  m->idx_off[m->rank] = off;
  if (upload_idx(*m) != 0) return nullptr;
  return m;
}

int meta_alloc_blocks(TMeta& m, int nblocks, size_t total) {
  m.nblocks = nblocks;
  m.total = total;
  m.d_qi = dalloc<int>((size_t)nblocks * std::max(m.rank, 1));
  m.d_off = dalloc<int64_t>((size_t)nblocks + 1);
  TTG_REQUIRE(m.d_qi && m.d_off, "meta: device allocation failed");
  return 0;
}

}  // namespace ttg

ttg_tensor_s::~ttg_tensor_s() {
  ttg::dfree(d_data);
  ttg::dfree(d_f32);
}

namespace ttg {

int tensor_from_meta(const MetaPtr& m, bool zero, ttg_tensor* out) {
  ttg_tensor t = new ttg_tensor_s();
  t->m = m;
  t->d_data = dalloc<double>(m->total);
  if (!t->d_data) {
    delete t;
    set_error(std::string("tensor: ") + ttg_last_error());   // the pool's reason
    return 1;
  }
  if (zero && m->total)
    fill_f64(t->d_data, m->total, 0.0);
  *out = t;
  return 0;
}

bool same_structure(ttg_tensor a, ttg_tensor b) { return same_meta(a->m, b->m); }

bool same_meta(const MetaPtr& a, const MetaPtr& b) {
  if (a == b) return true;
  if (!a || !b) return false;
  const TMeta &x = *a, &y = *b;
  if (x.nblocks != y.nblocks || x.total != y.total || !x.same_indices(y)) return false;
  // compare block lists on the host (small)
  std::vector<int> qa((size_t)x.nblocks * x.rank), qb(qa.size());
  if (!qa.empty()) {
    const D2HPiece pc[2] = {{qa.data(), x.d_qi, qa.size() * sizeof(int)},
                            {qb.data(), y.d_qi, qb.size() * sizeof(int)}};
    if (d2h_wait(pc, 2) != 0) {
      set_error("block list comparison: device copy failed");
      return false;
    }
  }
  return qa == qb;
}

// ---------------------------------------------------------------- kernels
__global__ void k_fill(double* __restrict__ d, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    d[i] = v;
}

__global__ void k_random(double* __restrict__ d, size_t n, uint64_t seed) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    d[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
}

__global__ void k_to_f32(const double* __restrict__ s, float* __restrict__ d, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    d[i] = (float)s[i];
}

__global__ void k_scale(double* __restrict__ d, size_t n, double a) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    d[i] *= a;
}

__global__ void k_axpy(double a, const double* __restrict__ x, double* __restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}

struct LinArgs {
  const double* x[8];
  double c[8];
  int n;
};
__global__ void k_lincomb(LinArgs a, double* __restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < a.n; k++) s += a.c[k] * a.x[k][i];
    y[i] = s;
  }
}

// Deterministic two-pass dot: fixed 512 partial sums, then one workgroup.
constexpr int kDotGrid = 512;
__global__ void __launch_bounds__(256) k_dot_partial(const double* __restrict__ x,
                                                     const double* __restrict__ y, size_t n,
                                                     double* __restrict__ part) {
  double s = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    s += x[i] * y[i];
  __shared__ double red[256];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void __launch_bounds__(512) k_dot_final(const double* __restrict__ part, int n,
                                                   double* __restrict__ out) {
  __shared__ double red[512];
  double s = threadIdx.x < n ? part[threadIdx.x] : 0.0;
  s = block_sum<512>(s, red);
  if (threadIdx.x == 0) *out = s;
}

static double* g_dot_scratch = nullptr;
int dev_dot(const double* x, const double* y, size_t n, double* d_out) {
  if (!g_dot_scratch) {
    g_dot_scratch = dalloc<double>(kDotGrid);
    TTG_REQUIRE(g_dot_scratch, "dot scratch allocation failed");
  }
  hipLaunchKernelGGL(k_dot_partial, dim3(kDotGrid), dim3(256), 0, stream(), x, y, n, g_dot_scratch);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(512), 0, stream(), g_dot_scratch, kDotGrid, d_out);
  return 0;
}

static unsigned grid_for(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void fill_f64(double* d, size_t n, double v) {
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, stream(), d, n, v);
}

// the same sectors and block list with every arrow flipped (qtensor::dag,
// qtensor.cpp:1575-1581): a new structure object, the blocks copied on the
// device
int meta_dag(const TMeta& o, MetaPtr* out) {
  auto m = std::make_shared<TMeta>();
  m->rank = o.rank;
  m->h_idx = o.h_idx;
  for (int j = 0; j <= MAXR; j++) m->idx_off[j] = o.idx_off[j];
  for (int j = 0; j < o.rank; j++) {
    m->arrow[j] = 1 - o.arrow[j];
    m->nsec[j] = o.nsec[j];
  }
  TTG_TRY(upload_idx(*m));
  TTG_TRY(meta_alloc_blocks(*m, o.nblocks, o.total));
  if (o.nblocks)
    TTG_HIP(hipMemcpyAsync(m->d_qi, o.d_qi, sizeof(int) * (size_t)o.nblocks * o.rank,
                           hipMemcpyDeviceToDevice, stream()));
  TTG_HIP(hipMemcpyAsync(m->d_off, o.d_off, sizeof(int64_t) * ((size_t)o.nblocks + 1),
                         hipMemcpyDeviceToDevice, stream()));
  *out = m;
  return 0;
}

int tensor_f32(ttg_tensor t, const float** out) {
  if (!t->d_f32) {
    t->d_f32 = dalloc<float>(t->m->total);
    TTG_REQUIRE(t->d_f32, "f32 shadow allocation failed");
  }
  if (t->f32_version != t->version) {
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(t->m->total)), dim3(256), 0, stream(), t->d_data,
                       t->d_f32, t->m->total);
    t->f32_version = t->version;
  }
  *out = t->d_f32;
  return 0;
}

// ---------------------------------------------------------------- initBlock
// One workgroup walks the sector combinations in chunks of 1024 (index 0
// fastest), flags zero-flux tuples and compacts them with a block scan.
// phase 0 counts (nblocks, total) into hdr; phase 1 writes qi and offsets.
__global__ void __launch_bounds__(1024) k_init_blocks(MetaView m, int64_t ncomb, int phase,
                                                      int* __restrict__ qi_out,
                                                      int64_t* __restrict__ off_out,
                                                      int64_t* __restrict__ hdr) {
  __shared__ int64_t scan_lds[2 * 32];
  int64_t carry_n = 0, carry_s = 0;
  for (int64_t base = 0; base < ncomb; base += 1024) {
    int64_t i = base + threadIdx.x;
    int legal = 0;
    int64_t size = 0;
    int pos[MAXR];
    if (i < ncomb) {
      int64_t r = i;
      int f0 = 0, f1 = 0;
      size = 1;
      for (int j = 0; j < m.rank; j++) {
        int p = (int)(r % m.nsec[j]);
        r /= m.nsec[j];
        pos[j] = p;
        const int* e = m.idx + 3 * (m.idx_off[j] + p);
        int sg = m.arrow[j] == 0 ? 1 : -1;
        f0 += sg * e[0];
        f1 += sg * e[1];
        size *= e[2];
      }
      legal = (f0 == 0 && f1 == 0);
      if (!legal) size = 0;
    }
    int64_t tot_n, tot_s;
    int64_t pn = block_excl_scan<1024>((int64_t)legal, scan_lds, &tot_n);
    int64_t ps = block_excl_scan<1024>(size, scan_lds + 32, &tot_s);
    if (phase == 1 && legal) {
      int64_t b = carry_n + pn;
      for (int j = 0; j < m.rank; j++) qi_out[b * m.rank + j] = pos[j];
      off_out[b] = carry_s + ps;
    }
    carry_n += tot_n;
    carry_s += tot_s;
  }
  if (threadIdx.x == 0) {
    hdr[0] = carry_n;
    hdr[1] = carry_s;
    if (phase == 1) off_out[carry_n] = carry_s;
  }
}

}  // namespace ttg

using namespace ttg;

extern "C" {

int ttg_tensor_create(int rank, const int* arrows, const int* nsec, const int* qn,
                      const int* qdim, int nblocks, const int* block_qi, ttg_tensor* out) {
  PhaseTimer _phase("ttg_tensor_create");
  MetaPtr m;
  TTG_TRY(meta_shell(rank, arrows, nsec, qn, qdim, &m));
  TTG_REQUIRE(nblocks >= 0, "nblocks < 0");
  std::vector<int64_t> off(nblocks + 1, 0);
  for (int b = 0; b < nblocks; b++) {
    int64_t sz = 1;
    for (int j = 0; j < rank; j++) {
      int p = block_qi[b * rank + j];
      TTG_REQUIRE(p >= 0 && p < nsec[j], "block sector position out of range");
      sz *= m->dim(j, p);
    }
    off[b + 1] = off[b] + sz;
  }
  TTG_TRY(meta_alloc_blocks(*m, nblocks, (size_t)off[nblocks]));
  if (nblocks) TTG_TRY(h2d(m->d_qi, block_qi, sizeof(int) * (size_t)nblocks * rank));
  TTG_TRY(h2d_vec(m->d_off, off));
  TTG_TRY(tensor_from_meta(m, true, out));
  return 0;
}

int ttg_tensor_init_blocks(int rank, const int* arrows, const int* nsec, const int* qn,
                           const int* qdim, ttg_tensor* out) {
  PhaseTimer _phase("ttg_tensor_init_blocks");
  MetaPtr m;
  TTG_TRY(meta_shell(rank, arrows, nsec, qn, qdim, &m));
  int64_t ncomb = 1;
  for (int j = 0; j < rank; j++) {
    ncomb *= nsec[j];
    TTG_REQUIRE(ncomb < ((int64_t)1 << 40), "initBlock: too many sector combinations");
  }
  Scratch sc;
  int64_t* hdr = sc.get<int64_t>(2);
  TTG_TRY(sc.ok("initBlock"));
  int64_t h[2];
  hipLaunchKernelGGL(k_init_blocks, dim3(1), dim3(1024), 0, stream(), m->view(), ncomb, 0,
                     (int*)nullptr, (int64_t*)nullptr, hdr);
  {
    const D2HPiece pc{h, hdr, sizeof(h)};
    TTG_TRY(d2h_wait(&pc, 1));
  }
  TTG_REQUIRE(h[0] < (1 << 30), "initBlock: too many blocks");
  TTG_TRY(meta_alloc_blocks(*m, (int)h[0], (size_t)h[1]));
  hipLaunchKernelGGL(k_init_blocks, dim3(1), dim3(1024), 0, stream(), m->view(), ncomb, 1,
                     m->d_qi, m->d_off, hdr);
  return tensor_from_meta(m, true, out);
}

int ttg_tensor_upload(ttg_tensor t, const double* host, size_t n) {
  PhaseTimer _phase("ttg_tensor_upload");
  TTG_REQUIRE(t && n == t->m->total, "upload: size mismatch");
  if (n) TTG_HIP(hipMemcpyAsync(t->d_data, host, n * sizeof(double), hipMemcpyHostToDevice, stream()));
  TTG_HIP(hipStreamSynchronize(stream()));
  mark_written(t);
  return 0;
}

int ttg_tensor_download(ttg_tensor t, double* host, size_t n) {
  PhaseTimer _phase("ttg_tensor_download");
  TTG_REQUIRE(t && n == t->m->total, "download: size mismatch");
  if (n) TTG_HIP(hipMemcpyAsync(host, t->d_data, n * sizeof(double), hipMemcpyDeviceToHost, stream()));
  TTG_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int ttg_tensor_info(ttg_tensor t, int* rank, int* nblocks, size_t* total) {
  TTG_REQUIRE(t, "null tensor");
  if (rank) *rank = t->m->rank;
  if (nblocks) *nblocks = t->m->nblocks;
  if (total) *total = t->m->total;
  return 0;
}

int ttg_tensor_meta_download(ttg_tensor t, int* block_qi, int64_t* offsets) {
  TTG_REQUIRE(t, "null tensor");
  const TMeta& m = *t->m;
  if (block_qi && m.nblocks)
    TTG_HIP(hipMemcpyAsync(block_qi, m.d_qi, sizeof(int) * (size_t)m.nblocks * m.rank,
                           hipMemcpyDeviceToHost, stream()));
  if (offsets)
    TTG_HIP(hipMemcpyAsync(offsets, m.d_off, sizeof(int64_t) * ((size_t)m.nblocks + 1),
                           hipMemcpyDeviceToHost, stream()));
  TTG_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int ttg_tensor_index_download(ttg_tensor t, int j, int* nsec, int* qn, int* qdim) {
  TTG_REQUIRE(t && j >= 0 && j < t->m->rank, "index_download: bad index");
  const TMeta& m = *t->m;
  if (nsec) *nsec = m.nsec[j];
  for (int s = 0; s < m.nsec[j]; s++) {
    if (qn) { qn[2 * s] = m.q0(j, s); qn[2 * s + 1] = m.q1(j, s); }
    if (qdim) qdim[s] = m.dim(j, s);
  }
  return 0;
}

int ttg_tensor_dag(ttg_tensor t) {
  PhaseTimer _phase("ttg_tensor_dag");
  TTG_REQUIRE(t, "null tensor");
  MetaPtr m;
  TTG_TRY(meta_dag(*t->m, &m));
  t->m = m;
  return 0;
}

int ttg_same_blocks(ttg_tensor a, ttg_tensor b, int* same) {
  TTG_REQUIRE(a && b && same, "same_blocks: null argument");
  *same = same_structure(a, b) ? 1 : 0;
  return 0;
}

int ttg_tensor_copy(ttg_tensor src, ttg_tensor* out) {
  PhaseTimer _phase("ttg_tensor_copy");
  TTG_REQUIRE(src, "null tensor");
  TTG_TRY(tensor_from_meta(src->m, false, out));
  if (src->m->total)
    TTG_HIP(hipMemcpyAsync((*out)->d_data, src->d_data, src->m->total * sizeof(double),
                           hipMemcpyDeviceToDevice, stream()));
  return 0;
}

int ttg_tensor_free(ttg_tensor t) {
  delete t;
  return 0;
}

int ttg_tensor_fill(ttg_tensor t, double v) {
  TTG_REQUIRE(t, "null tensor");
  fill_f64(t->d_data, t->m->total, v);
  mark_written(t);
  return 0;
}

int ttg_tensor_set_random(ttg_tensor t, uint64_t seed) {
  TTG_REQUIRE(t, "null tensor");
  size_t n = t->m->total;
  hipLaunchKernelGGL(k_random, dim3(grid_for(n)), dim3(256), 0, stream(), t->d_data, n, seed);
  mark_written(t);
  return 0;
}

// ---------------------------------------------------------------- vector ops
static double* g_scalar = nullptr;
static double* g_scalar_host = nullptr;

static int read_scalar(double* out) {
  if (!g_scalar_host) TTG_HIP(hipHostMalloc((void**)&g_scalar_host, sizeof(double)));
  TTG_HIP(hipMemcpyAsync(g_scalar_host, g_scalar, sizeof(double), hipMemcpyDeviceToHost, stream()));
  TTG_HIP(hipStreamSynchronize(stream()));
  *out = *g_scalar_host;
  return 0;
}

int ttg_dot(ttg_tensor x, ttg_tensor y, double* out) {
  PhaseTimer _phase("ttg_dot");
  TTG_REQUIRE(x && y, "null tensor");
  TTG_REQUIRE(same_structure(x, y), "dot: tensors have different block lists");
  if (!g_scalar) g_scalar = dalloc<double>(1);
  TTG_REQUIRE(g_scalar, "scalar allocation failed");
  TTG_TRY(dev_dot(x->d_data, y->d_data, x->m->total, g_scalar));
  return read_scalar(out);
}

int ttg_nrm2(ttg_tensor x, double* out) {
  PhaseTimer _phase("ttg_nrm2");
  TTG_REQUIRE(x, "null tensor");
  if (!g_scalar) g_scalar = dalloc<double>(1);
  TTG_REQUIRE(g_scalar, "scalar allocation failed");
  TTG_TRY(dev_dot(x->d_data, x->d_data, x->m->total, g_scalar));
  double s;
  TTG_TRY(read_scalar(&s));
  *out = std::sqrt(s);
  return 0;
}

int ttg_scale(ttg_tensor x, double a) {
  TTG_REQUIRE(x, "null tensor");
  hipLaunchKernelGGL(k_scale, dim3(grid_for(x->m->total)), dim3(256), 0, stream(), x->d_data,
                     x->m->total, a);
  mark_written(x);
  return 0;
}

int ttg_axpy(double a, ttg_tensor x, ttg_tensor y) {
  TTG_REQUIRE(x && y, "null tensor");
  TTG_REQUIRE(same_structure(x, y), "axpy: tensors have different block lists");
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(x->m->total)), dim3(256), 0, stream(), a, x->d_data,
                     y->d_data, x->m->total);
  mark_written(y);
  return 0;
}

int ttg_lincomb(int nx, const ttg_tensor* xs, const double* c, ttg_tensor y) {
  PhaseTimer _phase("ttg_lincomb");
  TTG_REQUIRE(nx >= 1 && nx <= 8, "lincomb: 1..8 operands");
  LinArgs a;
  a.n = nx;
  for (int k = 0; k < nx; k++) {
    TTG_REQUIRE(xs[k] && same_structure(xs[k], y), "lincomb: block lists differ");
    a.x[k] = xs[k]->d_data;
    a.c[k] = c[k];
  }
  hipLaunchKernelGGL(k_lincomb, dim3(grid_for(y->m->total)), dim3(256), 0, stream(), a, y->d_data,
                     y->m->total);
  mark_written(y);
  return 0;
}

}  // extern "C"
