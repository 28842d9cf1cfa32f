// checkpoint.cpp -- qTensorTrain::save / load (qtt.cpp:678-771 with
// qtensor::save / load, qtensor.cpp:1379-1415, 1477-1528): metadata in an
// HDF5 file, every block's raw data in one flat binary file at cumulative
// byte offsets.
//
// Layout (the reference's dataset names and element types; each vector is
// a 1-D dataset, each scalar a scalar dataset):
//   /wfn, /length, /phy_dim, /bond_dims, /center, /totalQ, /phy_qn, /id
//   /Tensor<i>/T                data file name (chars, NUL-terminated)
//   /Tensor<i>/offset_<l>       byte offset of block l in the data file
//   /Tensor<i>/num_blocks, rank, idx_arrows, idx_types, idx_levels
//   /Tensor<i>/idx_qn_<j>, idx_qdim_<j>, idx_name_<j>, block_<l>_qi
// A quantum number QN_t = (q0, q1) is stored as two native ints, so totalQ
// is an int[2], phy_qn and idx_qn_<j> are int[2n] (q0, q1 interleaved); the
// reference's ezh5 has no instantiation for its QN_t_2 type.  Blocks are
// written dense, column-major (first index fastest), as the device holds
// them, in block-list order; the offsets run on across the tensors of the
// train.  The data is moved through host memory one tensor at a time.
//
// Built into libttgpu_io.so with the HDF5 C library (the image's
// /opt/conda/lib/libhdf5, 1.10); libttgpu.so itself does not depend on it.
#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "tt_gpu.hpp"

namespace ttgpu {
namespace {

[[noreturn]] void io_fail(const std::string& what) { throw Error("checkpoint: " + what); }

// owning HDF5 identifiers
struct H5Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  H5Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  H5Id(const H5Id&) = delete;
  H5Id& operator=(const H5Id&) = delete;
  ~H5Id() {
    if (id >= 0 && close) close(id);
  }
};

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) fclose(f);
  }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

void write_raw(hid_t loc, const std::string& name, hid_t type, int ndim, hsize_t n, const void* v) {
  hsize_t dims[1] = {n};
  H5Id space(ndim == 0 ? H5Screate(H5S_SCALAR) : H5Screate_simple(1, dims, nullptr), H5Sclose);
  if (space.id < 0) io_fail("dataspace for " + name);
  H5Id ds(H5Dcreate2(loc, name.c_str(), type, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
          H5Dclose);
  if (ds.id < 0) io_fail("cannot create dataset " + name);
  if (n > 0 && H5Dwrite(ds.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0)
    io_fail("cannot write dataset " + name);
}
template <class T>
void put(hid_t loc, const std::string& name, hid_t type, const T& v) {
  write_raw(loc, name, type, 0, 1, &v);
}
template <class T>
void put_vec(hid_t loc, const std::string& name, hid_t type, const std::vector<T>& v) {
  write_raw(loc, name, type, 1, v.size(), v.data());
}

// element count of a dataset (1 for a scalar)
hsize_t extent(hid_t ds) {
  H5Id space(H5Dget_space(ds), H5Sclose);
  if (space.id < 0) io_fail("dataspace");
  const int nd = H5Sget_simple_extent_ndims(space.id);
  if (nd == 0) return 1;
  if (nd != 1) io_fail("dataset of rank " + std::to_string(nd));
  hsize_t d = 0;
  H5Sget_simple_extent_dims(space.id, &d, nullptr);
  return d;
}
template <class T>
std::vector<T> get_vec(hid_t loc, const std::string& name, hid_t type) {
  H5Id ds(H5Dopen2(loc, name.c_str(), H5P_DEFAULT), H5Dclose);
  if (ds.id < 0) io_fail("missing dataset " + name);
  std::vector<T> v(extent(ds.id));
  if (!v.empty() && H5Dread(ds.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) < 0)
    io_fail("cannot read dataset " + name);
  return v;
}
template <class T>
T get(hid_t loc, const std::string& name, hid_t type) {
  std::vector<T> v = get_vec<T>(loc, name, type);
  if (v.size() != 1) io_fail("dataset " + name + " is not a scalar");
  return v[0];
}

std::vector<char> cstr(const std::string& s) {
  std::vector<char> v(s.begin(), s.end());
  v.push_back(0);
  return v;
}

// qtensor::save(ezh5::Node&) (qtensor.cpp:1379-1411): indices and block list
void save_tensor_meta(hid_t g, const qtensor<double>& t,
                      const std::vector<std::vector<unsigned>>& blocks) {
  std::vector<unsigned> arrows, types, levels;
  for (const auto& ix : t.idx_set) {
    arrows.push_back((unsigned)ix.arrow());
    types.push_back((unsigned)ix.type());
    levels.push_back(ix.level());
  }
  put(g, "num_blocks", H5T_NATIVE_ULONG, (unsigned long)blocks.size());
  put(g, "rank", H5T_NATIVE_UINT, t.rank);
  put_vec(g, "idx_arrows", H5T_NATIVE_UINT, arrows);
  put_vec(g, "idx_types", H5T_NATIVE_UINT, types);
  put_vec(g, "idx_levels", H5T_NATIVE_UINT, levels);
  for (unsigned j = 0; j < t.rank; j++) {
    const auto& ix = t.idx_set[j];
    std::vector<int> qn, qd;
    for (unsigned s = 0; s < ix.size(); s++) {
      qn.push_back(ix.qn(s).q0);
      qn.push_back(ix.qn(s).q1);
      qd.push_back((int)ix.qdim(s));
    }
    put_vec(g, "idx_qn_" + std::to_string(j), H5T_NATIVE_INT, qn);
    put_vec(g, "idx_qdim_" + std::to_string(j), H5T_NATIVE_INT, qd);
  }
  for (unsigned j = 0; j < t.rank; j++) {
    const std::string& nm = t.idx_set[j].name();
    put_vec(g, "idx_name_" + std::to_string(j), H5T_NATIVE_CHAR, std::vector<char>(nm.begin(), nm.end()));
  }
  for (size_t l = 0; l < blocks.size(); l++)
    put_vec(g, "block_" + std::to_string(l) + "_qi", H5T_NATIVE_UINT, blocks[l]);
}

// qtensor::load(ezh5::Node&) (qtensor.cpp:1478-1528): indices and block list
qtensor<double> load_tensor_meta(hid_t g, std::vector<std::vector<unsigned>>& blocks) {
  const unsigned long nb = get<unsigned long>(g, "num_blocks", H5T_NATIVE_ULONG);
  const unsigned rank = get<unsigned>(g, "rank", H5T_NATIVE_UINT);
  auto arrows = get_vec<unsigned>(g, "idx_arrows", H5T_NATIVE_UINT);
  auto types = get_vec<unsigned>(g, "idx_types", H5T_NATIVE_UINT);
  auto levels = get_vec<unsigned>(g, "idx_levels", H5T_NATIVE_UINT);
  if (rank > TTG_MAX_RANK || arrows.size() != rank || types.size() != rank || levels.size() != rank)
    io_fail("inconsistent index tables");
  std::vector<qtensor_index> idx;
  for (unsigned j = 0; j < rank; j++) {
    auto nm = get_vec<char>(g, "idx_name_" + std::to_string(j), H5T_NATIVE_CHAR);
    if (arrows[j] > 1 || types[j] > 2) io_fail("bad index arrow or type");
    qtensor_index ix((arrow_type)arrows[j], std::string(nm.begin(), nm.end()), (index_type)types[j],
                     levels[j]);
    auto qn = get_vec<int>(g, "idx_qn_" + std::to_string(j), H5T_NATIVE_INT);
    auto qd = get_vec<int>(g, "idx_qdim_" + std::to_string(j), H5T_NATIVE_INT);
    if (qn.size() != 2 * qd.size()) io_fail("idx_qn / idx_qdim sizes differ");
    for (size_t s = 0; s < qd.size(); s++) {
      if (qd[s] <= 0) io_fail("non-positive sector dimension");
      ix.addQN(QN_t(qn[2 * s], qn[2 * s + 1]), (unsigned)qd[s]);
    }
    idx.push_back(ix);
  }
  blocks.assign(nb, {});
  for (unsigned long l = 0; l < nb; l++) {
    blocks[l] = get_vec<unsigned>(g, "block_" + std::to_string(l) + "_qi", H5T_NATIVE_UINT);
    if (blocks[l].size() != rank) io_fail("block tuple of the wrong length");
    for (unsigned j = 0; j < rank; j++)
      if (blocks[l][j] >= idx[j].size()) io_fail("block tuple outside the index's sectors");
  }
  return qtensor<double>(idx);
}

size_t block_elems(const qtensor<double>& t, const std::vector<unsigned>& qi) {
  size_t n = 1;
  for (unsigned j = 0; j < t.rank; j++) n *= t.idx_set[j].qdim(qi[j]);
  return n;
}

}  // namespace

// qtt.cpp:678-731
template <unsigned N>
void qTensorTrain<double, N>::save(std::string fn, std::string wfn) {
  if (!tensors_allocated || A.size() != length) fail("save: tensors not allocated");
  if (wfn.empty()) wfn = fn + "__T.bin";
  const std::vector<char> wfnC = cstr(wfn);
  H5Id f(H5Fcreate((fn + ".h5").c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) io_fail("cannot create " + fn + ".h5");
  FilePtr bin(fopen(wfn.c_str(), "wb"));
  if (!bin) io_fail("cannot create " + wfn);
  put_vec(f.id, "wfn", H5T_NATIVE_CHAR, wfnC);
  put(f.id, "length", H5T_NATIVE_UINT, length);
  put(f.id, "phy_dim", H5T_NATIVE_UINT, phy_dim);
  put_vec(f.id, "bond_dims", H5T_NATIVE_UINT, bond_dims);
  put(f.id, "center", H5T_NATIVE_INT, center);
  put_vec(f.id, "totalQ", H5T_NATIVE_INT, std::vector<int>{totalQ.q0, totalQ.q1});
  std::vector<int> pq;
  for (const auto& q : phy_qn) { pq.push_back(q.q0); pq.push_back(q.q1); }
  put_vec(f.id, "phy_qn", H5T_NATIVE_INT, pq);
  put(f.id, "id", H5T_NATIVE_UINT, _id);
  int64_t offset = 0;
  for (unsigned i = 0; i < length; i++) {
    const qtensor<double>& t = A[i];
    if (!t._initted || !t.dev) fail("save: tensor not initialised");
    H5Id g(H5Gcreate2(f.id, ("Tensor" + std::to_string(i)).c_str(), H5P_DEFAULT, H5P_DEFAULT,
                      H5P_DEFAULT),
           H5Gclose);
    if (g.id < 0) io_fail("cannot create group Tensor" + std::to_string(i));
    put_vec(g.id, "T", H5T_NATIVE_CHAR, wfnC);
    const auto blocks = t.block_list();
    for (size_t l = 0; l < blocks.size(); l++) {
      put(g.id, "offset_" + std::to_string(l), H5T_NATIVE_LONG, (long)offset);
      offset += (int64_t)(block_elems(t, blocks[l]) * sizeof(double));
    }
    save_tensor_meta(g.id, t, blocks);
    const std::vector<double> data = t.to_host();
    if (!data.empty() && fwrite(data.data(), sizeof(double), data.size(), bin.get()) != data.size())
      io_fail("short write to " + wfn);
  }
  if (fflush(bin.get()) != 0) io_fail("cannot flush " + wfn);
}

// qtt.cpp:738-771
template <unsigned N>
void qTensorTrain<double, N>::load(std::string fn, std::string dataPrefix) {
  H5Id f(H5Fopen(fn.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) io_fail("cannot open " + fn);
  const unsigned len = get<unsigned>(f.id, "length", H5T_NATIVE_UINT);
  const unsigned pd = get<unsigned>(f.id, "phy_dim", H5T_NATIVE_UINT);
  auto bd = get_vec<unsigned>(f.id, "bond_dims", H5T_NATIVE_UINT);
  const int ctr = get<int>(f.id, "center", H5T_NATIVE_INT);
  auto tq = get_vec<int>(f.id, "totalQ", H5T_NATIVE_INT);
  auto pq = get_vec<int>(f.id, "phy_qn", H5T_NATIVE_INT);
  const unsigned id = get<unsigned>(f.id, "id", H5T_NATIVE_UINT);
  auto wfnC = get_vec<char>(f.id, "wfn", H5T_NATIVE_CHAR);
  if (tq.size() != 2 || pq.size() % 2) io_fail("totalQ / phy_qn sizes");
  const std::string wfn = dataPrefix + std::string(wfnC.data(), strnlen(wfnC.data(), wfnC.size()));
  FilePtr bin(fopen(wfn.c_str(), "rb"));
  if (!bin) io_fail("bad data file name: " + wfn);
  std::vector<qtensor<double>> tens;
  for (unsigned i = 0; i < len; i++) {
    H5Id g(H5Gopen2(f.id, ("Tensor" + std::to_string(i)).c_str(), H5P_DEFAULT), H5Gclose);
    if (g.id < 0) io_fail("missing group Tensor" + std::to_string(i));
    std::vector<std::vector<unsigned>> blocks;
    qtensor<double> t = load_tensor_meta(g.id, blocks);
    size_t total = 0;
    for (const auto& qi : blocks) total += block_elems(t, qi);
    std::vector<double> data(total);
    size_t at = 0;
    for (size_t l = 0; l < blocks.size(); l++) {
      const long off = get<long>(g.id, "offset_" + std::to_string(l), H5T_NATIVE_LONG);
      const size_t n = block_elems(t, blocks[l]);
      if (off < 0 || fseek(bin.get(), off, SEEK_SET) != 0 ||
          fread(data.data() + at, sizeof(double), n, bin.get()) != n)
        io_fail("cannot read block " + std::to_string(l) + " of Tensor" + std::to_string(i) +
                " from " + wfn);
      at += n;
    }
    t.from_host(blocks, data);
    tens.push_back(std::move(t));
  }
  // commit only once everything has been read
  length = len;
  phy_dim = pd;
  bond_dims = bd;
  center = ctr;
  totalQ = QN_t(tq[0], tq[1]);
  phy_qn.clear();
  for (size_t k = 0; k + 1 < pq.size(); k += 2) phy_qn.emplace_back(pq[k], pq[k + 1]);
  _id = id;
  A = std::move(tens);
  tensors_allocated = true;
}

template void qTensorTrain<double, 1>::save(std::string, std::string);
template void qTensorTrain<double, 2>::save(std::string, std::string);
template void qTensorTrain<double, 1>::load(std::string, std::string);
template void qTensorTrain<double, 2>::load(std::string, std::string);

}  // namespace ttgpu

// C entry points for bindings (the host layer's tth_* convention: 0 = ok,
// -1 = failure with the message in tth_io_last_error())
namespace {
std::string g_io_err;
}

extern "C" {

const char* tth_io_last_error() { return g_io_err.c_str(); }

#define TTH_IO_GUARD(...)                  \
  try {                                    \
    g_io_err.clear();                      \
    __VA_ARGS__                            \
    return 0;                              \
  } catch (std::exception & e) {           \
    g_io_err = e.what();                   \
    return -1;                             \
  } catch (...) {                          \
    g_io_err = "unknown exception";        \
    return -1;                             \
  }

int tth_mps_save(void* q, const char* fn, const char* wfn) {
  TTH_IO_GUARD(((ttgpu::qMPS<double>*)q)->save(fn, wfn ? wfn : "");)
}
int tth_mps_load(void* q, const char* fn, const char* prefix) {
  TTH_IO_GUARD(((ttgpu::qMPS<double>*)q)->load(fn, prefix ? prefix : "");)
}
int tth_mpo_save(void* q, const char* fn, const char* wfn) {
  TTH_IO_GUARD(((ttgpu::qMPO<double>*)q)->save(fn, wfn ? wfn : "");)
}
int tth_mpo_load(void* q, const char* fn, const char* prefix) {
  TTH_IO_GUARD(((ttgpu::qMPO<double>*)q)->load(fn, prefix ? prefix : "");)
}

}  // extern "C"
