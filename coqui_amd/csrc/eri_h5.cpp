// Reader and writer of the THC ERI file (thc.eri.h5) that the reference's thc_reader_t
// writes after an ISDF build and reads back instead of rebuilding
// (src/methods/ERI/thc_reader_t.hpp:236-256 write, :507-574 read; the builder's datasets
// at src/methods/ERI/thc.cpp:361-377).  Host code over the HDF5 C library; the arrays land
// in caller-owned host buffers in the layouts of include/thcgw.h, ready for upload.
//
// On-disk layout (root group):
//   Np, number_of_bands                      int scalars
//   X_orbital_range, Y_orbital_range         int[2]  (first, last) orbital index
//   kpts (nk, 3), qpts (nq, 3)               double
//   nkpts_ibz, nqpts_ibz                     int scalars
//   collocation_matrix (ns*npol, nk, Np, nx) complex
//   Y_collocation_matrix (ns*npol, nk, Np, ny) complex, only when the ranges differ
//   interpolating_points (Np)                long
//   interpolating_vectors_G0, dual_interpolating_vectors_G0   complex (nq_ibz, Np)
//   coulomb_matrix (nq_ibz, Np, Np)          complex, including the N_k factor
// Complex arrays are stored as float64 arrays with a trailing dimension of 2 (re, im) and a
// string attribute __complex__ = "1", the convention of the nda HDF5 layer.
#include "../../include/thcgw_eri.h"

#include <hdf5.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct H5Err : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void fail(const std::string& m) { throw H5Err(m); }

// RAII handle of one HDF5 identifier type.
struct Hid {
  hid_t id;
  herr_t (*close)(hid_t);
  Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~Hid() {
    if (id >= 0) close(id);
  }
  Hid(const Hid&) = delete;
  Hid& operator=(const Hid&) = delete;
  operator hid_t() const { return id; }
};

bool has(hid_t f, const char* name) { return H5Lexists(f, name, H5P_DEFAULT) > 0; }

std::vector<hsize_t> dims_of(hid_t f, const char* name) {
  if (!has(f, name)) fail(std::string("dataset '") + name + "' is missing");
  Hid d(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
  if (d < 0) fail(std::string("cannot open dataset '") + name + "'");
  Hid sp(H5Dget_space(d), H5Sclose);
  const int rank = H5Sget_simple_extent_ndims(sp);
  if (rank < 0) fail(std::string("bad dataspace of '") + name + "'");
  std::vector<hsize_t> dims(static_cast<size_t>(rank));
  if (rank > 0) H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
  return dims;
}

// Reads the whole dataset into out, which holds exactly `count` values of mem_type.
template <class T>
void read_all(hid_t f, const char* name, hid_t mem_type, T* out, size_t count) {
  Hid d(H5Dopen2(f, name, H5P_DEFAULT), H5Dclose);
  if (d < 0) fail(std::string("cannot open dataset '") + name + "'");
  Hid sp(H5Dget_space(d), H5Sclose);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  if (n < 0 || static_cast<size_t>(n) != count)
    fail(std::string("dataset '") + name + "' has " + std::to_string(n) + " values, expected " + std::to_string(count));
  if (H5Dread(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) fail(std::string("cannot read '") + name + "'");
}

int read_int(hid_t f, const char* name) {
  int v = 0;
  read_all(f, name, H5T_NATIVE_INT, &v, 1);
  return v;
}

// Complex dataset: float dims (..., 2); returns the complex shape.
std::vector<hsize_t> complex_dims(hid_t f, const char* name) {
  std::vector<hsize_t> d = dims_of(f, name);
  if (d.empty() || d.back() != 2) fail(std::string("dataset '") + name + "' is not a complex array (..., 2)");
  d.pop_back();
  return d;
}

void write_ds(hid_t f, const char* name, hid_t file_type, hid_t mem_type, const std::vector<hsize_t>& dims,
              const void* data, bool complex_attr) {
  Hid sp(dims.empty() ? H5Screate(H5S_SCALAR) : H5Screate_simple(static_cast<int>(dims.size()), dims.data(), nullptr),
         H5Sclose);
  Hid d(H5Dcreate2(f, name, file_type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (d < 0) fail(std::string("cannot create dataset '") + name + "'");
  if (H5Dwrite(d, mem_type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) fail(std::string("cannot write '") + name + "'");
  if (complex_attr) {
    Hid st(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(st, 2);
    Hid asp(H5Screate(H5S_SCALAR), H5Sclose);
    Hid a(H5Acreate2(d, "__complex__", st, asp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    const char one[2] = {'1', '\0'};
    if (a < 0 || H5Awrite(a, st, one) < 0) fail(std::string("cannot tag '") + name + "' as complex");
  }
}

void write_complex(hid_t f, const char* name, std::vector<hsize_t> dims, const double* data) {
  dims.push_back(2);
  write_ds(f, name, H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, dims, data, true);
}

template <class F>
int guard(char* err, int errlen, F&& f) {
  H5E_auto2_t func = nullptr;
  void* cd = nullptr;
  H5Eget_auto2(H5E_DEFAULT, &func, &cd);
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // report through err, not stderr
  int rc = THCGW_ERI_OK;
  try {
    f();
  } catch (const std::exception& e) {
    if (err && errlen > 0) {
      std::strncpy(err, e.what(), static_cast<size_t>(errlen) - 1);
      err[errlen - 1] = '\0';
    }
    rc = THCGW_ERI_ERR;
  }
  H5Eset_auto2(H5E_DEFAULT, func, cd);
  return rc;
}

hid_t open_ro(const char* path) {
  if (!path) fail("null path");
  const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) fail(std::string("cannot open '") + path + "' as HDF5");
  return f;
}

}  // namespace

extern "C" {

int thcgw_eri_read_info(const char* path, thcgw_eri_info* info, char* err, int errlen) {
  return guard(err, errlen, [&] {
    if (!info) fail("null info");
    Hid f(open_ro(path), H5Fclose);
    thcgw_eri_info r{};
    r.Np = read_int(f, "Np");
    r.nbnd = has(f, "number_of_bands") ? read_int(f, "number_of_bands") : -1;
    int rng[2];
    read_all(f, "X_orbital_range", H5T_NATIVE_INT, rng, 2);
    r.x_first = rng[0];
    r.x_last = rng[1];
    read_all(f, "Y_orbital_range", H5T_NATIVE_INT, rng, 2);
    r.y_first = rng[0];
    r.y_last = rng[1];
    // thc_reader_t.hpp:519-524: ranges within [0, nbnd]
    if (r.x_first < 0 || r.x_last < r.x_first || r.y_first < 0 || r.y_last < r.y_first ||
        (r.nbnd >= 0 && (r.x_last > r.nbnd || r.y_last > r.nbnd)))
      fail("invalid X_orbital_range / Y_orbital_range");
    const std::vector<hsize_t> X = complex_dims(f, "collocation_matrix");
    if (X.size() != 4) fail("collocation_matrix must be (ns*npol, nk, Np, nx)");
    r.ns_npol = static_cast<int>(X[0]);
    r.nkpts = static_cast<int>(X[1]);
    if (static_cast<int>(X[2]) != r.Np) fail("collocation_matrix: dimension 2 differs from Np");
    if (static_cast<int>(X[3]) != r.x_last - r.x_first) fail("collocation_matrix: dimension 3 differs from the X range");
    r.has_Y = has(f, "Y_collocation_matrix") ? 1 : 0;
    if ((r.x_first != r.y_first || r.x_last != r.y_last) && !r.has_Y)
      fail("X and Y orbital ranges differ but Y_collocation_matrix is missing");
    if (r.has_Y) {
      const std::vector<hsize_t> Y = complex_dims(f, "Y_collocation_matrix");
      if (Y.size() != 4 || static_cast<int>(Y[0]) != r.ns_npol || static_cast<int>(Y[1]) != r.nkpts ||
          static_cast<int>(Y[2]) != r.Np || static_cast<int>(Y[3]) != r.y_last - r.y_first)
        fail("Y_collocation_matrix must be (ns*npol, nk, Np, y_last - y_first) like collocation_matrix");
    }
    r.nkpts_file = 0;
    if (has(f, "kpts")) {
      const std::vector<hsize_t> k = dims_of(f, "kpts");
      if (k.size() != 2 || k[1] != 3 || static_cast<int>(k[0]) != r.nkpts)
        fail("kpts must be (nk, 3) with the nk of collocation_matrix");
      r.nkpts_file = r.nkpts;
    }
    r.nqpts = 0;
    if (has(f, "qpts")) {
      const std::vector<hsize_t> q = dims_of(f, "qpts");
      if (q.size() != 2 || q[1] != 3) fail("qpts must be (nq, 3)");
      r.nqpts = static_cast<int>(q[0]);
    }
    const std::vector<hsize_t> V = complex_dims(f, "coulomb_matrix");
    if (V.size() != 3 || static_cast<int>(V[1]) != r.Np || static_cast<int>(V[2]) != r.Np)
      fail("coulomb_matrix must be (nq_ibz, Np, Np)");
    r.nqpts_ibz = static_cast<int>(V[0]);
    if (has(f, "nqpts_ibz") && read_int(f, "nqpts_ibz") != r.nqpts_ibz)
      fail("nqpts_ibz differs from the first dimension of coulomb_matrix");
    r.nkpts_ibz = has(f, "nkpts_ibz") ? read_int(f, "nkpts_ibz") : r.nkpts;
    if (has(f, "interpolating_points")) {
      const std::vector<hsize_t> rp = dims_of(f, "interpolating_points");
      // thc_reader_t.hpp:529-530
      if (rp.size() != 1 || static_cast<int>(rp[0]) != r.Np) fail("interpolating_points: shape differs from Np");
    }
    *info = r;
  });
}

// Values (doubles) of the dataset `name`; fails unless it equals the caller's capacity.
size_t sized_for(hid_t f, const char* name, long capacity, bool complex) {
  const std::vector<hsize_t> d = complex ? complex_dims(f, name) : dims_of(f, name);
  size_t n = complex ? 2 : 1;
  for (hsize_t v : d) n *= static_cast<size_t>(v);
  if (capacity < 0 || static_cast<size_t>(capacity) != n)
    fail(std::string("dataset '") + name + "' holds " + std::to_string(n) + " doubles, the output buffer " +
         std::to_string(capacity));
  return n;
}

int thcgw_eri_read_X(const char* path, int which, double* X, long capacity, char* err, int errlen) {
  return guard(err, errlen, [&] {
    if (!X) fail("null output");
    Hid f(open_ro(path), H5Fclose);
    const char* name = which ? "Y_collocation_matrix" : "collocation_matrix";
    read_all(f, name, H5T_NATIVE_DOUBLE, X, sized_for(f, name, capacity, true));
  });
}

int thcgw_eri_read_V(const char* path, double* V, long capacity, char* err, int errlen) {
  return guard(err, errlen, [&] {
    if (!V) fail("null output");
    Hid f(open_ro(path), H5Fclose);
    read_all(f, "coulomb_matrix", H5T_NATIVE_DOUBLE, V, sized_for(f, "coulomb_matrix", capacity, true));
  });
}

int thcgw_eri_read_kpts(const char* path, double* kpts, long kpts_capacity, double* qpts, long qpts_capacity,
                        char* err, int errlen) {
  return guard(err, errlen, [&] {
    Hid f(open_ro(path), H5Fclose);
    for (int w = 0; w < 2; ++w) {
      double* out = w ? qpts : kpts;
      const char* name = w ? "qpts" : "kpts";
      if (!out) continue;
      const std::vector<hsize_t> d = dims_of(f, name);
      if (d.size() != 2 || d[1] != 3) fail(std::string(name) + " must be (n, 3)");
      read_all(f, name, H5T_NATIVE_DOUBLE, out, sized_for(f, name, w ? qpts_capacity : kpts_capacity, false));
    }
  });
}

int thcgw_eri_write(const char* path, const thcgw_eri_info* info, const double* X, const double* Y,
                    const double* V, const double* kpts, const double* qpts, const long* interp_points,
                    const double* chi_head, const double* chi_bar_head, char* err, int errlen) {
  return guard(err, errlen, [&] {
    if (!path || !info || !X || !V) fail("null argument");
    const thcgw_eri_info& r = *info;
    const int nx = r.x_last - r.x_first, ny = r.y_last - r.y_first;
    if (r.Np <= 0 || nx <= 0 || ny <= 0 || r.ns_npol <= 0 || r.nkpts <= 0 || r.nqpts_ibz <= 0)
      fail("thcgw_eri_write: sizes must be positive");
    if ((r.x_first != r.y_first || r.x_last != r.y_last) && !Y) fail("thcgw_eri_write: Y is required when the ranges differ");
    Hid f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (f < 0) fail(std::string("cannot create '") + path + "'");
    const int nbnd = r.nbnd > 0 ? r.nbnd : std::max(r.x_last, r.y_last);
    write_ds(f, "Np", H5T_STD_I32LE, H5T_NATIVE_INT, {}, &r.Np, false);
    write_ds(f, "number_of_bands", H5T_STD_I32LE, H5T_NATIVE_INT, {}, &nbnd, false);
    const int xr[2] = {r.x_first, r.x_last}, yr[2] = {r.y_first, r.y_last};
    write_ds(f, "X_orbital_range", H5T_STD_I32LE, H5T_NATIVE_INT, {2}, xr, false);
    write_ds(f, "Y_orbital_range", H5T_STD_I32LE, H5T_NATIVE_INT, {2}, yr, false);
    if (kpts) write_ds(f, "kpts", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, {hsize_t(r.nkpts), 3}, kpts, false);
    if (qpts)
      write_ds(f, "qpts", H5T_IEEE_F64LE, H5T_NATIVE_DOUBLE, {hsize_t(r.nqpts > 0 ? r.nqpts : r.nkpts), 3}, qpts, false);
    const int nki = r.nkpts_ibz > 0 ? r.nkpts_ibz : r.nkpts;
    write_ds(f, "nkpts_ibz", H5T_STD_I32LE, H5T_NATIVE_INT, {}, &nki, false);
    write_ds(f, "nqpts_ibz", H5T_STD_I32LE, H5T_NATIVE_INT, {}, &r.nqpts_ibz, false);
    write_complex(f, "collocation_matrix", {hsize_t(r.ns_npol), hsize_t(r.nkpts), hsize_t(r.Np), hsize_t(nx)}, X);
    if (r.x_first != r.y_first || r.x_last != r.y_last)
      write_complex(f, "Y_collocation_matrix", {hsize_t(r.ns_npol), hsize_t(r.nkpts), hsize_t(r.Np), hsize_t(ny)}, Y);
    if (interp_points) write_ds(f, "interpolating_points", H5T_STD_I64LE, H5T_NATIVE_LONG, {hsize_t(r.Np)}, interp_points, false);
    if (chi_head) write_complex(f, "interpolating_vectors_G0", {hsize_t(r.nqpts_ibz), hsize_t(r.Np)}, chi_head);
    if (chi_bar_head) write_complex(f, "dual_interpolating_vectors_G0", {hsize_t(r.nqpts_ibz), hsize_t(r.Np)}, chi_bar_head);
    write_complex(f, "coulomb_matrix", {hsize_t(r.nqpts_ibz), hsize_t(r.Np), hsize_t(r.Np)}, V);
  });
}

}  // extern "C"
