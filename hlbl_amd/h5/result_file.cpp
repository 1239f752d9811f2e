// HDF5 result file of a run (include/hlbl_h5.h): the kernel_sum group, the chains
// table and the run attributes of the reference's ResultWriter (src/output.cpp:131-247,
// include/hlbl/output.hpp:45-86). Names, shapes, element types, the chunk policy and the
// attribute texts are those of that writer, so a reader of the reference's files reads
// these unchanged. Host code only; links HDF5 (1.10 C API).
#include "hlbl_h5.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

struct H5Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void ok(herr_t status, const char* what) {
  if (status < 0) throw H5Error(std::string("HDF5: ") + what + " failed");
}

// An HDF5 identifier and the call that releases it.
class Id {
 public:
  Id() = default;
  Id(hid_t id, herr_t (*close)(hid_t), const char* what) : id_(id), close_(close) {
    if (id < 0) throw H5Error(std::string("HDF5: ") + what + " failed");
  }
  Id(Id&& o) noexcept : id_(o.id_), close_(o.close_) { o.id_ = -1; }
  Id& operator=(Id&& o) noexcept {
    if (this != &o) {
      release();
      id_ = o.id_;
      close_ = o.close_;
      o.id_ = -1;
    }
    return *this;
  }
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  ~Id() { release(); }
  hid_t operator*() const { return id_; }
  void release() {
    if (id_ >= 0 && close_) close_(id_);
    id_ = -1;
  }

 private:
  hid_t id_ = -1;
  herr_t (*close_)(hid_t) = nullptr;
};

Id simple_space(const std::vector<hsize_t>& dims, const hsize_t* maxdims = nullptr) {
  return Id(H5Screate_simple((int)dims.size(), dims.data(), maxdims), H5Sclose, "create dataspace");
}

// A dataset of rows: shape [N][row...], N unlimited, appended to row by row. Chunks of
// 1 MiB of rows, at least 1 and at most 256 rows (the reference's chunk policy).
class RowSet {
 public:
  RowSet() = default;
  RowSet(hid_t loc, const char* name, hid_t type, std::vector<hsize_t> row)
      : type_(type), row_(std::move(row)) {
    hsize_t row_bytes = H5Tget_size(type);
    for (hsize_t d : row_) row_bytes *= d;
    const hsize_t chunk_rows = std::min<hsize_t>(256, std::max<hsize_t>(1, ((hsize_t)1 << 20) / row_bytes));
    std::vector<hsize_t> dims{0}, maxdims{H5S_UNLIMITED}, chunk{chunk_rows};
    for (hsize_t d : row_) {
      dims.push_back(d);
      maxdims.push_back(d);
      chunk.push_back(d);
    }
    Id sp = simple_space(dims, maxdims.data());
    Id dcpl(H5Pcreate(H5P_DATASET_CREATE), H5Pclose, "create dataset properties");
    ok(H5Pset_chunk(*dcpl, (int)chunk.size(), chunk.data()), "set chunk");
    ds_ = Id(H5Dcreate2(loc, name, type, *sp, H5P_DEFAULT, *dcpl, H5P_DEFAULT), H5Dclose, name);
  }
  // rows [first, first + n) from data (n rows of the row shape, row-major)
  void write(const void* data, hsize_t first, hsize_t n) {
    std::vector<hsize_t> extent{first + n}, start{first}, count{n};
    for (hsize_t d : row_) {
      extent.push_back(d);
      start.push_back(0);
      count.push_back(d);
    }
    ok(H5Dset_extent(*ds_, extent.data()), "extend dataset");
    Id fsp(H5Dget_space(*ds_), H5Sclose, "dataset space");
    ok(H5Sselect_hyperslab(*fsp, H5S_SELECT_SET, start.data(), nullptr, count.data(), nullptr),
       "select rows");
    Id msp = simple_space(count);
    ok(H5Dwrite(*ds_, type_, *msp, *fsp, H5P_DEFAULT, data), "write rows");
  }

 private:
  Id ds_;
  hid_t type_ = -1;
  std::vector<hsize_t> row_;
};

void attr_ints(hid_t loc, const char* name, const int* v, hsize_t n) {
  Id sp = simple_space({n});
  Id a(H5Acreate2(loc, name, H5T_NATIVE_INT, *sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, name);
  ok(H5Awrite(*a, H5T_NATIVE_INT, v), "write attribute");
}

void attr_double(hid_t loc, const char* name, double v) {
  Id sp(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
  Id a(H5Acreate2(loc, name, H5T_NATIVE_DOUBLE, *sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, name);
  ok(H5Awrite(*a, H5T_NATIVE_DOUBLE, &v), "write attribute");
}

// fixed-length, NUL-terminated string
void attr_text(hid_t loc, const char* name, const std::string& s) {
  Id type(H5Tcopy(H5T_C_S1), H5Tclose, "string type");
  ok(H5Tset_size(*type, s.size() + 1), "string size");
  Id sp(H5Screate(H5S_SCALAR), H5Sclose, "scalar space");
  Id a(H5Acreate2(loc, name, *type, *sp, H5P_DEFAULT, H5P_DEFAULT), H5Aclose, name);
  ok(H5Awrite(*a, *type, s.c_str()), "write attribute");
}

void coord4(const hlbl_coord& c, int* out) {
  out[0] = c.t;
  out[1] = c.x;
  out[2] = c.y;
  out[3] = c.z;
}

int as_int(double v, const char* what) {
  if (!(std::fabs(v) < 2147483648.0) || v != std::floor(v))
    throw H5Error(std::string("row field ") + what + " is not an integer");
  return (int)v;
}

// HDF5's automatic error printing off for the duration of one entry point (errors come
// back as status codes and a message here), the caller's handler restored afterwards
class QuietErrors {
 public:
  QuietErrors() {
    if (H5Eget_auto2(H5E_DEFAULT, &func_, &data_) >= 0) {
      saved_ = true;
      H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
  }
  ~QuietErrors() {
    if (saved_) H5Eset_auto2(H5E_DEFAULT, func_, data_);
  }
  QuietErrors(const QuietErrors&) = delete;
  QuietErrors& operator=(const QuietErrors&) = delete;

 private:
  H5E_auto2_t func_ = nullptr;
  void* data_ = nullptr;
  bool saved_ = false;
};

template <class F>
int guarded(F&& f) {
  QuietErrors quiet;
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
  } catch (...) {
    g_error = "unknown error";
  }
  return -1;
}

}  // namespace

struct hlbl_result_file {
  Id file, group, group_2p2;
  RowSet x, y, yv, chain, window, kernel_sum;
  RowSet w, chain_2p2, window_2p2, p1, pair_w, pair_y, pair_yv, p234;
  long long rows = 0, rows_2p2 = 0, pairs_2p2 = 0;
  int dims[4] = {0, 0, 0, 0};
  int lmax = 0;
  bool with_2p2 = false;
};

extern "C" {

hlbl_result_file* hlbl_result_open(const char* path, const hlbl_run_info* info) {
  hlbl_result_file* out = nullptr;
  guarded([&] {
    if (!path || !info) throw H5Error("path and run info are required");
    if (info->n_chains <= 0 || !info->chains) throw H5Error("no chains");
    auto f = std::make_unique<hlbl_result_file>();
    f->file = Id(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose, "create file");
    f->group = Id(H5Gcreate2(*f->file, "kernel_sum", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
                  "create group kernel_sum");
    const hid_t g = *f->group;
    f->x = RowSet(g, "x", H5T_NATIVE_INT, {4});
    f->y = RowSet(g, "y", H5T_NATIVE_INT, {4});
    f->yv = RowSet(g, "yv", H5T_NATIVE_INT, {4});
    f->chain = RowSet(g, "chain", H5T_NATIVE_INT, {});
    f->window = RowSet(g, "window", H5T_NATIVE_INT, {});
    f->kernel_sum = RowSet(g, "kernel_sum", H5T_NATIVE_DOUBLE, {2});
    coord4(info->lattice, f->dims);
    f->lmax = std::max(std::max(f->dims[0], f->dims[1]), std::max(f->dims[2], f->dims[3]));
    if (info->with_2p2) {
      f->with_2p2 = true;
      f->group_2p2 = Id(H5Gcreate2(*f->file, "2p2", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
                        "create group 2p2");
      const hid_t g2 = *f->group_2p2;
      const hsize_t lm = (hsize_t)f->lmax;
      f->w = RowSet(g2, "w", H5T_NATIVE_INT, {4});
      f->chain_2p2 = RowSet(g2, "chain", H5T_NATIVE_INT, {});
      f->window_2p2 = RowSet(g2, "window", H5T_NATIVE_INT, {});
      f->p1 = RowSet(g2, "P1", H5T_NATIVE_DOUBLE, {4, 4, 4, lm});
      f->pair_w = RowSet(g2, "pair_w", H5T_NATIVE_INT, {4});
      f->pair_y = RowSet(g2, "pair_y", H5T_NATIVE_INT, {4});
      f->pair_yv = RowSet(g2, "pair_yv", H5T_NATIVE_INT, {4});
      f->p234 = RowSet(g2, "P234", H5T_NATIVE_DOUBLE, {5, 4, 4, 4});
      attr_text(g2, "rows",
                "w, chain, window, P1: one row per 2+2 source, at its first solve; "
                "pair_w, pair_y, pair_yv, P234: one row per (source, target)");
      attr_text(g2, "P1_index",
                "P1[row][rho][sigma][nu][zeta]: Pi_{sigma nu}(z) summed over the sites z with "
                "(z - w)_rho mod L_rho = zeta; 0 for zeta >= L_rho");
      attr_text(g2, "P234_index",
                "P234[row][geometry][rho][sigma][nu], geometry P2_0, P2_1, P3, P4_0, P4_1; "
                "only rho < sigma is filled (P4_1: rho != sigma)");
      attr_text(g2, "flavour",
                "operator 0 only (iflavor 0 of compute_2p2_gpu, as cvc): the two-point function "
                "is the same for both flavours");
    }

    // chains[C][10]: start, direction, stride, steps
    const hsize_t nc = (hsize_t)info->n_chains;
    std::vector<int> table;
    for (hsize_t i = 0; i < nc; ++i) {
      const hlbl_chain_info& c = info->chains[i];
      int v[4];
      coord4(c.start, v);
      table.insert(table.end(), v, v + 4);
      coord4(c.direction, v);
      table.insert(table.end(), v, v + 4);
      table.push_back(c.stride);
      table.push_back(c.steps);
    }
    const hid_t file = *f->file;
    {
      Id sp = simple_space({nc, 10});
      Id ds(H5Dcreate2(file, "chains", H5T_NATIVE_INT, *sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
            H5Dclose, "create chains");
      ok(H5Dwrite(*ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, table.data()), "write chains");
    }
    attr_text(file, "chains_columns", "start0..3, dir0..3, stride, steps; row = chain index");
    int v[4];
    coord4(info->lattice, v);
    attr_ints(file, "lattice", v, 4);
    coord4(info->procs, v);
    attr_ints(file, "procs", v, 4);
    attr_ints(file, "nconf", &info->nconf, 1);
    attr_double(file, "muon_mass_MeV", info->muon_mass_mev);
    attr_double(file, "lattice_spacing_fm", info->lattice_spacing_fm);
    attr_double(file, "xunit", info->xunit);
    if (nc == 1) {  // the attributes of files written before chain lists
      coord4(info->chains[0].direction, v);
      attr_ints(file, "direction", v, 4);
      coord4(info->chains[0].start, v);
      attr_ints(file, "origin", v, 4);
      attr_ints(file, "stride", &info->chains[0].stride, 1);
    }
    attr_text(file, "kernel", info->kernel ? info->kernel : "LLambda0.4");
    attr_text(file, "flavour_order", "kernel_sum/kernel_sum[:,0] = operator 0, [:,1] = operator 1");
    attr_text(file, "coordinate_ranges",
              "x, y, w, pair_w, pair_y: [0, L); yv, pair_yv: minimum image of y - x, "
              "pair_y - pair_w in [-L/2, L/2)");
    ok(H5Fflush(file, H5F_SCOPE_GLOBAL), "flush");
    out = f.release();
  });
  return out;
}

int hlbl_result_append(hlbl_result_file* f, const double* rows, int n) {
  return guarded([&] {
    if (!f) throw H5Error("no result file");
    if (n < 0) throw H5Error("negative row count");
    if (n == 0) return;
    if (!rows) throw H5Error("rows is NULL");
    std::vector<int> x(4 * (size_t)n), y(4 * (size_t)n), yv(4 * (size_t)n), chain(n), window(n);
    std::vector<double> ks(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      const double* r = rows + 16 * (size_t)i;
      for (int k = 0; k < 4; ++k) {
        x[4 * i + k] = as_int(r[k], "x");
        y[4 * i + k] = as_int(r[4 + k], "y");
        yv[4 * i + k] = as_int(r[8 + k], "yv");
      }
      window[i] = as_int(r[12], "window");
      chain[i] = as_int(r[13], "chain");
      ks[2 * i] = r[14];
      ks[2 * i + 1] = r[15];
    }
    const hsize_t first = (hsize_t)f->rows, m = (hsize_t)n;
    f->x.write(x.data(), first, m);
    f->y.write(y.data(), first, m);
    f->yv.write(yv.data(), first, m);
    f->chain.write(chain.data(), first, m);
    f->window.write(window.data(), first, m);
    f->kernel_sum.write(ks.data(), first, m);
    f->rows += n;
    ok(H5Fflush(*f->file, H5F_SCOPE_GLOBAL), "flush");
  });
}

int hlbl_result_append_2p2(hlbl_result_file* f, const hlbl_source_2p2* rows, int n) {
  return guarded([&] {
    if (!f) throw H5Error("no result file");
    if (!f->with_2p2) throw H5Error("the file was opened without the 2p2 group (with_2p2)");
    if (n < 0) throw H5Error("negative row count");
    if (n == 0) return;
    if (!rows) throw H5Error("rows is NULL");
    const size_t p1_len = (size_t)64 * f->lmax;
    std::vector<int> w, chain, window, pw, py, pyv;
    std::vector<double> p1, p234;
    for (int i = 0; i < n; ++i) {
      const hlbl_source_2p2& r = rows[i];
      if (!r.p1 || r.n_targets < 0 || (r.n_targets > 0 && (!r.targets || !r.p234)))
        throw H5Error("a 2+2 row lacks its arrays");
      int wv[4];
      coord4(r.w, wv);
      w.insert(w.end(), wv, wv + 4);
      chain.push_back(r.chain);
      window.push_back(r.window);
      p1.insert(p1.end(), r.p1, r.p1 + p1_len);
      for (int t = 0; t < r.n_targets; ++t) {
        int yv[4];
        coord4(r.targets[t], yv);
        pw.insert(pw.end(), wv, wv + 4);
        py.insert(py.end(), yv, yv + 4);
        for (int k = 0; k < 4; ++k) {
          // minimum image of y - w in [-L/2, L/2)
          const int L = f->dims[k];
          int d = ((yv[k] - wv[k]) % L + L) % L;
          pyv.push_back(d >= L / 2 ? d - L : d);
        }
        p234.insert(p234.end(), r.p234 + (size_t)t * 320, r.p234 + (size_t)(t + 1) * 320);
      }
    }
    const hsize_t first = (hsize_t)f->rows_2p2, m = (hsize_t)n;
    f->w.write(w.data(), first, m);
    f->chain_2p2.write(chain.data(), first, m);
    f->window_2p2.write(window.data(), first, m);
    f->p1.write(p1.data(), first, m);
    const hsize_t npairs = (hsize_t)(pw.size() / 4);
    if (npairs > 0) {
      const hsize_t fp = (hsize_t)f->pairs_2p2;
      f->pair_w.write(pw.data(), fp, npairs);
      f->pair_y.write(py.data(), fp, npairs);
      f->pair_yv.write(pyv.data(), fp, npairs);
      f->p234.write(p234.data(), fp, npairs);
    }
    f->rows_2p2 += n;
    f->pairs_2p2 += (long long)npairs;
    ok(H5Fflush(*f->file, H5F_SCOPE_GLOBAL), "flush");
  });
}

int hlbl_result_append_rows(void* f, const double* rows, int n) {
  return hlbl_result_append(static_cast<hlbl_result_file*>(f), rows, n);
}

int hlbl_result_append_2p2_rows(void* f, const hlbl_source_2p2* rows, int n) {
  return hlbl_result_append_2p2(static_cast<hlbl_result_file*>(f), rows, n);
}

long long hlbl_result_rows(const hlbl_result_file* f) { return f ? f->rows : -1; }

int hlbl_result_close(hlbl_result_file* f) {
  return guarded([&] {
    if (!f) return;
    delete f;  // datasets, group, file closed in reverse order of opening
  });
}

const char* hlbl_h5_last_error(void) { return g_error.c_str(); }

}  // extern "C"
