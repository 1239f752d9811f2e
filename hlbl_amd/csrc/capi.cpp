// The C ABI: argument checks, status codes, and the host orchestration of
// the per-pair contraction (Contractor::kernel_sum, src/contract.cpp:104-114).
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hlbl_amd.h"
#include "context.hpp"
#include "contract_kernels.hpp"
#include "provider.hpp"
#include "qed_exact.hpp"
#include "qed_host.hpp"
#include "driver.hpp"
#include "schedule.hpp"

namespace hlbl {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

hipStream_t as_stream(hlbl_stream s) { return reinterpret_cast<hipStream_t>(s); }

void check_geom(const hlbl_geom& g, const char* what) {
  if (g.T <= 0 || g.LX <= 0 || g.LY <= 0 || g.LZ <= 0)
    throw Error(std::string(what) + ": extents must be positive");
}

void check_layout(const hlbl_geom& global, const hlbl_geom& local, const hlbl_coord& proc) {
  check_geom(global, "global geometry");
  check_geom(local, "local geometry");
  for (int r = 0; r < 4; ++r) {
    const int Lg = geom_dim(global, r), Ll = geom_dim(local, r), p = coord_dim(proc, r);
    if (Lg % Ll != 0) throw Error("local extents must divide the global extents");
    if (p < 0 || p * Ll >= Lg) throw Error("process coordinates out of range");
    if (Lg % 2 != 0) throw Error("global extents must be even");
  }
}

void check_ptr(const void* p, const char* what) {
  if (!p) throw Error(std::string(what) + " is null");
}

// m_mu a in muon units: the kernels scale site coordinates by it (and size the
// near-origin FP64 region by its inverse)
void check_xunit(double xunit) {
  if (!(xunit > 0) || !std::isfinite(xunit)) throw Error("xunit must be positive and finite");
}

int device_cu_count(int device) {
  int n = 0;
  check_hip(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device),
            "hipDeviceGetAttribute");
  return n;
}

}  // namespace

}  // namespace hlbl

using namespace hlbl;

namespace hlbl {
namespace {

// The committed table next to the library: <dir of libhlbl_amd.so>/data/qed_kernel_table.bin,
// unless HLBL_QED_TABLE names another file.
std::string default_table_path() {
  if (const char* env = std::getenv("HLBL_QED_TABLE")) return env;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&hlbl_qed_init), &info) == 0 || !info.dli_fname)
    throw Error("hlbl_qed_init: cannot locate the library's directory");
  std::string dir = info.dli_fname;
  const size_t slash = dir.find_last_of('/');
  dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
  return dir + "/data/qed_kernel_table.bin";
}

}  // namespace

void window_drop_graphs(hlbl_contractor* c);
}  // namespace hlbl

extern "C" {

const char* hlbl_last_error(void) { return g_last_error.c_str(); }

double hlbl_xunit(double muon_mass_mev, double lattice_spacing_fm) {
  return muon_mass_mev / qed::kHbarC * lattice_spacing_fm;
}

// ---- provider ------------------------------------------------------------------

uint64_t hlbl_source_seed(uint64_t seed, hlbl_coord gsx) {
  const int c[4] = {gsx.t, gsx.x, gsx.y, gsx.z};
  for (int i = 0; i < 4; ++i) seed = hash64(seed ^ (uint64_t)(int64_t)(c[i] + 1));
  return seed;
}

size_t hlbl_block_length(hlbl_geom local) { return (size_t)2 * 12 * 24 * geom_volume(local); }

int hlbl_fill_random_block_host(double* out, hlbl_geom local, hlbl_coord proc, uint64_t seed,
                                int n_threads) {
  return guarded([&] {
    check_ptr(out, "out");
    check_geom(local, "local geometry");
    const size_t V = geom_volume(local);
    int nt = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    auto work = [&](size_t s0, size_t s1) {
      for (size_t site = s0; site < s1; ++site) {
        int c[4];
        site_coords(site, local, c);
        const uint64_t g = site_key(proc.t * local.T + c[0], proc.x * local.LX + c[1],
                                    proc.y * local.LY + c[2], proc.z * local.LZ + c[3]);
        for (int f = 0; f < 2; ++f)
          for (int a = 0; a < 12; ++a) {
            double* dst = out + (((size_t)(f * 12 + a)) * V + site) * 24;
            for (int cc = 0; cc < 24; ++cc) dst[cc] = hash_to_unit(element_key(g, f, a, cc, seed));
          }
      }
    };
    std::vector<std::thread> pool;
    const size_t chunk = (V + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
      const size_t s0 = t * chunk, s1 = std::min(V, s0 + chunk);
      if (s0 < s1) pool.emplace_back(work, s0, s1);
    }
    for (auto& th : pool) th.join();
  });
}

int hlbl_fill_random_block_device(double* d_out, hlbl_geom local, hlbl_coord proc, uint64_t seed,
                                  hlbl_stream stream) {
  return guarded([&] {
    check_ptr(d_out, "d_out");
    check_geom(local, "local geometry");
    launch_fill_random(d_out, local, proc, seed, as_stream(stream));
  });
}

// ---- QED kernel -------------------------------------------------------------------

int hlbl_qed_init_file(hlbl_qed** out, int device, const char* table_path) {
  return guarded([&] {
    check_ptr(out, "out");
    check_ptr(table_path, "table_path");
    *out = nullptr;
    check_hip(hipSetDevice(device), "hipSetDevice");
    auto q = new hlbl_qed();
    try {
      q->device = device;
      qed::TermTable tt;
      qed::build_term_table(&tt);
      upload_terms_fp64(tt);
      upload_terms_fp32(tt);
      upload_terms_qed(tt);
      try {
        qed::load_table(table_path, q->host_table);
      } catch (const std::runtime_error& e) {
        throw Error(e.what());
      }
      const size_t n = q->host_table.size();
      q->table64.ensure(n);
      check_hip(hipMemcpy(q->table64.p, q->host_table.data(), n * sizeof(double),
                          hipMemcpyHostToDevice), "hipMemcpy table");
      std::vector<float> t32(q->host_table.begin(), q->host_table.end());
      q->table32.ensure(n);
      check_hip(hipMemcpy(q->table32.p, t32.data(), n * sizeof(float), hipMemcpyHostToDevice),
                "hipMemcpy table");
      q->noalloc_scratch = noalloc_scratch_setup(device);  // with `device` current
    } catch (...) {
      delete q;
      throw;
    }
    *out = q;
  });
}

int hlbl_qed_init(hlbl_qed** out, int device) {
  std::string path;
  const int rc = guarded([&] { path = hlbl::default_table_path(); });
  if (rc != 0) return rc;
  return hlbl_qed_init_file(out, device, path.c_str());
}

int hlbl_qed_exact(double a, double beta, double b, const int* quad, double* out) {
  return guarded([&] {
    check_ptr(out, "out");
    if (!(a >= 0 && b >= 0 && beta >= 0 && beta <= M_PI))
      throw Error("hlbl_qed_exact: need |x|, |y| >= 0 and beta in [0, pi]");
    qed_exact::Config cfg;
    cfg.nt = 16, cfg.nr = 28, cfg.nth = 8, cfg.nps = 16;
    if (quad) {
      if (quad[0] < 4 || quad[0] > 64 || quad[1] < 4 || quad[2] < 2 || quad[3] < 4)
        throw Error("hlbl_qed_exact: quadrature sizes out of range");
      cfg.nt = quad[0], cfg.nr = quad[1], cfg.nth = quad[2], cfg.nps = quad[3];
    }
    const qed_exact::Tframe f = qed_exact::Evaluator(cfg).eval(a, b, beta);
    for (int A = 0; A < 3; ++A)
      for (int c = 0; c < qed::kFrame; ++c) out[A * qed::kFrame + c] = f.c[A][c];
  });
}

int hlbl_qed_free(hlbl_qed* q) {
  return guarded([&] { delete q; });
}

int hlbl_qed_table_info(const hlbl_qed* q, int* n_channels, int* n_a, int* n_c, int* n_b,
                        double* radius) {
  return guarded([&] {
    check_ptr(q, "qed");
    if (n_channels) *n_channels = qed::kNCh;
    if (n_a) *n_a = qed::kNA;
    if (n_c) *n_c = qed::kNC;
    if (n_b) *n_b = qed::kNB;
    if (radius) *radius = qed::kRadius;
  });
}

int hlbl_qed_table_copy(const hlbl_qed* q, double* host_out, size_t n_doubles) {
  return guarded([&] {
    check_ptr(q, "qed");
    check_ptr(host_out, "host_out");
    if (n_doubles < q->host_table.size()) throw Error("hlbl_qed_table_copy: output too small");
    std::memcpy(host_out, q->host_table.data(), q->host_table.size() * sizeof(double));
  });
}

int hlbl_qed_kernel(hlbl_qed* q, int mode, const double* d_x, const double* d_y, int n,
                    double* d_kerv, hlbl_stream stream) {
  return guarded([&] {
    check_ptr(q, "qed");
    if (n < 0) throw Error("n < 0");
    if (n > 0) {
      check_ptr(d_x, "d_x");
      check_ptr(d_y, "d_y");
      check_ptr(d_kerv, "d_kerv");
    }
    launch_qed_kernel(q, mode, d_x, d_y, n, d_kerv, as_stream(stream));
  });
}

int hlbl_qed_channels(hlbl_qed* q, int mode, const double* d_acb, int n, double* d_out,
                      hlbl_stream stream) {
  return guarded([&] {
    check_ptr(q, "qed");
    if (n < 0) throw Error("n < 0");
    if (n > 0) {
      check_ptr(d_acb, "d_acb");
      check_ptr(d_out, "d_out");
    }
    launch_qed_channels(q, mode, d_acb, n, d_out, as_stream(stream));
  });
}

// ---- contraction ------------------------------------------------------------------

int hlbl_contractor_create(hlbl_contractor** out, hlbl_geom global, hlbl_geom local,
                           hlbl_coord proc_coords, hlbl_qed* qed) {
  return guarded([&] {
    check_ptr(out, "out");
    check_ptr(qed, "qed");
    *out = nullptr;
    check_layout(global, local, proc_coords);
    auto c = new hlbl_contractor();
    try {
      c->global = global;
      c->local = local;
      c->proc = proc_coords;
      c->qed = qed;
      c->volume = geom_volume(local);
      check_hip(hipSetDevice(qed->device), "hipSetDevice");
      c->n_cu = device_cu_count(qed->device);
      c->nwg2 = c->tune.fp64_pass2_wg_per_cu * c->n_cu;
      c->result.ensure(2);
      c->host_result.ensure(2);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  });
}

int hlbl_contractor_free(hlbl_contractor* c) {
  return guarded([&] { delete c; });
}

int hlbl_tuning_defaults(hlbl_tuning* out) {
  return guarded([&] {
    check_ptr(out, "out");
    *out = default_tuning();
  });
}

int hlbl_contractor_set_tuning(hlbl_contractor* c, const hlbl_tuning* t) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(t, "tuning");
    auto in = [](int v, int lo, int hi, const char* what) {
      if (v < lo || v > hi) throw Error(std::string("hlbl_contractor_set_tuning: ") + what + " out of range");
    };
    in(t->pass1_batch, 1, 1 << 20, "pass1_batch");
    if (t->field_waves != 4 && t->field_waves != 8) throw Error("hlbl_contractor_set_tuning: field_waves must be 4 or 8");
    in(t->field_wg_per_cu, 1, t->field_waves == 8 ? 2 : 3, "field_wg_per_cu");
    in(t->axes_wg_per_cu, 1, 16, "axes_wg_per_cu");
    in(t->near_wg_per_cu, 1, 8, "near_wg_per_cu");
    in(t->pass2_wg_per_cu, 1, 4, "pass2_wg_per_cu");
    in(t->fp64_field_wg_per_cu, 1, 16, "fp64_field_wg_per_cu");
    in(t->fp64_pass2_wg_per_cu, 1, 8, "fp64_pass2_wg_per_cu");
    c->tune = *t;
    c->nwg2 = t->fp64_pass2_wg_per_cu * c->n_cu;
    window_drop_graphs(c);
  });
}

int hlbl_contractor_get_tuning(const hlbl_contractor* c, hlbl_tuning* out) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(out, "out");
    *out = c->tune;
  });
}

int hlbl_kernel_sum(hlbl_contractor* c, const double* d_fwd_src, const double* d_fwd_y,
                    hlbl_coord gsx, hlbl_coord yv, double xunit, double out[2],
                    hlbl_stream stream) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(d_fwd_src, "d_fwd_src");
    check_ptr(d_fwd_y, "d_fwd_y");
    check_ptr(out, "out");
    check_xunit(xunit);
    hipStream_t st = as_stream(stream);
    launch_pass1(c, d_fwd_src, d_fwd_y, c->proc, gsx, c->global, c->local, 3, st);
    c->gdz.ensure(2 * 40 * 144 * 2);
    c->bmat.ensure(2 * 288 * 48);
    launch_gamma(c->dz.p, c->gdz.p, st);
    launch_bmat(c->gdz.p, c->bmat.p, st);
    launch_pass2(c, d_fwd_src, d_fwd_y, c->bmat.p, c->proc, gsx, yv, xunit, c->global, c->local,
                 3, c->result.p, 0, 0, st);
    check_hip(hipMemcpyAsync(c->host_result.p, c->result.p, 2 * sizeof(double),
                             hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    check_hip(hipStreamSynchronize(st), "kernel_sum");
    out[0] = c->host_result.p[0];
    out[1] = c->host_result.p[1];
  });
}

int hlbl_dzu_dzsu(hlbl_contractor* c, double* d_dzu, double* d_dzsu, const double* fwd_src,
                  const double* fwd_y, int iflavor, hlbl_coord proc_coords, hlbl_coord gsx,
                  hlbl_geom global, hlbl_geom local, hlbl_stream stream) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(d_dzu, "d_dzu");
    check_ptr(d_dzsu, "d_dzsu");
    check_ptr(fwd_src, "fwd_src");
    check_ptr(fwd_y, "fwd_y");
    if (iflavor != 0 && iflavor != 1) throw Error("iflavor must be 0 or 1");
    check_layout(global, local, proc_coords);
    hipStream_t st = as_stream(stream);
    launch_pass1(c, fwd_src, fwd_y, proc_coords, gsx, global, local, 1 << iflavor, st);
    // add this flavour's dzu / dzsu to the caller's buffers
    const double* dzf = c->dz.p + (size_t)iflavor * 10 * 144 * 2;
    launch_add(d_dzu, dzf, 6 * 144 * 2, st);
    launch_add(d_dzsu, dzf + 6 * 144 * 2, 4 * 144 * 2, st);
    check_hip(hipStreamSynchronize(st), "dzu_dzsu");
  });
}

int hlbl_pieces_2p2_p1(hlbl_contractor* c, const double* d_fwd, int iflavor, hlbl_coord gsw,
                       double* d_p1, hlbl_stream stream) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(d_fwd, "d_fwd");
    check_ptr(d_p1, "d_p1");
    if (iflavor != 0 && iflavor != 1) throw Error("iflavor must be 0 or 1");
    const hlbl_geom& g = c->global;
    if (gsw.t < 0 || gsw.t >= g.T || gsw.x < 0 || gsw.x >= g.LX || gsw.y < 0 || gsw.y >= g.LY ||
        gsw.z < 0 || gsw.z >= g.LZ)
      throw Error("gsw must be a global site in [0, L)");
    hipStream_t st = as_stream(stream);
    launch_p1_2p2(c, d_fwd, iflavor, gsw, d_p1, st);
    check_hip(hipStreamSynchronize(st), "pieces_2p2_p1");
  });
}

int hlbl_pieces_2p2(hlbl_contractor* c, const double* d_fwd, int iflavor, hlbl_coord gsw,
                    const hlbl_coord* targets, int n_targets, double xunit, double* d_p1,
                    double* d_p234, hlbl_stream stream) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(d_fwd, "d_fwd");
    check_ptr(d_p1, "d_p1");
    if (iflavor != 0 && iflavor != 1) throw Error("iflavor must be 0 or 1");
    if (n_targets < 0) throw Error("n_targets must be non-negative");
    if (n_targets > 0) {
      check_ptr(targets, "targets");
      check_ptr(d_p234, "d_p234");
    }
    if (!(xunit > 0)) throw Error("xunit must be positive");
    const hlbl_geom& g = c->global;
    auto inside = [&](const hlbl_coord& p) {
      return p.t >= 0 && p.t < g.T && p.x >= 0 && p.x < g.LX && p.y >= 0 && p.y < g.LY && p.z >= 0 &&
             p.z < g.LZ;
    };
    if (!inside(gsw)) throw Error("gsw must be a global site in [0, L)");
    for (int t = 0; t < n_targets; ++t)
      if (!inside(targets[t])) throw Error("targets must be global sites in [0, L)");
    hipStream_t st = as_stream(stream);
    launch_p1_2p2(c, d_fwd, iflavor, gsw, d_p1, st);
    if (n_targets > 0) launch_p234_2p2(c, c->pi2p2.p, gsw, targets, n_targets, xunit, d_p234, st);
    check_hip(hipStreamSynchronize(st), "pieces_2p2");
  });
}

int hlbl_insert_g5_gamma(const double* d_dzu, const double* d_dzsu, double* d_g_dzu,
                         double* d_g_dzsu, hlbl_stream stream) {
  return guarded([&] {
    check_ptr(d_dzu, "d_dzu");
    check_ptr(d_dzsu, "d_dzsu");
    check_ptr(d_g_dzu, "d_g_dzu");
    check_ptr(d_g_dzsu, "d_g_dzsu");
    hipStream_t st = as_stream(stream);
    launch_gamma_one(d_dzu, d_dzsu, d_g_dzu, d_g_dzsu, st);
    check_hip(hipStreamSynchronize(st), "insert_g5_gamma");
  });
}

int hlbl_4pt_contraction(hlbl_contractor* c, double* d_kernel_sum, const double* d_g_dzu,
                         const double* d_g_dzsu, const double* fwd_src, const double* fwd_y,
                         int iflavor, hlbl_coord proc_coords, hlbl_coord gsx, hlbl_pair xunit,
                         hlbl_coord yv, hlbl_geom global, hlbl_geom local,
                         hlbl_stream stream) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(d_kernel_sum, "d_kernel_sum");
    check_ptr(d_g_dzu, "d_g_dzu");
    check_ptr(d_g_dzsu, "d_g_dzsu");
    check_ptr(fwd_src, "fwd_src");
    check_ptr(fwd_y, "fwd_y");
    if (iflavor != 0 && iflavor != 1) throw Error("iflavor must be 0 or 1");
    check_layout(global, local, proc_coords);
    check_xunit(xunit.a);
    hipStream_t st = as_stream(stream);
    c->gdz.ensure(2 * 40 * 144 * 2);
    c->bmat.ensure(2 * 288 * 48);
    double* gf = c->gdz.p + (size_t)iflavor * 40 * 144 * 2;
    check_hip(hipMemcpyAsync(gf, d_g_dzu, 6 * 4 * 144 * 2 * sizeof(double), hipMemcpyDeviceToDevice, st), "copy");
    check_hip(hipMemcpyAsync(gf + 6 * 4 * 144 * 2, d_g_dzsu, 4 * 4 * 144 * 2 * sizeof(double), hipMemcpyDeviceToDevice, st), "copy");
    launch_bmat(c->gdz.p, c->bmat.p, st);
    launch_pass2(c, fwd_src, fwd_y, c->bmat.p, proc_coords, gsx, yv, xunit.a, global, local,
                 1 << iflavor, d_kernel_sum, 1, 1, st);
    check_hip(hipStreamSynchronize(st), "4pt_contraction");
  });
}

// ---- schedule -----------------------------------------------------------------------

}  // extern "C"

struct hlbl_schedule {
  hlbl::Schedule steps;
  // flattened sources in run order
  std::vector<std::pair<int, const hlbl::Source*>> flat;
};

namespace {
hlbl::Vec4 v4(const hlbl_coord& c) { return {c.t, c.x, c.y, c.z}; }
hlbl_coord c4(const hlbl::Vec4& v) { return {v[0], v[1], v[2], v[3]}; }
}  // namespace

extern "C" {

hlbl_schedule* hlbl_schedule_build(hlbl_coord h, int stride, int n_steps, hlbl_coord origin) {
  hlbl_schedule* s = nullptr;
  int rc = guarded([&] {
    for (int v : {h.t, h.x, h.y, h.z})
      if (v != 1 && v != -1) throw Error("hlbl_schedule_build: h must be a sign vector");
    if (stride <= 0) throw Error("hlbl_schedule_build: stride must be positive");
    if (n_steps < 0) throw Error("hlbl_schedule_build: n_steps must be >= 0");
    s = new hlbl_schedule();
    s->steps = build_schedule(v4(h), stride, n_steps, v4(origin));
    for (const auto& st : s->steps)
      for (const auto& src : st.sources) s->flat.emplace_back(st.index, &src);
  });
  return rc == 0 ? s : nullptr;
}

void hlbl_schedule_free(hlbl_schedule* s) { delete s; }

int hlbl_schedule_n_windows(const hlbl_schedule* s) { return s ? (int)s->steps.size() : -1; }
int hlbl_schedule_n_sources(const hlbl_schedule* s) { return s ? (int)s->flat.size() : -1; }

int hlbl_schedule_source(const hlbl_schedule* s, int i, int* window, hlbl_coord* pos,
                         int* n_partners) {
  return guarded([&] {
    check_ptr(s, "schedule");
    if (i < 0 || i >= (int)s->flat.size()) throw Error("source index out of range");
    if (window) *window = s->flat[i].first;
    if (pos) *pos = c4(s->flat[i].second->position);
    if (n_partners) *n_partners = (int)s->flat[i].second->partners.size();
  });
}

int hlbl_schedule_partner(const hlbl_schedule* s, int i, int j, hlbl_coord* pos) {
  return guarded([&] {
    check_ptr(s, "schedule");
    if (i < 0 || i >= (int)s->flat.size()) throw Error("source index out of range");
    const auto& p = s->flat[i].second->partners;
    if (j < 0 || j >= (int)p.size()) throw Error("partner index out of range");
    check_ptr(pos, "pos");
    *pos = c4(p[j]);
  });
}

int hlbl_schedule_check_complete(const hlbl_schedule* s, int stride) {
  return guarded([&] {
    check_ptr(s, "schedule");
    check_complete(s->steps, stride);
  });
}

int hlbl_schedule_simulate(const hlbl_schedule* s, int mode, long long stats[5]) {
  return guarded([&] {
    check_ptr(s, "schedule");
    check_ptr(stats, "stats");
    const ScheduleStats r = simulate(s->steps, mode == 1);
    stats[0] = r.n_propagators;
    stats[1] = r.n_solves;
    stats[2] = r.n_pairs;
    stats[3] = r.n_kernel_sums;
    stats[4] = r.peak_cache;
  });
}

int hlbl_max_steps_without_wrap(hlbl_coord h, int stride, hlbl_coord lattice, hlbl_coord origin,
                                int cap) {
  return max_steps_without_wrap(v4(h), stride, v4(lattice), v4(origin), cap);
}

}  // extern "C"

// ---- window entry, surrogate, driver ------------------------------------------------

namespace hlbl {

void window_run(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit, int form,
                double* out, hipStream_t st);
void window_judge_result(const hlbl_contractor* c, double* err, int* n);
void window_set_judge_bound(hlbl_contractor* c, double bound);
void window_fallback_result(const hlbl_contractor* c, int* fell_back);
void window_drop_fields(hlbl_contractor* c);
void window_reserve_fields(hlbl_contractor* c, int n);
int window_field_slots(const hlbl_contractor* c);
int window_fields_built(const hlbl_contractor* c);


// a row as the C ABI hands it out: x[4], y[4], yv[4], window, chain, ks0, ks1
void pack_row(const PairRow& p, int chain, double* r) {
  for (int k = 0; k < 4; ++k) {
    r[k] = p.x[k];
    r[4 + k] = p.y[k];
    r[8 + k] = p.yv[k];
  }
  r[12] = p.window;
  r[13] = chain;
  r[14] = p.ks[0];
  r[15] = p.ks[1];
}

// the window entry for every form; form 0 runs the FP64 pair path per pair
void window_entry(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit, int form,
                  double* out, hipStream_t st) {
  if (form < 0 || form > 2) throw Error("form must be 0, 1 or 2");
  check_xunit(xunit);
  for (int i = 0; i < n; ++i)
    if (!pairs[i].d_x || !pairs[i].d_y) throw Error("a pair has a null block");
  if (n == 0) return;
  if (form == 0) {
    for (int i = 0; i < n; ++i)
      if (hlbl_kernel_sum(c, pairs[i].d_x, pairs[i].d_y, pairs[i].gsx, pairs[i].yv, xunit, out + 2 * i, st) != 0)
        throw Error(hlbl_last_error());
    return;
  }
  window_run(c, pairs, n, xunit, form, out, st);
}

}  // namespace hlbl

#include "mlp.hpp"

extern "C" {

int hlbl_window_kernel_sum(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n_pairs,
                           double xunit, int form, double* out, hlbl_stream stream) {
  return guarded([&] {
    check_ptr(c, "contractor");
    if (n_pairs < 0) throw Error("n_pairs < 0");
    if (n_pairs > 0) {
      check_ptr(pairs, "pairs");
      check_ptr(out, "out");
    }
    window_entry(c, pairs, n_pairs, xunit, form, out, as_stream(stream));
  });
}

int hlbl_window_judge(const hlbl_contractor* c, double* max_rel_err, int* n_samples) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(max_rel_err, "max_rel_err");
    check_ptr(n_samples, "n_samples");
    window_judge_result(c, max_rel_err, n_samples);
  });
}

int hlbl_window_set_judge_bound(hlbl_contractor* c, double bound) {
  return guarded([&] {
    check_ptr(c, "contractor");
    window_set_judge_bound(c, bound);
  });
}

int hlbl_window_fallback(const hlbl_contractor* c, int* fell_back) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(fell_back, "fell_back");
    window_fallback_result(c, fell_back);
  });
}

int hlbl_window_drop_fields(hlbl_contractor* c) {
  return guarded([&] {
    check_ptr(c, "contractor");
    window_drop_fields(c);
  });
}

int hlbl_window_fields_built(const hlbl_contractor* c, int* n) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(n, "n");
    *n = window_fields_built(c);
  });
}

int hlbl_window_reserve_fields(hlbl_contractor* c, int n, int* slots) {
  return guarded([&] {
    check_ptr(c, "contractor");
    if (n < 0) throw Error("hlbl_window_reserve_fields: n < 0");
    if (n > 0) window_reserve_fields(c, n);
    if (slots) *slots = window_field_slots(c);
  });
}

int hlbl_mlp_param_count(const hlbl_qed* q) { return q ? mlp_layout().total : -1; }

int hlbl_mlp_init(hlbl_qed* q, uint64_t seed) {
  return guarded([&] {
    check_ptr(q, "qed");
    check_hip(hipSetDevice(q->device), "hipSetDevice");
    mlp_init_params(q, seed);
    q->mlp_ready = false;
    ++q->mlp_gen;
  });
}

int hlbl_mlp_train(hlbl_qed* q, int n_samples, int epochs, int batch, float lr, uint64_t seed,
                   float* loss, hlbl_stream stream) {
  return guarded([&] {
    check_ptr(q, "qed");
    if (n_samples <= 0 || epochs <= 0 || batch <= 0 || batch > n_samples)
      throw Error("hlbl_mlp_train: need 0 < batch <= n_samples and epochs > 0");
    check_hip(hipSetDevice(q->device), "hipSetDevice");
    mlp_train(q, n_samples, epochs, batch, lr, seed, loss, as_stream(stream));
  });
}

int hlbl_mlp_set_domain(hlbl_qed* q, double radius) {
  return guarded([&] {
    check_ptr(q, "qed");
    if (!(radius > 0.0) || radius > qed::kRadius * (1.0 + 1e-12))
      throw Error("hlbl_mlp_set_domain: need 0 < radius <= the kernel radius");
    if (!(radius > q->mlp_inner))
      throw Error("hlbl_mlp_set_domain: the radius must exceed the inner radius");
    q->mlp_radius = radius;
    q->mlp_ready = false;
    ++q->mlp_gen;
  });
}

int hlbl_mlp_set_inner_radius(hlbl_qed* q, double inner) {
  return guarded([&] {
    check_ptr(q, "qed");
    if (!(inner >= 0.0) || !(inner < q->mlp_radius))
      throw Error("hlbl_mlp_set_inner_radius: need 0 <= inner < the trained radius");
    q->mlp_inner = inner;
    q->mlp_ready = false;
    ++q->mlp_gen;
  });
}

int hlbl_mlp_get_params(const hlbl_qed* q, float* host, int n) {
  return guarded([&] {
    check_ptr(q, "qed");
    check_ptr(host, "host");
    const int total = mlp_layout().total;
    if (!q->mlp_w.p) throw Error("hlbl_mlp_get_params: the network is not initialised");
    if (n < total) throw Error("hlbl_mlp_get_params: buffer too small");
    check_hip(hipMemcpy(host, q->mlp_w.p, total * sizeof(float), hipMemcpyDeviceToHost), "copy params");
  });
}

int hlbl_mlp_set_params(hlbl_qed* q, const float* host, int n) {
  return guarded([&] {
    check_ptr(q, "qed");
    check_ptr(host, "host");
    const int total = mlp_layout().total;
    if (n != total) throw Error("hlbl_mlp_set_params: wrong parameter count");
    if (!q->mlp_w.p) mlp_init_params(q, 0);
    check_hip(hipMemcpy(q->mlp_w.p, host, total * sizeof(float), hipMemcpyHostToDevice), "copy params");
    mlp_requantise(q, nullptr);
    check_hip(hipDeviceSynchronize(), "requantise");
    ++q->mlp_gen;
  });
}

int hlbl_mlp_forward(hlbl_qed* q, const float* d_in, int n, float* d_out, int precision,
                     hlbl_stream stream) {
  return guarded([&] {
    check_ptr(q, "qed");
    if (!q->mlp_w.p) throw Error("hlbl_mlp_forward: the network is not initialised");
    if (n < 0) throw Error("n < 0");
    if (n == 0) return;
    check_ptr(d_in, "d_in");
    check_ptr(d_out, "d_out");
    if (precision == 0) mlp_forward_fp32(q, d_in, n, d_out, as_stream(stream));
    else mlp_infer(q, d_in, n, d_out, as_stream(stream));
    check_hip(hipStreamSynchronize(as_stream(stream)), "mlp forward");
  });
}

int hlbl_mlp_loss_grad(hlbl_qed* q, const float* d_in, const float* d_target, int n,
                       float* d_grad, float* loss_host, hlbl_stream stream) {
  return guarded([&] {
    check_ptr(q, "qed");
    if (!q->mlp_w.p) throw Error("hlbl_mlp_loss_grad: the network is not initialised");
    if (n <= 0) throw Error("n must be positive");
    check_ptr(d_in, "d_in");
    check_ptr(d_target, "d_target");
    check_ptr(d_grad, "d_grad");
    const double l = mlp_loss_grad(q, d_in, d_target, n, d_grad, as_stream(stream));
    if (loss_host) *loss_host = (float)l;
  });
}

int hlbl_run_chain_2p2(hlbl_contractor* c, const hlbl_schedule* s, uint64_t seed, double xunit,
                       int residency, int form, int batch, int chain, hlbl_window_rows_fn fn,
                       void* user, const hlbl_plan_2p2* plan, int n_plan, int iflavor,
                       hlbl_window_2p2_fn fn2, void* user2, double* rows, int max_rows,
                       double stats[5]) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(s, "schedule");
    check_ptr(stats, "stats");
    if (residency != 0 && residency != 1) throw Error("residency must be 0 (device) or 1 (host)");
    if (chain < 0) throw Error("chain must be non-negative");
    if (batch < 0 || batch > 2) throw Error("batch must be 0 (pairs), 1 (windows) or 2 (sources)");
    if (n_plan < 0) throw Error("n_plan < 0");
    check_xunit(xunit);
    if (n_plan > 0) check_ptr(plan, "plan");
    if (iflavor != 0 && iflavor != 1) throw Error("iflavor must be 0 or 1");
    std::map<Vec4, Row2p2> plan2;
    const hlbl_geom& g = c->global;
    for (int i = 0; i < n_plan; ++i) {
      const hlbl_plan_2p2& e = plan[i];
      const Vec4 w = {e.w.t, e.w.x, e.w.y, e.w.z};
      if (w[0] < 0 || w[0] >= g.T || w[1] < 0 || w[1] >= g.LX || w[2] < 0 || w[2] >= g.LY || w[3] < 0 ||
          w[3] >= g.LZ)
        throw Error("hlbl_run_chain_2p2: a planned site is outside [0, L)");
      if (e.n_targets < 0 || (e.n_targets > 0 && !e.targets)) throw Error("hlbl_run_chain_2p2: bad targets");
      if (plan2.count(w)) throw Error("hlbl_run_chain_2p2: a site is planned twice");
      Row2p2 r;
      r.w = w;
      r.window = e.window;
      r.targets.assign(e.targets, e.targets + e.n_targets);
      plan2.emplace(w, std::move(r));
    }
    std::vector<PairRow> out;
    std::vector<double> packed;
    std::function<void(size_t, size_t)> on_window;
    if (fn)
      on_window = [&](size_t first, size_t n) {
        packed.assign(16 * n, 0.0);
        for (size_t i = 0; i < n; ++i) pack_row(out[first + i], chain, packed.data() + 16 * i);
        if (fn(user, packed.data(), (int)n) != 0)
          throw Error("hlbl_run_chain: the window callback failed");
      };
    std::function<void(std::vector<Row2p2>&)> on_2p2;
    if (fn2)
      on_2p2 = [&](std::vector<Row2p2>& rs) {
        std::vector<hlbl_source_2p2> v(rs.size());
        for (size_t i = 0; i < rs.size(); ++i)
          v[i] = {{rs[i].w[0], rs[i].w[1], rs[i].w[2], rs[i].w[3]}, chain, rs[i].window, rs[i].p1.data(),
                  rs[i].targets.data(), (int)rs[i].targets.size(), rs[i].p234.data()};
        if (fn2(user2, v.data(), (int)v.size()) != 0)
          throw Error("hlbl_run_chain_2p2: the 2+2 callback failed");
      };
    run_chain(c, s->steps, seed, xunit, residency == 1, form, batch, out, stats, on_window, plan2, iflavor,
              on_2p2);
    if ((int)out.size() > max_rows) throw Error("hlbl_run_chain: rows buffer too small");
    if (!out.empty()) check_ptr(rows, "rows");
    for (size_t i = 0; i < out.size(); ++i) pack_row(out[i], chain, rows + 16 * i);
  });
}

int hlbl_run_chain_each_window(hlbl_contractor* c, const hlbl_schedule* s, uint64_t seed,
                               double xunit, int residency, int form, int batch, int chain,
                               hlbl_window_rows_fn fn, void* user, double* rows, int max_rows,
                               double stats[5]) {
  return hlbl_run_chain_2p2(c, s, seed, xunit, residency, form, batch, chain, fn, user, nullptr, 0, 0,
                            nullptr, nullptr, rows, max_rows, stats);
}

int hlbl_run_chain(hlbl_contractor* c, const hlbl_schedule* s, uint64_t seed, double xunit,
                   int residency, int form, int batch, double* rows, int max_rows,
                   double stats[5]) {
  return hlbl_run_chain_each_window(c, s, seed, xunit, residency, form, batch, 0, nullptr, nullptr,
                                    rows, max_rows, stats);
}

}  // extern "C"

namespace hlbl {
void window_profile(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit, int form,
                    double* out, double ms[4]);
}

extern "C" int hlbl_window_profile(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n_pairs,
                                   double xunit, int form, double* out, double ms[4]) {
  return guarded([&] {
    check_ptr(c, "contractor");
    check_ptr(pairs, "pairs");
    check_ptr(out, "out");
    check_ptr(ms, "ms");
    if (n_pairs <= 0) throw Error("n_pairs must be positive");
    window_profile(c, pairs, n_pairs, xunit, form, out, ms);
  });
}
