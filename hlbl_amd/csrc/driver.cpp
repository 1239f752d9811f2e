// The schedule loop on random fields and the propagator cache
// (src/driver.cpp:23-165, src/propagator_cache.cpp:23-92).
#include <hip/hip_runtime.h>

#include <chrono>
#include <functional>
#include <map>
#include <memory>
#include <optional>
#include <stdexcept>
#include <vector>

#include "../../include/hlbl_amd.h"
#include "context.hpp"
#include "contract_kernels.hpp"
#include "driver.hpp"
#include "schedule.hpp"

namespace hlbl {

// Blocks the schedule still needs. Device tier: blocks on the GPU, read in
// place. Host tier: pinned host blocks, a device solve target and two device
// staging slots (0 for y, 1 for x). A position is freed after its last pair.
class PropagatorCache {
 public:
  PropagatorCache(bool host, size_t block_len, std::map<Vec4, int> pending)
      : host_(host), len_(block_len), pending_(std::move(pending)) {}

  double* begin_insert(const Vec4& p) {
    if (entries_.count(p)) throw Error("PropagatorCache: position inserted twice");
    if (!host_) {
      Entry& e = entries_[p];
      e.dev = std::make_unique<DevBuf<double>>();
      e.dev->ensure(len_);
      return e.dev->p;
    }
    scratch_.ensure(len_);
    return scratch_.p;
  }

  void end_insert(const Vec4& p, hipStream_t st) {
    if (host_) {
      Entry& e = entries_[p];
      if (pool_.empty()) {
        e.host = std::make_unique<PinBuf<double>>();
        e.host->ensure(len_);
      } else {
        e.host = std::move(pool_.back());
        pool_.pop_back();
      }
      check_hip(hipMemcpyAsync(e.host->p, scratch_.p, len_ * sizeof(double), hipMemcpyDeviceToHost, st), "D2H block");
      check_hip(hipStreamSynchronize(st), "D2H block");
    }
    peak_ = std::max(peak_, (int)entries_.size());
    if (pending_[p] == 0) evict(p);
  }

  const double* device_ptr(const Vec4& p, int slot, hipStream_t st) {
    auto it = entries_.find(p);
    if (it == entries_.end()) throw Error("PropagatorCache: position is not cached");
    if (!host_) return it->second.dev->p;
    if (slot < 0 || slot > 1) throw Error("PropagatorCache: bad slot");
    staging_[slot].ensure(len_);
    if (!staged_[slot] || *staged_[slot] != p) {
      check_hip(hipMemcpyAsync(staging_[slot].p, it->second.host->p, len_ * sizeof(double), hipMemcpyHostToDevice, st), "H2D block");
      staged_[slot] = p;
    }
    return staging_[slot].p;
  }

  void pair_done(const Vec4& a, const Vec4& b) {
    for (const Vec4& p : {a, b})
      if (--pending_[p] == 0) evict(p);
  }
  int size() const { return (int)entries_.size(); }
  int peak() const { return peak_; }

 private:
  struct Entry {
    std::unique_ptr<DevBuf<double>> dev;
    std::unique_ptr<PinBuf<double>> host;
  };
  void evict(const Vec4& p) {
    auto it = entries_.find(p);
    if (it != entries_.end()) {
      if (host_ && it->second.host) pool_.push_back(std::move(it->second.host));
      entries_.erase(it);
    }
    for (auto& s : staged_)
      if (s && *s == p) s.reset();
  }
  bool host_;
  size_t len_;
  std::map<Vec4, int> pending_;
  std::map<Vec4, Entry> entries_;
  int peak_ = 0;
  std::vector<std::unique_ptr<PinBuf<double>>> pool_;
  DevBuf<double> scratch_;
  DevBuf<double> staging_[2];
  std::optional<Vec4> staged_[2];
};

void window_entry(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit, int form,
                  double* out, hipStream_t st);

namespace {

Vec4 wrap4(const Vec4& p, const hlbl_geom& g) {
  return {mod_pos(p[0], g.T), mod_pos(p[1], g.LX), mod_pos(p[2], g.LY), mod_pos(p[3], g.LZ)};
}
hlbl_coord to_c(const Vec4& v) { return {v[0], v[1], v[2], v[3]}; }
Vec4 min_image(const Vec4& a, const Vec4& b, const hlbl_geom& g) {
  return {coord_map_diff(a[0], b[0], g.T), coord_map_diff(a[1], b[1], g.LX),
          coord_map_diff(a[2], b[2], g.LY), coord_map_diff(a[3], b[3], g.LZ)};
}

}  // namespace




void run_chain(hlbl_contractor* c, const Schedule& s, uint64_t seed, double xunit, bool host_tier,
               int form, int batch, std::vector<PairRow>& rows, double stats[5],
               const std::function<void(size_t, size_t)>& on_window,
               const std::map<Vec4, Row2p2>& plan_2p2, int iflavor_2p2,
               const std::function<void(std::vector<Row2p2>&)>& on_2p2) {
  if (c->local.T != c->global.T || c->local.LX != c->global.LX || c->local.LY != c->global.LY ||
      c->local.LZ != c->global.LZ)
    throw Error("hlbl_run_chain: one process holds the whole lattice here");
  std::map<Vec4, int> pending;
  for (const Step& st : s)
    for (const Source& src : st.sources)
      for (const Vec4& q : src.partners) {
        ++pending[q];
        ++pending[src.position];
      }
  const size_t len = (size_t)2 * 12 * 24 * c->volume;
  PropagatorCache cache(host_tier, len, pending);
  hipStream_t stm = nullptr;
  double t_contract = 0, t_solve = 0;
  int solved = 0;
  using clk = std::chrono::steady_clock;

  std::vector<Row2p2> window_2p2;  // the current window's 2+2 rows
  const int lmax = std::max(std::max(c->global.T, c->global.LX), std::max(c->global.LY, c->global.LZ));
  auto pieces_2p2 = [&](int window, const Vec4& site, const double* d_fwd) {
    const auto it = plan_2p2.find(site);
    if (it == plan_2p2.end() || it->second.window != window) return;
    Row2p2 r = it->second;
    const int nt = (int)r.targets.size();
    c->p1buf.ensure((size_t)64 * lmax);
    c->p234buf.ensure((size_t)std::max(1, nt) * 320);
    launch_p1_2p2(c, d_fwd, iflavor_2p2, to_c(site), c->p1buf.p, stm);
    if (nt > 0) launch_p234_2p2(c, c->pi2p2.p, to_c(site), r.targets.data(), nt, xunit, c->p234buf.p, stm);
    r.p1.resize((size_t)64 * lmax);
    r.p234.resize((size_t)nt * 320);
    check_hip(hipMemcpyAsync(r.p1.data(), c->p1buf.p, r.p1.size() * sizeof(double), hipMemcpyDeviceToHost, stm),
              "copy P1");
    if (nt > 0)
      check_hip(hipMemcpyAsync(r.p234.data(), c->p234buf.p, r.p234.size() * sizeof(double),
                               hipMemcpyDeviceToHost, stm), "copy P234");
    check_hip(hipStreamSynchronize(stm), "2+2 pieces");
    window_2p2.push_back(std::move(r));
  };

  auto solve = [&](int window, const Vec4& p) {
    const auto t0 = clk::now();
    double* dst = cache.begin_insert(p);
    const Vec4 site = wrap4(p, c->global);
    const uint64_t sd = hlbl_source_seed(seed, to_c(site));
    launch_fill_random(dst, c->local, c->proc, sd, stm);
    check_hip(hipStreamSynchronize(stm), "solve");
    ++solved;
    t_solve += std::chrono::duration<double>(clk::now() - t0).count();
    // 2+2 before the block is stored: dst is the solve target (the cache entry, or
    // the host tier's device scratch block), as src/driver.cpp:88-93
    pieces_2p2(window, site, dst);
    cache.end_insert(p, stm);
  };

  for (const Step& st : s) {
    const size_t window_first = rows.size();
    if (batch == 2) {
      if (host_tier) throw Error("hlbl_run_chain: source batches read every block in place (device tier)");
      for (const Source& src : st.sources) {
        solve(st.index, src.position);
        if (src.partners.empty()) continue;
        const double* d_y = cache.device_ptr(src.position, 0, stm);
        const size_t first = rows.size();
        std::vector<hlbl_pair_desc> descs;
        for (const Vec4& q : src.partners) {
          PairRow r{wrap4(q, c->global), wrap4(src.position, c->global),
                    min_image(src.position, q, c->global), st.index, {0, 0}};
          descs.push_back({cache.device_ptr(q, 1, stm), d_y, to_c(r.x), to_c(r.yv)});
          rows.push_back(r);
        }
        std::vector<double> out(2 * descs.size());
        const auto t0 = clk::now();
        window_entry(c, descs.data(), (int)descs.size(), xunit, form, out.data(), stm);
        t_contract += std::chrono::duration<double>(clk::now() - t0).count();
        for (size_t i = 0; i < descs.size(); ++i) {
          rows[first + i].ks[0] = out[2 * i];
          rows[first + i].ks[1] = out[2 * i + 1];
        }
        for (const Vec4& q : src.partners) cache.pair_done(q, src.position);
      }
    } else if (batch == 0) {
      for (const Source& src : st.sources) {
        solve(st.index, src.position);
        if (src.partners.empty()) continue;
        const double* d_y = cache.device_ptr(src.position, 0, stm);
        for (const Vec4& q : src.partners) {
          const double* d_x = cache.device_ptr(q, 1, stm);
          PairRow r{wrap4(q, c->global), wrap4(src.position, c->global),
                    min_image(src.position, q, c->global), st.index, {0, 0}};
          const auto t0 = clk::now();
          if (form == 0) {
            if (hlbl_kernel_sum(c, d_x, d_y, to_c(r.x), to_c(r.yv), xunit, r.ks, stm) != 0)
              throw Error(hlbl_last_error());
          } else {
            hlbl_pair_desc d{d_x, d_y, to_c(r.x), to_c(r.yv)};
            window_entry(c, &d, 1, xunit, form, r.ks, stm);
          }
          t_contract += std::chrono::duration<double>(clk::now() - t0).count();
          rows.push_back(r);
          cache.pair_done(q, src.position);
        }
      }
    } else {
      // the whole window: solve its sources, contract its pairs in one call
      if (host_tier) throw Error("hlbl_run_chain: window batches read every block in place (device tier)");
      for (const Source& src : st.sources) solve(st.index, src.position);
      std::vector<hlbl_pair_desc> descs;
      std::vector<std::pair<Vec4, Vec4>> done;
      const size_t first = rows.size();
      for (const Source& src : st.sources)
        for (const Vec4& q : src.partners) {
          PairRow r{wrap4(q, c->global), wrap4(src.position, c->global),
                    min_image(src.position, q, c->global), st.index, {0, 0}};
          descs.push_back({cache.device_ptr(q, 1, stm), cache.device_ptr(src.position, 0, stm),
                           to_c(r.x), to_c(r.yv)});
          rows.push_back(r);
          done.emplace_back(q, src.position);
        }
      if (!descs.empty()) {
        std::vector<double> out(2 * descs.size());
        const auto t0 = clk::now();
        window_entry(c, descs.data(), (int)descs.size(), xunit, form, out.data(), stm);
        t_contract += std::chrono::duration<double>(clk::now() - t0).count();
        for (size_t i = 0; i < descs.size(); ++i) {
          rows[first + i].ks[0] = out[2 * i];
          rows[first + i].ks[1] = out[2 * i + 1];
        }
      }
      for (const auto& d : done) cache.pair_done(d.first, d.second);
    }
    if (on_window) on_window(window_first, rows.size() - window_first);
    if (!window_2p2.empty()) {
      if (on_2p2) on_2p2(window_2p2);
      window_2p2.clear();
    }
  }
  if (cache.size() != 0) throw Error("run_chain: propagators left in the cache");
  stats[0] = (double)rows.size();
  stats[1] = solved;
  stats[2] = cache.peak();
  stats[3] = t_contract;
  stats[4] = t_solve;
}

}  // namespace hlbl
