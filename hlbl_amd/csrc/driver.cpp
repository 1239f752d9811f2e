// The schedule loop on random fields and the propagator cache
// (src/driver.cpp:23-165, src/propagator_cache.cpp:23-92).
#include <hip/hip_runtime.h>

#include <chrono>
#include <functional>
#include <map>
#include <memory>
#include <optional>
#include <set>
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
// staging slots (0 for y, 1 for x). A position is freed after its last pair; its
// memory (device or pinned) is kept for the next position solved, so a run allocates
// no more blocks than its cache peak.
class PropagatorCache {
 public:
  PropagatorCache(bool host, size_t block_len, std::map<Vec4, int> pending)
      : host_(host), len_(block_len), pending_(std::move(pending)) {}

  double* begin_insert(const Vec4& p) {
    if (entries_.count(p)) throw Error("PropagatorCache: position inserted twice");
    if (!host_) {
      Entry& e = entries_[p];
      if (dev_pool_.empty()) {
        e.dev = std::make_unique<DevBuf<double>>();
        e.dev->ensure(len_);
      } else {
        e.dev = std::move(dev_pool_.back());
        dev_pool_.pop_back();
      }
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

  // the pinned host copy of a cached position (host tier)
  const double* host_ptr(const Vec4& p) const {
    auto it = entries_.find(p);
    if (it == entries_.end() || !it->second.host) throw Error("PropagatorCache: position is not cached on the host");
    return it->second.host->p;
  }
  // the device solve target of the host tier (holds the block last solved)
  const double* scratch() const { return scratch_.p; }
  bool host_tier() const { return host_; }

  void pair_done(const Vec4& a, const Vec4& b) {
    for (const Vec4& p : {a, b})
      if (--pending_[p] == 0) evict(p);
  }
  bool has(const Vec4& p) const { return entries_.count(p) != 0; }
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
      if (!host_ && it->second.dev) dev_pool_.push_back(std::move(it->second.dev));
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
  std::vector<std::unique_ptr<DevBuf<double>>> dev_pool_;
  DevBuf<double> scratch_;
  DevBuf<double> staging_[2];
  std::optional<Vec4> staged_[2];
};

// Device blocks that feed the host tier's batched calls (one call per window or per
// new source): the blocks a call reads are staged here from pinned host memory. A
// block just solved is copied from the solve target on the device; every other block
// is copied host to device on a stream of its own, issued while the previous call
// computes (the calls are synchronous, so a slot that no pending call reads can be
// refilled at once). Slots hold positions until they are needed for others or their
// position leaves the cache.
class StagingPool {
 public:
  StagingPool(size_t block_len, int n_slots) : len_(block_len), slots_(n_slots) {
    for (Slot& s : slots_) s.buf.ensure(len_);
    check_hip(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking), "copy stream");
    check_hip(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "copy event");
  }
  ~StagingPool() {
    if (done_) (void)hipEventDestroy(done_);
    if (copy_) {
      (void)hipStreamSynchronize(copy_);
      (void)hipStreamDestroy(copy_);
    }
  }
  StagingPool(const StagingPool&) = delete;
  StagingPool& operator=(const StagingPool&) = delete;

  // the block at p, solved into `src` on the device, copied into a slot on st
  void put_solved(const Vec4& p, const double* src, const std::set<Vec4>& keep, hipStream_t st) {
    Slot& s = take(p, keep);
    check_hip(hipMemcpyAsync(s.buf.p, src, len_ * sizeof(double), hipMemcpyDeviceToDevice, st), "D2D block");
  }
  // issue host-to-device copies of the positions in `want` that no slot holds; slots
  // holding `keep` or `want` are not overwritten
  void prefetch(const std::vector<Vec4>& want, const std::set<Vec4>& keep, const PropagatorCache& cache) {
    std::set<Vec4> hold(keep);
    hold.insert(want.begin(), want.end());
    bool issued = false;
    for (const Vec4& p : want) {
      if (find(p)) continue;
      Slot& s = take(p, hold);
      check_hip(hipMemcpyAsync(s.buf.p, cache.host_ptr(p), len_ * sizeof(double), hipMemcpyHostToDevice, copy_),
                "H2D block");
      issued = true;
      ++copies_;
    }
    if (issued) {
      check_hip(hipEventRecord(done_, copy_), "copy event");
      pending_ = true;
    }
  }
  // st waits for the copies issued so far
  void wait(hipStream_t st) {
    if (!pending_) return;
    check_hip(hipStreamWaitEvent(st, done_, 0), "wait for copies");
    pending_ = false;
  }
  const double* ptr(const Vec4& p) {
    Slot* s = find(p);
    if (!s) throw Error("StagingPool: position is not staged");
    s->used = ++clock_;
    return s->buf.p;
  }
  void forget(const Vec4& p) {
    if (Slot* s = find(p)) s->pos.reset();
  }
  long copies() const { return copies_; }

 private:
  struct Slot {
    DevBuf<double> buf;
    std::optional<Vec4> pos;
    long used = 0;
  };
  Slot* find(const Vec4& p) {
    for (Slot& s : slots_)
      if (s.pos && *s.pos == p) return &s;
    return nullptr;
  }
  // a slot for p: the one holding p, else a free one, else the least recently used
  // one whose position is not held
  Slot& take(const Vec4& p, const std::set<Vec4>& hold) {
    if (Slot* s = find(p)) return *s;
    Slot* best = nullptr;
    for (Slot& s : slots_) {
      if (!s.pos) {
        best = &s;
        break;
      }
      if (hold.count(*s.pos)) continue;
      if (!best || s.used < best->used) best = &s;
    }
    if (!best) throw Error("StagingPool: no free slot (too few staging blocks)");
    best->pos = p;
    best->used = ++clock_;
    return *best;
  }
  size_t len_;
  std::vector<Slot> slots_;
  hipStream_t copy_ = nullptr;
  hipEvent_t done_ = nullptr;
  bool pending_ = false;
  long clock_ = 0, copies_ = 0;
};

void window_entry(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit, int form,
                  double* out, hipStream_t st);
void window_reserve_fields(hlbl_contractor* c, int n);

namespace {

Vec4 wrap4(const Vec4& p, const hlbl_geom& g) {
  return {mod_pos(p[0], g.T), mod_pos(p[1], g.LX), mod_pos(p[2], g.LY), mod_pos(p[3], g.LZ)};
}
hlbl_coord to_c(const Vec4& v) { return {v[0], v[1], v[2], v[3]}; }
Vec4 min_image(const Vec4& a, const Vec4& b, const hlbl_geom& g) {
  return {coord_map_diff(a[0], b[0], g.T), coord_map_diff(a[1], b[1], g.LX),
          coord_map_diff(a[2], b[2], g.LY), coord_map_diff(a[3], b[3], g.LZ)};
}

// the source after source i of step `step` in solve order, or null
const Source* next_source(const Schedule& s, int step, size_t i) {
  for (int k = step; k < (int)s.size(); ++k) {
    const auto& src = s[k].sources;
    const size_t j = k == step ? i + 1 : 0;
    if (j < src.size()) return &src[j];
  }
  return nullptr;
}

// staging blocks the host tier's batched calls need: the most distinct positions of
// one window (batch 1), or of one source's call together with the next source's
// partners (batch 2)
int staging_slots(const Schedule& s, int batch) {
  size_t most = 0;
  for (int k = 0; k < (int)s.size(); ++k) {
    if (batch == 1) {
      std::set<Vec4> w;
      for (const Source& src : s[k].sources)
        for (const Vec4& q : src.partners) {
          w.insert(q);
          w.insert(src.position);
        }
      most = std::max(most, w.size());
      continue;
    }
    for (size_t i = 0; i < s[k].sources.size(); ++i) {
      const Source& src = s[k].sources[i];
      if (src.partners.empty()) continue;
      std::set<Vec4> w(src.partners.begin(), src.partners.end());
      w.insert(src.position);
      if (const Source* nx = next_source(s, k, i)) w.insert(nx->partners.begin(), nx->partners.end());
      most = std::max(most, w.size());
    }
  }
  return (int)most;
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
  // the run's own stream: the window calls wait for this stream only, not for the
  // staging copies of the next call on the copy stream
  struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() {
      if (s) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
      }
    }
  } run_stream;
  check_hip(hipStreamCreateWithFlags(&run_stream.s, hipStreamNonBlocking), "run stream");
  PropagatorCache cache(host_tier, len, pending);
  // the host tier's batched calls read their blocks from a device staging pool sized for
  // the largest call (windows) or the largest two consecutive calls (sources, whose next
  // call's blocks are copied while one computes)
  std::unique_ptr<StagingPool> pool;
  if (host_tier && batch != 0) {
    const int n_slots = staging_slots(s, batch);
    if (n_slots > 0) pool = std::make_unique<StagingPool>(len, n_slots);
  }
  // The window forms keep one coefficient field per distinct yv across calls; a chain
  // has few (16 at 24^3x48 n=3 K=8 and 32^3x64 n=4 K=4), so the store gets a slot for
  // each when the device memory left beside the chain's blocks allows: otherwise
  // per-source calls (up to 8 yv each) evict fields that later calls build again.
  if (form != 0 && batch != 0) {
    std::set<Vec4> yvs;
    for (const Step& st : s)
      for (const Source& src : st.sources)
        for (const Vec4& q : src.partners) yvs.insert(min_image(src.position, q, c->global));
    size_t free_b = 0, total_b = 0;
    check_hip(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const size_t block_b = len * sizeof(double);
    // the cache peak of the call mode: a window's sources are solved before its call
    // (serial), a source's call follows its solve (interleaved)
    const size_t peak = (size_t)simulate(s, batch == 2).peak_cache;
    // the host tier's staging pool is already allocated; its solve target is not yet
    const size_t blocks_b = (host_tier ? 1 : peak) * block_b;
    const size_t margin = total_b / 8;  // window buffers, axis rows, allocator slack
    const size_t field_b = c->volume * 640 * sizeof(float);
    const size_t avail = free_b > blocks_b + margin ? free_b - blocks_b - margin : 0;
    const size_t fit = avail / field_b;
    // the surrogate form keeps its own fields and, when its judge sends a window back to
    // the plain kernel, form 1's too
    const int want = (int)std::min<size_t>(yvs.size() * (form == 2 ? 2 : 1), fit);
    if (want > 0) window_reserve_fields(c, want);
  }
  auto pair_done = [&](const Vec4& a, const Vec4& b) {
    cache.pair_done(a, b);
    if (pool)
      for (const Vec4& p : {a, b})
        if (!cache.has(p)) pool->forget(p);
  };
  hipStream_t stm = run_stream.s;
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
      for (size_t i = 0; i < st.sources.size(); ++i) {
        const Source& src = st.sources[i];
        solve(st.index, src.position);
        if (src.partners.empty()) continue;
        std::set<Vec4> keep(src.partners.begin(), src.partners.end());
        keep.insert(src.position);
        if (pool) {
          // this call's blocks: the one just solved from the solve target, the partners
          // (prefetched during the previous call, else now); then the next call's
          // partners, copied while this one computes
          if (cache.has(src.position)) pool->put_solved(src.position, cache.scratch(), keep, stm);
          pool->prefetch(src.partners, keep, cache);
          pool->wait(stm);
          const Source* nx = next_source(s, st.index, i);
          if (nx && !nx->partners.empty()) pool->prefetch(nx->partners, keep, cache);
        }
        auto dptr = [&](const Vec4& p, int slot) { return pool ? pool->ptr(p) : cache.device_ptr(p, slot, stm); };
        const double* d_y = dptr(src.position, 0);
        const size_t first = rows.size();
        std::vector<hlbl_pair_desc> descs;
        for (const Vec4& q : src.partners) {
          PairRow r{wrap4(q, c->global), wrap4(src.position, c->global),
                    min_image(src.position, q, c->global), st.index, {0, 0}};
          descs.push_back({dptr(q, 1), d_y, to_c(r.x), to_c(r.yv)});
          rows.push_back(r);
        }
        std::vector<double> out(2 * descs.size());
        const auto t0 = clk::now();
        window_entry(c, descs.data(), (int)descs.size(), xunit, form, out.data(), stm);
        t_contract += std::chrono::duration<double>(clk::now() - t0).count();
        for (size_t i2 = 0; i2 < descs.size(); ++i2) {
          rows[first + i2].ks[0] = out[2 * i2];
          rows[first + i2].ks[1] = out[2 * i2 + 1];
        }
        for (const Vec4& q : src.partners) pair_done(q, src.position);
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
          pair_done(q, src.position);
        }
      }
    } else {
      // the whole window: solve its sources, contract its pairs in one call
      std::set<Vec4> keep;
      for (const Source& src : st.sources)
        for (const Vec4& q : src.partners) {
          keep.insert(q);
          keep.insert(src.position);
        }
      for (const Source& src : st.sources) {
        solve(st.index, src.position);
        if (pool && cache.has(src.position)) pool->put_solved(src.position, cache.scratch(), keep, stm);
      }
      if (pool) {
        pool->prefetch(std::vector<Vec4>(keep.begin(), keep.end()), keep, cache);
        pool->wait(stm);
      }
      auto dptr = [&](const Vec4& p, int slot) { return pool ? pool->ptr(p) : cache.device_ptr(p, slot, stm); };
      std::vector<hlbl_pair_desc> descs;
      std::vector<std::pair<Vec4, Vec4>> done;
      const size_t first = rows.size();
      for (const Source& src : st.sources)
        for (const Vec4& q : src.partners) {
          PairRow r{wrap4(q, c->global), wrap4(src.position, c->global),
                    min_image(src.position, q, c->global), st.index, {0, 0}};
          descs.push_back({dptr(q, 1), dptr(src.position, 0), to_c(r.x), to_c(r.yv)});
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
      for (const auto& d : done) pair_done(d.first, d.second);
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
