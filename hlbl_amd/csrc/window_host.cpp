// Host side of the window entry: distinct-yv slots, buffers, and one hipGraph
// per window shape (pairs, distinct yv, form, xunit), replayed for later windows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <tuple>
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/hlbl_amd.h"
#include "context.hpp"
#include "contract_kernels.hpp"
#include "pass1.hpp"
#include "window.hpp"

namespace hlbl {

namespace {

// canonical (min-image) yv, the key of a coefficient field
hlbl_coord canonical_yv(const hlbl_geom& g, const hlbl_coord& yv) {
  hlbl_coord r;
  r.t = coord_map(mod_pos(yv.t, g.T), g.T);
  r.x = coord_map(mod_pos(yv.x, g.LX), g.LX);
  r.y = coord_map(mod_pos(yv.y, g.LY), g.LY);
  r.z = coord_map(mod_pos(yv.z, g.LZ), g.LZ);
  return r;
}

bool same(const hlbl_coord& a, const hlbl_coord& b) {
  return a.t == b.t && a.x == b.x && a.y == b.y && a.z == b.z;
}

constexpr int kJudgeSamples = 64;

}  // namespace

void pass1_order_units(const WinPair* pairs, int n_pairs, int batch, int* order) {
  const int nu = 2 * n_pairs, G = (nu + batch - 1) / batch, unit_slots = (nu + G - 1) / G;
  // the two (block, flavour) halves a unit reads: S[f] of x and Y[1-f] of y
  std::map<std::pair<const double*, int>, int> ids;
  auto hid = [&](const double* b, int f) {
    return ids.emplace(std::make_pair(b, f), (int)ids.size()).first->second;
  };
  std::vector<std::array<int, 2>> uh(nu);
  for (int i = 0; i < n_pairs; ++i)
    for (int f = 0; f < 2; ++f) uh[2 * i + f] = {hid(pairs[i].S, f), hid(pairs[i].Y, 1 - f)};
  const int nh = (int)ids.size();
  std::vector<int> cap(G), cnt((size_t)G * nh, 0);
  for (int k = 0; k < G; ++k) cap[k] = (nu % G == 0 || k < nu % G) ? unit_slots : unit_slots - 1;
  std::vector<std::vector<int>> grp(G);
  auto put = [&](int k, int u, int s) {
    cnt[(size_t)k * nh + uh[u][0]] += s;
    cnt[(size_t)k * nh + uh[u][1]] += s;
  };
  // halves of u that group k lacks
  auto lacks = [&](int k, int u) {
    const int h0 = uh[u][0], h1 = uh[u][1];
    return (cnt[(size_t)k * nh + h0] == 0) + (h1 != h0 && cnt[(size_t)k * nh + h1] == 0);
  };
  for (int u = 0; u < nu; ++u) {
    int best = -1, bc = 0;
    for (int k = 0; k < G; ++k) {
      if ((int)grp[k].size() >= cap[k]) continue;
      const int c = lacks(k, u);
      if (best < 0 || c < bc || (c == bc && grp[k].size() < grp[best].size())) best = k, bc = c;
    }
    grp[best].push_back(u);
    put(best, u, 1);
  }
  // swaps between lists while one lowers the number of distinct halves
  for (int sweep = 0; sweep < 20; ++sweep) {
    bool improved = false;
    for (int a = 0; a < G; ++a)
      for (int b = a + 1; b < G; ++b)
        for (size_t ia = 0; ia < grp[a].size(); ++ia)
          for (size_t ib = 0; ib < grp[b].size(); ++ib) {
            const int u = grp[a][ia], v = grp[b][ib];
            put(a, u, -1);
            put(b, v, -1);
            const int before = lacks(a, u) + lacks(b, v), after = lacks(a, v) + lacks(b, u);
            if (after < before) {
              put(a, v, 1);
              put(b, u, 1);
              grp[a][ia] = v;
              grp[b][ib] = u;
              improved = true;
            } else {
              put(a, u, 1);
              put(b, v, 1);
            }
          }
    if (!improved) break;
  }
  std::vector<int> seen(nu, 0);
  int j = 0;
  for (int k = 0; k < G; ++k)
    for (const int u : grp[k]) {
      if (j == nu) throw Error("pass1_order_units: more units listed than there are");
      order[j++] = u;
      ++seen[u];
    }
  for (int u = 0; u < nu; ++u)
    if (seen[u] != 1) throw Error("pass1_order_units: a unit is not listed exactly once");
}

// Default gate of the surrogate form: a form-2 window whose judge (the largest
// deviation of the surrogate's coefficients from the plain kernel's at the sampled
// sites, relative to the site's largest coefficient) exceeds it is recomputed with
// the plain kernel (form 1).
constexpr double kDefaultJudgeBound = 1e-3;

// Coefficient fields kept across calls. A field depends on (yv, xunit, form) and, in the
// surrogate form, on the network (its generation counter) only, so the windows of a
// chain, which see the same distinct yv (the schedule is the same hypercube shifted),
// build them once. Slots are reused least recently used first.
struct FieldStore {
  struct Key {
    hlbl_coord yv;
    double xunit;
    int form;
    uint64_t gen;
  };
  DevBuf<float> coef;  // [cap][V][640]
  int cap = 0;
  std::vector<Key> key;
  std::vector<char> valid;
  std::vector<long> used;
  std::vector<double> judge;  // the surrogate judge of each form-2 field
  long clock = 0;
  bool matches(int s, const Key& k) const {
    const Key& a = key[s];
    return valid[s] && same(a.yv, k.yv) && a.xunit == k.xunit && a.form == k.form && a.gen == k.gen;
  }
};

// the surrogate's judge of a field, remembered after the field itself is evicted:
// (yv, xunit, surrogate generation)
using JudgeKey = std::tuple<int, int, int, int, double, uint64_t>;

struct WindowCache {
  std::vector<std::unique_ptr<WindowPlan>> plans;
  FieldStore fields;
  std::map<JudgeKey, double> judged;  // judges of the surrogate's fields built so far
  hipStream_t stream = nullptr;
  double last_judge = -1.0;
  int last_judge_n = 0;
  double judge_bound = kDefaultJudgeBound;
  int last_fallback = 0;  // the last form-2 window returned form 1's results
  int last_built = 0;     // fields the last call built
  ~WindowCache() {
    plans.clear();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

WindowCache* window_cache(hlbl_contractor* c) {
  if (!c->window) c->window = new WindowCache();
  return c->window;
}

void window_cache_free(WindowCache* w) { delete w; }

void window_judge_result(const hlbl_contractor* c, double* err, int* n) {
  if (!c->window || c->window->last_judge < 0) throw Error("hlbl_window_judge: no surrogate window has run");
  *err = c->window->last_judge;
  *n = c->window->last_judge_n;
}

void window_set_judge_bound(hlbl_contractor* c, double bound) {
  if (!(bound > 0)) throw Error("hlbl_window_set_judge_bound: the bound must be positive");
  window_cache(c)->judge_bound = bound;
}

// forget the kept fields and give their memory back (the captured graphs hold its
// address: they are captured again)
void window_drop_fields(hlbl_contractor* c) {
  if (!c->window) return;
  FieldStore& fs = c->window->fields;
  fs.coef.release();
  fs.cap = 0;
  fs.key.clear();
  fs.valid.clear();
  fs.used.clear();
  fs.judge.clear();
  for (auto& pl : c->window->plans) pl->drop_graph();
}

// at least n slots in the field store (growing it drops the kept fields and every
// captured graph, which hold its address)
void window_reserve_fields(hlbl_contractor* c, int n) {
  WindowCache* cache = window_cache(c);
  FieldStore& fs = cache->fields;
  if (fs.cap >= n) return;
  // an empty store first, so that a failed allocation leaves no slots behind
  window_drop_fields(c);
  fs.coef.ensure((size_t)n * c->volume * 640);
  fs.cap = n;
  fs.key.assign(n, FieldStore::Key{});
  fs.valid.assign(n, 0);
  fs.used.assign(n, 0);
  fs.judge.assign(n, 0.0);
}

int window_field_slots(const hlbl_contractor* c) { return c->window ? c->window->fields.cap : 0; }

int window_fields_built(const hlbl_contractor* c) { return c->window ? c->window->last_built : 0; }

// launch shapes changed: every window graph is captured again
void window_drop_graphs(hlbl_contractor* c) {
  if (c->window)
    for (auto& pl : c->window->plans) pl->drop_graph();
}

void window_fallback_result(const hlbl_contractor* c, int* fell_back) {
  if (!c->window || c->window->last_judge < 0) throw Error("hlbl_window_fallback: no surrogate window has run");
  *fell_back = c->window->last_fallback;
}

// the window's plan (reused or built) with the pair table filled in
WindowPlan* window_prepare(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit,
                           int form) {
  const hlbl_geom g = c->global;
  if (g.T != c->local.T || g.LX != c->local.LX || g.LY != c->local.LY || g.LZ != c->local.LZ)
    throw Error("hlbl_window_kernel_sum: the mixed-precision forms need the whole lattice on this process");
  if (form == 2 && !c->qed->mlp_ready)
    throw Error("hlbl_window_kernel_sum: form 2 needs a trained surrogate (hlbl_mlp_train)");
  WindowCache* cache = window_cache(c);

  // distinct yv
  std::vector<hlbl_coord> yvs;
  std::vector<int> slot(n);
  for (int i = 0; i < n; ++i) {
    const hlbl_coord cy = canonical_yv(g, pairs[i].yv);
    int s = -1;
    for (size_t u = 0; u < yvs.size(); ++u)
      if (same(yvs[u], cy)) s = (int)u;
    if (s < 0) {
      s = (int)yvs.size();
      yvs.push_back(cy);
    }
    slot[i] = s;
  }
  const int nu = (int)yvs.size();
  const size_t V = c->volume;

  // the field store: at least one slot per distinct yv of the call (growing it drops
  // the cached fields and every captured graph, which hold its address)
  FieldStore& fs = cache->fields;
  window_reserve_fields(c, nu);
  const uint64_t gen = form == 2 ? c->qed->mlp_gen : 0;
  std::vector<int> call_slot(nu, -1);
  std::vector<char> taken(fs.cap, 0);
  for (int u = 0; u < nu; ++u) {
    const FieldStore::Key k{yvs[u], xunit, form, gen};
    for (int s = 0; s < fs.cap; ++s)
      if (!taken[s] && fs.matches(s, k)) {
        call_slot[u] = s;
        taken[s] = 1;
        break;
      }
  }
  std::vector<int> build;  // distinct yv (indices u) whose fields this call builds
  for (int u = 0; u < nu; ++u) {
    if (call_slot[u] >= 0) continue;
    int best = -1;
    for (int s = 0; s < fs.cap; ++s) {
      if (taken[s]) continue;
      if (!fs.valid[s]) {
        best = s;
        break;
      }
      if (best < 0 || fs.used[s] < fs.used[best]) best = s;
    }
    call_slot[u] = best;
    taken[best] = 1;
    fs.valid[best] = 0;  // valid again once its build has run
    fs.key[best] = FieldStore::Key{yvs[u], xunit, form, gen};
    build.push_back(u);
  }
  for (int u = 0; u < nu; ++u) fs.used[call_slot[u]] = ++fs.clock;
  const int nb = (int)build.size();
  cache->last_built = nb;

  // One plan whose buffers only grow: a window of another size reuses them and its graph
  // is captured again, so the footprint is that of the largest window so far and a run
  // of differently sized batches does not reallocate. The graph (pass 1 and pass 2) does
  // not depend on the form or xunit, which only the field builds read.
  if (cache->plans.empty()) cache->plans.push_back(std::make_unique<WindowPlan>());
  WindowPlan* w = cache->plans.front().get();
  w->form = form;
  w->xunit = xunit;
  if (!(w->n_pairs == n && w->coef == fs.coef.p && w->exec)) {
    w->drop_graph();
    w->n_pairs = n;
    w->n_samples = kJudgeSamples;
    w->coef = fs.coef.p;
    w->d_pairs.ensure(n);
    w->h_pairs.ensure(n);
    w->batch = c->tune.pass1_batch;
    w->d_order.ensure((size_t)2 * n);
    w->h_order.ensure((size_t)2 * n);
    w->part1.ensure(pass1_part_len(c->global, n));
    w->red1.ensure(pass1_red_len(c->global, n));
    w->xi.ensure((size_t)n * 2 * 8 * 5 * 144 * 2);
    w->dz.ensure((size_t)n * 2 * 10 * 144 * 2);
    w->gdz.ensure((size_t)n * 2 * 40 * 144 * 2);
    w->bsplit.ensure(window_bsplit_len(n));
    w->bexp.ensure(window_bexp_len(n));
    w->part2.ensure(window_part2_len(c, n));
    w->result.ensure((size_t)2 * n);
    w->h_result.ensure((size_t)2 * n);
    if (w->d_samples.n == 0) {
      // judge samples: a fixed spread of sites
      std::vector<size_t> smp(kJudgeSamples);
      for (int k = 0; k < kJudgeSamples; ++k) smp[k] = (size_t)(((unsigned long long)k * 2654435761ULL) % V);
      w->d_samples.ensure(kJudgeSamples);
      check_hip(hipMemcpy(w->d_samples.p, smp.data(), smp.size() * sizeof(size_t), hipMemcpyHostToDevice),
                "hipMemcpy samples");
    }
  }
  // the builds: their yv, slots and work buffers (grown as needed, outside the graph)
  w->n_unique = nb;
  if (nb > 0) {
    w->d_yvs.ensure(nb);
    w->h_yvs.ensure(nb);
    w->d_slots.ensure(nb);
    w->h_slots.ensure(nb);
    w->ax.ensure((size_t)2 * V * qed::kNL);
    w->cst64.ensure((size_t)nb * 2 * qed::kNL);
    w->cst32.ensure((size_t)nb * 2 * qed::kNL);
    if (form == 2) w->chbuf.ensure((size_t)(2 + 3 * nb) * V * qed::kNCh);
    w->judge.ensure((size_t)nb * kJudgeSamples);
    w->h_judge.ensure((size_t)nb * kJudgeSamples);
    for (int j = 0; j < nb; ++j) {
      w->h_yvs.p[j] = yvs[build[j]];
      w->h_slots.p[j] = call_slot[build[j]];
    }
  }
  w->call_slots = call_slot;
  for (int i = 0; i < n; ++i) {
    // gsx inside the lattice: the kernels step site coordinates relative to it
    const hlbl_coord gs = {mod_pos(pairs[i].gsx.t, g.T), mod_pos(pairs[i].gsx.x, g.LX),
                           mod_pos(pairs[i].gsx.y, g.LY), mod_pos(pairs[i].gsx.z, g.LZ)};
    w->h_pairs.p[i] = WinPair{pairs[i].d_x, pairs[i].d_y, gs, pairs[i].yv, call_slot[slot[i]], 0};
  }
  pass1_order_units(w->h_pairs.p, n, w->batch, w->h_order.p);
  if (!cache->stream) check_hip(hipStreamCreateWithFlags(&cache->stream, hipStreamNonBlocking), "hipStreamCreate");
  return w;
}

namespace {
// the fields the call built are valid, with their judges (surrogate form); the call's
// judge is the largest of its fields' (built now or before)
void window_record_fields(hlbl_contractor* c, WindowPlan* w) {
  FieldStore& fs = c->window->fields;
  for (int j = 0; j < w->n_unique; ++j) {
    const int s = w->h_slots.p[j];
    double e = 0.0;
    if (w->form == 2)
      for (int k = 0; k < w->n_samples; ++k) e = std::max(e, w->h_judge.p[(size_t)j * w->n_samples + k]);
    fs.judge[s] = e;
    fs.valid[s] = 1;
    if (w->form == 2) {
      const hlbl_coord& y = w->h_yvs.p[j];
      c->window->judged[JudgeKey{y.t, y.x, y.y, y.z, w->xunit, c->qed->mlp_gen}] = e;
    }
  }
  if (w->form == 2) {
    double e = 0.0;
    for (int s : w->call_slots) e = std::max(e, fs.judge[s]);
    c->window->last_judge = e;
    c->window->last_judge_n = (int)w->call_slots.size() * w->n_samples;
  }
}
void window_results(WindowPlan* w, double* out) {
  for (int i = 0; i < 2 * w->n_pairs; ++i) out[i] = w->h_result.p[i];
}
void window_tail_copies(WindowPlan* w, hipStream_t st) {
  check_hip(hipMemcpyAsync(w->h_result.p, w->result.p, 2 * w->n_pairs * sizeof(double), hipMemcpyDeviceToHost, st), "copy result");
}
void window_head_copies(WindowPlan* w, hipStream_t st) {
  check_hip(hipMemcpyAsync(w->d_pairs.p, w->h_pairs.p, w->n_pairs * sizeof(WinPair), hipMemcpyHostToDevice, st), "copy pairs");
  check_hip(hipMemcpyAsync(w->d_order.p, w->h_order.p, (size_t)2 * w->n_pairs * sizeof(int),
                           hipMemcpyHostToDevice, st), "copy unit order");
}
// the fields the store lacks: their list, the surrogate's channels (form 2), the fields
// and the judge, launched directly on st (not part of the window's graph)
void window_builds(hlbl_contractor* c, WindowPlan* w, hipStream_t st, hipEvent_t mid) {
  const int nb = w->n_unique;
  if (nb == 0) {
    if (mid) check_hip(hipEventRecord(mid, st), "event");
    return;
  }
  check_hip(hipMemcpyAsync(w->d_yvs.p, w->h_yvs.p, nb * sizeof(hlbl_coord), hipMemcpyHostToDevice, st), "copy yv");
  check_hip(hipMemcpyAsync(w->d_slots.p, w->h_slots.p, nb * sizeof(int), hipMemcpyHostToDevice, st), "copy slots");
  if (w->form == 2) launch_surrogate_channels(c, *w, st);
  if (mid) check_hip(hipEventRecord(mid, st), "event");
  window_stage(c, *w, 1, st);
  if (w->form == 2)
    check_hip(hipMemcpyAsync(w->h_judge.p, w->judge.p, (size_t)nb * w->n_samples * sizeof(double),
                             hipMemcpyDeviceToHost, st), "copy judge");
}
// a new window's first call leaves the store without the fields it was building when
// anything fails before they are complete
struct BuildGuard {
  hlbl_contractor* c;
  WindowPlan* w;
  bool done = false;
  ~BuildGuard() {
    if (done) return;
    FieldStore& fs = c->window->fields;
    for (int j = 0; j < w->n_unique; ++j) fs.valid[w->h_slots.p[j]] = 0;
  }
};
}  // namespace

void window_run(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit, int form,
                double* out, hipStream_t user_stream) {
  if (form == 2 && c->qed->mlp_ready) {
    // a yv whose surrogate field has failed the judge before (kept or since evicted)
    // fails it again: the call goes to the plain kernel without building any surrogate
    // field
    WindowCache* cache = window_cache(c);
    double known = -1.0;
    int n_known = 0;
    std::vector<hlbl_coord> seen;
    for (int i = 0; i < n; ++i) {
      const hlbl_coord y = canonical_yv(c->global, pairs[i].yv);
      bool dup = false;
      for (const hlbl_coord& s : seen) dup = dup || same(s, y);
      if (dup) continue;
      seen.push_back(y);
      const auto it = cache->judged.find(JudgeKey{y.t, y.x, y.y, y.z, xunit, c->qed->mlp_gen});
      if (it != cache->judged.end()) {
        known = std::max(known, it->second);
        ++n_known;
      }
    }
    if (known > cache->judge_bound) {
      window_run(c, pairs, n, xunit, 1, out, user_stream);
      cache->last_judge = known;
      cache->last_judge_n = n_known * kJudgeSamples;
      cache->last_fallback = 1;
      return;
    }
  }
  WindowPlan* w = window_prepare(c, pairs, n, xunit, form);
  BuildGuard guard{c, w};
  hipStream_t st = c->window->stream;
  // capture once per plan (pass 1 and pass 2); graphs need a stream that is not the
  // null stream
  if (!w->exec) {
    check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      window_head_copies(w, st);
      window_stage(c, *w, 0, st);
      window_stage(c, *w, 2, st);
      window_tail_copies(w, st);
    } catch (...) {
      hipGraph_t gr = nullptr;
      (void)hipStreamEndCapture(st, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      throw;
    }
    check_hip(hipStreamEndCapture(st, &w->graph), "hipStreamEndCapture");
    check_hip(hipGraphInstantiate(&w->exec, w->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  }
  // order after the caller's earlier work on its stream
  if (user_stream) check_hip(hipStreamSynchronize(user_stream), "user stream");
  else check_hip(hipDeviceSynchronize(), "device");
  window_builds(c, w, st, nullptr);
  if (form == 2) {
    // the judge as a gate, before the window runs: past the bound (the judge of the
    // fields built now, or recorded when the kept ones were built), the plain kernel's
    // fields and results instead, and the surrogate's passes are not run at all
    check_hip(hipStreamSynchronize(st), "window fields");
    window_record_fields(c, w);
    guard.done = true;
    WindowCache* cache = c->window;
    cache->last_fallback = cache->last_judge > cache->judge_bound ? 1 : 0;
    if (cache->last_fallback) {
      const double judge = cache->last_judge;
      const int judge_n = cache->last_judge_n;
      window_run(c, pairs, n, xunit, 1, out, user_stream);
      cache->last_judge = judge;
      cache->last_judge_n = judge_n;
      cache->last_fallback = 1;
      return;
    }
  }
  check_hip(hipGraphLaunch(w->exec, st), "hipGraphLaunch");
  check_hip(hipStreamSynchronize(st), "window");
  guard.done = true;
  if (form != 2) window_record_fields(c, w);
  window_results(w, out);
}

void window_profile(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit, int form,
                    double* out, double ms[4]) {
  if (form != 1 && form != 2) throw Error("hlbl_window_profile: form must be 1 or 2");
  WindowPlan* w = window_prepare(c, pairs, n, xunit, form);
  BuildGuard guard{c, w};
  hipStream_t st = c->window->stream;
  check_hip(hipDeviceSynchronize(), "device");
  hipEvent_t ev[5];
  for (auto& e : ev) check_hip(hipEventCreate(&e), "hipEventCreate");
  try {
    window_head_copies(w, st);
    check_hip(hipEventRecord(ev[0], st), "event");
    window_stage(c, *w, 0, st);
    check_hip(hipEventRecord(ev[1], st), "event");
    window_builds(c, w, st, ev[2]);
    check_hip(hipEventRecord(ev[3], st), "event");
    window_stage(c, *w, 2, st);
    check_hip(hipEventRecord(ev[4], st), "event");
    window_tail_copies(w, st);
    check_hip(hipStreamSynchronize(st), "window");
    for (int i = 0; i < 4; ++i) {
      float t = 0.f;
      check_hip(hipEventElapsedTime(&t, ev[i], ev[i + 1]), "hipEventElapsedTime");
      ms[i] = t;
    }
  } catch (...) {
    for (auto& e : ev) (void)hipEventDestroy(e);
    throw;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  guard.done = true;
  window_record_fields(c, w);
  window_results(w, out);
}

}  // namespace hlbl
