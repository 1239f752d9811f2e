// Host side of the window entry: distinct-yv slots, buffers, and one hipGraph
// per window shape (pairs, distinct yv, form, xunit), replayed for later windows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
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

void pass1_group_units(const WinPair* pairs, int n_pairs, int* units, int unit_slots) {
  const int nu = 2 * n_pairs, G = kPass1Groups;
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
  for (int k = 0; k < G; ++k) {
    if ((int)grp[k].size() > unit_slots) throw Error("pass1_group_units: a list overflows its slots");
    for (int j = 0; j < unit_slots; ++j) {
      const int u = j < (int)grp[k].size() ? grp[k][j] : -1;
      units[k * unit_slots + j] = u;
      if (u >= 0) ++seen[u];
    }
  }
  for (int u = 0; u < nu; ++u)
    if (seen[u] != 1) throw Error("pass1_group_units: a unit is not listed exactly once");
}

// Default gate of the surrogate form: a form-2 window whose judge (the largest
// deviation of the surrogate's coefficients from the plain kernel's at the sampled
// sites, relative to the site's largest coefficient) exceeds it is recomputed with
// the plain kernel (form 1).
constexpr double kDefaultJudgeBound = 1e-3;

struct WindowCache {
  std::vector<std::unique_ptr<WindowPlan>> plans;
  hipStream_t stream = nullptr;
  double last_judge = -1.0;
  int last_judge_n = 0;
  double judge_bound = kDefaultJudgeBound;
  int last_fallback = 0;  // the last form-2 window returned form 1's results
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

  // One plan whose buffers only grow: a window of another size or form reuses them and
  // its graph is captured again, so the footprint is that of the largest window so far
  // and a run of differently sized batches does not reallocate.
  if (cache->plans.empty()) cache->plans.push_back(std::make_unique<WindowPlan>());
  WindowPlan* w = cache->plans.front().get();
  if (!(w->n_pairs == n && w->n_unique == nu && w->form == form && w->xunit == xunit)) {
    w->drop_graph();
    w->n_pairs = n;
    w->n_unique = nu;
    w->form = form;
    w->xunit = xunit;
    w->n_samples = kJudgeSamples;
    const size_t V = c->volume;
    w->d_pairs.ensure(n);
    w->h_pairs.ensure(n);
    w->unit_slots = (2 * n + kPass1Groups - 1) / kPass1Groups;
    w->d_units.ensure((size_t)kPass1Groups * w->unit_slots);
    w->h_units.ensure((size_t)kPass1Groups * w->unit_slots);
    w->d_yvs.ensure(nu);
    w->h_yvs.ensure(nu);
    w->part1.ensure(pass1_part_len(c->global, n));
    w->red1.ensure(pass1_red_len(c->global, n));
    w->xi.ensure((size_t)n * 2 * 8 * 5 * 144 * 2);
    w->dz.ensure((size_t)n * 2 * 10 * 144 * 2);
    w->gdz.ensure((size_t)n * 2 * 40 * 144 * 2);
    w->bsplit.ensure(window_bsplit_len(n));
    w->bexp.ensure(window_bexp_len(n));
    w->coef.ensure((size_t)nu * V * 640);
    w->ax.ensure((size_t)2 * V * qed::kNL);
    w->cst64.ensure((size_t)nu * 2 * qed::kNL);
    w->cst32.ensure((size_t)nu * 2 * qed::kNL);
    if (form == 2) w->chbuf.ensure((size_t)(2 + 3 * nu) * V * qed::kNCh);
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
    w->judge.ensure((size_t)nu * kJudgeSamples);
    w->h_judge.ensure((size_t)nu * kJudgeSamples);
  }
  for (int i = 0; i < n; ++i) {
    // gsx inside the lattice: the kernels step site coordinates relative to it
    const hlbl_coord gs = {mod_pos(pairs[i].gsx.t, g.T), mod_pos(pairs[i].gsx.x, g.LX),
                           mod_pos(pairs[i].gsx.y, g.LY), mod_pos(pairs[i].gsx.z, g.LZ)};
    w->h_pairs.p[i] = WinPair{pairs[i].d_x, pairs[i].d_y, gs, pairs[i].yv, slot[i], 0};
  }
  for (int u = 0; u < nu; ++u) w->h_yvs.p[u] = yvs[u];
  pass1_group_units(w->h_pairs.p, n, w->h_units.p, w->unit_slots);
  if (!cache->stream) check_hip(hipStreamCreateWithFlags(&cache->stream, hipStreamNonBlocking), "hipStreamCreate");
  return w;
}

namespace {
void window_finish(hlbl_contractor* c, WindowPlan* w, double* out) {
  for (int i = 0; i < 2 * w->n_pairs; ++i) out[i] = w->h_result.p[i];
  if (w->form == 2) {
    double e = 0.0;
    for (int k = 0; k < w->n_unique * w->n_samples; ++k) e = std::max(e, w->h_judge.p[k]);
    c->window->last_judge = e;
    c->window->last_judge_n = w->n_unique * w->n_samples;
  }
}
void window_tail_copies(WindowPlan* w, hipStream_t st) {
  check_hip(hipMemcpyAsync(w->h_result.p, w->result.p, 2 * w->n_pairs * sizeof(double), hipMemcpyDeviceToHost, st), "copy result");
  if (w->form == 2)
    check_hip(hipMemcpyAsync(w->h_judge.p, w->judge.p, (size_t)w->n_unique * w->n_samples * sizeof(double),
                             hipMemcpyDeviceToHost, st), "copy judge");
}
void window_head_copies(WindowPlan* w, hipStream_t st) {
  check_hip(hipMemcpyAsync(w->d_pairs.p, w->h_pairs.p, w->n_pairs * sizeof(WinPair), hipMemcpyHostToDevice, st), "copy pairs");
  check_hip(hipMemcpyAsync(w->d_yvs.p, w->h_yvs.p, w->n_unique * sizeof(hlbl_coord), hipMemcpyHostToDevice, st), "copy yv");
  check_hip(hipMemcpyAsync(w->d_units.p, w->h_units.p, (size_t)kPass1Groups * w->unit_slots * sizeof(int),
                           hipMemcpyHostToDevice, st), "copy units");
}
}  // namespace

void window_run(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit, int form,
                double* out, hipStream_t user_stream) {
  WindowPlan* w = window_prepare(c, pairs, n, xunit, form);
  hipStream_t st = c->window->stream;
  // capture once per plan; graphs need a stream that is not the null stream
  if (!w->exec) {
    check_hip(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    try {
      window_head_copies(w, st);
      if (form == 2) launch_surrogate_channels(c, *w, st);
      window_record(c, *w, st);
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
  check_hip(hipGraphLaunch(w->exec, st), "hipGraphLaunch");
  check_hip(hipStreamSynchronize(st), "window");
  window_finish(c, w, out);
  if (form == 2) {
    // the judge as a gate: past the bound, the plain kernel's results instead
    WindowCache* cache = c->window;
    cache->last_fallback = cache->last_judge > cache->judge_bound ? 1 : 0;
    if (cache->last_fallback) {
      const double judge = cache->last_judge;
      const int judge_n = cache->last_judge_n;
      window_run(c, pairs, n, xunit, 1, out, user_stream);
      cache->last_judge = judge;
      cache->last_judge_n = judge_n;
      cache->last_fallback = 1;
    }
  }
}

void window_profile(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n, double xunit, int form,
                    double* out, double ms[4]) {
  if (form != 1 && form != 2) throw Error("hlbl_window_profile: form must be 1 or 2");
  WindowPlan* w = window_prepare(c, pairs, n, xunit, form);
  hipStream_t st = c->window->stream;
  check_hip(hipDeviceSynchronize(), "device");
  hipEvent_t ev[5];
  for (auto& e : ev) check_hip(hipEventCreate(&e), "hipEventCreate");
  try {
    window_head_copies(w, st);
    check_hip(hipEventRecord(ev[0], st), "event");
    window_stage(c, *w, 0, st);
    check_hip(hipEventRecord(ev[1], st), "event");
    if (form == 2) launch_surrogate_channels(c, *w, st);
    check_hip(hipEventRecord(ev[2], st), "event");
    window_stage(c, *w, 1, st);
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
  window_finish(c, w, out);
}

}  // namespace hlbl
