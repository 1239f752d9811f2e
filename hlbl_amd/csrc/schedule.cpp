#include "schedule.hpp"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <set>
#include <utility>

#include "context.hpp"

namespace hlbl {

namespace {

Vec4 add(const Vec4& a, const Vec4& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2], a[3] + b[3]}; }
Vec4 scaled(int s, const Vec4& a) { return {s * a[0], s * a[1], s * a[2], s * a[3]}; }

bool diagonal(const Vec4& p, const Vec4& q, int stride) {
  for (int i = 0; i < 4; ++i)
    if (std::abs(p[i] - q[i]) != stride) return false;
  return true;
}

Vec4 wrap(const Vec4& p, const Vec4& L) {
  Vec4 r;
  for (int i = 0; i < 4; ++i) r[i] = ((p[i] % L[i]) + L[i]) % L[i];
  return r;
}

std::pair<Vec4, Vec4> unordered(const Vec4& a, const Vec4& b) {
  return a < b ? std::make_pair(a, b) : std::make_pair(b, a);
}

Vec4 window_point(const Vec4& origin, int stride, int k, const Vec4& h, const Vec4& a) {
  return add(origin, scaled(stride, add(scaled(k, h), a)));
}

}  // namespace

// the 16 sign vectors, lexicographic in (-1, +1) per component
std::array<Vec4, 16> sign_vectors() {
  std::array<Vec4, 16> g{};
  for (int i = 0; i < 16; ++i)
    for (int d = 0; d < 4; ++d) g[i][d] = ((i >> (3 - d)) & 1) ? 1 : -1;
  return g;
}

// {0} then the sign vectors other than (1,1,1,1), each reflected by h: the
// solve order of every direction mirrors that of (1,1,1,1)
std::vector<Vec4> window_offsets(const Vec4& h) {
  std::vector<Vec4> out{Vec4{0, 0, 0, 0}};
  for (const Vec4& g : sign_vectors()) {
    if (g == Vec4{1, 1, 1, 1}) continue;
    out.push_back({h[0] * g[0], h[1] * g[1], h[2] * g[2], h[3] * g[3]});
  }
  return out;
}

Schedule build_schedule(const Vec4& h, int stride, int n_steps, const Vec4& origin) {
  const std::vector<Vec4> offsets = window_offsets(h);
  Schedule out;
  std::vector<Vec4> prev;
  for (int k = 0; k <= n_steps; ++k) {
    std::vector<Vec4> cur;
    for (const Vec4& a : offsets) cur.push_back(window_point(origin, stride, k, h, a));
    std::vector<Vec4> seen = prev;  // available partners: the previous window, then new ones
    Step step;
    step.index = k;
    for (const Vec4& p : cur) {
      if (std::find(seen.begin(), seen.end(), p) != seen.end()) continue;
      Source src;
      src.position = p;
      for (const Vec4& q : seen)
        if (diagonal(p, q, stride)) src.partners.push_back(q);
      step.sources.push_back(src);
      seen.push_back(p);
    }
    out.push_back(std::move(step));
    prev = std::move(cur);
  }
  return out;
}

void check_complete(const Schedule& s, int stride) {
  std::vector<Vec4> pos;
  for (const Step& st : s)
    for (const Source& src : st.sources) pos.push_back(src.position);
  std::set<Vec4> unique(pos.begin(), pos.end());
  if (unique.size() != pos.size()) throw Error("check_complete: a propagator is scheduled twice");
  std::set<std::pair<Vec4, Vec4>> want, have;
  for (size_t i = 0; i < pos.size(); ++i)
    for (size_t j = i + 1; j < pos.size(); ++j)
      if (diagonal(pos[i], pos[j], stride)) want.insert(unordered(pos[i], pos[j]));
  for (const Step& st : s)
    for (const Source& src : st.sources)
      for (const Vec4& q : src.partners)
        if (!have.insert(unordered(q, src.position)).second)
          throw Error("check_complete: a pair is scheduled twice");
  if (have != want)
    throw Error("check_complete: scheduled pairs do not match every diagonal pair among the sources");
}

ScheduleStats simulate(const Schedule& s, bool interleaved) {
  std::map<Vec4, int> pending;
  for (const Step& st : s)
    for (const Source& src : st.sources)
      for (const Vec4& q : src.partners) {
        ++pending[q];
        ++pending[src.position];
      }
  std::set<Vec4> cache;
  ScheduleStats r;
  auto solve = [&](const Vec4& p) {
    cache.insert(p);
    r.peak_cache = std::max<int>(r.peak_cache, (int)cache.size());
    r.n_propagators++;
    r.n_solves += 24;
  };
  auto contract = [&](const Vec4& q, const Vec4& p) {
    r.n_pairs++;
    r.n_kernel_sums += 2;
    for (const Vec4& x : {q, p})
      if (--pending[x] == 0) cache.erase(x);
  };
  for (const Step& st : s) {
    if (interleaved) {
      for (const Source& src : st.sources) {
        solve(src.position);
        for (const Vec4& q : src.partners) contract(q, src.position);
      }
    } else {
      for (const Source& src : st.sources) solve(src.position);
      for (const Source& src : st.sources)
        for (const Vec4& q : src.partners) contract(q, src.position);
    }
  }
  if (!cache.empty()) throw Error("simulate: propagators left in the cache after the last pair");
  return r;
}

int max_steps_without_wrap(const Vec4& h, int stride, const Vec4& lattice, const Vec4& origin,
                           int cap) {
  const std::vector<Vec4> offsets = window_offsets(h);
  std::set<Vec4> seen;
  for (int k = 0; k < cap; ++k) {
    std::set<Vec4> fresh;
    const Vec4 centre_prev = wrap(window_point(origin, stride, k - 1, h, Vec4{0, 0, 0, 0}), lattice);
    for (const Vec4& a : offsets) {
      const Vec4 p = wrap(window_point(origin, stride, k, h, a), lattice);
      if (k == 0 || p != centre_prev) fresh.insert(p);
    }
    for (const Vec4& p : fresh)
      if (seen.count(p)) return k - 1;
    seen.insert(fresh.begin(), fresh.end());
  }
  return cap;
}

}  // namespace hlbl
