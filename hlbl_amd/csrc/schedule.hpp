// The sliding-hypercube schedule: windows A_k = origin + stride*(k h + A_0),
// A_0 = {0} u (G \ {h}) in the solve order of h = (1,1,1,1) reflected by h,
// and for each new source its partners at separation stride*(+-1,+-1,+-1,+-1)
// (src/schedule.cpp:34-214, include/hlbl/schedule.hpp:17-83).
#pragma once

#include <array>
#include <vector>

namespace hlbl {

using Vec4 = std::array<int, 4>;

struct Source {
  Vec4 position;               // unwrapped
  std::vector<Vec4> partners;  // earlier sources it pairs with
};
struct Step {
  int index = 0;
  std::vector<Source> sources;
};
using Schedule = std::vector<Step>;

std::array<Vec4, 16> sign_vectors();
std::vector<Vec4> window_offsets(const Vec4& h);
Schedule build_schedule(const Vec4& h, int stride, int n_steps, const Vec4& origin);
// throws Error when a position or a pair is repeated or a pair is missing
void check_complete(const Schedule& s, int stride);

struct ScheduleStats {
  long long n_propagators = 0, n_solves = 0, n_pairs = 0, n_kernel_sums = 0;
  int peak_cache = 0;
};
// interleaved: contract each source right after its solve; serial: solve a window first
ScheduleStats simulate(const Schedule& s, bool interleaved);
int max_steps_without_wrap(const Vec4& h, int stride, const Vec4& lattice, const Vec4& origin,
                           int cap);

}  // namespace hlbl
