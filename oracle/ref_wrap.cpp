// ORACLE — test infrastructure only. A C ABI over the reference's own
// schedule and gamma-table sources (src/schedule.cpp, src/gamma.cpp of
// jing2li/hlbl), compiled from where they lie by oracle/Makefile into
// oracle/_ref/libhlbl_ref.so. The tests compare the product's schedule and
// gamma basis with these, bit for bit.
#include <cstring>
#include <stdexcept>
#include <vector>

#include "hlbl/gamma.hpp"
#include "hlbl/schedule.hpp"

using namespace hlbl;

namespace {
thread_local std::vector<int> g_rows;
}

extern "C" {

// gamma_permutation / gamma_sign row g (0..3, 5)
void ref_gamma_row(int g, int* perm, int* sign) {
  for (int j = 0; j < 24; ++j) {
    perm[j] = gamma_permutation[g][j];
    sign[j] = gamma_sign[g][j];
  }
}

// build_schedule flattened: rows of 10 ints per (source, partner):
// window, p[4], q[4], n_partners of p; a source without partners gives one row
// with q = p and n_partners = 0. Returns the row count; rows via ref_schedule_rows.
int ref_build_schedule(const int* h, int stride, int n_steps, const int* origin) {
  const Schedule s = build_schedule(Vec4{h[0], h[1], h[2], h[3]}, stride, n_steps,
                                    Vec4{origin[0], origin[1], origin[2], origin[3]});
  g_rows.clear();
  for (const Step& st : s)
    for (const Source& src : st.sources) {
      const int np = (int)src.partners.size();
      if (np == 0) {
        g_rows.push_back(st.index);
        for (int i = 0; i < 4; ++i) g_rows.push_back(src.position[i]);
        for (int i = 0; i < 4; ++i) g_rows.push_back(src.position[i]);
        g_rows.push_back(0);
      }
      for (const Vec4& q : src.partners) {
        g_rows.push_back(st.index);
        for (int i = 0; i < 4; ++i) g_rows.push_back(src.position[i]);
        for (int i = 0; i < 4; ++i) g_rows.push_back(q[i]);
        g_rows.push_back(np);
      }
    }
  return (int)(g_rows.size() / 10);
}

void ref_schedule_rows(int* out) { std::memcpy(out, g_rows.data(), g_rows.size() * sizeof(int)); }

// simulate(): {n_propagators, n_solves, n_pairs, n_kernel_sums, peak}; mode 1 interleaved
int ref_simulate(const int* h, int stride, int n_steps, const int* origin, int mode, long long* out) {
  try {
    const Schedule s = build_schedule(Vec4{h[0], h[1], h[2], h[3]}, stride, n_steps,
                                      Vec4{origin[0], origin[1], origin[2], origin[3]});
    const ScheduleStats st = simulate(s, mode == 1 ? Mode::kInterleaved : Mode::kSerial);
    out[0] = st.n_propagators;
    out[1] = st.n_solves;
    out[2] = st.n_pairs;
    out[3] = st.n_kernel_sums;
    out[4] = st.peak_cache;
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

int ref_max_steps_without_wrap(const int* h, int stride, const int* lattice, const int* origin,
                               int cap) {
  return max_steps_without_wrap(Vec4{h[0], h[1], h[2], h[3]}, stride,
                                Vec4{lattice[0], lattice[1], lattice[2], lattice[3]},
                                Vec4{origin[0], origin[1], origin[2], origin[3]}, cap);
}

}  // extern "C"
