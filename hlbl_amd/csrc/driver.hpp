// The schedule loop of a run (src/driver.cpp:23-165): rows, the 2+2 step, run_chain.
#pragma once

#include <functional>
#include <map>
#include <vector>

#include "../../include/hlbl_amd.h"
#include "context.hpp"
#include "schedule.hpp"

namespace hlbl {

struct PairRow {
  Vec4 x, y, yv;
  int window;
  double ks[2];
};

// one planned 2+2 source of the chain, and its computed pieces (host copies)
struct Row2p2 {
  Vec4 w;
  int window;
  std::vector<hlbl_coord> targets;
  std::vector<double> p1, p234;
};

// Rows in the reference's order: windows, new sources in solve order, partners;
// on_window(first, n) after each window with its rows [first, first + n), as the
// reference writes each window's rows when it is done (src/driver.cpp:134-137).
// batch 0: one pair per call; 1: one call per window (its sources solved first);
// 2: one call per new source with all its partners (the solve order and cache peak of
// batch 0, for lattices where a whole window's blocks do not fit). host_tier: pinned
// host blocks; batched calls read them from device staging blocks filled on a second
// stream (batch 2: the next call's blocks while one computes).
// plan_2p2: the chain's planned 2+2 sources by site (in [0, L)); a planned site solved
// in its planned window gets its pieces (flavour iflavor_2p2) from the solved block
// before it is stored (src/driver.cpp:58-95), and on_2p2(rows) receives each window's
// 2+2 rows after its kernel-sum rows.
void run_chain(hlbl_contractor* c, const Schedule& s, uint64_t seed, double xunit, bool host_tier,
               int form, int batch, std::vector<PairRow>& rows, double stats[5],
               const std::function<void(size_t, size_t)>& on_window,
               const std::map<Vec4, Row2p2>& plan_2p2, int iflavor_2p2,
               const std::function<void(std::vector<Row2p2>&)>& on_2p2);

}  // namespace hlbl
