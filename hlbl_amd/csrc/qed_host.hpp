// Host-only parts of the QED kernel: table builders.
#pragma once

#include <vector>

#include "qed.hpp"

namespace hlbl {
namespace qed {

// the 31 channels (plus pad) at (u, v, w) = (x^2, y^2, x.y)
void weight_channels(double u, double v, double w, double ch[kNCh]);
// the full table [kNA][kNC][kNB][kNCh]
void build_weight_table(std::vector<double>& table);

}  // namespace qed
}  // namespace hlbl
