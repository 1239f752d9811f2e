// Host-only parts of the QED kernel.
#pragma once

#include <string>
#include <vector>

#include "qed.hpp"

namespace hlbl {
namespace qed {
// the tabulated T tensors (qed_grid.hpp file format) as [kNA][kNC][kNB][kNCh] doubles
// (channels kNChUsed..kNCh-1 zero); throws std::runtime_error on a missing, truncated
// or differently gridded file
void load_table(const std::string& path, std::vector<double>& table);
}  // namespace qed
}  // namespace hlbl
