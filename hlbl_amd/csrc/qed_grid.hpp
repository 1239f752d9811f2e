// Grid and file format of the tabulated plain QED kernel (no HIP dependency: shared
// by the library, its table generator and the tests' readers).
//
// The table holds, at each grid point (|x|, beta, |y|), the 42 frame components of
// the tensors T^I, T^II, T^III of the kernel (qed_exact.hpp): channel A*14 + c.
// Nodes: |x| = R i / (kNA-1) (uniform), beta = pi j / (kNBeta-1) (uniform; beta the
// angle between x and y), |y| = R (k / (kNB-1))^2 (clustered at |y| = 0, where the
// tensors behave like |y| log|y|), R = 12 fm in muon units. The kernel is 0 beyond R.
// Interpolation: cubic Lagrange in each of (|x|/h_a, beta/h_beta, sqrt(|y|/R) (kNB-1))
// on the four nodes around the point (the first or last four at the edges).
//
// File: little-endian; 64-byte header
//   uint32 magic 'HLQT', uint32 version, uint32 na, nbeta, nb, nch,
//   float64 R, uint32 quadrature nt, nr, ntheta, npsi, uint32 reserved[4]
// then float32 values [na][nbeta][nb][nch].
#pragma once
#include <cstdint>

namespace hlbl {
namespace qed_grid {

constexpr double kRadius = 12.0 * 105.6583745 / 197.3269804;  // 12 fm in muon units
constexpr int kNA = 33, kNBeta = 33, kNB = 49;
constexpr int kNChan = 42;  // 3 tensors x 14 frame components
constexpr uint32_t kMagic = 0x54514C48u;  // "HLQT"
constexpr uint32_t kVersion = 1;

struct FileHeader {
  uint32_t magic, version, na, nbeta, nb, nch;
  double radius;
  uint32_t nt, nr, ntheta, npsi;
  uint32_t reserved[4];
};
static_assert(sizeof(FileHeader) == 64, "64-byte header");

inline double node_a(int i) { return kRadius * i / (kNA - 1); }
inline double node_beta(int j) { return 3.14159265358979323846 * j / (kNBeta - 1); }
inline double node_b(int k) {
  const double s = (double)k / (kNB - 1);
  return kRadius * s * s;
}

}  // namespace qed_grid
}  // namespace hlbl
