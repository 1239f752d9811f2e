// Site indexing, coordinate maps and the cvc gamma basis, shared by host and
// device code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/hlbl_amd.h"

#define HLBL_HD __host__ __device__ inline

namespace hlbl {

// a mod L folded into [0, L) (include/hlbl/vec4.hpp:21-24)
HLBL_HD int mod_pos(int a, int L) {
  int r = a % L;
  return r < 0 ? r + L : r;
}
// x in [0, L) -> [-L/2, L/2), L/2 going to -L/2 (cuda_lattice.cu:128-130)
HLBL_HD int coord_map(int x, int L) { return x >= L / 2 ? x - L : x; }
// as coord_map but L/2 -> 0; only for the odd weights x_rho, z_rho
// (cuda_lattice.cu:136-138)
HLBL_HD int coord_map_zerohalf(int x, int L) {
  return x > L / 2 ? x - L : (x < L / 2 ? x : 0);
}
// minimum-image a - b in [-L/2, L/2) (cuda_lattice.cu:149-154)
HLBL_HD int coord_map_diff(int a, int b, int L) { return coord_map(mod_pos(a - b, L), L); }

HLBL_HD int geom_dim(const hlbl_geom& g, int mu) {
  return mu == 0 ? g.T : (mu == 1 ? g.LX : (mu == 2 ? g.LY : g.LZ));
}
HLBL_HD int coord_dim(const hlbl_coord& c, int mu) {
  return mu == 0 ? c.t : (mu == 1 ? c.x : (mu == 2 ? c.y : c.z));
}
HLBL_HD size_t geom_volume(const hlbl_geom& g) {
  return (size_t)g.T * g.LX * g.LY * g.LZ;
}

// Local site ((t*LX+x)*LY+y)*LZ+z -> local coordinates (tmLQCD order).
HLBL_HD void site_coords(size_t s, const hlbl_geom& l, int c[4]) {
  c[3] = (int)(s % l.LZ);
  s /= l.LZ;
  c[2] = (int)(s % l.LY);
  s /= l.LY;
  c[1] = (int)(s % l.LX);
  c[0] = (int)(s / l.LX);
}

// The (rho, sigma) pair of antisymmetric index k, rho < sigma, in the order
// (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) (cuda_lattice.cu:52-60).
HLBL_HD int comb_rho(int k) { return k < 3 ? 0 : (k < 5 ? 1 : 2); }
HLBL_HD int comb_sigma(int k) { return k < 3 ? k + 1 : (k < 5 ? k - 1 : 3); }

// ---- cvc gamma basis ------------------------------------------------------
// gamma_mu (mu = 0..3) and gamma_5 are monomial 4x4 spin matrices: row s has
// one entry, i^phase at column col. Explicitly (rows 0..3):
//   gamma_0 = -[[0,1],[1,0]] in 2x2 spin blocks
//   gamma_1: (-i at col 3, -i at col 2, +i at col 1, +i at col 0)
//   gamma_2: (-1 at col 3, +1 at col 2, +1 at col 1, -1 at col 0)
//   gamma_3: (-i at col 2, +i at col 3, +i at col 0, -i at col 1)
//   gamma_5 = diag(1, 1, -1, -1) = gamma_0 gamma_1 gamma_2 gamma_3
// Spin index s, colour kept: component c = 3*s + colour.
// Index 4 is gamma_5 here.
HLBL_HD int gamma_col(int g, int s) {
  return g == 4 ? s : ((g == 0 || g == 3) ? (s ^ 2) : (3 - s));
}
// phase as a power of i (0: 1, 1: i, 2: -1, 3: -i)
HLBL_HD int gamma_phase(int g, int s) {
  switch (g) {
    case 0: return 2;
    case 1: return s < 2 ? 3 : 1;
    case 2: return (s == 0 || s == 3) ? 2 : 0;
    case 3: return (s == 0 || s == 3) ? 3 : 1;
    default: return s < 2 ? 0 : 2;
  }
}

// Double-level action on a 24-double spinor, out[j] = in[perm[j]] * sign[j],
// derived from the monomial form above: a phase i^p on complex component
// (s, colour) maps (re, im) -> (re, im) * p=0, (-im, re) p=1, (-re, -im) p=2,
// (im, -re) p=3.
HLBL_HD int gamma_perm24(int g, int j) {
  int comp = j >> 1, part = j & 1;
  int s = comp / 3, colour = comp % 3;
  int src = 3 * gamma_col(g, s) + colour;
  int p = gamma_phase(g, s);
  int src_part = (p & 1) ? (part ^ 1) : part;
  return 2 * src + src_part;
}
HLBL_HD int gamma_sign24(int g, int j) {
  int comp = j >> 1, part = j & 1;
  int p = gamma_phase(g, comp / 3);
  // p=0: (+,+); p=1: (-,+); p=2: (-,-); p=3: (+,-)
  if (p == 0) return 1;
  if (p == 2) return -1;
  if (p == 1) return part == 0 ? -1 : 1;
  return part == 0 ? 1 : -1;
}

// The reference's gamma index: 0..3 are gamma_mu, 5 is gamma_5 (src/gamma.cpp).
HLBL_HD int gamma_internal(int ref_index) { return ref_index == 5 ? 4 : ref_index; }

}  // namespace hlbl
