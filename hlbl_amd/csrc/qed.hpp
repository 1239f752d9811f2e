// The position-space QED kernel L^(Lambda)_{[rho sigma]; mu nu lambda}(x, y),
// Lambda = 0.4, arguments in muon units, as used by ker_4pt_contraction
// (kernels/cuda_lattice.cu:266-275, 403-476) through KQED's QED_Mkernel_L2.
//
// Structure (position-space Mainz kernel):
//   L0_{k mu nu lam}(x,y) = sum_{A, al, be, de} G^A_{k mu nu lam; al be de} T^A_{al be de}(x,y)
// with k the antisymmetric pair (rho, sigma) and
//   G^I   = 1/8 Tr(P_{de rho sigma} g_mu g_al g_nu g_be g_lam)
//   G^II  = -1/4 Tr(P_{de rho sigma} g_mu g_al g_nu) delta_{be lam}
//   G^III = -1/4 Tr(P_{de rho sigma} g_nu g_be g_lam) delta_{al mu}
//   P_{de rho sigma} = g_de [g_rho, g_sigma] + 2 (delta_{sigma de} g_rho - delta_{rho de} g_sigma)
//   T^I_{al be de}   = d^x_al (d^x_be + d^y_be) V_de
//   T^II_{al be de}  = d^x_al (T_{be de} + delta_{be de} S / 4)
//   T^III_{al be de} = (d^x_be + d^y_be)(T_{al de} + delta_{al de} S / 4)
//   S = g0,  V_de = x_de g1 + y_de g2,
//   T_{al be} = (x_al x_be - delta x^2/4) l1 + (y_al y_be - delta y^2/4) l2
//             + (x_al y_be + y_al x_be - delta x.y/2) l3
// and the Lambda subtraction, which makes the kernel vanish at x = 0 and y = 0:
//   L(x,y) = L0(x,y) - d^x_mu(x_al e^{-Lambda x^2/2}) L0_{k al nu lam}(0,y)
//                    - d^y_nu(y_al e^{-Lambda y^2/2}) L0_{k mu al lam}(x,0)
// The kernel is 0 when |x| or |y| exceeds kRadius (12 fm).
//
// The scalar weight functions g0, g1, g2, l1, l2, l3 of (u, v, w) = (x^2, y^2, x.y)
// are tabulated with the derivatives the tensors need, on a grid in
// (|x|, cos beta, |y|), and the plain kernel interpolates the table trilinearly.
// Functions of (u, v, w) keep every derivative regular at x = 0 and y = 0.
#pragma once

#include <cmath>

#include "lattice.hpp"

namespace hlbl {
namespace qed {

constexpr double kLambda = 0.4;
constexpr double kHbarC = 197.3269804;        // MeV fm
constexpr double kMuonMass = 105.6583745;     // MeV
constexpr double kRadiusFm = 12.0;
constexpr double kRadius = kRadiusFm * kMuonMass / kHbarC;  // 6.4253... muon units

// table grid
constexpr int kNA = 49;   // |x| in [0, kRadius]
constexpr int kNC = 25;   // cos beta in [-1, 1]
constexpr int kNB = 49;   // |y| in [0, kRadius]
constexpr int kNCh = 32;  // channels per grid point (31 used, 1 pad)
constexpr int kNChUsed = 31;

// channels
enum Channel {
  G0U = 0, G0V, G0W,
  L1 = 3, L1U, L1V, L1W,
  L2 = 7, L2U, L2V, L2W,
  L3 = 11, L3U, L3V, L3W,
  G1U = 15, G1V, G1W, G1UU, G1UV, G1UW, G1VW, G1WW,
  G2U = 23, G2V, G2W, G2UU, G2UV, G2UW, G2VW, G2WW,
};

constexpr int kNT = 192;   // T^I, T^II, T^III: 3 x 64 components (al*4+be)*4+de
constexpr int kNL = 384;   // L: [6][4][4][4]
constexpr int kMaxTerms = 9;

// Sparse form of the G^A tensors: output o = ((k*4+mu)*4+nu)*4+lam has nterm[o]
// terms; term t is T index (low byte) and a coefficient in {-2,-1,1,2} (high byte,
// signed).
struct TermTable {
  unsigned short term[kNL][kMaxTerms];
  unsigned char nterm[kNL];
  unsigned short order[kNL];  // outputs sorted by decreasing nterm
};
// built on the host from gamma-matrix traces (qed_tables.cpp)
void build_term_table(TermTable* out);

HLBL_HD int term_index(unsigned short t) { return t & 0xff; }
HLBL_HD int term_coef(unsigned short t) { return (int)(signed char)(t >> 8); }

// ---- table interpolation --------------------------------------------------

template <class R>
struct Cell {
  int base;  // flat index of the (ia, ic, ib) corner, in grid points
  R fa, fc, fb;
};

// Locate (a, c, b) on the grid. a, b >= 0 (clamped to the last cell), c in [-1, 1].
template <class R>
HLBL_HD Cell<R> locate(R a, R c, R b) {
  const R ha = R(kNA - 1) / R(kRadius);
  const R hb = R(kNB - 1) / R(kRadius);
  const R hc = R(kNC - 1) / R(2);
  R ta = a * ha, tb = b * hb, tc = (c + R(1)) * hc;
  int ia = (int)floor(ta), ib = (int)floor(tb), ic = (int)floor(tc);
  ia = ia < 0 ? 0 : (ia > kNA - 2 ? kNA - 2 : ia);
  ib = ib < 0 ? 0 : (ib > kNB - 2 ? kNB - 2 : ib);
  ic = ic < 0 ? 0 : (ic > kNC - 2 ? kNC - 2 : ic);
  Cell<R> cell;
  cell.base = (ia * kNC + ic) * kNB + ib;
  cell.fa = ta - R(ia);
  cell.fc = tc - R(ic);
  cell.fb = tb - R(ib);
  return cell;
}

// Trilinear interpolation of channel ch. table is [kNA][kNC][kNB][kNCh].
template <class R, class TT>
HLBL_HD R interp(const TT* __restrict__ table, const Cell<R>& q, int ch) {
  const int sa = kNC * kNB * kNCh, sc = kNB * kNCh, sb = kNCh;
  const TT* p = table + (size_t)q.base * kNCh + ch;
  R v000 = (R)p[0], v001 = (R)p[sb], v010 = (R)p[sc], v011 = (R)p[sc + sb];
  R v100 = (R)p[sa], v101 = (R)p[sa + sb], v110 = (R)p[sa + sc], v111 = (R)p[sa + sc + sb];
  R v00 = v000 + q.fb * (v001 - v000);
  R v01 = v010 + q.fb * (v011 - v010);
  R v10 = v100 + q.fb * (v101 - v100);
  R v11 = v110 + q.fb * (v111 - v110);
  R v0 = v00 + q.fc * (v01 - v00);
  R v1 = v10 + q.fc * (v11 - v10);
  return v0 + q.fa * (v1 - v0);
}

// (|x|, cos beta, |y|) of an argument pair; cos beta = 0 when either is 0
template <class R>
HLBL_HD void acb(const R x[4], const R y[4], R& a, R& c, R& b) {
  R u = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
  R v = y[0] * y[0] + y[1] * y[1] + y[2] * y[2] + y[3] * y[3];
  R w = x[0] * y[0] + x[1] * y[1] + x[2] * y[2] + x[3] * y[3];
  a = sqrt(u);
  b = sqrt(v);
  c = (a > R(0) && b > R(0)) ? w / (a * b) : R(0);
  c = c > R(1) ? R(1) : (c < R(-1) ? R(-1) : c);
}

// ---- T tensors -------------------------------------------------------------

// v[i] for a runtime i without indexing the array (keeps v in registers)
template <class R>
HLBL_HD R pick4(const R v[4], int i) {
  return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3]));
}

// d^x_al F = 2 F_u x_al + F_w y_al
template <class R>
HLBL_HD R dx(R fu, R fw, R xa, R ya) {
  return R(2) * fu * xa + fw * ya;
}
// (d^x_be + d^y_be) F = (2 F_u + F_w) x_be + (2 F_v + F_w) y_be
template <class R>
HLBL_HD R dsum(R fu, R fv, R fw, R xb, R yb) {
  return (R(2) * fu + fw) * xb + (R(2) * fv + fw) * yb;
}
// d^x_al (d^x_be + d^y_be) F, from F's channels starting at base (U V W UU UV UW VW WW)
template <class R>
HLBL_HD R dxdsum(const R* f, R xa, R ya, R xb, R yb, bool ab) {
  R fu = f[0], fw = f[2], fuu = f[3], fuv = f[4], fuw = f[5], fvw = f[6], fww = f[7];
  R r = ab ? (R(2) * fu + fw) : R(0);
  r += (R(2) * (R(2) * fuu + fuw) * xa + (R(2) * fuw + fww) * ya) * xb;
  r += (R(2) * (R(2) * fuv + fuw) * xa + (R(2) * fvw + fww) * ya) * yb;
  return r;
}

// What the three T components of one index j = (al*4+be)*4+de share at the argument
// pair (x, y): the picked coordinates, the Kronecker deltas and the invariants.
template <class R>
struct TArgs {
  R xa, xb, xd, ya, yb, yd, dab, dad, dbd, u, v, w;
  bool ab;
};
template <class R>
HLBL_HD TArgs<R> t_args(int j, const R x[4], const R y[4]) {
  const int al = (j >> 4) & 3, be = (j >> 2) & 3, de = j & 3;
  TArgs<R> t;
  t.ab = al == be;
  t.dab = al == be ? R(1) : R(0);
  t.dad = al == de ? R(1) : R(0);
  t.dbd = be == de ? R(1) : R(0);
  t.xa = pick4(x, al), t.xb = pick4(x, be), t.xd = pick4(x, de);
  t.ya = pick4(y, al), t.yb = pick4(y, be), t.yd = pick4(y, de);
  t.u = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
  t.v = y[0] * y[0] + y[1] * y[1] + y[2] * y[2] + y[3] * y[3];
  t.w = x[0] * y[0] + x[1] * y[1] + x[2] * y[2] + x[3] * y[3];
  return t;
}

// Component A*64 + j of the T tensors from the channels ch, with t = t_args(j, x, y).
template <class R>
HLBL_HD R t_value(int A, const TArgs<R>& t, const R* ch) {
  const R xa = t.xa, xb = t.xb, xd = t.xd, ya = t.ya, yb = t.yb, yd = t.yd;
  const R dab = t.dab, dad = t.dad, dbd = t.dbd;
  if (A == 0) {
    R r = dbd * (dx(ch[G1U], ch[G1W], xa, ya) + dx(ch[G2U], ch[G2W], xa, ya));
    r += dad * dsum(ch[G1U], ch[G1V], ch[G1W], xb, yb);
    r += xd * dxdsum(ch + G1U, xa, ya, xb, yb, t.ab);
    r += yd * dxdsum(ch + G2U, xa, ya, xb, yb, t.ab);
    return r;
  }
  const R u = t.u, v = t.v, w = t.w;
  const R q = R(0.25), h = R(0.5);
  if (A == 1) {
    R r = (dab * xd + xb * dad - h * dbd * xa) * ch[L1];
    r += (xb * xd - q * dbd * u) * dx(ch[L1U], ch[L1W], xa, ya);
    r += (yb * yd - q * dbd * v) * dx(ch[L2U], ch[L2W], xa, ya);
    r += (dab * yd + yb * dad - h * dbd * ya) * ch[L3];
    r += (xb * yd + yb * xd - h * dbd * w) * dx(ch[L3U], ch[L3W], xa, ya);
    r += q * dbd * dx(ch[G0U], ch[G0W], xa, ya);
    return r;
  }
  R r = (dab * xd + xa * dbd - h * dad * xb) * ch[L1];
  r += (xa * xd - q * dad * u) * dsum(ch[L1U], ch[L1V], ch[L1W], xb, yb);
  r += (dab * yd + ya * dbd - h * dad * yb) * ch[L2];
  r += (ya * yd - q * dad * v) * dsum(ch[L2U], ch[L2V], ch[L2W], xb, yb);
  r += (dab * (xd + yd) + dbd * (xa + ya) - h * dad * (xb + yb)) * ch[L3];
  r += (xa * yd + ya * xd - h * dad * w) * dsum(ch[L3U], ch[L3V], ch[L3W], xb, yb);
  r += q * dad * dsum(ch[G0U], ch[G0V], ch[G0W], xb, yb);
  return r;
}

// Component i = A*64 + (al*4+be)*4+de of the T tensors, from the channels ch
// at the argument pair (x, y).
template <class R>
HLBL_HD R t_component(int i, const R x[4], const R y[4], const R* ch) {
  return t_value(i >> 6, t_args(i & 63, x, y), ch);
}

// L0 output o from the T tensors
template <class R>
HLBL_HD R l0_component(const TermTable& tt, int o, const R* t) {
  R r = R(0);
  const int n = tt.nterm[o];
  for (int j = 0; j < n; ++j) {
    unsigned short e = tt.term[o][j];
    r += R(term_coef(e)) * t[term_index(e)];
  }
  return r;
}

// D_{mu al}(x) = d^x_mu(x_al e^{-Lambda x^2/2}) = (delta_{mu al} - Lambda x_mu x_al) e^{-Lambda x^2/2}
template <class R>
HLBL_HD void subtraction_matrix(const R x[4], R d[16]) {
  R u = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
  R e = exp(-R(kLambda) * u * R(0.5));
  for (int m = 0; m < 4; ++m)
    for (int a = 0; a < 4; ++a) d[m * 4 + a] = ((m == a ? R(1) : R(0)) - R(kLambda) * x[m] * x[a]) * e;
}

// L(x,y)[o] = L0xy[o] - sum_al Dx[mu][al] L0(0,y)[k al nu lam] - sum_al Dy[nu][al] L0(x,0)[k mu al lam]
// (P: a pointer, or any type indexed by [] as one, such as a lane-strided view)
template <class R, class P>
HLBL_HD R lambda_subtract(int o, P l0xy, P l00y, P l0x0, const R dxm[16], const R dyn[16]) {
  const int lam = o & 3, nu = (o >> 2) & 3, mu = (o >> 4) & 3, k = o >> 6;
  R r = l0xy[o];
  for (int al = 0; al < 4; ++al) {
    r -= dxm[mu * 4 + al] * l00y[((k * 4 + al) * 4 + nu) * 4 + lam];
    r -= dyn[nu * 4 + al] * l0x0[((k * 4 + mu) * 4 + al) * 4 + lam];
  }
  return r;
}

HLBL_HD bool inside(double a, double b) { return a <= kRadius && b <= kRadius; }

}  // namespace qed
}  // namespace hlbl
