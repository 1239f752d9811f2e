// The position-space QED kernel L^(Lambda)_{[rho sigma]; mu nu lambda}(x, y),
// Lambda = 0.4, arguments in muon units, as used by ker_4pt_contraction
// (kernels/cuda_lattice.cu:266-275, 403-476) through KQED's QED_Mkernel_L2.
//
//   L0_{k mu nu lam}(x,y) = sum_{A, al, be, de} G^A_{k mu nu lam; al be de} T^A_{al be de}(x,y)
// with k the antisymmetric pair (rho, sigma) and
//   G^I   = 1/8 Tr(P_{de rho sigma} g_mu g_al g_nu g_be g_lam)
//   G^II  = -1/4 Tr(P_{de rho sigma} g_mu g_al g_nu) delta_{be lam}
//   G^III = -1/4 Tr(P_{de rho sigma} g_nu g_be g_lam) delta_{al mu}
//   P_{de rho sigma} = g_de [g_rho, g_sigma] + 2 (delta_{sigma de} g_rho - delta_{rho de} g_sigma)
// T^I, T^II, T^III the tensors of qed_exact.hpp (computed there from the defining
// integrals, tabulated on the grid of qed_grid.hpp), and the Lambda subtraction,
// which makes the kernel vanish at x = 0 and y = 0:
//   L(x,y) = L0(x,y) - d^x_mu(x_al e^{-Lambda x^2/2}) L0_{k al nu lam}(0,y)
//                    - d^y_nu(y_al e^{-Lambda y^2/2}) L0_{k mu al lam}(x,0)
// The kernel is 0 when |x| or |y| exceeds kRadius (12 fm).
//
// The table holds the T tensors in the frame of the argument pair: e0 = x^,
// e1 = the unit part of y orthogonal to x, P = the projector on the plane
// orthogonal to both; per tensor 14 components (channel A*14 + c): t_{ijk},
// i,j,k in {0,1}, at c = (i*2+j)*2+k, and t_{PPd}, t_{PdP}, t_{dPP} at 8+d, 10+d, 12+d:
//   T_{ijk} = sum t_{abd} e_a,i e_b,j e_d,k
//           + sum_d (t_{PPd} P_ij e_d,k + t_{PdP} P_ik e_d,j + t_{dPP} e_d,i P_jk).
// The components are interpolated cubically in (|x|, beta, sqrt|y|), beta the
// angle between x and y.
#pragma once

#include <cmath>

#include "lattice.hpp"
#include "qed_grid.hpp"

namespace hlbl {
namespace qed {

constexpr double kLambda = 0.4;
constexpr double kHbarC = 197.3269804;     // MeV fm
constexpr double kMuonMass = 105.6583745;  // MeV
constexpr double kRadius = qed_grid::kRadius;  // 12 fm in muon units (6.4253...)

// table grid (qed_grid.hpp)
constexpr int kNA = qed_grid::kNA;        // |x|
constexpr int kNC = qed_grid::kNBeta;     // beta
constexpr int kNB = qed_grid::kNB;        // |y|, clustered at 0
constexpr int kNChUsed = qed_grid::kNChan;  // 42 frame components
constexpr int kNCh = 48;                  // channel stride of a grid point (6 zero pad)
constexpr int kFrame = 14;                // components per tensor

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

// Cubic Lagrange weights of the four nodes i0..i0+3 around the position p (in node
// units) on a grid of n nodes: i0 = floor(p) - 1, clamped to [0, n-4].
template <class R>
HLBL_HD int lagrange4(R p, int n, R w[4]) {
  int i0 = (int)floor(p) - 1;
  i0 = i0 < 0 ? 0 : (i0 > n - 4 ? n - 4 : i0);
  const R t = p - R(i0);
  const R t1 = t - R(1), t2 = t - R(2), t3 = t - R(3);
  w[0] = -t1 * t2 * t3 * R(1.0 / 6.0);
  w[1] = t * t2 * t3 * R(0.5);
  w[2] = -t * t1 * t3 * R(0.5);
  w[3] = t * t1 * t2 * R(1.0 / 6.0);
  return i0;
}

template <class R>
struct Cell {
  int base;  // flat index of the first stencil node (ia0, ic0, ib0), in grid points
  R wa[4], wc[4], wb[4];
};

// Stencil of the grid coordinates (a, beta, b): a, b in [0, kRadius] (clamped),
// beta in [0, pi].
template <class R>
HLBL_HD Cell<R> locate(R a, R beta, R b) {
  const R ra = R(kNA - 1) / R(kRadius);
  const R rc = R(kNC - 1) / R(3.14159265358979323846);
  const R rb = R(1) / R(kRadius);
  a = a < R(0) ? R(0) : (a > R(kRadius) ? R(kRadius) : a);
  b = b < R(0) ? R(0) : (b > R(kRadius) ? R(kRadius) : b);
  Cell<R> cell;
  const int ia = lagrange4<R>(a * ra, kNA, cell.wa);
  const int ic = lagrange4<R>(beta * rc, kNC, cell.wc);
  const int ib = lagrange4<R>(R(sqrt(b * rb)) * R(kNB - 1), kNB, cell.wb);
  cell.base = (ia * kNC + ic) * kNB + ib;
  return cell;
}

// Tricubic interpolation of channel ch. table is [kNA][kNC][kNB][kNCh].
template <class R, class TT>
HLBL_HD R interp(const TT* __restrict__ table, const Cell<R>& q, int ch) {
  const int sa = kNC * kNB * kNCh, sc = kNB * kNCh, sb = kNCh;
  const TT* p0 = table + (size_t)q.base * kNCh + ch;
  R r = R(0);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R ri = R(0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const TT* p = p0 + i * sa + j * sc;
      const R s = q.wb[0] * (R)p[0] + q.wb[1] * (R)p[sb] + q.wb[2] * (R)p[2 * sb] + q.wb[3] * (R)p[3 * sb];
      ri += q.wc[j] * s;
    }
    r += q.wa[i] * ri;
  }
  return r;
}

// (|x|, beta, |y|) of an argument pair, beta = atan2(|x ^ y|, x.y) (accurate for
// nearly parallel arguments); beta = 0 when either argument is 0
template <class R>
HLBL_HD void grid_coords(const R x[4], const R y[4], R& a, R& beta, R& b) {
  const R u = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
  const R v = y[0] * y[0] + y[1] * y[1] + y[2] * y[2] + y[3] * y[3];
  const R w = x[0] * y[0] + x[1] * y[1] + x[2] * y[2] + x[3] * y[3];
  R s2 = R(0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i + 1; j < 4; ++j) {
      const R m = x[i] * y[j] - x[j] * y[i];
      s2 += m * m;
    }
  a = sqrt(u);
  b = sqrt(v);
  beta = (u > R(0) && v > R(0)) ? atan2(sqrt(s2), w) : R(0);
}

// ---- T tensors from the frame components ------------------------------------

// v[i] for a runtime i without indexing the array (keeps v in registers)
template <class R>
HLBL_HD R pick4(const R v[4], int i) {
  return i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3]));
}

// |x ^ y|^2 / (|x|^2 |y|^2) at or below which x and y count as parallel: far above
// the rounding of |x ^ y| for parallel arguments (with fused multiply-adds its
// terms x_i y_j - x_j y_i are the rounding error of one product, not 0), so that
// y - (y.e0) e0 is never a rounding remainder; the components that depend on e1
// vanish linearly in the angle, so below the threshold its choice is immaterial.
template <class R>
HLBL_HD constexpr R parallel_sin2() { return R(1e-24); }  // FP64: |sin beta| <= 1e-12
template <>
HLBL_HD constexpr float parallel_sin2<float>() { return 1e-10f; }  // FP32: <= 1e-5

// The frame of an argument pair: e0 = x^ (y^ when x = 0, axis 0 when both are 0),
// e1 = the unit part of y orthogonal to e0, or, when y is parallel to e0 (see
// parallel_sin2) or 0, the unit part orthogonal to e0 of the axis on which e0 is
// smallest (the first such axis).
template <class R>
HLBL_HD void frame(const R x[4], const R y[4], R e0[4], R e1[4]) {
  const R u = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + x[3] * x[3];
  const R v = y[0] * y[0] + y[1] * y[1] + y[2] * y[2] + y[3] * y[3];
  if (u > R(0)) {
    const R inv = R(1) / sqrt(u);
#pragma unroll
    for (int i = 0; i < 4; ++i) e0[i] = x[i] * inv;
  } else if (v > R(0)) {
    const R inv = R(1) / sqrt(v);
#pragma unroll
    for (int i = 0; i < 4; ++i) e0[i] = y[i] * inv;
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i) e0[i] = i == 0 ? R(1) : R(0);
  }
  R s2 = R(0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i + 1; j < 4; ++j) {
      const R m = x[i] * y[j] - x[j] * y[i];
      s2 += m * m;
    }
  const bool parallel = !(u > R(0) && v > R(0)) || s2 <= parallel_sin2<R>() * u * v;
  R c = R(0);
  R t[4];
  if (!parallel) {
#pragma unroll
    for (int i = 0; i < 4; ++i) c += y[i] * e0[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = y[i] - c * e0[i];
  } else {
    int m = 0;
    R am = fabs(e0[0]);
#pragma unroll
    for (int i = 1; i < 4; ++i)
      if (fabs(e0[i]) < am) {
        am = fabs(e0[i]);
        m = i;
      }
    const R em = pick4(e0, m);
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = (i == m ? R(1) : R(0)) - em * e0[i];
  }
  const R n = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + t[3] * t[3]);
  const R inv = R(1) / n;
#pragma unroll
  for (int i = 0; i < 4; ++i) e1[i] = t[i] * inv;
}

// What the three T components of one index j = (al*4+be)*4+de share: the frame
// vectors' entries at al, be, de and the projector entries.
template <class R>
struct TArgs {
  R p0[3], p1[3];  // e0 and e1 at al, be, de
  R Pab, Pad, Pbd;
};
template <class R>
HLBL_HD TArgs<R> t_args(int j, const R e0[4], const R e1[4]) {
  const int al = (j >> 4) & 3, be = (j >> 2) & 3, de = j & 3;
  TArgs<R> t;
  t.p0[0] = pick4(e0, al), t.p0[1] = pick4(e0, be), t.p0[2] = pick4(e0, de);
  t.p1[0] = pick4(e1, al), t.p1[1] = pick4(e1, be), t.p1[2] = pick4(e1, de);
  const R dab = al == be ? R(1) : R(0), dad = al == de ? R(1) : R(0), dbd = be == de ? R(1) : R(0);
  t.Pab = dab - t.p0[0] * t.p0[1] - t.p1[0] * t.p1[1];
  t.Pad = dad - t.p0[0] * t.p0[2] - t.p1[0] * t.p1[2];
  t.Pbd = dbd - t.p0[1] * t.p0[2] - t.p1[1] * t.p1[2];
  return t;
}

// Component A*64 + j of the T tensors from the 14 frame components c of tensor A
// (c = ch + 14 A), with t = t_args(j, e0, e1).
template <class R>
HLBL_HD R t_value(const TArgs<R>& t, const R* c) {
  // sum_{a,b,d} c[(a*2+b)*2+d] E_a[al] E_b[be] E_d[de]
  R r = R(0);
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const R ea = a == 0 ? t.p0[0] : t.p1[0];
    R rb = R(0);
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const R eb = b == 0 ? t.p0[1] : t.p1[1];
      rb += eb * (c[(a * 2 + b) * 2] * t.p0[2] + c[(a * 2 + b) * 2 + 1] * t.p1[2]);
    }
    r += ea * rb;
  }
  r += t.Pab * (c[8] * t.p0[2] + c[9] * t.p1[2]);
  r += t.Pad * (c[10] * t.p0[1] + c[11] * t.p1[1]);
  r += t.Pbd * (c[12] * t.p0[0] + c[13] * t.p1[0]);
  return r;
}

// Component i = A*64 + (al*4+be)*4+de of the T tensors from the channels ch[42]
// at the argument pair (x, y).
template <class R>
HLBL_HD R t_component(int i, const R x[4], const R y[4], const R* ch) {
  R e0[4], e1[4];
  frame(x, y, e0, e1);
  return t_value(t_args(i & 63, e0, e1), ch + kFrame * (i >> 6));
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
