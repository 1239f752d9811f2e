// Device helpers shared by the FP64 and the mixed-precision passes.
#pragma once

#include "lattice.hpp"

namespace hlbl {

// multiply (re, im) by i^q
template <class R>
HLBL_HD void mul_ipow(R& re, R& im, int q) {
  const R r = re, m = im;
  switch (q & 3) {
    case 0: break;
    case 1: re = -m; im = r; break;
    case 2: re = -r; im = -m; break;
    default: re = m; im = -r; break;
  }
}

// dxu_mu[r][a] for mu = 0..3 of one site, from S[1-f][a] (sa) and Y[f][r] (ur),
// 24 elements each (complex interleaved):
//   dxu_mu[r][a] = - sum_c conj((g5 gamma_mu S_a)[c]) Y_r[c]        (cuda_lattice.cu:324-344)
//                = - sum_s eps_s i^{-ph_mu(s)} Xi[cls(mu)][s],
//   Xi[cls][s] = sum_colour conj(S_a[(s', colour)]) Y_r[(s, colour)],
// s' = s^2 (cls 0: gamma_0, gamma_3) or 3-s (cls 1: gamma_1, gamma_2).
// out[mu * stride] receives (re, im). Inputs R, arithmetic A, output O.
template <class R, class A, class O = R>
__device__ inline void dxu_site(const R* __restrict__ sa, const R* __restrict__ ur, O* out,
                                int stride) {
  A xr[2][4], xim[2][4];
#pragma unroll
  for (int cls = 0; cls < 2; ++cls)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int spn = cls == 0 ? (s ^ 2) : (3 - s);
      A re = 0, im = 0;
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        const A tr = sa[2 * (3 * spn + cc)], ti = sa[2 * (3 * spn + cc) + 1];
        const A vr = ur[2 * (3 * s + cc)], vi = ur[2 * (3 * s + cc) + 1];
        re += tr * vr + ti * vi;
        im += tr * vi - ti * vr;
      }
      xr[cls][s] = re;
      xim[cls][s] = im;
    }
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    const int cls = (mu == 0 || mu == 3) ? 0 : 1;
    A re = 0, im = 0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      A a_r = xr[cls][s], a_i = xim[cls][s];
      mul_ipow(a_r, a_i, 4 - gamma_phase(mu, s));
      const A eps = s < 2 ? A(1) : A(-1);
      re += eps * a_r;
      im += eps * a_i;
    }
    out[mu * stride] = (O)(-re);
    out[mu * stride + 1] = (O)(-im);
  }
}

// dxu_mu[r][a] of dxu_site, mu = 0..3, returned in (re[mu], im[mu]); the sink
// spins unrolled, so the phases i^q are constants.
template <class R, class A>
__device__ inline void dxu_vals(const R* __restrict__ sa, const R* __restrict__ ur, A (&re)[4],
                                A (&im)[4]) {
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) re[mu] = im[mu] = A(0);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int s0 = s ^ 2, s1 = 3 - s;
    A x0r = 0, x0i = 0, x1r = 0, x1i = 0;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const A vr = ur[2 * (3 * s + cc)], vi = ur[2 * (3 * s + cc) + 1];
      const A t0r = sa[2 * (3 * s0 + cc)], t0i = sa[2 * (3 * s0 + cc) + 1];
      const A t1r = sa[2 * (3 * s1 + cc)], t1i = sa[2 * (3 * s1 + cc) + 1];
      // conj(t) v as four fused multiply-adds each
      x0r = fma(t0r, vr, x0r);
      x0r = fma(t0i, vi, x0r);
      x0i = fma(t0r, vi, x0i);
      x0i = fma(-t0i, vr, x0i);
      x1r = fma(t1r, vr, x1r);
      x1r = fma(t1i, vi, x1r);
      x1i = fma(t1r, vi, x1i);
      x1i = fma(-t1i, vr, x1i);
    }
    // dxu_mu = -sum_s eps_s i^{-ph_mu(s)} Xi[cls(mu)][s]: the sign -eps_s folded in
    const bool neg = s < 2;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      const bool c0 = (mu == 0 || mu == 3);
      A ar = c0 ? x0r : x1r, ai = c0 ? x0i : x1i;
      mul_ipow(ar, ai, 4 - gamma_phase(mu, s));
      re[mu] += neg ? -ar : ar;
      im[mu] += neg ? -ai : ai;
    }
  }
}

// dxu_vals with the spin sums paired: the phases of gamma_3 differ from those of gamma_0
// by +-i spin by spin, as those of gamma_1 from gamma_2, so with the signed sums
//   A+ = Xi0(1) - Xi0(2), A- = Xi0(0) - Xi0(3), B+ = -Xi1(1) - Xi1(3), B- = Xi1(0) + Xi1(2)
// (each six colour-contracted products, the signs folded into the multiply-adds)
//   dxu_0 = A+ + A-, dxu_3 = i (A+ - A-), dxu_2 = B+ + B-, dxu_1 = i (B+ - B-):
// 96 multiply-adds and 8 adds, where dxu_vals spends 32 adds on the phases.
template <class R, class A>
HLBL_HD void dxu_vals_paired(const R* __restrict__ sa, const R* __restrict__ ur, A (&re)[4],
                             A (&im)[4]) {
  A ap_r = 0, ap_i = 0, am_r = 0, am_i = 0, bp_r = 0, bp_i = 0, bm_r = 0, bm_i = 0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int s0 = s ^ 2, s1 = 3 - s;
    // class 0 (Xi0, S spin s ^ 2): s = 1 -> A+ (+), 2 -> A+ (-), 0 -> A- (+), 3 -> A- (-)
    // class 1 (Xi1, S spin 3 - s): s = 0, 2 -> B- (+), s = 1, 3 -> B+ (-)
    const A g0 = (s == 2 || s == 3) ? A(-1) : A(1);
    const A g1 = (s == 1 || s == 3) ? A(-1) : A(1);
    A& c0r = (s == 1 || s == 2) ? ap_r : am_r;
    A& c0i = (s == 1 || s == 2) ? ap_i : am_i;
    A& c1r = (s == 1 || s == 3) ? bp_r : bm_r;
    A& c1i = (s == 1 || s == 3) ? bp_i : bm_i;
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      const A vr = ur[2 * (3 * s + cc)], vi = ur[2 * (3 * s + cc) + 1];
      const A t0r = g0 * A(sa[2 * (3 * s0 + cc)]), t0i = g0 * A(sa[2 * (3 * s0 + cc) + 1]);
      const A t1r = g1 * A(sa[2 * (3 * s1 + cc)]), t1i = g1 * A(sa[2 * (3 * s1 + cc) + 1]);
      // +- conj(t) v as four fused multiply-adds each (the signs are constants)
      c0r = fma(t0r, vr, c0r);
      c0r = fma(t0i, vi, c0r);
      c0i = fma(t0r, vi, c0i);
      c0i = fma(-t0i, vr, c0i);
      c1r = fma(t1r, vr, c1r);
      c1r = fma(t1i, vi, c1r);
      c1i = fma(t1r, vi, c1i);
      c1i = fma(-t1i, vr, c1i);
    }
  }
  re[0] = ap_r + am_r;
  im[0] = ap_i + am_i;
  re[3] = am_i - ap_i;
  im[3] = ap_r - am_r;
  re[2] = bp_r + bm_r;
  im[2] = bp_i + bm_i;
  re[1] = bm_i - bp_i;
  im[1] = bp_r - bm_r;
}

// Element (kidx = r*24 + j, col < 40) of flavour f of the pass-2 B operand, from
// gdz[f][10][4][12][12] (complex): j = 2i: X_re[i][r], j = 2i+1: -X_im[i][r],
// X = G[k][nu] (col = k*4+nu < 24, gdz rows 0..5: dzu) or H[s][nu]
// (col = 24 + s*4+nu, gdz rows 6..9: dzsu).
HLBL_HD double bmat_value(const double* gdz, int f, int kidx, int col) {
  const int r = kidx / 24, jj = kidx % 24, ii = jj >> 1, part = jj & 1;
  const int jrow = col < 24 ? (col >> 2) : 6 + ((col - 24) >> 2);
  const int nu = col & 3;
  const double* x = gdz + (((((size_t)f * 10 + jrow) * 4 + nu) * 12 + ii) * 12 + r) * 2;
  return part ? -x[1] : x[0];
}

typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
// (x, y) = hi + lo in FP16, each packed as (x bits, y bits) (v_cvt_pk_f16_f32: two
// values per conversion, nearest, ties to even)
__device__ inline void f16_split2(float x, float y, unsigned& hi, unsigned& lo) {
  const f2v v = {x, y};
  const h2v h = __builtin_convertvector(v, h2v);
  const h2v l = __builtin_convertvector(v - __builtin_convertvector(h, f2v), h2v);
  __builtin_memcpy(&hi, &h, 4);
  __builtin_memcpy(&lo, &l, 4);
}

// (x, y) * sc (sc a power of two) = hi + lo in FP16, each packed as (x bits, y bits):
// hi by one conversion of the scaled pair; lo = x sc - hi by a multiply-add that reads hi
// as FP16 (a mixed-precision multiply-add), then one conversion. The same values as
// f16_split2(x sc, y sc).
__device__ inline void f16_split2_scaled(float x, float y, float sc, unsigned& hi, unsigned& lo) {
  const f2v v = {x * sc, y * sc};
  const h2v h = __builtin_convertvector(v, h2v);
  const f2v r = {__builtin_fmaf(x, sc, -(float)h[0]), __builtin_fmaf(y, sc, -(float)h[1])};
  const h2v l = __builtin_convertvector(r, h2v);
  __builtin_memcpy(&hi, &h, 4);
  __builtin_memcpy(&lo, &l, 4);
}

// 2^e as a float, e in [-126, 127]
__device__ inline float pow2f(int e) { return __int_as_float((e + 127) << 23); }

// Kahan-compensated accumulation
struct Kahan {
  double s = 0.0, c = 0.0;
  __device__ inline void add(double v) {
    const double y = v - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

}  // namespace hlbl
