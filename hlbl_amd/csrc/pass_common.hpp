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
