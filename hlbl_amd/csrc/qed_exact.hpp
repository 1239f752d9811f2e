// The plain position-space QED kernel from its defining integrals (host only).
//
// The reference contracts with KQED's L^(Lambda) (QED_Mkernel_L2(0.4, x, y, ...),
// kernels/cuda_lattice.cu:266-275, called at :435, :448, :461; KQED itself is a
// third-party library absent from the reference tree). This file computes the same
// kernel, in muon units (m = 1), from the Mainz position-space representation:
//
//   L0_{[rho sigma]; mu nu lam}(x, y) = sum_A G^A_{de [rho sigma] mu al nu be lam} T^A_{al be de}(x, y)
//
// (G^A: gamma-matrix traces, built in qed_tables.cpp), with the three T tensors
//
//   T^I_{al be de}   = d^x_al (d^x_be + d^y_be) < eps_de I >
//   T^II_{al be de}  = d^x_al < eps_be eps_de I >
//   T^III_{al be de} = (d^x_be + d^y_be) < eps_al eps_de I >
//
// of the master function (< . > the average over the unit 4-vector eps, the
// direction of the muon momentum p = i m eps)
//
//   I(eps; x, y) = int d^4u G0(y - u) J(eps; x - u) J(eps; u),
//   J(eps; z)    = int d^4w G0(z - w) e^{eps.w} G_m(w)                       (1)
//               = 1/(8 pi^2) int_0^1 dt e^{t eps.z} K0(t |z|),                (2)
//
// G0(z) = 1/(4 pi^2 z^2), G_m(z) = K1(|z|)/(4 pi^2 |z|): the muon line with the photons
// of x, y and 0 attached in that order, whose two inner muon propagators e^{eps.v}
// (-gamma.d + m) G_m(v) give the derivatives (d^x on the first, d^x + d^y = -d^0 on
// the second) and the mass terms (the projector turns the adjacent gamma into
// 2 eps). (2) follows from (1) by a Feynman parameter: the Fourier transform of J is
// 1/(k^2 (k^2 + 2i eps.k)).
//
// With the derivatives moved under the u integral,
//   T^A(x, y) = int d^4u G0(y - u) Q^A(x - u, u)
//   Q^I_{al be de}(a, b)   = < eps_de dJ_al(eps; a) dJ_be(eps; b) >
//   Q^II_{al be de}(a, b)  = < eps_be eps_de dJ_al(eps; a) J(eps; b) >
//   Q^III_{al be de}(a, b) = < eps_al eps_de J(eps; a) dJ_be(eps; b) >
// and, from (2), dJ_al(eps; z) = 1/(8 pi^2) int dt t e^{t eps.z} [eps_al K0(t|z|) - z^_al K1(t|z|)].
// The eps averages of e^{eps.A}, A = t a + s b, are closed forms,
//   <eps_i e> = A_i Psi1,  <eps_i eps_j e> = delta_ij Psi1 + A_i A_j Psi2,
//   <eps_i eps_j eps_k e> = (delta_ij A_k + delta_ik A_j + delta_jk A_i) Psi2 + A_i A_j A_k Psi3,
//   Psi_k(R) = 2 I_{k+1}(R) / R^{k+1},
// so each Q is a two-dimensional integral over the Feynman parameters (t, s),
// evaluated here by Gauss-Legendre in t = tau^3 (which tames the log of K0 at t = 0).
// The u integral runs over the half space (u0, u1, rho >= 0) of the plane of x and y
// (the two directions orthogonal to it averaged analytically), split by a smooth
// partition of unity (Becke's) into cells around the three points 0, x, y where the
// integrand is singular, each cell in spherical coordinates about its point with a
// radial map r = R0 z/(1-z) to infinity.
//
// The result is given in the frame e0 = x^, e1 = the unit part of y orthogonal to x,
// P = the projector on the two directions orthogonal to both: per tensor A the
// 14 components T_{ijk}, i,j,k in {0,1} ((i*2+j)*2+k), T_{PPd} (8+d), T_{PdP}
// (10+d), T_{dPP} (12+d); the full tensor is
//   T_{ijk} = sum t_{abd} e_a e_b e_d + sum_d (t_{PPd} P_ij e_d,k + t_{PdP} P_ik e_d,j + t_{dPP} e_d,i P_jk).
// The normalisation and the orientation of eps (the sign of T^I against T^II and
// T^III) are those for which the reference's golden kernel sums and 2+2 sums
// (sanity_check/test_contract_random.cpp:58, :126-127) are reproduced.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

namespace hlbl {
namespace qed_exact {


constexpr double kPi = 3.14159265358979323846;
constexpr double kEuler = 0.57721566490153286061;

// e^x K_0(x), e^x K_1(x), x > 0
inline void kbessel_scaled(double x, double& k0, double& k1) {
  if (x <= 2.0) {
    const double y = x * x / 4, l = std::log(x / 2);
    double term = 1, i0 = 0, i1 = 0, s0 = 0, s1 = 0, h = 0;
    // term_k = y^k / (k!)^2 ; i1 term = (x/2) y^k/(k!(k+1)!)
    for (int k = 0; k < 40; ++k) {
      if (k > 0) { term *= y / ((double)k * k); h += 1.0 / k; }
      i0 += term;
      const double t1 = term / (k + 1);
      i1 += t1;
      s0 += term * h;
      const double psi1 = -kEuler + h, psi2 = -kEuler + h + 1.0 / (k + 1);
      s1 += t1 * (psi1 + psi2);
      if (term < 1e-18 * i0 && k > 2) break;
    }
    i1 *= x / 2;
    const double K0 = -(l + kEuler) * i0 + s0;
    const double K1 = 1.0 / x + l * i1 - (x / 4) * s1;
    const double e = std::exp(x);
    k0 = K0 * e;
    k1 = K1 * e;
    return;
  }
  // trapezoid on e^x K_nu(x) = int_0^inf exp(-x (cosh t - 1)) cosh(nu t) dt
  double h = std::min(0.25, 0.7 / std::sqrt(x));
  double a0 = 0.5, a1 = 0.5;  // t = 0
  for (int i = 1; i < 400; ++i) {
    const double t = i * h, c = std::cosh(t);
    const double e = std::exp(-x * (c - 1));
    a0 += e;
    a1 += e * c;
    if (e < 1e-18) break;
  }
  k0 = a0 * h;
  k1 = a1 * h;
}

// e^{-R} Psi_k(R) for k = 1, 2, 3, Psi_k(R) = 2 I_{k+1}(R) / R^{k+1}
inline void psi_scaled(double R, double& p1, double& p2, double& p3) {
  if (R <= 25.0) {
    const double y = R * R / 4;
    // Psi_k = 2^{-k} sum_j y^j / (j! (j+k+1)!)
    double t = 1.0;  // y^j / j!
    double s1 = 0, s2 = 0, s3 = 0;
    double f2 = 2, f3 = 6, f4 = 24;  // (j+2)!, (j+3)!, (j+4)!
    for (int j = 0; j < 120; ++j) {
      if (j > 0) {
        t *= y / j;
        f2 *= (j + 2);
        f3 *= (j + 3);
        f4 *= (j + 4);
      }
      s1 += t / f2;
      s2 += t / f3;
      s3 += t / f4;
      if (t / f2 < 1e-19 * s1 && j > y) break;
    }
    const double e = std::exp(-R);
    p1 = s1 / 2 * e;
    p2 = s2 / 4 * e;
    p3 = s3 / 8 * e;
    return;
  }
  // asymptotic e^{-R} I_nu(R) = (2 pi R)^{-1/2} sum_m (-1)^m a_m(nu) / R^m
  auto ive = [&](int nu) {
    const double mu = 4.0 * nu * nu;
    double term = 1, sum = 1;
    for (int m = 1; m < 60; ++m) {
      term *= -(mu - (2.0 * m - 1) * (2.0 * m - 1)) / (8.0 * m * R);
      sum += term;
      if (std::fabs(term) < 1e-18 * std::fabs(sum)) break;
    }
    return sum / std::sqrt(2 * kPi * R);
  };
  p1 = 2 * ive(2) / (R * R);
  p2 = 2 * ive(3) / (R * R * R);
  p3 = 2 * ive(4) / (R * R * R * R);
}

inline void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5)), pp = 0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1, p2 = 0;
      for (int j = 0; j < n; ++j) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2 * j + 1) * z * p2 - j * p3) / (j + 1);
      }
      pp = n * (z * p1 - p2) / (z * z - 1);
      double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) < 1e-16) break;
    }
    x[i] = (1 - z) / 2;  // on [0,1]
    w[i] = 1.0 / ((1 - z * z) * pp * pp);  // 2/((1-z^2)pp^2) * 1/2
  }
}

struct Config {
  int nt = 20;       // Gauss nodes in t and in s
  int tpow = 3;      // t = tau^tpow
  int nr = 48, nth = 12, nps = 24;  // per-centre radial, polar, azimuthal nodes
  double rscale = 1.0;
};

// 42 frame components: T^A, A = 0..2, in the frame x = |x| e0, y in the (e0, e1) plane:
// [A][0..7]  = T_{ijk}, i,j,k in {0,1}, index (i*2+j)*2+k
// [A][8+d]   = T_{PPd}, [A][10+d] = T_{PdP}, [A][12+d] = T_{dPP}
struct Tframe {
  double c[3][14];
};

struct Accum {
  // E3 sets for W1 (I), W5 (II), W7 (III): psi2 p, psi2 q, psi3 ppp, ppq, pqq, qqq
  double e3[3][6];
  // E2 sets for W2 (I, b_be), W3 (I, a_al), W6 (II, a_al), W8 (III, b_be): psi1, psi2 pp, pq, qq
  double e2[4][4];
  // E1 for W4: psi1 p, psi1 q
  double e1[2];
};

class Evaluator {
 public:
  explicit Evaluator(const Config& c) : cfg_(c) {
    std::vector<double> x, w;
    gauss_legendre(cfg_.nt, x, w);
    tn_.resize(cfg_.nt);
    tw_.resize(cfg_.nt);
    for (int i = 0; i < cfg_.nt; ++i) {
      const double tau = x[i];
      tn_[i] = std::pow(tau, cfg_.tpow);
      tw_[i] = w[i] * cfg_.tpow * std::pow(tau, cfg_.tpow - 1);
    }
    gauss_legendre(cfg_.nr, rx_, rw_);
    gauss_legendre(cfg_.nth, thx_, thw_);
  }

  // Q^A(a, b) components at the u-point, accumulated as scalar integrals
  void q_scalars(double ra, double rb, double gam, Accum& acc) const {
    const int n = cfg_.nt;
    double K0a[64], K1a[64], K0b[64], K1b[64];
    for (int i = 0; i < n; ++i) {
      kbessel_scaled(tn_[i] * ra, K0a[i], K1a[i]);
      kbessel_scaled(tn_[i] * rb, K0b[i], K1b[i]);
    }
    std::fill(&acc.e3[0][0], &acc.e3[0][0] + 18, 0.0);
    std::fill(&acc.e2[0][0], &acc.e2[0][0] + 16, 0.0);
    acc.e1[0] = acc.e1[1] = 0;
    for (int i = 0; i < n; ++i) {
      const double t = tn_[i], p = t * ra;
      for (int j = 0; j < n; ++j) {
        const double s = tn_[j], q = s * rb;
        double R2 = p * p + q * q + 2 * p * q * gam;
        if (R2 < 0) R2 = 0;
        const double R = std::sqrt(R2);
        double s1, s2, s3;
        psi_scaled(R, s1, s2, s3);
        const double sc = std::exp(R - p - q) * tw_[i] * tw_[j];
        const double k00 = K0a[i] * K0b[j] * sc, k01 = K0a[i] * K1b[j] * sc;
        const double k10 = K1a[i] * K0b[j] * sc, k11 = K1a[i] * K1b[j] * sc;
        const double pp = p * p, pq = p * q, qq = q * q;
        const double m3[6] = {s2 * p, s2 * q, s3 * pp * p, s3 * pp * q, s3 * p * qq, s3 * qq * q};
        const double m2[4] = {s1, s2 * pp, s2 * pq, s2 * qq};
        const double w1 = t * s * k00, w5 = t * k00, w7 = s * k00;
        const double w2 = t * s * k01, w3 = t * s * k10, w6 = t * k10, w8 = s * k01;
        const double w4 = t * s * k11;
        for (int m = 0; m < 6; ++m) {
          acc.e3[0][m] += w1 * m3[m];
          acc.e3[1][m] += w5 * m3[m];
          acc.e3[2][m] += w7 * m3[m];
        }
        for (int m = 0; m < 4; ++m) {
          acc.e2[0][m] += w2 * m2[m];
          acc.e2[1][m] += w3 * m2[m];
          acc.e2[2][m] += w6 * m2[m];
          acc.e2[3][m] += w8 * m2[m];
        }
        acc.e1[0] += w4 * s1 * p;
        acc.e1[1] += w4 * s1 * q;
      }
    }
  }

  // full Q tensors [3][4][4][4] from the accumulators and unit vectors ah, bh
  static void q_tensors(const Accum& acc, const double ah[4], const double bh[4], double Q[3][64]) {
    auto d = [](int i, int j) { return i == j ? 1.0 : 0.0; };
    auto E3 = [&](const double* c, int i, int j, int k) {
      auto A = [&](int l) { return c[0] * ah[l] + c[1] * bh[l]; };
      double r = d(i, j) * A(k) + d(i, k) * A(j) + d(j, k) * A(i);
      r += c[2] * ah[i] * ah[j] * ah[k];
      r += c[3] * (ah[i] * ah[j] * bh[k] + ah[i] * bh[j] * ah[k] + bh[i] * ah[j] * ah[k]);
      r += c[4] * (ah[i] * bh[j] * bh[k] + bh[i] * ah[j] * bh[k] + bh[i] * bh[j] * ah[k]);
      r += c[5] * bh[i] * bh[j] * bh[k];
      return r;
    };
    auto E2 = [&](const double* c, int i, int j) {
      return d(i, j) * c[0] + c[1] * ah[i] * ah[j] + c[2] * (ah[i] * bh[j] + bh[i] * ah[j]) + c[3] * bh[i] * bh[j];
    };
    for (int al = 0; al < 4; ++al)
      for (int be = 0; be < 4; ++be)
        for (int de = 0; de < 4; ++de) {
          const int o = (al * 4 + be) * 4 + de;
          Q[0][o] = E3(acc.e3[0], al, be, de) - bh[be] * E2(acc.e2[0], al, de) - ah[al] * E2(acc.e2[1], be, de) +
                    ah[al] * bh[be] * (acc.e1[0] * ah[de] + acc.e1[1] * bh[de]);
          Q[1][o] = E3(acc.e3[1], al, be, de) - ah[al] * E2(acc.e2[2], be, de);
          Q[2][o] = E3(acc.e3[2], al, be, de) - bh[be] * E2(acc.e2[3], al, de);
        }
  }

  // T tensors at x = (xa, 0, 0, 0), y = yb (cos beta, sin beta, 0, 0)
  Tframe eval(double xa, double yb, double beta) const {
    const double X[2] = {xa, 0.0};
    const double Y[2] = {yb * std::cos(beta), yb * std::sin(beta)};
    // centres (in-plane), merged when they coincide
    double C[3][2];
    int nc = 0;
    auto add = [&](double c0, double c1) {
      for (int i = 0; i < nc; ++i)
        if (std::hypot(C[i][0] - c0, C[i][1] - c1) < 1e-12) return;
      C[nc][0] = c0;
      C[nc][1] = c1;
      ++nc;
    };
    add(0, 0);
    add(X[0], X[1]);
    add(Y[0], Y[1]);
    double dist[3][3];
    for (int i = 0; i < nc; ++i)
      for (int j = 0; j < nc; ++j) dist[i][j] = std::hypot(C[i][0] - C[j][0], C[i][1] - C[j][1]);
    double acc[3][14] = {};
    const double norm = 1.0 / (64 * kPi * kPi * kPi * kPi);
    for (int c = 0; c < nc; ++c) {
      double dmin = 1e30;
      for (int j = 0; j < nc; ++j)
        if (j != c) dmin = std::min(dmin, dist[c][j]);
      const double R0 = cfg_.rscale * (nc == 1 ? 1.0 : std::clamp(0.5 * dmin, 0.25, 1.5));
      for (int ir = 0; ir < cfg_.nr; ++ir) {
        const double z = rx_[ir];
        const double r = R0 * z / (1 - z);
        const double wr = rw_[ir] * R0 / ((1 - z) * (1 - z));
        for (int it = 0; it < cfg_.nth; ++it) {
          const double th = thx_[it] * kPi / 2;
          const double wth = thw_[it] * kPi / 2;
          const double rho = r * std::cos(th), rin = r * std::sin(th);
          const double wmeas = 2 * kPi * r * r * r * std::cos(th) * std::sin(th) * wr * wth;
          for (int ip = 0; ip < cfg_.nps; ++ip) {
            const double ps = 2 * kPi * ip / cfg_.nps;
            const double u0 = C[c][0] + rin * std::cos(ps), u1 = C[c][1] + rin * std::sin(ps);
            // Becke weight
            double P[3];
            double dc[3];
            for (int i = 0; i < nc; ++i) dc[i] = std::sqrt((u0 - C[i][0]) * (u0 - C[i][0]) + (u1 - C[i][1]) * (u1 - C[i][1]) + rho * rho);
            double Psum = 0;
            for (int i = 0; i < nc; ++i) {
              P[i] = 1;
              for (int j = 0; j < nc; ++j) {
                if (j == i) continue;
                double mu = (dc[i] - dc[j]) / dist[i][j];
                for (int k = 0; k < 3; ++k) mu = 1.5 * mu - 0.5 * mu * mu * mu;
                P[i] *= 0.5 * (1 - mu);
              }
              Psum += P[i];
            }
            const double wb = P[c] / Psum;
            if (wb < 1e-300) continue;
            const double w = wmeas * 2 * kPi / cfg_.nps * wb;
            // integrand
            const double u[4] = {u0, u1, rho, 0};
            double a[4] = {X[0] - u[0], X[1] - u[1], -u[2], 0};
            double ra = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            double rb = std::sqrt(u0 * u0 + u1 * u1 + rho * rho);
            double ah[4], bh[4];
            for (int l = 0; l < 4; ++l) {
              ah[l] = a[l] / ra;
              bh[l] = u[l] / rb;
            }
            const double gam = ah[0] * bh[0] + ah[1] * bh[1] + ah[2] * bh[2];
            const double dy2 = (Y[0] - u0) * (Y[0] - u0) + (Y[1] - u1) * (Y[1] - u1) + rho * rho;
            const double g0 = 1.0 / (4 * kPi * kPi * dy2);
            Accum A;
            q_scalars(ra, rb, gam, A);
            double Q[3][64];
            q_tensors(A, ah, bh, Q);
            const double f = w * g0 * norm;
            for (int B = 0; B < 3; ++B) {
              const double* q = Q[B];
              auto at = [&](int i, int j, int k) { return q[(i * 4 + j) * 4 + k]; };
              for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                  for (int k = 0; k < 2; ++k) acc[B][(i * 2 + j) * 2 + k] += f * at(i, j, k);
              for (int dd = 0; dd < 2; ++dd) {
                acc[B][8 + dd] += f * 0.5 * (at(2, 2, dd) + at(3, 3, dd));
                acc[B][10 + dd] += f * 0.5 * (at(2, dd, 2) + at(3, dd, 3));
                acc[B][12 + dd] += f * 0.5 * (at(dd, 2, 2) + at(dd, 3, 3));
              }
            }
          }
        }
      }
    }
    Tframe out;
    for (int B = 0; B < 3; ++B)
      for (int i = 0; i < 14; ++i) out.c[B][i] = acc[B][i];
    return out;
  }

 private:
  Config cfg_;
  std::vector<double> tn_, tw_, rx_, rw_, thx_, thw_;
};

// full T^A[3][64] at lattice vectors x, y from frame components
inline void assemble(const Tframe& f, const double x[4], const double y[4], double T[3][64]) {
  double nx = 0, ny = 0;
  for (int i = 0; i < 4; ++i) {
    nx += x[i] * x[i];
    ny += y[i] * y[i];
  }
  nx = std::sqrt(nx);
  ny = std::sqrt(ny);
  double e0[4], e1[4];
  // e0 = x^ (or y^ if x = 0, or axis 0 if both 0)
  if (nx > 0) {
    for (int i = 0; i < 4; ++i) e0[i] = x[i] / nx;
  } else if (ny > 0) {
    for (int i = 0; i < 4; ++i) e0[i] = y[i] / ny;
  } else {
    for (int i = 0; i < 4; ++i) e0[i] = i == 0;
  }
  double c = 0;
  for (int i = 0; i < 4; ++i) c += y[i] * e0[i];
  double yp[4], n1 = 0;
  for (int i = 0; i < 4; ++i) {
    yp[i] = y[i] - c * e0[i];
    n1 += yp[i] * yp[i];
  }
  n1 = std::sqrt(n1);
  if (n1 > 1e-13 * std::max(1.0, ny)) {
    for (int i = 0; i < 4; ++i) e1[i] = yp[i] / n1;
  } else {
    // any unit vector orthogonal to e0
    int m = 0;
    for (int i = 1; i < 4; ++i)
      if (std::fabs(e0[i]) < std::fabs(e0[m])) m = i;
    double v[4] = {0, 0, 0, 0};
    v[m] = 1;
    double d = e0[m];
    double nn = 0;
    for (int i = 0; i < 4; ++i) {
      v[i] -= d * e0[i];
      nn += v[i] * v[i];
    }
    nn = std::sqrt(nn);
    for (int i = 0; i < 4; ++i) e1[i] = v[i] / nn;
  }
  double P[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) P[i][j] = (i == j) - e0[i] * e0[j] - e1[i] * e1[j];
  const double* E[2] = {e0, e1};
  for (int B = 0; B < 3; ++B)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) {
          double r = 0;
          for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
              for (int d = 0; d < 2; ++d) r += f.c[B][(a * 2 + b) * 2 + d] * E[a][i] * E[b][j] * E[d][k];
          for (int d = 0; d < 2; ++d) {
            r += f.c[B][8 + d] * P[i][j] * E[d][k];
            r += f.c[B][10 + d] * P[i][k] * E[d][j];
            r += f.c[B][12 + d] * E[d][i] * P[j][k];
          }
          T[B][(i * 4 + j) * 4 + k] = r;
        }
}

}  // namespace qed_exact
}  // namespace hlbl
