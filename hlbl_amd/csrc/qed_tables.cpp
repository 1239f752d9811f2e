// Host side of the QED kernel: the G^A term table from gamma-matrix traces,
// and the weight-function table.
#include <cmath>
#include <complex>
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "qed.hpp"
#include "qed_host.hpp"

namespace hlbl {
namespace qed {

namespace {

// A monomial 4x4 matrix: row s has i^ph[s] at column col[s].
struct Mono {
  int col[4];
  int ph[4];
};

Mono gamma_mono(int g) {
  Mono m;
  for (int s = 0; s < 4; ++s) {
    m.col[s] = gamma_col(g, s);
    m.ph[s] = gamma_phase(g, s);
  }
  return m;
}

Mono mul(const Mono& a, const Mono& b) {
  Mono m;
  for (int s = 0; s < 4; ++s) {
    int mid = a.col[s];
    m.col[s] = b.col[mid];
    m.ph[s] = (a.ph[s] + b.ph[mid]) & 3;
  }
  return m;
}

// trace as a Gaussian integer (re, im)
void trace(const Mono& m, int& re, int& im) {
  re = im = 0;
  static const int pr[4] = {1, 0, -1, 0}, pi[4] = {0, 1, 0, -1};
  for (int s = 0; s < 4; ++s) {
    if (m.col[s] == s) {
      re += pr[m.ph[s]];
      im += pi[m.ph[s]];
    }
  }
}

Mono product(std::initializer_list<int> gs) {
  Mono m;
  for (int s = 0; s < 4; ++s) {  // identity
    m.col[s] = s;
    m.ph[s] = 0;
  }
  for (int g : gs) m = mul(m, gamma_mono(g));
  return m;
}

// Tr(P_{de rho sigma} g_a g_b g_c ...) for the list `rest`, times 2 (integer);
// P = 2 g_de g_rho g_sigma + 2 delta_{sigma de} g_rho - 2 delta_{rho de} g_sigma for rho != sigma
int trace_p(int de, int rho, int sigma, std::initializer_list<int> rest) {
  int total_re = 0, total_im = 0, re, im;
  auto add = [&](int coef, std::initializer_list<int> head) {
    Mono m = product(head);
    for (int g : rest) m = mul(m, gamma_mono(g));
    trace(m, re, im);
    total_re += coef * re;
    total_im += coef * im;
  };
  add(2, {de, rho, sigma});
  if (sigma == de) add(2, {rho});
  if (rho == de) add(-2, {sigma});
  if (total_im != 0) throw std::logic_error("QED kernel: complex trace");
  return total_re;
}

}  // namespace

void build_term_table(TermTable* out) {
  std::memset(out, 0, sizeof(TermTable));
  for (int k = 0; k < 6; ++k) {
    const int rho = comb_rho(k), sigma = comb_sigma(k);
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = 0; nu < 4; ++nu)
        for (int lam = 0; lam < 4; ++lam) {
          const int o = ((k * 4 + mu) * 4 + nu) * 4 + lam;
          for (int A = 0; A < 3; ++A)
            for (int al = 0; al < 4; ++al)
              for (int be = 0; be < 4; ++be)
                for (int de = 0; de < 4; ++de) {
                  // coefficients as multiples of 1/8: G^I = Tr/8, G^II,III = -Tr/4
                  int c8 = 0;
                  if (A == 0) {
                    c8 = trace_p(de, rho, sigma, {mu, al, nu, be, lam});
                  } else if (A == 1) {
                    if (be == lam) c8 = -2 * trace_p(de, rho, sigma, {mu, al, nu});
                  } else {
                    if (al == mu) c8 = -2 * trace_p(de, rho, sigma, {nu, be, lam});
                  }
                  if (c8 == 0) continue;
                  if (c8 % 8 != 0) throw std::logic_error("QED kernel: non-integer G");
                  int n = out->nterm[o];
                  if (n >= kMaxTerms) throw std::logic_error("QED kernel: too many terms");
                  int idx = A * 64 + (al * 4 + be) * 4 + de;
                  out->term[o][n] = (unsigned short)(idx | ((unsigned)(signed char)(c8 / 8) & 0xff) << 8);
                  out->nterm[o] = (unsigned char)(n + 1);
                }
        }
  }
  // outputs by decreasing term count (stable): the lanes of a wave that evaluate
  // consecutive entries of `order` loop over similar term counts
  for (int o = 0; o < kNL; ++o) out->order[o] = (unsigned short)o;
  std::stable_sort(out->order, out->order + kNL,
                   [&](unsigned short x, unsigned short y) { return out->nterm[x] > out->nterm[y]; });
}

// ---- weight functions ------------------------------------------------------
// Second-order jet in (u, v, w): value, gradient, Hessian (uu uv uw vv vw ww).
struct Jet {
  double f, g[3], h[6];
};

namespace {

inline int hidx(int i, int j) {
  if (i > j) std::swap(i, j);
  static const int t[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  return t[i][j];
}

Jet jconst(double c) {
  Jet r{};
  r.f = c;
  return r;
}
Jet jvar(double x, int i) {
  Jet r{};
  r.f = x;
  r.g[i] = 1.0;
  return r;
}
Jet operator+(const Jet& a, const Jet& b) {
  Jet r;
  r.f = a.f + b.f;
  for (int i = 0; i < 3; ++i) r.g[i] = a.g[i] + b.g[i];
  for (int i = 0; i < 6; ++i) r.h[i] = a.h[i] + b.h[i];
  return r;
}
Jet operator*(double s, const Jet& a) {
  Jet r;
  r.f = s * a.f;
  for (int i = 0; i < 3; ++i) r.g[i] = s * a.g[i];
  for (int i = 0; i < 6; ++i) r.h[i] = s * a.h[i];
  return r;
}
Jet operator+(double s, const Jet& a) {
  Jet r = a;
  r.f += s;
  return r;
}
Jet operator*(const Jet& a, const Jet& b) {
  Jet r;
  r.f = a.f * b.f;
  for (int i = 0; i < 3; ++i) r.g[i] = a.f * b.g[i] + a.g[i] * b.f;
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) {
      int k = hidx(i, j);
      r.h[k] = a.f * b.h[k] + a.h[k] * b.f + a.g[i] * b.g[j] + a.g[j] * b.g[i];
    }
  return r;
}
// apply a scalar function with derivatives (f0, f1, f2) at a.f
Jet chain(const Jet& a, double f0, double f1, double f2) {
  Jet r;
  r.f = f0;
  for (int i = 0; i < 3; ++i) r.g[i] = f1 * a.g[i];
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) {
      int k = hidx(i, j);
      r.h[k] = f1 * a.h[k] + f2 * a.g[i] * a.g[j];
    }
  return r;
}
Jet jexp(const Jet& a) {
  double e = std::exp(a.f);
  return chain(a, e, e, e);
}
Jet jsqrt(const Jet& a) {
  double s = std::sqrt(a.f);
  return chain(a, s, 0.5 / s, -0.25 / (s * a.f));
}
Jet jinv(const Jet& a) {
  double i = 1.0 / a.f;
  return chain(a, i, -i * i, 2.0 * i * i * i);
}

}  // namespace

// The weight functions g0, g1, g2, l1, l2, l3 at (u, v, w) = (x^2, y^2, x.y).
void weight_jets(double u, double v, double w, Jet out[6]) {
  Jet U = jvar(u, 0), V = jvar(v, 1), W = jvar(w, 2);
  Jet sx = jsqrt(1.0 + U), sy = jsqrt(1.0 + V);
  Jet sd = jsqrt(4.0 + (U + V + (-2.0) * W));
  Jet E = jexp(1.5 + ((-0.5) * (sx + sy) + (-0.25) * sd));
  out[0] = E * (1.0 + 0.25 * W) * jinv(1.0 + 0.5 * (U + V));         // g0
  out[1] = 0.5 * E * jinv(sx * sd);                                    // g1
  out[2] = (-0.3) * E * jinv(sy * sd);                                 // g2
  out[3] = 0.2 * E * jinv(1.0 + U);                                    // l1
  out[4] = (-0.15) * E * jinv(1.0 + V);                                // l2
  out[5] = 0.1 * E * (1.0 + (-0.1) * W) * jinv(sx * sy * sd);         // l3
}

void weight_channels(double u, double v, double w, double ch[kNCh]) {
  Jet j[6];
  weight_jets(u, v, w, j);
  const Jet &g0 = j[0], &g1 = j[1], &g2 = j[2], &l1 = j[3], &l2 = j[4], &l3 = j[5];
  for (int i = 0; i < kNCh; ++i) ch[i] = 0.0;
  ch[G0U] = g0.g[0]; ch[G0V] = g0.g[1]; ch[G0W] = g0.g[2];
  const Jet* ls[3] = {&l1, &l2, &l3};
  for (int q = 0; q < 3; ++q) {
    ch[L1 + 4 * q] = ls[q]->f;
    ch[L1 + 4 * q + 1] = ls[q]->g[0];
    ch[L1 + 4 * q + 2] = ls[q]->g[1];
    ch[L1 + 4 * q + 3] = ls[q]->g[2];
  }
  const Jet* gs[2] = {&g1, &g2};
  for (int q = 0; q < 2; ++q) {
    double* c = ch + G1U + 8 * q;
    const Jet& g = *gs[q];
    c[0] = g.g[0]; c[1] = g.g[1]; c[2] = g.g[2];
    c[3] = g.h[0];  // uu
    c[4] = g.h[1];  // uv
    c[5] = g.h[2];  // uw
    c[6] = g.h[4];  // vw
    c[7] = g.h[5];  // ww
  }
}

void build_weight_table(std::vector<double>& table) {
  table.assign((size_t)kNA * kNC * kNB * kNCh, 0.0);
  for (int ia = 0; ia < kNA; ++ia) {
    double a = kRadius * ia / (kNA - 1);
    for (int ic = 0; ic < kNC; ++ic) {
      double c = -1.0 + 2.0 * ic / (kNC - 1);
      for (int ib = 0; ib < kNB; ++ib) {
        double b = kRadius * ib / (kNB - 1);
        double* p = &table[(((size_t)ia * kNC + ic) * kNB + ib) * kNCh];
        weight_channels(a * a, b * b, a * b * c, p);
      }
    }
  }
}

}  // namespace qed
}  // namespace hlbl
