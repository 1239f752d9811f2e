// ORACLE — test infrastructure only. A CPU restatement (FP64, fixed-order
// sums, OpenMP over sites) of the reference's per-pair connected kernel sum,
// used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
// the checker. Nothing in the product links or calls it.
//
// Restated from the reference (read as text; jing2li/hlbl):
//   random fields       src/provider.cpp:10-69
//   gamma basis         src/gamma.cpp:7-45 (as complex 4x4 matrices here)
//   minimum-image maps  kernels/cuda_lattice.cu:128-154, include/hlbl/vec4.hpp:21-44
//   pass 1 (dzu, dzsu)  kernels/cuda_lattice.cu:160-262
//   gamma insertion     src/contract.cpp:19-24, 58-77
//   pass 2 (4-pt)       kernels/cuda_lattice.cu:277-484
//   kernel_sum          src/contract.cpp:104-114
//   2+2 Pi and P1       kernels/hlbl_kernels/kernels.cu:241-340 (no QED kernel: its
//                       sum is the reference's golden value, a pin of the fields and
//                       gammas; the binning pinned by its invariants in the tests)
//   2+2 P2, P3, P4      kernels/hlbl_kernels/kernels.cu:374-760 with this build's plain
//                       kernel: parity unpinned against the reference (its golden sums
//                       involve KQED); only the vanishing of P2, P3 at the source is
//                       fixed independently
// The QED kernel L^(Lambda) is KQED's in the reference (third party, absent here).
// This oracle restates the build's own plain kernel: the Mainz tensor structure
// (trace tensors G^A from gamma matrices, T^A from the scalar weight functions,
// Lambda = 0.4 subtraction, 12 fm cutoff) with tabulated stand-in weight
// functions, trilinearly interpolated. Its individual values are not KQED's:
// parity of L^(Lambda) against KQED is unpinned (no KQED in this environment).
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

using cd = std::complex<double>;

// ---------------------------------------------------------------- fields --
uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double unit(uint64_t key) {
  const uint64_t m = mix64(key) >> 11;  // 53 random bits
  return std::ldexp((double)m, -53) * 2.0 - 1.0;
}

// --------------------------------------------------------------- geometry --
struct Geo {
  int L[4];    // global extents (t, x, y, z)
  int Ll[4];   // local extents
  int pc[4];   // process coordinates
  size_t V;
};
Geo make_geo(const int* global, const int* local, const int* proc) {
  Geo g;
  g.V = 1;
  for (int i = 0; i < 4; ++i) {
    g.L[i] = global[i];
    g.Ll[i] = local[i];
    g.pc[i] = proc ? proc[i] : 0;
    g.V *= (size_t)local[i];
  }
  return g;
}
void coords_of(const Geo& g, size_t s, int c[4]) {
  for (int i = 3; i >= 0; --i) {
    c[i] = (int)(s % g.Ll[i]);
    s /= g.Ll[i];
  }
}
int modp(int a, int L) { return ((a % L) + L) % L; }
int cmap(int x, int L) { return (x >= L / 2) ? x - L : x; }
int cmap_zh(int x, int L) { return (x > L / 2) ? x - L : ((x < L / 2) ? x : 0); }
int cmap_diff(int a, int b, int L) { return cmap(modp(a - b, L), L); }

// ------------------------------------------------------------------ gammas --
// cvc basis; index 5 is gamma_5.
struct Gammas {
  cd m[6][4][4];
  Gammas() {
    const cd I(0, 1);
    std::memset((void*)m, 0, sizeof(m));
    // gamma_0
    m[0][0][2] = -1; m[0][1][3] = -1; m[0][2][0] = -1; m[0][3][1] = -1;
    // gamma_1
    m[1][0][3] = -I; m[1][1][2] = -I; m[1][2][1] = I; m[1][3][0] = I;
    // gamma_2
    m[2][0][3] = -1; m[2][1][2] = 1; m[2][2][1] = 1; m[2][3][0] = -1;
    // gamma_3
    m[3][0][2] = -I; m[3][1][3] = I; m[3][2][0] = I; m[3][3][1] = -I;
    // gamma_5
    m[5][0][0] = 1; m[5][1][1] = 1; m[5][2][2] = -1; m[5][3][3] = -1;
  }
};
const Gammas& gam() {
  static Gammas G;
  return G;
}
// out = gamma_g * in on a spin-colour vector (12 complex, c = 3*spin + colour)
void gmul(int g, const cd* in, cd* out) {
  const Gammas& G = gam();
  for (int s = 0; s < 4; ++s)
    for (int col = 0; col < 3; ++col) {
      cd acc = 0;
      for (int t = 0; t < 4; ++t) acc += G.m[g][s][t] * in[3 * t + col];
      out[3 * s + col] = acc;
    }
}
void load12(const double* p, cd* v) {
  for (int i = 0; i < 12; ++i) v[i] = cd(p[2 * i], p[2 * i + 1]);
}

// ------------------------------------------------------------------ QED ---
constexpr double kLambda = 0.4;
constexpr double kRadius = 12.0 * 105.6583745 / 197.3269804;
constexpr int NA = 49, NC = 25, NB = 49, NCH = 32;

// second-order hyper-dual number in three directions
struct HD {
  double v;
  double d[3];
  double h[3][3];
};
HD hd_c(double c) {
  HD r{};
  r.v = c;
  return r;
}
HD hd_var(double x, int i) {
  HD r{};
  r.v = x;
  r.d[i] = 1;
  return r;
}
HD operator+(HD a, const HD& b) {
  a.v += b.v;
  for (int i = 0; i < 3; ++i) {
    a.d[i] += b.d[i];
    for (int j = 0; j < 3; ++j) a.h[i][j] += b.h[i][j];
  }
  return a;
}
HD operator*(double s, HD a) {
  a.v *= s;
  for (int i = 0; i < 3; ++i) {
    a.d[i] *= s;
    for (int j = 0; j < 3; ++j) a.h[i][j] *= s;
  }
  return a;
}
HD operator*(const HD& a, const HD& b) {
  HD r;
  r.v = a.v * b.v;
  for (int i = 0; i < 3; ++i) r.d[i] = a.v * b.d[i] + a.d[i] * b.v;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      r.h[i][j] = a.v * b.h[i][j] + a.h[i][j] * b.v + a.d[i] * b.d[j] + a.d[j] * b.d[i];
  return r;
}
HD fn(const HD& a, double f0, double f1, double f2) {
  HD r;
  r.v = f0;
  for (int i = 0; i < 3; ++i) r.d[i] = f1 * a.d[i];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.h[i][j] = f1 * a.h[i][j] + f2 * a.d[i] * a.d[j];
  return r;
}
HD hexp(const HD& a) {
  const double e = std::exp(a.v);
  return fn(a, e, e, e);
}
HD hsqrt(const HD& a) {
  const double s = std::sqrt(a.v);
  return fn(a, s, 0.5 / s, -0.25 / (a.v * s));
}
HD hrecip(const HD& a) {
  const double r = 1.0 / a.v;
  return fn(a, r, -r * r, 2 * r * r * r);
}

// the six stand-in weight functions g0, g1, g2, l1, l2, l3 of (u, v, w) = (x^2, y^2, x.y)
void weights(double u, double v, double w, HD out[6]) {
  const HD U = hd_var(u, 0), V = hd_var(v, 1), W = hd_var(w, 2);
  const HD one = hd_c(1.0);
  const HD sx = hsqrt(one + U), sy = hsqrt(one + V);
  const HD sd = hsqrt(hd_c(4.0) + U + V + (-2.0) * W);
  const HD E = hexp(hd_c(1.5) + (-0.5) * sx + (-0.5) * sy + (-0.25) * sd);
  out[0] = E * (one + 0.25 * W) * hrecip(one + 0.5 * U + 0.5 * V);
  out[1] = 0.5 * E * hrecip(sx * sd);
  out[2] = -0.3 * E * hrecip(sy * sd);
  out[3] = 0.2 * E * hrecip(one + U);
  out[4] = -0.15 * E * hrecip(one + V);
  out[5] = 0.1 * E * (one + (-0.1) * W) * hrecip(sx * sy * sd);
}

// channel layout: g0 (u v w) | l1 l2 l3 (f u v w) | g1 g2 (u v w uu uv uw vw ww)
void channels_exact(double u, double v, double w, double* ch) {
  HD f[6];
  weights(u, v, w, f);
  for (int i = 0; i < NCH; ++i) ch[i] = 0;
  for (int i = 0; i < 3; ++i) ch[i] = f[0].d[i];
  for (int q = 0; q < 3; ++q) {
    const HD& l = f[3 + q];
    ch[3 + 4 * q] = l.v;
    for (int i = 0; i < 3; ++i) ch[4 + 4 * q + i] = l.d[i];
  }
  for (int q = 0; q < 2; ++q) {
    const HD& g = f[1 + q];
    double* c = ch + 15 + 8 * q;
    c[0] = g.d[0]; c[1] = g.d[1]; c[2] = g.d[2];
    c[3] = g.h[0][0]; c[4] = g.h[0][1]; c[5] = g.h[0][2]; c[6] = g.h[1][2]; c[7] = g.h[2][2];
  }
}

struct Table {
  std::vector<double> t;
  Table() {
    t.resize((size_t)NA * NC * NB * NCH);
    for (int i = 0; i < NA; ++i)
      for (int j = 0; j < NC; ++j)
        for (int k = 0; k < NB; ++k) {
          const double a = kRadius * i / (NA - 1), c = -1.0 + 2.0 * j / (NC - 1), b = kRadius * k / (NB - 1);
          channels_exact(a * a, b * b, a * b * c, &t[(((size_t)i * NC + j) * NB + k) * NCH]);
        }
  }
  // trilinear interpolation at (|x|, cos beta, |y|)
  void interp(double a, double c, double b, double* ch) const {
    const double ga = a / (kRadius / (NA - 1)), gb = b / (kRadius / (NB - 1)), gc = (c + 1.0) / (2.0 / (NC - 1));
    int i = std::min(std::max((int)std::floor(ga), 0), NA - 2);
    int j = std::min(std::max((int)std::floor(gc), 0), NC - 2);
    int k = std::min(std::max((int)std::floor(gb), 0), NB - 2);
    const double fa = ga - i, fc = gc - j, fb = gb - k;
    for (int q = 0; q < NCH; ++q) ch[q] = 0;
    for (int di = 0; di < 2; ++di)
      for (int dj = 0; dj < 2; ++dj)
        for (int dk = 0; dk < 2; ++dk) {
          const double wgt = (di ? fa : 1 - fa) * (dj ? fc : 1 - fc) * (dk ? fb : 1 - fb);
          const double* p = &t[(((size_t)(i + di) * NC + (j + dj)) * NB + (k + dk)) * NCH];
          for (int q = 0; q < 31; ++q) ch[q] += wgt * p[q];
        }
  }
};
const Table& table() {
  static Table T;
  return T;
}

// T^I, T^II, T^III [3][4][4][4] (al, be, de) from channels at (x, y); m_mu = 1
void t_tensors(const double* x, const double* y, const double* ch, double T[3][4][4][4]) {
  double u = 0, v = 0, w = 0;
  for (int i = 0; i < 4; ++i) {
    u += x[i] * x[i];
    v += y[i] * y[i];
    w += x[i] * y[i];
  }
  auto kd = [](int a, int b) { return a == b ? 1.0 : 0.0; };
  // gradient vectors of a function with channels (Fu, Fv, Fw)
  auto gx = [&](int fu, int fw, double out[4]) {
    for (int i = 0; i < 4; ++i) out[i] = 2 * ch[fu] * x[i] + ch[fw] * y[i];
  };
  auto gs = [&](int fu, int fv, int fw, double out[4]) {
    for (int i = 0; i < 4; ++i) out[i] = (2 * ch[fu] + ch[fw]) * x[i] + (2 * ch[fv] + ch[fw]) * y[i];
  };
  // d^x_al (d^x_be + d^y_be) F for g1/g2 blocks (u v w uu uv uw vw ww) at `base`
  auto hxs = [&](int base, double M[4][4]) {
    const double Fu = ch[base], Fw = ch[base + 2], Fuu = ch[base + 3], Fuv = ch[base + 4],
                 Fuw = ch[base + 5], Fvw = ch[base + 6], Fww = ch[base + 7];
    // d^x_al of the coefficients (2Fu+Fw) and (2Fv+Fw)
    for (int al = 0; al < 4; ++al) {
      const double dA = 2 * (2 * Fuu + Fuw) * x[al] + (2 * Fuw + Fww) * y[al];
      const double dB = 2 * (2 * Fuv + Fuw) * x[al] + (2 * Fvw + Fww) * y[al];
      for (int be = 0; be < 4; ++be) M[al][be] = kd(al, be) * (2 * Fu + Fw) + dA * x[be] + dB * y[be];
    }
  };
  double g1x[4], g2x[4], g1s[4], H1[4][4], H2[4][4];
  gx(15, 17, g1x);
  gx(23, 25, g2x);
  gs(15, 16, 17, g1s);
  hxs(15, H1);
  hxs(23, H2);
  double l1x[4], l2x[4], l3x[4], g0x[4], l1s[4], l2s[4], l3s[4], g0s[4];
  gx(4, 6, l1x);
  gx(8, 10, l2x);
  gx(12, 14, l3x);
  gx(0, 2, g0x);
  gs(4, 5, 6, l1s);
  gs(8, 9, 10, l2s);
  gs(12, 13, 14, l3s);
  gs(0, 1, 2, g0s);
  const double L1 = ch[3], L2 = ch[7], L3 = ch[11];
  for (int al = 0; al < 4; ++al)
    for (int be = 0; be < 4; ++be)
      for (int de = 0; de < 4; ++de) {
        // T^I = d^x_al (d^x_be + d^y_be) (x_de g1 + y_de g2)
        T[0][al][be][de] = kd(be, de) * (g1x[al] + g2x[al]) + kd(al, de) * g1s[be] +
                           x[de] * H1[al][be] + y[de] * H2[al][be];
        // T^II = d^x_al (T_{be de} + delta_{be de} S / 4)
        double t2 = (kd(al, be) * x[de] + kd(al, de) * x[be] - 0.5 * kd(be, de) * x[al]) * L1;
        t2 += (x[be] * x[de] - 0.25 * kd(be, de) * u) * l1x[al];
        t2 += (y[be] * y[de] - 0.25 * kd(be, de) * v) * l2x[al];
        t2 += (kd(al, be) * y[de] + kd(al, de) * y[be] - 0.5 * kd(be, de) * y[al]) * L3;
        t2 += (x[be] * y[de] + y[be] * x[de] - 0.5 * kd(be, de) * w) * l3x[al];
        t2 += 0.25 * kd(be, de) * g0x[al];
        T[1][al][be][de] = t2;
        // T^III = (d^x_be + d^y_be)(T_{al de} + delta_{al de} S / 4)
        double t3 = (kd(be, al) * x[de] + kd(be, de) * x[al] - 0.5 * kd(al, de) * x[be]) * L1;
        t3 += (x[al] * x[de] - 0.25 * kd(al, de) * u) * l1s[be];
        t3 += (kd(be, al) * y[de] + kd(be, de) * y[al] - 0.5 * kd(al, de) * y[be]) * L2;
        t3 += (y[al] * y[de] - 0.25 * kd(al, de) * v) * l2s[be];
        t3 += (kd(be, al) * (x[de] + y[de]) + kd(be, de) * (x[al] + y[al]) - 0.5 * kd(al, de) * (x[be] + y[be])) * L3;
        t3 += (x[al] * y[de] + y[al] * x[de] - 0.5 * kd(al, de) * w) * l3s[be];
        t3 += 0.25 * kd(al, de) * g0s[be];
        T[2][al][be][de] = t3;
      }
}

const int kComb[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

// G^A[k][mu][nu][lam][al][be][de] from dense complex gamma traces, as a sparse list
struct GTerms {
  struct Term {
    int o, i;
    double c;
  };
  std::vector<Term> terms;
  GTerms() {
    const Gammas& G = gam();
    auto mm = [](const cd A[4][4], const cd B[4][4], cd C[4][4]) {
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          cd s = 0;
          for (int k = 0; k < 4; ++k) s += A[i][k] * B[k][j];
          C[i][j] = s;
        }
    };
    auto tr_prod = [&](const cd P[4][4], std::initializer_list<int> gs) {
      cd M[4][4], N[4][4];
      std::memcpy((void*)M, (const void*)P, sizeof(M));
      for (int g : gs) {
        mm(M, G.m[g], N);
        std::memcpy((void*)M, (const void*)N, sizeof(M));
      }
      return M[0][0] + M[1][1] + M[2][2] + M[3][3];
    };
    for (int k = 0; k < 6; ++k) {
      const int rho = kComb[k][0], sig = kComb[k][1];
      for (int de = 0; de < 4; ++de) {
        // P = g_de [g_rho, g_sig] + 2 (delta_{sig de} g_rho - delta_{rho de} g_sig)
        cd A[4][4], B[4][4], C[4][4], P[4][4];
        mm(G.m[rho], G.m[sig], A);
        mm(G.m[sig], G.m[rho], B);
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) C[i][j] = A[i][j] - B[i][j];
        mm(G.m[de], C, P);
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j)
            P[i][j] += 2.0 * ((sig == de ? 1.0 : 0.0) * G.m[rho][i][j] - (rho == de ? 1.0 : 0.0) * G.m[sig][i][j]);
        for (int mu = 0; mu < 4; ++mu)
          for (int nu = 0; nu < 4; ++nu)
            for (int lam = 0; lam < 4; ++lam) {
              const int o = ((k * 4 + mu) * 4 + nu) * 4 + lam;
              for (int al = 0; al < 4; ++al)
                for (int be = 0; be < 4; ++be) {
                  const cd g1 = tr_prod(P, {mu, al, nu, be, lam}) / 8.0;
                  const cd g2 = be == lam ? -tr_prod(P, {mu, al, nu}) / 4.0 : cd(0);
                  const cd g3 = al == mu ? -tr_prod(P, {nu, be, lam}) / 4.0 : cd(0);
                  const cd gg[3] = {g1, g2, g3};
                  for (int A_ = 0; A_ < 3; ++A_)
                    if (std::abs(gg[A_]) > 1e-12)
                      terms.push_back({o, A_ * 64 + (al * 4 + be) * 4 + de, gg[A_].real()});
                }
            }
      }
    }
  }
};
const GTerms& gterms() {
  static GTerms G;
  return G;
}

void l0_eval(const double* x, const double* y, double* L0) {
  double a = 0, b = 0, w = 0;
  for (int i = 0; i < 4; ++i) {
    a += x[i] * x[i];
    b += y[i] * y[i];
    w += x[i] * y[i];
  }
  a = std::sqrt(a);
  b = std::sqrt(b);
  double c = (a > 0 && b > 0) ? w / (a * b) : 0.0;
  c = std::min(1.0, std::max(-1.0, c));
  double ch[NCH];
  table().interp(a, c, b, ch);
  double T[3][4][4][4];
  t_tensors(x, y, ch, T);
  const double* Tf = &T[0][0][0][0];
  for (int o = 0; o < 384; ++o) L0[o] = 0;
  for (const auto& t : gterms().terms) L0[t.o] += t.c * Tf[t.i];
}

// L^(Lambda)(x, y) into kerv[6][4][4][4]
void kernel_plain(const double* x, const double* y, double* kerv) {
  double nx = 0, ny = 0;
  for (int i = 0; i < 4; ++i) {
    nx += x[i] * x[i];
    ny += y[i] * y[i];
  }
  if (nx > kRadius * kRadius || ny > kRadius * kRadius) {
    for (int o = 0; o < 384; ++o) kerv[o] = 0;
    return;
  }
  const double zero[4] = {0, 0, 0, 0};
  double Lxy[384], L0y[384], Lx0[384];
  l0_eval(x, y, Lxy);
  l0_eval(zero, y, L0y);
  l0_eval(x, zero, Lx0);
  double Dx[4][4], Dy[4][4];
  const double ex = std::exp(-kLambda * nx / 2), ey = std::exp(-kLambda * ny / 2);
  for (int m = 0; m < 4; ++m)
    for (int a = 0; a < 4; ++a) {
      Dx[m][a] = ((m == a) - kLambda * x[m] * x[a]) * ex;
      Dy[m][a] = ((m == a) - kLambda * y[m] * y[a]) * ey;
    }
  for (int k = 0; k < 6; ++k)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = 0; nu < 4; ++nu)
        for (int lam = 0; lam < 4; ++lam) {
          double r = Lxy[((k * 4 + mu) * 4 + nu) * 4 + lam];
          for (int al = 0; al < 4; ++al) {
            r -= Dx[mu][al] * L0y[((k * 4 + al) * 4 + nu) * 4 + lam];
            r -= Dy[nu][al] * Lx0[((k * 4 + mu) * 4 + al) * 4 + lam];
          }
          kerv[((k * 4 + mu) * 4 + nu) * 4 + lam] = r;
        }
}

// ------------------------------------------------------------ contraction --
// spinor (flavour f, source a) at site s of a block
const double* spinor(const double* blk, size_t V, int f, int a, size_t s) {
  return blk + (((size_t)(f * 12 + a)) * V + s) * 24;
}

// pass 1: dzu[6][12][12], dzsu[4][12][12] complex
void pass1(const double* S, const double* Y, int f, const Geo& g, const int* gsx, cd* dzu,
           cd* dzsu) {
  const int nt = omp_get_max_threads();
  std::vector<cd> part((size_t)nt * 1440, cd(0));
#pragma omp parallel num_threads(nt)
  {
    cd* mine = &part[(size_t)omp_get_thread_num() * 1440];
#pragma omp for schedule(static)
    for (long long z = 0; z < (long long)g.V; ++z) {
      int c[4];
      coords_of(g, z, c);
      int zz[4];
      for (int r = 0; r < 4; ++r) zz[r] = cmap_zh(modp(c[r] + g.pc[r] * g.Ll[r] - gsx[r], g.L[r]), g.L[r]);
      for (int a = 0; a < 12; ++a) {
        cd s0[12], tg[4][12], t[4][12];
        load12(spinor(S, g.V, f, a, z), s0);
        for (int sg = 0; sg < 4; ++sg) {
          gmul(sg, s0, tg[sg]);
          gmul(5, tg[sg], t[sg]);
        }
        for (int b = 0; b < 12; ++b) {
          cd y[12];
          load12(spinor(Y, g.V, 1 - f, b, z), y);
          for (int k = 0; k < 6; ++k) {
            const int rho = kComb[k][0], sig = kComb[k][1];
            cd acc = 0;
            for (int i = 0; i < 12; ++i) acc += std::conj(y[i]) * (t[sig][i] * (double)zz[rho] - t[rho][i] * (double)zz[sig]);
            mine[(k * 12 + a) * 12 + b] += acc;
          }
          for (int sg = 0; sg < 4; ++sg) {
            cd acc = 0;
            for (int i = 0; i < 12; ++i) acc += std::conj(y[i]) * t[sg][i];
            mine[864 + (sg * 12 + a) * 12 + b] += acc;
          }
        }
      }
    }
  }
  for (int i = 0; i < 864; ++i) dzu[i] = 0;
  for (int i = 0; i < 576; ++i) dzsu[i] = 0;
  for (int t = 0; t < nt; ++t) {
    for (int i = 0; i < 864; ++i) dzu[i] += part[(size_t)t * 1440 + i];
    for (int i = 0; i < 576; ++i) dzsu[i] += part[(size_t)t * 1440 + 864 + i];
  }
}

// gamma_mu gamma_5 on the b index: g[j][mu][a][.] = gamma_mu gamma_5 d[j][a][.]
void insert_gamma(const cd* d, int nj, cd* gd) {
  for (int j = 0; j < nj; ++j)
    for (int a = 0; a < 12; ++a) {
      cd t[12];
      gmul(5, d + (j * 12 + a) * 12, t);
      for (int mu = 0; mu < 4; ++mu) gmul(mu, t, gd + ((j * 4 + mu) * 12 + a) * 12);
    }
}

typedef void (*KernelFn)(const double*, const double*, double*);

// pass 2 for flavour f: sum over sites of the 4-point contraction times the kernels
double pass2(const double* S, const double* Y, int f, const Geo& g, const int* gsx, const int* yv,
             double xunit, const cd* gdzu, const cd* gdzsu, KernelFn kern) {
  const int nt = omp_get_max_threads();
  std::vector<double> part(nt, 0.0);
#pragma omp parallel num_threads(nt)
  {
    double mine = 0.0;
    std::vector<double> kv(384);
#pragma omp for schedule(static)
    for (long long x = 0; x < (long long)g.V; ++x) {
      int c[4], xrel[4], xv[4], xzh[4];
      coords_of(g, x, c);
      for (int r = 0; r < 4; ++r) {
        xrel[r] = modp(c[r] + g.pc[r] * g.Ll[r] - gsx[r], g.L[r]);
        xv[r] = cmap(xrel[r], g.L[r]);
        xzh[r] = cmap_zh(xrel[r], g.L[r]);
      }
      double corrI[384], corrII[384];
      for (int mu = 0; mu < 4; ++mu) {
        // dxu[r][a] = - sum_c conj((g5 gamma_mu S[1-f][a])[c]) Y[f][r][c]
        cd dxu[12][12];
        for (int a = 0; a < 12; ++a) {
          cd s0[12], t1[12], t[12];
          load12(spinor(S, g.V, 1 - f, a, x), s0);
          gmul(mu, s0, t1);
          gmul(5, t1, t);
          for (int r = 0; r < 12; ++r) {
            cd u[12];
            load12(spinor(Y, g.V, f, r, x), u);
            cd acc = 0;
            for (int i = 0; i < 12; ++i) acc += std::conj(t[i]) * u[i];
            dxu[r][a] = -acc;
          }
        }
        for (int k = 0; k < 6; ++k) {
          const int rho = kComb[k][0], sig = kComb[k][1];
          for (int nu = 0; nu < 4; ++nu)
            for (int lam = 0; lam < 4; ++lam) {
              double cI = 0, cII = 0;
              for (int r = 0; r < 12; ++r) {
                cd t5[12], q[12];
                gmul(5, dxu[r], t5);
                gmul(lam, t5, q);
                for (int i = 0; i < 12; ++i) {
                  const cd vI = gdzu[((k * 4 + nu) * 12 + i) * 12 + r];
                  const cd vII = (double)xzh[rho] * gdzsu[((sig * 4 + nu) * 12 + i) * 12 + r] -
                                 (double)xzh[sig] * gdzsu[((rho * 4 + nu) * 12 + i) * 12 + r];
                  cI -= (q[i] * vI).real();
                  cII -= (q[i] * vII).real();
                }
              }
              corrI[((k * 4 + mu) * 4 + nu) * 4 + lam] = cI;
              corrII[((k * 4 + mu) * 4 + nu) * 4 + lam] = cII;
            }
        }
      }
      double xm[4], ym[4], dm[4];
      for (int r = 0; r < 4; ++r) {
        xm[r] = xv[r] * xunit;
        ym[r] = cmap(modp(yv[r], g.L[r]), g.L[r]) * xunit;
        dm[r] = cmap_diff(xrel[r], yv[r], g.L[r]) * xunit;
      }
      double dtmp = 0.0;
      auto idx = [](int k, int a, int b, int cc) { return ((k * 4 + a) * 4 + b) * 4 + cc; };
      kern(xm, ym, kv.data());
      for (int k = 0; k < 6; ++k)
        for (int mu = 0; mu < 4; ++mu)
          for (int nu = 0; nu < 4; ++nu)
            for (int lam = 0; lam < 4; ++lam) dtmp += kv[idx(k, mu, nu, lam)] * corrI[idx(k, mu, nu, lam)];
      kern(ym, xm, kv.data());
      for (int k = 0; k < 6; ++k)
        for (int mu = 0; mu < 4; ++mu)
          for (int nu = 0; nu < 4; ++nu)
            for (int lam = 0; lam < 4; ++lam) dtmp += kv[idx(k, nu, mu, lam)] * corrI[idx(k, mu, nu, lam)];
      kern(xm, dm, kv.data());
      for (int k = 0; k < 6; ++k)
        for (int mu = 0; mu < 4; ++mu)
          for (int nu = 0; nu < 4; ++nu)
            for (int lam = 0; lam < 4; ++lam) {
              dtmp -= kv[idx(k, lam, nu, mu)] * corrI[idx(k, mu, nu, lam)];
              dtmp += kv[idx(k, lam, nu, mu)] * corrII[idx(k, mu, nu, lam)];
            }
      mine += dtmp;
    }
    part[omp_get_thread_num()] = mine;
  }
  double s = 0;
  for (double v : part) s += v;
  return s;
}

void kernel_zero(const double*, const double*, double* k) {
  for (int o = 0; o < 384; ++o) k[o] = 0;
}
// a smooth analytic test kernel (not physics): used to check the contraction
// structure independently of the QED tables
void kernel_probe(const double* x, const double* y, double* k) {
  for (int kk = 0; kk < 6; ++kk)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = 0; nu < 4; ++nu)
        for (int lam = 0; lam < 4; ++lam)
          k[((kk * 4 + mu) * 4 + nu) * 4 + lam] =
              std::sin(0.3 * kk + 0.7 * mu + 1.1 * nu + 1.3 * lam + x[mu] - 0.5 * y[nu] + 0.2 * x[lam] * y[lam]);
}

KernelFn pick(int mode) {
  switch (mode) {
    case 1: return kernel_zero;
    case 2: return kernel_probe;
    default: return kernel_plain;
  }
}

}  // namespace

extern "C" {

uint64_t oracle_hash64(uint64_t x) { return mix64(x); }

uint64_t oracle_source_seed(uint64_t seed, const int* gsx) {
  for (int i = 0; i < 4; ++i) seed = mix64(seed ^ (uint64_t)(int64_t)(gsx[i] + 1));
  return seed;
}

// fill_random_block: element ((f*12+a)*V + s)*24 + c of the local block
void oracle_fill_block(double* out, const int* local, const int* proc, uint64_t seed) {
  const int pc0[4] = {0, 0, 0, 0};
  const Geo g = make_geo(local, local, proc ? proc : pc0);
#pragma omp parallel for schedule(static)
  for (long long s = 0; s < (long long)g.V; ++s) {
    int c[4];
    coords_of(g, s, c);
    uint64_t key = 0;
    for (int r = 0; r < 4; ++r) key = key * 1000003u + (uint64_t)(g.pc[r] * g.Ll[r] + c[r]);
    for (int f = 0; f < 2; ++f)
      for (int a = 0; a < 12; ++a)
        for (int cc = 0; cc < 24; ++cc)
          out[(((size_t)(f * 12 + a)) * g.V + s) * 24 + cc] =
              unit(((key * 2 + f) * 12 + a) * 24 + cc + seed * 0x2545F4914F6CDD1DULL);
  }
}

// the gamma action as double-level (perm, sign) tables: out[j] = in[perm[j]] * sign[j];
// index 0..3 gamma_mu, 5 gamma_5
void oracle_gamma_tables(int g, int* perm, int* sign) {
  for (int j = 0; j < 24; ++j) {
    cd in[12] = {}, out[12];
    in[j / 2] = (j & 1) ? cd(0, 1) : cd(1, 0);  // basis element j
    gmul(g, in, out);
    for (int i = 0; i < 12; ++i) {
      const double v[2] = {out[i].real(), out[i].imag()};
      for (int part = 0; part < 2; ++part)
        if (v[part] != 0.0) {
          perm[2 * i + part] = j;
          sign[2 * i + part] = v[part] > 0 ? 1 : -1;
        }
    }
  }
}

// pass 1 of flavour f; dzu [6][12][12] and dzsu [4][12][12] complex (interleaved doubles)
void oracle_dzu_dzsu(const double* S, const double* Y, int f, const int* global, const int* local,
                     const int* proc, const int* gsx, double* dzu, double* dzsu) {
  const Geo g = make_geo(global, local, proc);
  std::vector<cd> u(864), s(576);
  pass1(S, Y, f, g, gsx, u.data(), s.data());
  std::memcpy(dzu, u.data(), 864 * sizeof(cd));
  std::memcpy(dzsu, s.data(), 576 * sizeof(cd));
}

// gamma_mu gamma_5 insertion: g_dzu [6][4][12][12], g_dzsu [4][4][12][12]
void oracle_insert_gamma(const double* dzu, const double* dzsu, double* g_dzu, double* g_dzsu) {
  std::vector<cd> gu(6 * 4 * 144), gs(4 * 4 * 144);
  insert_gamma(reinterpret_cast<const cd*>(dzu), 6, gu.data());
  insert_gamma(reinterpret_cast<const cd*>(dzsu), 4, gs.data());
  std::memcpy(g_dzu, gu.data(), gu.size() * sizeof(cd));
  std::memcpy(g_dzsu, gs.data(), gs.size() * sizeof(cd));
}

// the 4-point pass for flavour f, given g_dzu/g_dzsu; kernel_mode 0 plain, 1 zero, 2 probe
double oracle_4pt(const double* g_dzu, const double* g_dzsu, const double* S, const double* Y,
                  int f, const int* global, const int* local, const int* proc, const int* gsx,
                  const int* yv, double xunit, int kernel_mode) {
  const Geo g = make_geo(global, local, proc);
  return pass2(S, Y, f, g, gsx, yv, xunit, reinterpret_cast<const cd*>(g_dzu),
               reinterpret_cast<const cd*>(g_dzsu), pick(kernel_mode));
}

// Contractor::kernel_sum on one process: out[2] for flavours 0, 1
void oracle_kernel_sum(const double* S, const double* Y, const int* global, const int* gsx,
                       const int* yv, double xunit, int kernel_mode, double* out) {
  const int proc[4] = {0, 0, 0, 0};
  const Geo g = make_geo(global, global, proc);
  for (int f = 0; f < 2; ++f) {
    std::vector<cd> u(864), s(576), gu(6 * 4 * 144), gs(4 * 4 * 144);
    pass1(S, Y, f, g, gsx, u.data(), s.data());
    insert_gamma(u.data(), 6, gu.data());
    insert_gamma(s.data(), 4, gs.data());
    out[f] = pass2(S, Y, f, g, gsx, yv, xunit, gu.data(), gs.data(), pick(kernel_mode));
  }
}

// plain L^(Lambda) at n argument pairs: x, y [n][4], kerv [n][384]
void oracle_qed_kernel(const double* x, const double* y, int n, double* kerv) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) kernel_plain(x + 4 * i, y + 4 * i, kerv + (size_t)i * 384);
}

// L0 (no subtraction, no cutoff) at n argument pairs
void oracle_qed_l0(const double* x, const double* y, int n, double* l0) {
  for (int i = 0; i < n; ++i) l0_eval(x + 4 * i, y + 4 * i, l0 + (size_t)i * 384);
}

// exact (not interpolated) channels at (u, v, w), and interpolated ones at (a, cos beta, b)
void oracle_channels_exact(double u, double v, double w, double* ch) { channels_exact(u, v, w, ch); }
void oracle_channels_interp(double a, double c, double b, double* ch) { table().interp(a, c, b, ch); }
void oracle_table(double* out) {
  const Table& t = table();
  std::memcpy(out, t.t.data(), t.t.size() * sizeof(double));
}
size_t oracle_table_size(void) { return table().t.size(); }

// T tensors [3][4][4][4] from given channels
void oracle_t_tensors(const double* x, const double* y, const double* ch, double* T) {
  double t[3][4][4][4];
  t_tensors(x, y, ch, t);
  std::memcpy(T, t, sizeof(t));
}

// S, V[4], T[4][4] of the weight functions at (x, y), exact (for derivative checks)
void oracle_svt(const double* x, const double* y, double* S, double* V, double* T) {
  double u = 0, v = 0, w = 0;
  for (int i = 0; i < 4; ++i) {
    u += x[i] * x[i];
    v += y[i] * y[i];
    w += x[i] * y[i];
  }
  HD f[6];
  weights(u, v, w, f);
  *S = f[0].v;
  for (int i = 0; i < 4; ++i) V[i] = x[i] * f[1].v + y[i] * f[2].v;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const double d = a == b ? 1.0 : 0.0;
      T[a * 4 + b] = (x[a] * x[b] - d * u / 4) * f[3].v + (y[a] * y[b] - d * v / 4) * f[4].v +
                     (x[a] * y[b] + y[a] * x[b] - d * w / 2) * f[5].v;
    }
}

// number of terms and the dense G^A value at (o, i) (for checks of the trace tensors)
int oracle_g_terms(int* o, int* i, double* c, int max) {
  const auto& t = gterms().terms;
  const int n = (int)t.size();
  for (int j = 0; j < n && j < max; ++j) {
    o[j] = t[j].o;
    i[j] = t[j].i;
    c[j] = t[j].c;
  }
  return n;
}

}  // extern "C"

// Pi_{mu nu}(z) of the 2+2 pieces (kernels.cu:241-294) at local site z:
//   Pi_{mu nu}(z) = sum_a Re[(gamma_nu gamma_5 w_a)[a]],
//   w_a[b] = sum_c conj(D_b[c]) (g5 gamma_mu U_a)[c], U = block[f], D = block[1-f]
static void pi_site(const double* blk, size_t V, int f, size_t z, double pi[16]) {
  cd u[12][12], d[12][12];
  for (int a = 0; a < 12; ++a) {
    load12(spinor(blk, V, f, a, z), u[a]);
    load12(spinor(blk, V, 1 - f, a, z), d[a]);
  }
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = 0; nu < 4; ++nu) {
      double acc = 0;
      for (int a = 0; a < 12; ++a) {
        cd t1[12], gu[12], wv[12], t5[12], q[12];
        gmul(mu, u[a], t1);
        gmul(5, t1, gu);
        for (int b = 0; b < 12; ++b) {
          cd w = 0;
          for (int c = 0; c < 12; ++c) w += std::conj(d[b][c]) * gu[c];
          wv[b] = w;
        }
        gmul(5, wv, t5);
        gmul(nu, t5, q);
        acc += q[a].real();
      }
      pi[mu * 4 + nu] = acc;
    }
}

extern "C" {

// sum over all entries of the 2+2 piece P1 of one block (kernels.cu:241-340):
// P1[rho][sigma][nu][zeta] bins Pi_{sigma nu}(z) along rho, so its total is
// 4 * sum_z sum_{sigma nu} Pi_{sigma nu}(z)
double oracle_sum_p1(const double* blk, const int* local, int f) {
  const int pc0[4] = {0, 0, 0, 0};
  const Geo g = make_geo(local, local, pc0);
  const int nt = omp_get_max_threads();
  std::vector<double> part(nt, 0.0);
#pragma omp parallel num_threads(nt)
  {
    double mine = 0;
#pragma omp for schedule(static)
    for (long long z = 0; z < (long long)g.V; ++z) {
      double pi[16];
      pi_site(blk, g.V, f, (size_t)z, pi);
      for (int k = 0; k < 16; ++k) mine += pi[k];
    }
    part[omp_get_thread_num()] = mine;
  }
  double s = 0;
  for (double v : part) s += v;
  return 4.0 * s;
}

// The 2+2 piece P1 of one block on one process (kernel_p1, kernels.cu:296-340):
// out[rho][sigma][nu][zeta] = sum of Pi_{sigma nu}(z) over the sites with
// (z - w)_rho mod L_rho = zeta, zeta < lmax (0 for zeta >= L_rho); w = gsw.
void oracle_p1(const double* blk, const int* dims, int f, const int* gsw, int lmax, double* out) {
  const int pc0[4] = {0, 0, 0, 0};
  const Geo g = make_geo(dims, dims, pc0);
  std::vector<double> pi(g.V * 16);
#pragma omp parallel for schedule(static)
  for (long long z = 0; z < (long long)g.V; ++z) pi_site(blk, g.V, f, (size_t)z, pi.data() + 16 * z);
  std::memset(out, 0, sizeof(double) * 64 * lmax);
  for (size_t z = 0; z < g.V; ++z) {
    int c[4];
    size_t q = z;
    for (int r = 3; r >= 0; --r) {
      c[r] = (int)(q % dims[r]);
      q /= dims[r];
    }
    for (int rho = 0; rho < 4; ++rho) {
      const int zeta = ((c[rho] - gsw[rho]) % dims[rho] + dims[rho]) % dims[rho];
      for (int k = 0; k < 16; ++k) out[(rho * 16 + k) * lmax + zeta] += pi[16 * z + k];
    }
  }
}

// The 2+2 pieces P2_0, P2_1, P3, P4_0, P4_1 of one block for one target (kernel_p20,
// kernel_p21, kernel_p3, kernel_p4, kernels.cu:374-760), on one process:
// out[5][rho][sigma][nu]. With x = z - w and y = target - w (mod L), each argument
// difference min-imaged on its own (site_map, [-L/2, L/2)) and scaled by xunit, and k
// over (r, s) = (0,1),(0,2),(0,3),(1,2),(1,3),(2,3):
//   P2_0[r][s][nu] += L[k](x, y)[mu][nu][lam]     Pi_{mu lam}(z)
//   P2_1[r][s][nu] += L[k](y, x)[nu][mu][lam]     Pi_{mu lam}(z)
//   P3  [r][s][nu] += L[k](x - y, -y)[mu][lam][nu] Pi_{mu lam}(z)
//   P4_0[r][s][nu] += L[k](y - x, -x)[nu][lam][mu] Pi_{mu lam}(z)
//   P4_1[r][s][nu] += xv_r (that term), P4_1[s][r][nu] -= xv_s (that term), xv the
//   integer min-image of x.
void oracle_p234(const double* blk, const int* dims, int f, const int* gsw, const int* target,
                 double xunit, double* out) {
  const int pc0[4] = {0, 0, 0, 0};
  const Geo g = make_geo(dims, dims, pc0);
  auto smap = [&](int v, int r) {
    v = ((v % dims[r]) + dims[r]) % dims[r];
    return v >= dims[r] / 2 ? v - dims[r] : v;
  };
  static const int kr[6] = {0, 0, 0, 1, 1, 2}, ks[6] = {1, 2, 3, 2, 3, 3};
  int y[4], yv[4], ymv[4];
  for (int r = 0; r < 4; ++r) {
    y[r] = ((target[r] - gsw[r]) % dims[r] + dims[r]) % dims[r];
    yv[r] = smap(y[r], r);
    ymv[r] = smap(-y[r], r);
  }
  const int nt = omp_get_max_threads();
  std::vector<double> part((size_t)nt * 320, 0.0);
#pragma omp parallel num_threads(nt)
  {
    double* acc = part.data() + (size_t)omp_get_thread_num() * 320;
    std::vector<double> K(384);
#pragma omp for schedule(static)
    for (long long z = 0; z < (long long)g.V; ++z) {
      double pi[16];
      pi_site(blk, g.V, f, (size_t)z, pi);
      int c[4];
      long long q = z;
      for (int r = 3; r >= 0; --r) {
        c[r] = (int)(q % dims[r]);
        q /= dims[r];
      }
      int x[4], xv[4];
      double xm[4], ym[4], a[4], b[4];
      for (int r = 0; r < 4; ++r) {
        x[r] = ((c[r] - gsw[r]) % dims[r] + dims[r]) % dims[r];
        xv[r] = smap(x[r], r);
        xm[r] = xv[r] * xunit;
        ym[r] = yv[r] * xunit;
      }
      auto L = [&](int k, int i, int j, int l) { return K[((size_t)k * 4 + i) * 16 + j * 4 + l]; };
      for (int geo = 0; geo < 4; ++geo) {
        if (geo == 0) kernel_plain(xm, ym, K.data());
        if (geo == 1) kernel_plain(ym, xm, K.data());
        if (geo == 2) {
          for (int r = 0; r < 4; ++r) {
            a[r] = smap(x[r] - y[r], r) * xunit;
            b[r] = ymv[r] * xunit;
          }
          kernel_plain(a, b, K.data());
        }
        if (geo == 3) {
          for (int r = 0; r < 4; ++r) {
            a[r] = smap(y[r] - x[r], r) * xunit;
            b[r] = smap(-x[r], r) * xunit;
          }
          kernel_plain(a, b, K.data());
        }
        for (int k = 0; k < 6; ++k)
          for (int nu = 0; nu < 4; ++nu) {
            double t = 0;
            for (int mu = 0; mu < 4; ++mu)
              for (int lam = 0; lam < 4; ++lam) {
                const double v = geo == 0 ? L(k, mu, nu, lam)
                               : geo == 1 ? L(k, nu, mu, lam)
                               : geo == 2 ? L(k, mu, lam, nu)
                                          : L(k, nu, lam, mu);
                t += v * pi[mu * 4 + lam];
              }
            const int rs = kr[k] * 16 + ks[k] * 4 + nu, sr = ks[k] * 16 + kr[k] * 4 + nu;
            acc[geo * 64 + rs] += t;
            if (geo == 3) {
              acc[4 * 64 + rs] += xv[kr[k]] * t;
              acc[4 * 64 + sr] -= xv[ks[k]] * t;
            }
          }
      }
    }
  }
  for (int j = 0; j < 320; ++j) {
    double s = 0;
    for (int t = 0; t < nt; ++t) s += part[(size_t)t * 320 + j];
    out[j] = s;
  }
}

void oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

}  // extern "C"

// ---------------------------------------------------------------- MLP ----
// The surrogate network in FP64: 3 -> 4 x 128 GELU -> 31, parameters packed
// layer by layer as W[out][in] then b[out], weights rounded to BF16 (nearest,
// ties to even) before use; loss = mean over rows and outputs of (y - t)^2.
namespace {
float bf16(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}
const int kIn[5] = {3, 128, 128, 128, 128}, kOut[5] = {128, 128, 128, 128, 31};
}  // namespace

extern "C" {

int oracle_mlp_param_count(void) {
  int n = 0;
  for (int l = 0; l < 5; ++l) n += kIn[l] * kOut[l] + kOut[l];
  return n;
}

// loss of n rows and (if grad) its gradient w.r.t. the (rounded) parameters; y gets outputs
double oracle_mlp(const float* params, const float* x, const float* t, int n, double* grad, double* y) {
  const int np = oracle_mlp_param_count();
  std::vector<double> w(np);
  for (int i = 0; i < np; ++i) w[i] = bf16(params[i]);
  int offw[5], offb[5], o = 0;
  for (int l = 0; l < 5; ++l) {
    offw[l] = o;
    o += kIn[l] * kOut[l];
    offb[l] = o;
    o += kOut[l];
  }
  if (grad) std::fill(grad, grad + np, 0.0);
  double loss = 0.0;
  std::vector<double> z[5], a[6];
  for (int r = 0; r < n; ++r) {
    a[0].assign(x + 3 * r, x + 3 * r + 3);
    for (int l = 0; l < 5; ++l) {
      z[l].assign(kOut[l], 0.0);
      for (int j = 0; j < kOut[l]; ++j) {
        double s = w[offb[l] + j];
        for (int k = 0; k < kIn[l]; ++k) s += w[offw[l] + j * kIn[l] + k] * a[l][k];
        z[l][j] = s;
      }
      a[l + 1].resize(kOut[l]);
      for (int j = 0; j < kOut[l]; ++j)
        a[l + 1][j] = l < 4 ? 0.5 * z[l][j] * (1.0 + std::erf(z[l][j] / std::sqrt(2.0))) : z[l][j];
    }
    std::vector<double> d(31);
    for (int c = 0; c < 31; ++c) {
      const double e = a[5][c] - t[31 * r + c];
      loss += e * e;
      d[c] = 2.0 * e / ((double)n * 31);
      if (y) y[31 * r + c] = a[5][c];
    }
    if (!grad) continue;
    for (int l = 4; l >= 0; --l) {
      // d holds dL/dz[l]
      for (int j = 0; j < kOut[l]; ++j) {
        grad[offb[l] + j] += d[j];
        for (int k = 0; k < kIn[l]; ++k) grad[offw[l] + j * kIn[l] + k] += d[j] * a[l][k];
      }
      if (l == 0) break;
      std::vector<double> dn(kIn[l], 0.0);
      for (int k = 0; k < kIn[l]; ++k) {
        double s = 0;
        for (int j = 0; j < kOut[l]; ++j) s += d[j] * w[offw[l] + j * kIn[l] + k];
        const double zz = z[l - 1][k];
        const double gp = 0.5 * (1.0 + std::erf(zz / std::sqrt(2.0))) + zz * std::exp(-0.5 * zz * zz) / std::sqrt(2.0 * M_PI);
        dn[k] = s * gp;
      }
      d = dn;
    }
  }
  return loss / ((double)n * 31);
}

// Forward of the BF16 inference: BF16 weights, every hidden activation rounded
// to BF16 after the GELU, FP64 arithmetic otherwise. y[n][31].
void oracle_mlp_forward_bf16(const float* params, const float* x, int n, double* y) {
  const int np = oracle_mlp_param_count();
  std::vector<double> w(np);
  for (int i = 0; i < np; ++i) w[i] = bf16(params[i]);
  int offw[5], offb[5], o = 0;
  for (int l = 0; l < 5; ++l) {
    offw[l] = o;
    o += kIn[l] * kOut[l];
    offb[l] = o;
    o += kOut[l];
  }
  std::vector<double> a, z;
  for (int r = 0; r < n; ++r) {
    a.assign(x + 3 * r, x + 3 * r + 3);
    for (int l = 0; l < 5; ++l) {
      z.assign(kOut[l], 0.0);
      for (int j = 0; j < kOut[l]; ++j) {
        double s = w[offb[l] + j];
        for (int k = 0; k < kIn[l]; ++k) s += w[offw[l] + j * kIn[l] + k] * a[k];
        z[j] = l < 4 ? (double)bf16((float)(0.5 * s * (1.0 + std::erf(s / std::sqrt(2.0))))) : s;
      }
      a = z;
    }
    for (int c = 0; c < 31; ++c) y[31 * r + c] = a[c];
  }
}

}  // extern "C"
