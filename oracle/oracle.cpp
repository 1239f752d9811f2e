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
//   2+2 P2, P3, P4      kernels/hlbl_kernels/kernels.cu:374-760 with the library's plain
//                       kernel: pinned by the reference's golden sums of P2 + P3 and of
//                       P4 (sanity_check/test_contract_random.cpp:126-127, which involve
//                       KQED) to 2e-2, and by the vanishing of P2, P3 at the source
// The QED kernel L^(Lambda) is KQED's in the reference (third party, absent here).
// This oracle restates the library's own plain kernel: the Mainz tensor structure
// (trace tensors G^A from gamma matrices, the T^A tensors read from the library's
// table, which its generator computes from the kernel's defining integrals,
// interpolated cubically, Lambda = 0.4 subtraction, 12 fm cutoff). Against KQED it
// is pinned only through the reference's four golden sums (the two kernel sums and
// the two 2+2 sums of sanity_check/test_contract_random.cpp), which it meets to about
// 1%: individual L^(Lambda) values of KQED are not available here. The ~1% is not the
// table's: with the kernel's integrals evaluated at every argument point a case meets
// (oracle_channel_mode 1 and 2: record the points, replay with values given for them)
// the golden kernel sum moves by 3.6e-4 only.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include <dlfcn.h>

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
// The plain kernel: the T tensors of the Mainz representation, tabulated by the
// library's generator from the kernel's defining integrals (file format and grid:
// a 64-byte header, then float32 [33 |x|][33 beta][49 |y|][42] frame components;
// |x| = R i/32, beta = pi j/32, |y| = R (k/48)^2, R = 12 fm in muon units),
// interpolated cubically in (|x|, beta, sqrt|y|) and assembled in the frame of the
// argument pair; G^A from gamma traces; the Lambda = 0.4 subtraction; 0 beyond R.
constexpr double kLambda = 0.4;
constexpr double kRadius = 12.0 * 105.6583745 / 197.3269804;
constexpr int NA = 33, NBETA = 33, NB = 49, NCH = 42;
constexpr double kPi = 3.14159265358979323846;

struct Table {
  std::vector<double> t;  // [NA][NBETA][NB][NCH]
  std::string path;
  explicit Table(const std::string& p) : path(p) {
    std::FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) throw std::runtime_error("oracle: cannot open QED table " + p);
    unsigned char h[64];
    const bool ok_h = std::fread(h, 1, 64, f) == 64;
    uint32_t hu[6];
    double rad;
    std::memcpy(hu, h, sizeof(hu));
    std::memcpy(&rad, h + 24, 8);
    std::vector<float> raw((size_t)NA * NBETA * NB * NCH);
    const bool ok_d = ok_h && std::fread(raw.data(), sizeof(float), raw.size(), f) == raw.size();
    std::fclose(f);
    if (!ok_d || hu[0] != 0x54514C48u || hu[1] != 1u || hu[2] != (uint32_t)NA || hu[3] != (uint32_t)NBETA ||
        hu[4] != (uint32_t)NB || hu[5] != (uint32_t)NCH || rad != kRadius)
      throw std::runtime_error("oracle: " + p + " is not the expected QED table");
    t.assign(raw.begin(), raw.end());
  }
  // four-node Lagrange weights around the node position p on a grid of n nodes
  static int lag(double p, int n, double w[4]) {
    int i0 = (int)std::floor(p) - 1;
    i0 = std::min(std::max(i0, 0), n - 4);
    const double u = p - i0;
    w[0] = -(u - 1) * (u - 2) * (u - 3) * (1.0 / 6.0);
    w[1] = u * (u - 2) * (u - 3) * 0.5;
    w[2] = -u * (u - 1) * (u - 3) * 0.5;
    w[3] = u * (u - 1) * (u - 2) * (1.0 / 6.0);
    return i0;
  }
  void interp(double a, double be, double b, double* ch) const {
    a = std::min(std::max(a, 0.0), kRadius);
    b = std::min(std::max(b, 0.0), kRadius);
    double wa[4], wc[4], wb[4];
    const int ia = lag(a * ((NA - 1) / kRadius), NA, wa);
    const int ic = lag(be * ((NBETA - 1) / kPi), NBETA, wc);
    const int ib = lag(std::sqrt(b * (1.0 / kRadius)) * (NB - 1), NB, wb);
    for (int q = 0; q < NCH; ++q) {
      double r = 0;
      for (int i = 0; i < 4; ++i) {
        double ri = 0;
        for (int j = 0; j < 4; ++j) {
          const double* p = &t[((((size_t)(ia + i) * NBETA + (ic + j)) * NB) + ib) * NCH + q];
          ri += wc[j] * (wb[0] * p[0] + wb[1] * p[NCH] + wb[2] * p[2 * NCH] + wb[3] * p[3 * NCH]);
        }
        r += wa[i] * ri;
      }
      ch[q] = r;
    }
  }
};
std::string default_table_path() {
  if (const char* env = std::getenv("HLBL_QED_TABLE")) return env;
  Dl_info info;
  std::string dir = ".";
  if (dladdr(reinterpret_cast<void*>(&default_table_path), &info) != 0 && info.dli_fname) {
    dir = info.dli_fname;
    const size_t slash = dir.find_last_of('/');
    dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
  }
  return dir + "/../hlbl_amd/data/qed_kernel_table.bin";
}
std::unique_ptr<Table> g_table;
const Table& table() {
  if (!g_table) g_table.reset(new Table(default_table_path()));
  return *g_table;
}

// (|x|, beta, |y|), beta = atan2(|x ^ y|, x.y), 0 when either argument is 0
void grid_coords(const double* x, const double* y, double& a, double& be, double& b) {
  double u = 0, v = 0, w = 0, s2 = 0;
  for (int i = 0; i < 4; ++i) {
    u += x[i] * x[i];
    v += y[i] * y[i];
    w += x[i] * y[i];
  }
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) {
      const double m = x[i] * y[j] - x[j] * y[i];
      s2 += m * m;
    }
  a = std::sqrt(u);
  b = std::sqrt(v);
  be = (u > 0 && v > 0) ? std::atan2(std::sqrt(s2), w) : 0.0;
}

// e0 = x^ (y^ if x = 0, axis 0 if both are 0); e1 = unit part of y orthogonal to e0,
// or, if y is parallel to e0 (|x ^ y| <= 1e-12 |x||y|) or 0, of the first axis on
// which |e0| is smallest
void frame(const double* x, const double* y, double* e0, double* e1) {
  double u = 0, v = 0, s2 = 0;
  for (int i = 0; i < 4; ++i) {
    u += x[i] * x[i];
    v += y[i] * y[i];
  }
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) {
      const double m = x[i] * y[j] - x[j] * y[i];
      s2 += m * m;
    }
  for (int i = 0; i < 4; ++i)
    e0[i] = u > 0 ? x[i] / std::sqrt(u) : (v > 0 ? y[i] / std::sqrt(v) : (i == 0 ? 1.0 : 0.0));
  double t[4];
  if (u > 0 && v > 0 && s2 > 1e-24 * u * v) {
    double c = 0;
    for (int i = 0; i < 4; ++i) c += y[i] * e0[i];
    for (int i = 0; i < 4; ++i) t[i] = y[i] - c * e0[i];
  } else {
    int m = 0;
    for (int i = 1; i < 4; ++i)
      if (std::fabs(e0[i]) < std::fabs(e0[m])) m = i;
    for (int i = 0; i < 4; ++i) t[i] = (i == m ? 1.0 : 0.0) - e0[m] * e0[i];
  }
  const double n = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + t[3] * t[3]);
  for (int i = 0; i < 4; ++i) e1[i] = t[i] / n;
}

// T^I, T^II, T^III [3][64], index (al*4+be)*4+de, from the 42 frame components
void t_tensors(const double* x, const double* y, const double* ch, double T[3][64]) {
  double e0[4], e1[4];
  frame(x, y, e0, e1);
  const double* E[2] = {e0, e1};
  double P[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) P[i][j] = (i == j) - e0[i] * e0[j] - e1[i] * e1[j];
  for (int A = 0; A < 3; ++A) {
    const double* c = ch + 14 * A;
    for (int al = 0; al < 4; ++al)
      for (int be = 0; be < 4; ++be)
        for (int de = 0; de < 4; ++de) {
          double r = 0;
          for (int p = 0; p < 2; ++p)
            for (int q = 0; q < 2; ++q)
              for (int s = 0; s < 2; ++s) r += c[(p * 2 + q) * 2 + s] * E[p][al] * E[q][be] * E[s][de];
          for (int d = 0; d < 2; ++d)
            r += c[8 + d] * P[al][be] * E[d][de] + c[10 + d] * P[al][de] * E[d][be] + c[12 + d] * E[d][al] * P[be][de];
          T[A][(al * 4 + be) * 4 + de] = r;
        }
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

// Where the T tensors' frame components come from: 0 the table (default); 1 the table,
// recording each argument point (|x|, beta, |y|) met; 2 values given per point (the
// kernel's integrals evaluated there), the table for a point without one. Points are
// keyed by their three coordinates rounded to 1e-12.
int g_channel_mode = 0;
std::mutex g_channel_mu;
std::map<std::array<long long, 3>, std::vector<double>> g_channel_points;
long long g_channel_misses = 0;
std::array<long long, 3> channel_key(double a, double be, double b) {
  return {std::llround(a * 1e12), std::llround(be * 1e12), std::llround(b * 1e12)};
}
void channels(double a, double be, double b, double* ch) {
  if (g_channel_mode == 2) {
    const auto it = g_channel_points.find(channel_key(a, be, b));
    if (it != g_channel_points.end()) {
      std::memcpy(ch, it->second.data(), NCH * sizeof(double));
      return;
    }
    std::lock_guard<std::mutex> l(g_channel_mu);
    ++g_channel_misses;
  }
  table().interp(a, be, b, ch);
  if (g_channel_mode == 1) {
    std::lock_guard<std::mutex> l(g_channel_mu);
    g_channel_points.emplace(channel_key(a, be, b), std::vector<double>{a, be, b});
  }
}

void l0_eval(const double* x, const double* y, double* L0) {
  double a, be, b, ch[NCH];
  grid_coords(x, y, a, be, b);
  channels(a, be, b, ch);
  double T[3][64];
  t_tensors(x, y, ch, T);
  const double* Tf = &T[0][0];
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

// use the QED table at `path` (else HLBL_QED_TABLE, else ../hlbl_amd/data beside this
// library); 0 on success, -1 if it cannot be read
int oracle_load_table(const char* path) {
  try {
    g_table.reset(new Table(path));
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}
// interpolated frame components (42) at (a = |x|, beta, b = |y|)
void oracle_channels_interp(double a, double be, double b, double* ch) { table().interp(a, be, b, ch); }

// channel source of the kernel evaluations: 0 table, 1 table recording the points met,
// 2 the values set by oracle_channel_points_set (table where none); switching to 1
// clears the recorded points
void oracle_channel_mode(int mode) {
  if (mode == 1) g_channel_points.clear();
  g_channel_misses = 0;
  g_channel_mode = mode;
}
// the recorded points: n = their count; with out, [n][3] = (|x|, beta, |y|)
int oracle_channel_points(double* out) {
  int i = 0;
  for (const auto& kv : g_channel_points) {
    if (out)
      for (int j = 0; j < 3; ++j) out[3 * i + j] = kv.second[j];
    ++i;
  }
  return i;
}
// values for n points: abb [n][3], ch [n][42]
void oracle_channel_points_set(const double* abb, const double* ch, int n) {
  g_channel_points.clear();
  for (int i = 0; i < n; ++i)
    g_channel_points[channel_key(abb[3 * i], abb[3 * i + 1], abb[3 * i + 2])] =
        std::vector<double>(ch + (size_t)NCH * i, ch + (size_t)NCH * (i + 1));
}
// evaluations in mode 2 that found no value
long long oracle_channel_misses(void) { return g_channel_misses; }
// the table [33][33][49][42] as doubles
void oracle_table(double* out) {
  const Table& t = table();
  std::memcpy(out, t.t.data(), t.t.size() * sizeof(double));
}
size_t oracle_table_size(void) { return table().t.size(); }
// (|x|, beta, |y|) and the frame (e0, e1) of an argument pair
void oracle_grid_coords(const double* x, const double* y, double* abb, double* e0, double* e1) {
  grid_coords(x, y, abb[0], abb[1], abb[2]);
  frame(x, y, e0, e1);
}

// T tensors [3][4][4][4] at (x, y) from given frame components ch[42]
void oracle_t_tensors(const double* x, const double* y, const double* ch, double* T) {
  double t[3][64];
  t_tensors(x, y, ch, t);
  std::memcpy(T, t, sizeof(t));
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
// The surrogate network in FP64: 3 -> 4 x 128 GELU -> 42, parameters packed
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
const int kIn[5] = {3, 128, 128, 128, 128}, kOut[5] = {128, 128, 128, 128, 42};
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
    std::vector<double> d(42);
    for (int c = 0; c < 42; ++c) {
      const double e = a[5][c] - t[42 * r + c];
      loss += e * e;
      d[c] = 2.0 * e / ((double)n * 42);
      if (y) y[42 * r + c] = a[5][c];
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
  return loss / ((double)n * 42);
}

// Forward of the BF16 inference: BF16 weights, every hidden activation rounded
// to BF16 after the GELU, FP64 arithmetic otherwise. y[n][42].
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
    for (int c = 0; c < 42; ++c) y[42 * r + c] = a[c];
  }
}

}  // extern "C"
