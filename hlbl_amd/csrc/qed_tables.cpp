// Host side of the QED kernel: the G^A term table from gamma-matrix traces,
// and the reader of the tabulated T tensors.
#include <cmath>
#include <complex>
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <fstream>
#include <string>
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

// ---- the tabulated T tensors --------------------------------------------------

void load_table(const std::string& path, std::vector<double>& table) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("QED kernel table: cannot open " + path);
  qed_grid::FileHeader h;
  in.read(reinterpret_cast<char*>(&h), sizeof(h));
  if (!in || h.magic != qed_grid::kMagic || h.version != qed_grid::kVersion)
    throw std::runtime_error("QED kernel table: " + path + " is not a version-1 table");
  if ((int)h.na != kNA || (int)h.nbeta != kNC || (int)h.nb != kNB || (int)h.nch != kNChUsed ||
      h.radius != kRadius)
    throw std::runtime_error("QED kernel table: " + path + " has another grid");
  const size_t npts = (size_t)kNA * kNC * kNB;
  std::vector<float> raw(npts * kNChUsed);
  in.read(reinterpret_cast<char*>(raw.data()), (std::streamsize)(raw.size() * sizeof(float)));
  if (!in) throw std::runtime_error("QED kernel table: " + path + " is truncated");
  table.assign(npts * kNCh, 0.0);
  for (size_t p = 0; p < npts; ++p)
    for (int c = 0; c < kNChUsed; ++c) table[p * kNCh + c] = (double)raw[p * kNChUsed + c];
}

}  // namespace qed
}  // namespace hlbl
