// Writes the tabulated plain QED kernel (qed_grid.hpp format) from the defining
// integrals (qed_exact.hpp), OpenMP over the grid points.
//
//   gen_qed_table OUT [nt nr ntheta npsi]
//   gen_qed_table --symmetrize FILE     (the last step alone, on a written table)
//
// The committed table was made with the default quadrature (16 28 8 16): its
// components agree with a four-times finer quadrature to about 2e-4 of each
// tensor's largest component.
//
// Last step: at beta = 0 and pi (x parallel to y; also every pair with x = 0 or
// y = 0, which reads the beta = 0 nodes) the configuration is symmetric under O(3)
// about e0, so each tensor is t000 e0e0e0 + A Q_ij e0_k + B Q_ik e0_j + C e0_i Q_jk,
// Q the projector orthogonal to e0 (which holds e1): t110 = tPP0, t101 = tP0P,
// t011 = t0PP, and every component with an odd number of e1 or P indices is 0.
// Those nodes get the mean of each pair and zeroes, so the kernel at such pairs
// does not depend on the choice of e1, which is arbitrary there.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <omp.h>

#include "../csrc/qed_exact.hpp"
#include "../csrc/qed_grid.hpp"

using namespace hlbl;
using namespace hlbl::qed_grid;

namespace {

constexpr size_t kValues = (size_t)kNA * kNBeta * kNB * kNChan;

void symmetrize(std::vector<float>& data) {
  const int pairs[3][2] = {{6, 8}, {5, 10}, {3, 12}};
  const int odd[7] = {1, 2, 4, 7, 9, 11, 13};
  for (int i = 0; i < kNA; ++i)
    for (int j : {0, kNBeta - 1})
      for (int k = 0; k < kNB; ++k) {
        float* q = &data[(((size_t)i * kNBeta + j) * kNB + k) * kNChan];
        for (int A = 0; A < 3; ++A) {
          float* t = q + 14 * A;
          for (const auto& pr : pairs) {
            const float m = (float)(0.5 * ((double)t[pr[0]] + (double)t[pr[1]]));
            t[pr[0]] = t[pr[1]] = m;
          }
          for (int c : odd) t[c] = 0.0f;
        }
      }
}

bool write_table(const char* path, const FileHeader& h, const std::vector<float>& data) {
  std::FILE* fp = std::fopen(path, "wb");
  if (!fp) {
    std::perror(path);
    return false;
  }
  const bool ok = std::fwrite(&h, sizeof(h), 1, fp) == 1 &&
                  std::fwrite(data.data(), sizeof(float), data.size(), fp) == data.size();
  std::fclose(fp);
  if (!ok) std::fprintf(stderr, "%s: write failed\n", path);
  return ok;
}

int symmetrize_file(const char* path) {
  std::FILE* fp = std::fopen(path, "rb");
  if (!fp) {
    std::perror(path);
    return 1;
  }
  FileHeader h;
  std::vector<float> data(kValues);
  const bool ok = std::fread(&h, sizeof(h), 1, fp) == 1 &&
                  std::fread(data.data(), sizeof(float), data.size(), fp) == data.size();
  std::fclose(fp);
  if (!ok || h.magic != kMagic || h.version != kVersion || h.na != (uint32_t)kNA ||
      h.nbeta != (uint32_t)kNBeta || h.nb != (uint32_t)kNB || h.nch != (uint32_t)kNChan) {
    std::fprintf(stderr, "%s: not a kernel table of this grid\n", path);
    return 1;
  }
  symmetrize(data);
  return write_table(path, h, data) ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc == 3 && std::strcmp(argv[1], "--symmetrize") == 0) return symmetrize_file(argv[2]);
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUT [nt nr ntheta npsi] | --symmetrize FILE\n", argv[0]);
    return 2;
  }
  qed_exact::Config cfg;
  cfg.nt = 16;
  cfg.nr = 28;
  cfg.nth = 8;
  cfg.nps = 16;
  if (argc >= 6) {
    cfg.nt = std::atoi(argv[2]);
    cfg.nr = std::atoi(argv[3]);
    cfg.nth = std::atoi(argv[4]);
    cfg.nps = std::atoi(argv[5]);
  }
  const size_t npts = (size_t)kNA * kNBeta * kNB;
  std::vector<float> data(kValues, 0.0f);
  const qed_exact::Evaluator ev(cfg);
  size_t done = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (long long p = 0; p < (long long)npts; ++p) {
    const int i = (int)(p / ((size_t)kNBeta * kNB));
    const int j = (int)((p / kNB) % kNBeta);
    const int k = (int)(p % kNB);
    const double a = node_a(i), b = node_b(k);
    // at |x| = 0 or |y| = 0 the angle is immaterial up to the frame; the evaluation
    // at the node's own angle keeps the table continuous in |x| and |y|
    const qed_exact::Tframe f = ev.eval(a, b, node_beta(j));
    float* out = &data[(size_t)p * kNChan];
    for (int A = 0; A < 3; ++A)
      for (int c = 0; c < 14; ++c) out[A * 14 + c] = (float)f.c[A][c];
#pragma omp atomic
    ++done;
    if (omp_get_thread_num() == 0 && p % 2000 == 0) {
      std::fprintf(stderr, "%zu / %zu points\n", done, npts);
      std::fflush(stderr);
    }
  }
  symmetrize(data);
  FileHeader h;
  std::memset(&h, 0, sizeof(h));
  h.magic = kMagic;
  h.version = kVersion;
  h.na = kNA;
  h.nbeta = kNBeta;
  h.nb = kNB;
  h.nch = kNChan;
  h.radius = kRadius;
  h.nt = cfg.nt;
  h.nr = cfg.nr;
  h.ntheta = cfg.nth;
  h.npsi = cfg.nps;
  return write_table(argv[1], h, data) ? 0 : 1;
}
