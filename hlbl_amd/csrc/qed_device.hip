// QED kernel entries for single points (tests and the surrogate's judge), and
// the device random fill of propagator blocks.
#include <hip/hip_runtime.h>

#include "context.hpp"
#include "contract_kernels.hpp"
#include "kpart.hpp"
#include "provider.hpp"
#include "qed.hpp"

namespace hlbl {

__constant__ qed::TermTable c_terms_q;

void upload_terms_qed(const qed::TermTable& t) {
  check_hip(hipMemcpyToSymbol(HIP_SYMBOL(c_terms_q), &t, sizeof(t)), "upload terms");
}

// L0(u, v) of one argument pair, FP64, into l0[384]
__device__ void l0_point(const double* table, const double u[4], const double v[4], double* l0) {
  double a, c, b, ch[qed::kNCh], t[qed::kNT];
  qed::acb(u, v, a, c, b);
  const qed::Cell<double> cell = qed::locate(a, c, b);
  for (int i = 0; i < qed::kNChUsed; ++i) ch[i] = qed::interp(table, cell, i);
  ch[qed::kNCh - 1] = 0.0;
  for (int i = 0; i < qed::kNT; ++i) t[i] = qed::t_component(i, u, v, ch);
  for (int o = 0; o < qed::kNL; ++o) l0[o] = qed::l0_component(c_terms_q, o, t);
}

// L^(Lambda)(x, y) at n pairs, one thread each (plain form)
__global__ void qed_kernel_points(const double* __restrict__ table, const double* __restrict__ X,
                                  const double* __restrict__ Yv, int n, double* __restrict__ kerv,
                                  double* __restrict__ scratch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[4], y[4], zero[4] = {0, 0, 0, 0};
  for (int r = 0; r < 4; ++r) {
    x[r] = X[4 * i + r];
    y[r] = Yv[4 * i + r];
  }
  double* out = kerv + (size_t)i * qed::kNL;
  const double rr = qed::kRadius * qed::kRadius;
  if (norm2(x) > rr || norm2(y) > rr) {
    for (int o = 0; o < qed::kNL; ++o) out[o] = 0.0;
    return;
  }
  double* l0xy = scratch + (size_t)i * 3 * qed::kNL;
  double* l00y = l0xy + qed::kNL;
  double* l0x0 = l00y + qed::kNL;
  l0_point(table, x, y, l0xy);
  l0_point(table, zero, y, l00y);
  l0_point(table, x, zero, l0x0);
  double dx[16], dy[16];
  qed::subtraction_matrix(x, dx);
  qed::subtraction_matrix(y, dy);
  for (int o = 0; o < qed::kNL; ++o) out[o] = qed::lambda_subtract(o, l0xy, l00y, l0x0, dx, dy);
}

__global__ void qed_channels_points(const double* __restrict__ table, const double* __restrict__ acb,
                                    int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const qed::Cell<double> cell = qed::locate(acb[3 * i], acb[3 * i + 1], acb[3 * i + 2]);
  for (int c = 0; c < qed::kNChUsed; ++c) out[(size_t)i * qed::kNChUsed + c] = qed::interp(table, cell, c);
}

void launch_qed_kernel(const hlbl_qed* q, int mode, const double* x, const double* y, int n,
                       double* kerv, hipStream_t st) {
  if (mode != 0) throw Error("hlbl_qed_kernel: only the plain form (mode 0) evaluates single points");
  if (n <= 0) return;
  double* scratch = nullptr;
  check_hip(hipMallocAsync((void**)&scratch, (size_t)n * 3 * qed::kNL * sizeof(double), st),
            "hipMallocAsync");
  hipLaunchKernelGGL(qed_kernel_points, dim3((n + 63) / 64), dim3(64), 0, st, q->table64.p, x, y,
                     n, kerv, scratch);
  hipError_t e = hipGetLastError();
  (void)hipFreeAsync(scratch, st);
  check_hip(e, "qed_kernel_points launch");
}

void launch_surrogate_points(hlbl_qed* q, const double* acb, int n, double* out, hipStream_t st);

void launch_qed_channels(const hlbl_qed* q, int mode, const double* acb, int n, double* out,
                         hipStream_t st) {
  if (mode != 0 && mode != 1) throw Error("hlbl_qed_channels: mode must be 0 (table) or 1 (surrogate)");
  if (n <= 0) return;
  if (mode == 1) {
    launch_surrogate_points(const_cast<hlbl_qed*>(q), acb, n, out, st);
    return;
  }
  hipLaunchKernelGGL(qed_channels_points, dim3((n + 127) / 128), dim3(128), 0, st, q->table64.p,
                     acb, n, out);
  check_hip(hipGetLastError(), "qed_channels_points launch");
}

// ---- device RandomProvider --------------------------------------------------
// one thread per (site, f, a): 24 doubles, written as 12 16-byte stores
__global__ void fill_random_kernel(double* __restrict__ out, hlbl_geom local, hlbl_coord proc,
                                   uint64_t seed, size_t V) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V * 24) return;
  const size_t site = i % V;
  const int fa = (int)(i / V);
  const int f = fa / 12, a = fa % 12;
  int c[4];
  site_coords(site, local, c);
  const uint64_t g = site_key(coord_dim(proc, 0) * local.T + c[0], coord_dim(proc, 1) * local.LX + c[1],
                              coord_dim(proc, 2) * local.LY + c[2], coord_dim(proc, 3) * local.LZ + c[3]);
  double2* dst = reinterpret_cast<double2*>(out + ((size_t)fa * V + site) * 24);
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    double2 v;
    v.x = hash_to_unit(element_key(g, f, a, 2 * k, seed));
    v.y = hash_to_unit(element_key(g, f, a, 2 * k + 1, seed));
    dst[k] = v;
  }
}

void launch_fill_random(double* d_out, hlbl_geom local, hlbl_coord proc, uint64_t seed,
                        hipStream_t st) {
  const size_t V = geom_volume(local);
  const size_t n = V * 24;
  hipLaunchKernelGGL(fill_random_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d_out,
                     local, proc, seed, V);
  check_hip(hipGetLastError(), "fill_random_kernel launch");
}

}  // namespace hlbl
