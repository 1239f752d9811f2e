// QED kernel entries for single points (tests and the surrogate's judge), and
// the device random fill of propagator blocks.
#include <hip/hip_runtime.h>

#include <algorithm>

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
// Element j of a lane's scratch at p[j * s]: with p = base + lane and s the number of
// lanes, a wave's accesses to one element are contiguous.
struct LaneView {
  double* p;
  size_t s;
  __device__ double& operator[](size_t j) const { return p[j * s]; }
  __device__ LaneView operator+(size_t j) const { return {p + j * s, s}; }
};

// P: double* or LaneView
template <class P>
__device__ void l0_point(const double* table, const double u[4], const double v[4], P l0) {
  double a, c, b, ch[qed::kNCh], t[qed::kNT];
  qed::acb(u, v, a, c, b);
  const qed::Cell<double> cell = qed::locate(a, c, b);
  for (int i = 0; i < qed::kNChUsed; ++i) ch[i] = qed::interp(table, cell, i);
  ch[qed::kNCh - 1] = 0.0;
  for (int i = 0; i < qed::kNT; ++i) t[i] = qed::t_component(i, u, v, ch);
  for (int o = 0; o < qed::kNL; ++o) l0[o] = qed::l0_component(c_terms_q, o, t);
}

// L^(Lambda)(x, y) into out[384] (plain form, FP64; 0 beyond the radius), with
// 3 x 384 doubles of scratch
template <class P>
__device__ void kernel_point(const double* table, const double x[4], const double y[4], P scratch, P out) {
  const double zero[4] = {0, 0, 0, 0};
  const double rr = qed::kRadius * qed::kRadius;
  if (norm2(x) > rr || norm2(y) > rr) {
    for (int o = 0; o < qed::kNL; ++o) out[o] = 0.0;
    return;
  }
  const P l0xy = scratch;
  const P l00y = l0xy + qed::kNL;
  const P l0x0 = l00y + qed::kNL;
  l0_point(table, x, y, l0xy);
  l0_point(table, zero, y, l00y);
  l0_point(table, x, zero, l0x0);
  double dx[16], dy[16];
  qed::subtraction_matrix(x, dx);
  qed::subtraction_matrix(y, dy);
  for (int o = 0; o < qed::kNL; ++o) out[o] = qed::lambda_subtract(o, l0xy, l00y, l0x0, dx, dy);
}

// L^(Lambda)(x, y) at n pairs, one thread each (plain form)
__global__ void qed_kernel_points(const double* __restrict__ table, const double* __restrict__ X,
                                  const double* __restrict__ Yv, int n, double* __restrict__ kerv,
                                  double* __restrict__ scratch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[4], y[4];
  for (int r = 0; r < 4; ++r) {
    x[r] = X[4 * i + r];
    y[r] = Yv[4 * i + r];
  }
  kernel_point(table, x, y, scratch + (size_t)i * 3 * qed::kNL, kerv + (size_t)i * qed::kNL);
}

// The 2+2 pieces P2_0, P2_1, P3, P4_0, P4_1 (kernel_p20, kernel_p21, kernel_p3,
// kernel_p4 of compute_2p2_gpu, kernels/hlbl_kernels/kernels.cu:374-760) of sites
// [s0, s0 + n) for one target: contrib[5][rho][sigma][nu][site]. With x = z - w and
// y = target - w (mod L; y given in [0, L)), each argument difference min-imaged on its
// own and scaled by xunit, k over (r, s) = (0,1),(0,2),(0,3),(1,2),(1,3),(2,3):
//   P2_0[r][s][nu] = L[k](x, y)[mu][nu][lam] Pi_{mu lam}(z)
//   P2_1[r][s][nu] = L[k](y, x)[nu][mu][lam] Pi_{mu lam}(z)
//   P3  [r][s][nu] = L[k](x - y, -y)[mu][lam][nu] Pi_{mu lam}(z)
//   P4_0[r][s][nu] = L[k](y - x, -x)[nu][lam][mu] Pi_{mu lam}(z)
//   P4_1[r][s][nu] = xv_r P4_0 term, P4_1[s][r][nu] = -xv_s P4_0 term (xv the integer
//   min-image of x). One thread per site; scratch 4 x 384 doubles per thread, lane-strided
//   (element j of thread i at scratch[j n + i]); contrib[j][V], site-minor.
__global__ void p234_sites(const double* __restrict__ table, const double* __restrict__ pi,
                           hlbl_geom local, hlbl_geom global, hlbl_coord proc, hlbl_coord gsw,
                           hlbl_coord y, double xunit, size_t s0, size_t n, size_t V, double* __restrict__ scratch,
                           double* __restrict__ contrib) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t z = s0 + i;
  int c[4];
  site_coords(z, local, c);
  int x[4], xv[4], yi[4];
  double xm[4], ym[4], a[4], b[4];
  for (int r = 0; r < 4; ++r) {
    const int L = geom_dim(global, r);
    x[r] = mod_pos(c[r] + geom_dim(local, r) * coord_dim(proc, r) - coord_dim(gsw, r), L);
    xv[r] = coord_map(x[r], L);
    yi[r] = coord_dim(y, r);
    xm[r] = xv[r] * xunit;
    ym[r] = coord_map(yi[r], L) * xunit;
  }
  const LaneView K{scratch + i, n};
  const LaneView tmp = K + qed::kNL;
  const LaneView out{contrib + z, V};
  for (int j = 0; j < 320; ++j) out[j] = 0.0;
  double p[16];
  for (int j = 0; j < 16; ++j) p[j] = pi[z * 16 + j];
  for (int geo = 0; geo < 4; ++geo) {
    if (geo == 0) kernel_point(table, xm, ym, tmp, K);
    if (geo == 1) kernel_point(table, ym, xm, tmp, K);
    if (geo == 2) {
      for (int r = 0; r < 4; ++r) {
        const int L = geom_dim(global, r);
        a[r] = coord_map(mod_pos(x[r] - yi[r], L), L) * xunit;
        b[r] = coord_map(mod_pos(-yi[r], L), L) * xunit;
      }
      kernel_point(table, a, b, tmp, K);
    }
    if (geo == 3) {
      for (int r = 0; r < 4; ++r) {
        const int L = geom_dim(global, r);
        a[r] = coord_map(mod_pos(yi[r] - x[r], L), L) * xunit;
        b[r] = coord_map(mod_pos(-x[r], L), L) * xunit;
      }
      kernel_point(table, a, b, tmp, K);
    }
    for (int k = 0; k < 6; ++k) {
      const int r0 = comb_rho(k), s1 = comb_sigma(k);
      for (int nu = 0; nu < 4; ++nu) {
        double t = 0.0;
        for (int mu = 0; mu < 4; ++mu)
          for (int lam = 0; lam < 4; ++lam) {
            const int o = geo == 0 ? (mu * 4 + nu) * 4 + lam
                        : geo == 1 ? (nu * 4 + mu) * 4 + lam
                        : geo == 2 ? (mu * 4 + lam) * 4 + nu
                                   : (nu * 4 + lam) * 4 + mu;
            t = fma(K[k * 64 + o], p[mu * 4 + lam], t);
          }
        out[geo * 64 + r0 * 16 + s1 * 4 + nu] = t;
        if (geo == 3) {
          out[256 + r0 * 16 + s1 * 4 + nu] = xv[r0] * t;
          out[256 + s1 * 16 + r0 * 4 + nu] = -xv[s1] * t;
        }
      }
    }
  }
}

// out[j] = sum over the V sites of contrib[j][site], j < 320: one workgroup per j,
// per-thread strided sums and a fixed tree
__global__ __launch_bounds__(256) void p234_reduce(const double* __restrict__ contrib, size_t V,
                                                   double* __restrict__ out) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double s = 0.0;
  for (size_t z = threadIdx.x; z < V; z += 256) s += contrib[(size_t)j * V + z];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

__global__ void qed_channels_points(const double* __restrict__ table, const double* __restrict__ acb,
                                    int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const qed::Cell<double> cell = qed::locate(acb[3 * i], acb[3 * i + 1], acb[3 * i + 2]);
  for (int c = 0; c < qed::kNChUsed; ++c) out[(size_t)i * qed::kNChUsed + c] = qed::interp(table, cell, c);
}

void launch_p234_2p2(hlbl_contractor* c, const double* pi, hlbl_coord gsw, const hlbl_coord* targets,
                     int n_targets, double xunit, double* d_p234, hipStream_t st) {
  const size_t V = c->volume;
  // sites per launch, one thread each with 4 x 384 doubles (12 KiB) of scratch, strided
  // by the launch's site count: 2^19 sites are 8192 waves, 8 per SIMD, for 6 GiB
  constexpr size_t kChunk = size_t(1) << 19;
  double *scratch = nullptr, *contrib = nullptr;
  check_hip(hipMallocAsync((void**)&scratch, std::min(V, kChunk) * 4 * qed::kNL * sizeof(double), st),
            "hipMallocAsync");
  check_hip(hipMallocAsync((void**)&contrib, V * 320 * sizeof(double), st), "hipMallocAsync");
  hipError_t e = hipSuccess;
  for (int t = 0; t < n_targets && e == hipSuccess; ++t) {
    const hlbl_geom& g = c->global;
    const hlbl_coord y = {mod_pos(targets[t].t - gsw.t, g.T), mod_pos(targets[t].x - gsw.x, g.LX),
                          mod_pos(targets[t].y - gsw.y, g.LY), mod_pos(targets[t].z - gsw.z, g.LZ)};
    for (size_t s0 = 0; s0 < V && e == hipSuccess; s0 += kChunk) {
      const size_t n = std::min(kChunk, V - s0);
      hipLaunchKernelGGL(p234_sites, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, c->qed->table64.p,
                         pi, c->local, c->global, c->proc, gsw, y, xunit, s0, n, V, scratch, contrib);
      e = hipGetLastError();
    }
    if (e == hipSuccess) {
      hipLaunchKernelGGL(p234_reduce, dim3(320), dim3(256), 0, st, contrib, V, d_p234 + (size_t)t * 320);
      e = hipGetLastError();
    }
  }
  (void)hipFreeAsync(scratch, st);
  (void)hipFreeAsync(contrib, st);
  check_hip(e, "p234 launch");
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
