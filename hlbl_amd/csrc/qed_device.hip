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
  double a, be, b, ch[qed::kNChUsed], t[qed::kNT], e0[4], e1[4];
  qed::grid_coords(u, v, a, be, b);
  const qed::Cell<double> cell = qed::locate(a, be, b);
  for (int i = 0; i < qed::kNChUsed; ++i) ch[i] = qed::interp(table, cell, i);
  qed::frame(u, v, e0, e1);
  for (int i = 0; i < qed::kNT; ++i) t[i] = qed::t_value(qed::t_args(i & 63, e0, e1), ch + qed::kFrame * (i >> 6));
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
// kernel_p4 of compute_2p2_gpu, kernels/hlbl_kernels/kernels.cu:374-760) for one
// target, summed over the sites. With x = z - w and y = target - w (mod L; y given in
// [0, L)), each argument difference min-imaged on its own and scaled by xunit, k over
// (r, s) = (0,1),(0,2),(0,3),(1,2),(1,3),(2,3):
//   P2_0[r][s][nu] += L[k](x, y)[mu][nu][lam] Pi_{mu lam}(z)
//   P2_1[r][s][nu] += L[k](y, x)[nu][mu][lam] Pi_{mu lam}(z)
//   P3  [r][s][nu] += L[k](x - y, -y)[mu][lam][nu] Pi_{mu lam}(z)
//   P4_0[r][s][nu] += L[k](y - x, -x)[nu][lam][mu] Pi_{mu lam}(z)
//   P4_1[r][s][nu] += xv_r (that term), P4_1[s][r][nu] -= xv_s (that term) (xv the integer
//   min-image of x).
// One wave per site (a fixed set of sites per wave): the three L0 evaluations of each
// kernel L = L0(u,v) - D(u) L0(0,v) - D(v) L0(u,0) are computed by the wave in its own
// LDS area (channels, T tensors, L0 from the term table), the lanes forming their own
// rows of the subtraction matrices; each wave accumulates its sites' 320 sums in LDS,
// the workgroup adds its waves in order into part[workgroup][320], and p234_reduce adds
// the workgroups in order. Deterministic: the site-to-wave map depends on V and the
// grid only.
constexpr int kP234Waves = 4;
// v[i] of an integer 4-vector for a runtime i, without indexing (keeps v in registers)
__device__ inline double blend4i(const int v[4], int i) {
  return (double)(i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3])));
}
constexpr int kP234Chan = 48;                                     // channel row of one evaluation
constexpr int kP234Lds = kP234Chan + qed::kNT + 3 * qed::kNL;     // per wave, doubles

// L0 of the argument pair (u, v) into dst[384], by the calling wave (ch: kP234Chan,
// T: kNT doubles of its LDS area)
__device__ __forceinline__ void wave_l0_pair(const double* __restrict__ table, const TermLds<double>& tt,
                                             const double u[4], const double v[4], double* ch, double* T,
                                             double* dst) {
  const int lane = threadIdx.x & 63;
  if (lane < qed::kNChUsed) {
    double a, be, b;
    qed::grid_coords(u, v, a, be, b);
    ch[lane] = qed::interp(table, qed::locate(a, be, b), lane);
  }
  wave_sync();
  {
    double e0[4], e1[4];
    qed::frame(u, v, e0, e1);
    const qed::TArgs<double> t = qed::t_args(lane, e0, e1);
#pragma unroll
    for (int A = 0; A < 3; ++A) T[64 * A + lane] = qed::t_value(t, ch + qed::kFrame * A);
  }
  wave_sync();
  for (int t = lane; t < qed::kNL; t += 64) {
    const int o = l0_output(tt, t);
    dst[o] = l0_lds(tt, o, T);
  }
  wave_sync();
}

__global__ __launch_bounds__(64 * kP234Waves, 2) void p234_waves(
    const double* __restrict__ table, const double* __restrict__ pi, hlbl_geom local, hlbl_geom global,
    hlbl_coord proc, hlbl_coord gsw, hlbl_coord y, double xunit, double* __restrict__ part) {
  __shared__ TermLds<double> tt;
  __shared__ double wlds[kP234Waves][kP234Lds];
  __shared__ double acc[kP234Waves][320];
  __shared__ double spi[kP234Waves][16];  // Pi_{mu lam}(z) of the wave's site
  stage_terms(c_terms_q, &tt);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int j = lane; j < 320; j += 64) acc[wave][j] = 0.0;
  __syncthreads();
  double* ch = wlds[wave];
  double* T = ch + kP234Chan;
  double* L0 = T + qed::kNT;  // [3][384]: (u, v) then the kernel in place, (0, v), (u, 0)
  const size_t V = geom_volume(local);
  const double rr = qed::kRadius * qed::kRadius, lam = qed::kLambda;
  // the lane's L0 outputs o = lane + 64 kk share (mu, nu, lam) = ((o>>4)&3, (o>>2)&3, o&3)
  const int om = (lane >> 4) & 3, on = (lane >> 2) & 3, ol = lane & 3;
  for (size_t z = (size_t)blockIdx.x * kP234Waves + wave; z < V; z += (size_t)gridDim.x * kP234Waves) {
    int c[4], x[4], xv[4], yi[4];
    site_coords(z, local, c);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int L = geom_dim(global, r);
      x[r] = mod_pos(c[r] + geom_dim(local, r) * coord_dim(proc, r) - coord_dim(gsw, r), L);
      xv[r] = coord_map(x[r], L);
      yi[r] = coord_dim(y, r);
    }
    if (lane < 16) spi[wave][lane] = pi[z * 16 + lane];
    wave_sync();
    const double* p = spi[wave];
#pragma unroll 1
    for (int geo = 0; geo < 4; ++geo) {
      double u[4], v[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int L = geom_dim(global, r);
        const double xm = xv[r] * xunit, ym = coord_map(yi[r], L) * xunit;
        u[r] = geo == 0 ? xm : geo == 1 ? ym : geo == 2 ? coord_map(mod_pos(x[r] - yi[r], L), L) * xunit
                                                        : coord_map(mod_pos(yi[r] - x[r], L), L) * xunit;
        v[r] = geo == 0 ? ym : geo == 1 ? xm : geo == 2 ? coord_map(mod_pos(-yi[r], L), L) * xunit
                                                        : coord_map(mod_pos(-x[r], L), L) * xunit;
      }
      if (norm2(u) > rr || norm2(v) > rr) continue;  // the kernel is 0 (wave-uniform)
      const double zero[4] = {0.0, 0.0, 0.0, 0.0};
      wave_l0_pair(table, tt, u, v, ch, T, L0);
      wave_l0_pair(table, tt, zero, v, ch, T, L0 + qed::kNL);
      wave_l0_pair(table, tt, u, zero, ch, T, L0 + 2 * qed::kNL);
      {
        // L = L0(u,v) - D(u) L0(0,v) - D(v) L0(u,0), D_{m a}(p) = (delta - Lambda p_m p_a) e^{-Lambda p^2/2}
        const double eu = exp(-lam * norm2(u) * 0.5), ev = exp(-lam * norm2(v) * 0.5);
        const double um = blend4(u, om), vn = blend4(v, on);
        double dm[4], dn[4];
#pragma unroll
        for (int al = 0; al < 4; ++al) {
          dm[al] = ((om == al ? 1.0 : 0.0) - lam * um * u[al]) * eu;
          dn[al] = ((on == al ? 1.0 : 0.0) - lam * vn * v[al]) * ev;
        }
        const double* l00v = L0 + qed::kNL;
        const double* l0u0 = L0 + 2 * qed::kNL;
#pragma unroll 2
        for (int kk = 0; kk < 6; ++kk) {
          const int o = lane + 64 * kk;
          double k = L0[o];
#pragma unroll
          for (int al = 0; al < 4; ++al) {
            k -= dm[al] * l00v[((kk * 4 + al) * 4 + on) * 4 + ol];
            k -= dn[al] * l0u0[((kk * 4 + om) * 4 + al) * 4 + ol];
          }
          L0[o] = k;
        }
      }
      wave_sync();
      if (lane < 24) {
        const int k = lane >> 2, nu = lane & 3, r0 = comb_rho(k), s1 = comb_sigma(k);
        double t = 0.0;
#pragma unroll
        for (int mu = 0; mu < 4; ++mu)
#pragma unroll
          for (int la = 0; la < 4; ++la) {
            const int o = geo == 0 ? (mu * 4 + nu) * 4 + la
                        : geo == 1 ? (nu * 4 + mu) * 4 + la
                        : geo == 2 ? (mu * 4 + la) * 4 + nu
                                   : (nu * 4 + la) * 4 + mu;
            t = fma(L0[k * 64 + o], p[mu * 4 + la], t);
          }
        acc[wave][geo * 64 + r0 * 16 + s1 * 4 + nu] += t;
        if (geo == 3) {
          acc[wave][256 + r0 * 16 + s1 * 4 + nu] += blend4i(xv, r0) * t;
          acc[wave][256 + s1 * 16 + r0 * 4 + nu] -= blend4i(xv, s1) * t;
        }
      }
      wave_sync();
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < 320; j += 64 * kP234Waves) {
    double s = 0.0;
#pragma unroll
    for (int wv = 0; wv < kP234Waves; ++wv) s += acc[wv][j];
    part[(size_t)blockIdx.x * 320 + j] = s;
  }
}

// out[j] = sum over the n workgroup partials part[wg][j], in workgroup order
__global__ void p234_reduce(const double* __restrict__ part, int n, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= 320) return;
  double s = 0.0;
  for (int g = 0; g < n; ++g) s += part[(size_t)g * 320 + j];
  out[j] = s;
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
  const int nwg = (int)std::min<size_t>((V + kP234Waves - 1) / kP234Waves, (size_t)4 * c->n_cu);
  c->part234.ensure((size_t)nwg * 320);
  for (int t = 0; t < n_targets; ++t) {
    const hlbl_geom& g = c->global;
    const hlbl_coord y = {mod_pos(targets[t].t - gsw.t, g.T), mod_pos(targets[t].x - gsw.x, g.LX),
                          mod_pos(targets[t].y - gsw.y, g.LY), mod_pos(targets[t].z - gsw.z, g.LZ)};
    hipLaunchKernelGGL(p234_waves, dim3(nwg), dim3(64 * kP234Waves), 0, st, c->qed->table64.p, pi, c->local,
                       c->global, c->proc, gsw, y, xunit, c->part234.p);
    check_hip(hipGetLastError(), "p234_waves launch");
    hipLaunchKernelGGL(p234_reduce, dim3(5), dim3(64), 0, st, c->part234.p, nwg, d_p234 + (size_t)t * 320);
    check_hip(hipGetLastError(), "p234_reduce launch");
  }
}

// L^(Lambda) at n points with the caller's scratch (n * 3 * kNL doubles)
void launch_qed_points(const double* table, const double* x, const double* y, int n, double* kerv,
                       double* scratch, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(qed_kernel_points, dim3((n + 63) / 64), dim3(64), 0, st, table, x, y, n, kerv, scratch);
  check_hip(hipGetLastError(), "qed_kernel_points launch");
}

void launch_qed_kernel(const hlbl_qed* q, int mode, const double* x, const double* y, int n,
                       double* kerv, hipStream_t st) {
  if (mode != 0) throw Error("hlbl_qed_kernel: only the plain form (mode 0) evaluates single points");
  if (n <= 0) return;
  double* scratch = nullptr;
  check_hip(hipMallocAsync((void**)&scratch, (size_t)n * 3 * qed::kNL * sizeof(double), st),
            "hipMallocAsync");
  try {
    launch_qed_points(q->table64.p, x, y, n, kerv, scratch, st);
  } catch (...) {
    (void)hipFreeAsync(scratch, st);
    throw;
  }
  (void)hipFreeAsync(scratch, st);
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
