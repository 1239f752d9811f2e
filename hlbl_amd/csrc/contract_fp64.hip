// Plain FP64 contraction of one pair (both flavours): pass 1 (the z-sums dzu,
// dzsu, cuda_lattice.cu:160-262), its deterministic reduction and the
// gamma_mu gamma_5 insertion (src/contract.cpp:58-77), then pass 2 (the
// per-site 4-point contraction times the QED kernel, cuda_lattice.cu:277-484)
// with the 640 traces per site on FP64 MFMA.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "context.hpp"
#include "contract_kernels.hpp"
#include "kpart.hpp"
#include "lattice.hpp"
#include "pass1.hpp"
#include "pass_common.hpp"
#include "qed.hpp"
#include "window.hpp"

namespace hlbl {

__constant__ qed::TermTable c_terms64;

void upload_terms_fp64(const qed::TermTable& t) {
  check_hip(hipMemcpyToSymbol(HIP_SYMBOL(c_terms64), &t, sizeof(t)), "upload terms");
}

typedef double d4v __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- pass 1 --
// Pass 1 itself runs in pass1.hip; xi per pair is [2][8][5][144][2].
constexpr int kXiLen = 2 * 8 * 5 * 144 * 2;

// Per pair (blockIdx.y), with strides kXiLen (xi), kDzLen (dz), kGdzLen (gdz),
// kBmatLen (bmat):
// dz[f][j][a][b] (complex), j < 6: dzu[k = j], j >= 6: dzsu[sigma = j-6]:
//   M^{sigma,W}[a][b] = sum_s eps_s i^{ph_sigma(s)} Xi^W[cls(sigma)*4+s][a][b]
//   dzsu[sigma] = M^{sigma,1},  dzu[(rho,sigma)] = M^{sigma,xz_rho} - M^{rho,xz_sigma}
constexpr int kDzLen = 2 * 10 * 144 * 2;
constexpr int kGdzLen = 2 * 40 * 144 * 2;
constexpr int kBmatLen = 2 * 288 * kCoefCols;

__global__ void form_dz(const double* __restrict__ xi, double* __restrict__ dz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (f, j, ab)
  if (i >= 2 * 10 * 144) return;
  xi += (size_t)blockIdx.y * kXiLen;
  dz += (size_t)blockIdx.y * kDzLen;
  const int f = i / 1440, j = (i / 144) % 10, ab = i % 144;
  auto m = [&](int sigma, int w, double& re, double& im) {
    const int cls = (sigma == 0 || sigma == 3) ? 0 : 1;
    re = im = 0.0;
    for (int s = 0; s < 4; ++s) {
      const double* x = xi + ((((size_t)f * 8 + cls * 4 + s) * 5 + w) * 144 + ab) * 2;
      double xr = x[0], xim = x[1];
      mul_ipow(xr, xim, gamma_phase(sigma, s));
      const double eps = s < 2 ? 1.0 : -1.0;
      re += eps * xr;
      im += eps * xim;
    }
  };
  double re, im;
  if (j < 6) {
    const int rho = comb_rho(j), sigma = comb_sigma(j);
    double r1, i1, r2, i2;
    m(sigma, 1 + rho, r1, i1);
    m(rho, 1 + sigma, r2, i2);
    re = r1 - r2;
    im = i1 - i2;
  } else {
    m(j - 6, 0, re, im);
  }
  dz[(size_t)i * 2] = re;
  dz[(size_t)i * 2 + 1] = im;
}

// gdz[f][j][nu][a][b] = (gamma_nu gamma_5 dz[f][j][a][.])[b]:
//   out[(s,colour)] = i^{ph_nu(s)} eps_{col_nu(s)} v[(col_nu(s), colour)]
__global__ void form_gdz(const double* __restrict__ dz, double* __restrict__ gdz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (f, j, nu, a, b)
  if (i >= 2 * 10 * 4 * 144) return;
  dz += (size_t)blockIdx.y * kDzLen;
  gdz += (size_t)blockIdx.y * kGdzLen;
  const int b = i % 12, a = (i / 12) % 12, nu = (i / 144) % 4, j = (i / 576) % 10, f = i / 5760;
  const int s = b / 3, colour = b % 3;
  const int sc = gamma_col(nu, s);
  const double* v = dz + ((((size_t)f * 10 + j) * 12 + a) * 12 + 3 * sc + colour) * 2;
  double re = v[0], im = v[1];
  if (sc >= 2) {
    re = -re;
    im = -im;
  }
  mul_ipow(re, im, gamma_phase(nu, s));
  gdz[(size_t)i * 2] = re;
  gdz[(size_t)i * 2 + 1] = im;
}

// One flavour straight from and to the caller's arrays (the seam-2 insertion):
// g_dzu[j][nu][a][b] (j < 6) and g_dzsu[j-6][nu][a][b] from dzu[j][a][.], dzsu[j-6][a][.],
// the same arithmetic as form_gdz.
__global__ void form_gdz_one(const double* __restrict__ dzu, const double* __restrict__ dzsu,
                             double* __restrict__ g_dzu, double* __restrict__ g_dzsu) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (j, nu, a, b)
  if (i >= 10 * 4 * 144) return;
  const int b = i % 12, a = (i / 12) % 12, nu = (i / 144) % 4, j = i / 576;
  const int s = b / 3, colour = b % 3;
  const int sc = gamma_col(nu, s);
  const double* src = j < 6 ? dzu + (size_t)j * 288 : dzsu + (size_t)(j - 6) * 288;
  const double* v = src + ((size_t)a * 12 + 3 * sc + colour) * 2;
  double re = v[0], im = v[1];
  if (sc >= 2) {
    re = -re;
    im = -im;
  }
  mul_ipow(re, im, gamma_phase(nu, s));
  double* dst = j < 6 ? g_dzu + (size_t)j * 4 * 288 : g_dzsu + (size_t)(j - 6) * 4 * 288;
  const size_t o = (((size_t)nu * 12 + a) * 12 + b) * 2;
  dst[o] = re;
  dst[o + 1] = im;
}

// B operand of the pass-2 traces: bmat[f][r*24 + j][col] = bmat_value (pad columns 0)
__global__ void form_bmat(const double* __restrict__ gdz, double* __restrict__ bmat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (f, kidx, col)
  if (i >= kBmatLen) return;
  gdz += (size_t)blockIdx.y * kGdzLen;
  bmat += (size_t)blockIdx.y * kBmatLen;
  const int col = i % kCoefCols, kidx = (i / kCoefCols) % 288, f = i / (288 * kCoefCols);
  bmat[i] = col < 40 ? bmat_value(gdz, f, kidx, col) : 0.0;
}

// ---------------------------------------------------------------- pass 2 --
// (a) the coefficient field of the pair (the QED kernel of every site, both
//     flavours at once), (b) the traces on MFMA weighted by it.
constexpr int kKfThreads = 256;  // 4 waves, one site each at a time

// the pair constants [L0(0,y) | L0(y,0)] into cst[768], one wave
__global__ __launch_bounds__(64) void kfield_consts64(const double* __restrict__ table, hlbl_geom global,
                                                      hlbl_coord yv, double xunit, double* __restrict__ cst) {
  __shared__ TermLds<double> tt;
  __shared__ double w[kKwLen];
  __shared__ double c[2 * qed::kNL];
  stage_terms(c_terms64, &tt);
  __syncthreads();
  const hlbl_coord zero = {0, 0, 0, 0};
  SiteGeom<double> g;
  site_geom(0, global, global, zero, zero, yv, xunit, g);
  kfield_wave_constants<double, double>(table, tt, g, w, c);
  for (int i = threadIdx.x; i < 2 * qed::kNL; i += 64) cst[i] = c[i];
}

// coef[site][640] of every local site of the pair; wave w of workgroup b takes
// sites b*4 + w, then strides by the grid's wave count
__global__ __launch_bounds__(kKfThreads) void kfield_fp64(
    const double* __restrict__ table, hlbl_geom global, hlbl_geom local, hlbl_coord proc,
    hlbl_coord gsx, hlbl_coord yv, double xunit, const double* __restrict__ cst,
    double* __restrict__ coef) {
  __shared__ TermLds<double> tt;
  __shared__ double w[kKfThreads / 64][kKwLen];
  stage_terms(c_terms64, &tt);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: site math on the scalar unit
  const size_t V = geom_volume(local);
  const size_t nw = (size_t)gridDim.x * (kKfThreads / 64);
  for (size_t site = (size_t)blockIdx.x * (kKfThreads / 64) + wave; site < V; site += nw) {
    SiteGeom<double> g;
    site_geom(site, global, local, proc, gsx, yv, xunit, g);
    kfield_wave<double, double, double>(table, tt, g, ChannelRows::none(), cst, w[wave],
                                        coef + site * kCoefField);
  }
}

// The site-only rows of the plain field, FP64, at every relative site (gsx = 0):
// ax[0][V][384] = L0(x, 0), ax[1][V][384] = L0(0, x); one wave per site. They serve
// every pair of the contractor's lattice and xunit (kept on the contractor).
__global__ __launch_bounds__(kKfThreads) void kfield_axes64(const double* __restrict__ table, hlbl_geom global,
                                                           double xunit, double* __restrict__ ax) {
  __shared__ TermLds<double> tt;
  __shared__ double w[kKfThreads / 64][kKwLen];
  stage_terms(c_terms64, &tt);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const size_t V = geom_volume(global);
  const size_t nw = (size_t)gridDim.x * (kKfThreads / 64);
  const hlbl_coord zero = {0, 0, 0, 0};
  for (size_t site = (size_t)blockIdx.x * (kKfThreads / 64) + wave; site < V; site += nw) {
    SiteGeom<double> g;
    site_geom(site, global, global, zero, zero, zero, xunit, g);
    kfield_wave_axis_rows<double, double, double>(table, tt, g, ChannelRows::none(), w[wave],
                                                  ax + site * qed::kNL, ax + (V + site) * qed::kNL);
  }
}

// kfield_fp64 on one process (local = global) with the site-only rows read from ax:
// per site only L0(x, y), L0(y, x) and L0(x, x - y) are evaluated; L0(x, 0) and L0(0, x)
// are the rows of the relative site x - gsx, L0(0, x - y) the (0, x) row of the
// relative site x - gsx - yv.
__global__ __launch_bounds__(kKfThreads) void kfield_fp64_axes(
    const double* __restrict__ table, hlbl_geom global, hlbl_coord gsx, hlbl_coord yv, double xunit,
    const double* __restrict__ cst, const double* __restrict__ ax, double* __restrict__ coef) {
  __shared__ TermLds<double> tt;
  __shared__ double w[kKfThreads / 64][kKwLen64A];
  stage_terms(c_terms64, &tt);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const size_t V = geom_volume(global);
  const size_t nw = (size_t)gridDim.x * (kKfThreads / 64);
  const hlbl_coord zero = {0, 0, 0, 0};
  for (size_t site = (size_t)blockIdx.x * (kKfThreads / 64) + wave; site < V; site += nw) {
    SiteGeom<double> g;
    site_geom(site, global, global, zero, gsx, yv, xunit, g);
    int crd[4], rel[4], dsh[4];
    site_coords(site, global, crd);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int L = geom_dim(global, r);
      rel[r] = mod_pos(crd[r] - coord_dim(gsx, r), L);
      dsh[r] = mod_pos(rel[r] - coord_dim(yv, r), L);
    }
    const size_t rs = (((size_t)rel[0] * global.LX + rel[1]) * global.LY + rel[2]) * global.LZ + rel[3];
    const size_t sd = (((size_t)dsh[0] * global.LX + dsh[1]) * global.LY + dsh[2]) * global.LZ + dsh[3];
    kfield_wave_axes64<double, double>(table, tt, g, cst, ax + rs * qed::kNL, ax + (V + rs) * qed::kNL,
                                       ax + (V + sd) * qed::kNL, w[wave], coef + site * kCoefField);
  }
}

// 8 waves: wave = (flavour, quarter of the K range); each covers the three
// 16-column blocks, so one A fragment feeds three MFMAs
constexpr int kP2Threads = 512;
constexpr int kP2Waves = 8;
constexpr int kKSteps = 18;  // MFMA k-steps per wave (a quarter of 288 / 4)
constexpr int kP2Sites = 2;  // sites staged per iteration

// Per site and flavour f the traces
//   C[(mu lam)][col] = Re sum_{r,i} q^{mu lam}_r[i] X_col[i][r],  q^{mu lam}_r = gamma_lam gamma_5 dxu_mu[r],
//   dxu_mu[r][a] = - sum_c conj((g5 gamma_mu S[1-f][a])[c]) Y[f][r][c]   (cuda_lattice.cu:324-344)
// as a 16 x 48 x 288 real MFMA product, A = q (from dxu in LDS), B = X (held in
// registers), then the kernel sum += sum coef * C.
__global__ __launch_bounds__(kP2Threads) void pass2_fp64(
    const double* __restrict__ S, const double* __restrict__ Y, const double* __restrict__ bmat,
    const double* __restrict__ coef, hlbl_geom local, int flav_mask, size_t sites_per_wg,
    double* __restrict__ part) {
  __shared__ double sSY[kP2Sites][2][2][12][24];   // [site][S/Y][flavour][a][24]
  __shared__ double sDxu[kP2Sites][2][4][12][24];  // [site][flavour][mu][r][24 over a]

  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int f = wave / 4, kq = wave & 3;
  const bool mfma_on = (flav_mask >> f) & 1;
  const size_t V = geom_volume(local);

  // B fragments of this wave: k-step st covers K indices 4*st .. 4*st+3,
  // st in [kq*18, kq*18 + 18), for the column blocks cb = 0, 1, 2
  double breg[3][kKSteps];
#pragma unroll
  for (int cb = 0; cb < 3; ++cb)
#pragma unroll
    for (int i = 0; i < kKSteps; ++i)
      breg[cb][i] = bmat[((size_t)f * 288 + 4 * (kq * kKSteps + i) + (lane >> 4)) * kCoefCols +
                         16 * cb + (lane & 15)];
  // A fragments: row (mu, lam) = lane%16, K index r*24 + j, j = 4*jq + lane/16
  const int arow = lane & 15, amu = arow >> 2, alam = arow & 3;
  int aoff[6];
  double asg[6];
#pragma unroll
  for (int jq = 0; jq < 6; ++jq) {
    const int j = 4 * jq + (lane >> 4);
    const int pj = gamma_perm24(alam, j);
    aoff[jq] = pj;
    asg[jq] = (double)(gamma_sign24(4, pj) * gamma_sign24(alam, j));
  }
  // C rows of this lane: (mu, lam) = 4*i + lane/16 for its element i; columns 16*cb + lane%16

  double dacc = 0.0;
  const size_t begin = (size_t)blockIdx.x * sites_per_wg;
  const size_t end = begin + sites_per_wg < V ? begin + sites_per_wg : V;
  for (size_t z0 = begin; z0 < end; z0 += kP2Sites) {
    const int ns = (end - z0) < (size_t)kP2Sites ? (int)(end - z0) : kP2Sites;
    // P0: stage the spinors of both blocks and flavours
    for (int i = tid; i < kP2Sites * 1152; i += kP2Threads) {
      const int zi = i / 1152, rr = i % 1152;
      const int bl = rr / 576, rem = rr % 576, ff = rem / 288, a = (rem / 24) % 12, c = rem % 24;
      double v = 0.0;
      if (zi < ns) v = (bl ? Y : S)[(((size_t)(ff * 12 + a)) * V + z0 + zi) * 24 + c];
      sSY[zi][bl][ff][a][c] = v;
    }
    __syncthreads();
    // P1: dxu of both flavours, task (site, f, r, a)
    for (int t = tid; t < kP2Sites * 288; t += kP2Threads) {
      const int zi = t / 288, f2 = (t / 144) % 2, r = (t / 12) % 12, a = t % 12;
      dxu_site<double, double>(sSY[zi][0][1 - f2][a], sSY[zi][1][f2][r], &sDxu[zi][f2][0][r][2 * a], 12 * 24);
    }
    __syncthreads();
    // P2: traces on MFMA, weighted by the coefficient field
    if (mfma_on) {
      for (int zi = 0; zi < ns; ++zi) {
        const double* dx = &sDxu[zi][f][amu][0][0];
        d4v c0 = {0.0, 0.0, 0.0, 0.0}, c1 = c0, c2 = c0;
#pragma unroll
        for (int i = 0; i < kKSteps; ++i) {
          const int st = kq * kKSteps + i;  // kq*18 is a multiple of 6: jq = i % 6
          const int r = st / 6, jq = i % 6;
          const double a = asg[jq] * dx[r * 24 + aoff[jq]];
          c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, breg[0][i], c0, 0, 0, 0);
          c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, breg[1][i], c1, 0, 0, 0);
          c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, breg[2][i], c2, 0, 0, 0);
        }
        // FP64 16x16x4 result: element i of lane l is row 4*i + l/16, column l%16
        const double* cs = coef + (z0 + zi) * kCoefField + (lane >> 4) * kCoefUsed + (lane & 15);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          dacc += c0[i] * cs[4 * i * kCoefUsed] + c1[i] * cs[4 * i * kCoefUsed + 16];
          if ((lane & 15) < kCoefUsed - 32) dacc += c2[i] * cs[4 * i * kCoefUsed + 32];
        }
      }
    }
    // the next P0 writes sSY only; P1 rewrites sDxu after the barrier
  }
  // wave sum, one partial per wave: part[wg][wave], wave = f*4 + kq
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) dacc += __shfl_xor(dacc, off);
  if (lane == 0) part[(size_t)blockIdx.x * kP2Waves + wave] = mfma_on ? dacc : 0.0;
}

// kernel_sum[f] (+)= sum over workgroups and column blocks, fixed order
__global__ void reduce_pass2(const double* __restrict__ part, int nwg, int flav_mask,
                             double* __restrict__ out, int accumulate, int single_flavour) {
  const int f = threadIdx.x;
  if (f >= 2 || !((flav_mask >> f) & 1)) return;
  double s = 0.0;
  for (int w = 0; w < nwg; ++w)
    for (int j = 0; j < kP2Waves / 2; ++j) s += part[(size_t)w * kP2Waves + f * (kP2Waves / 2) + j];
  const int o = single_flavour ? 0 : f;
  out[o] = accumulate ? out[o] + s : s;
}

__global__ void add_kernel(double* __restrict__ dst, const double* __restrict__ src, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] += src[i];
}

// ------------------------------------------------------------- launchers --

namespace {
size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }
}  // namespace

void launch_pass1(hlbl_contractor* c, const double* S, const double* Y, hlbl_coord proc,
                  hlbl_coord gsx, hlbl_geom global, hlbl_geom local, int flav_mask,
                  hipStream_t st) {
  c->part1.ensure(pass1_part_len(local, 1));
  c->red1.ensure(pass1_red_len(local, 1));
  c->xi.ensure(kXiLen);
  WinPair single{};
  single.S = S;
  single.Y = Y;
  single.gsx = gsx;
  launch_pass1_planes<double>(nullptr, single, 1, global, local, proc, flav_mask, c->part1.p,
                              c->red1.p, c->xi.p, st);
  c->dz.ensure(2 * 10 * 144 * 2);
  hipLaunchKernelGGL(form_dz, dim3(div_up(2 * 10 * 144, 256), 1), dim3(256), 0, st, c->xi.p, c->dz.p);
  check_hip(hipGetLastError(), "pass1 reduction launch");
}

void launch_add(double* dst, const double* src, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(add_kernel, dim3((unsigned)div_up(n, 256)), dim3(256), 0, st, dst, src, n);
  check_hip(hipGetLastError(), "add_kernel launch");
}

void launch_forms_batched(const double* xi, double* dz, double* gdz, int n_pairs, hipStream_t st) {
  hipLaunchKernelGGL(form_dz, dim3(div_up(2 * 10 * 144, 256), n_pairs), dim3(256), 0, st, xi, dz);
  hipLaunchKernelGGL(form_gdz, dim3(div_up(2 * 10 * 4 * 144, 256), n_pairs), dim3(256), 0, st, dz, gdz);
  check_hip(hipGetLastError(), "form kernels launch");
}

void launch_gamma(const double* dz, double* gdz, hipStream_t st) {
  hipLaunchKernelGGL(form_gdz, dim3(div_up(2 * 10 * 4 * 144, 256), 1), dim3(256), 0, st, dz, gdz);
  check_hip(hipGetLastError(), "form_gdz launch");
}

void launch_gamma_one(const double* dzu, const double* dzsu, double* g_dzu, double* g_dzsu, hipStream_t st) {
  hipLaunchKernelGGL(form_gdz_one, dim3(div_up(10 * 4 * 144, 256)), dim3(256), 0, st, dzu, dzsu, g_dzu, g_dzsu);
  check_hip(hipGetLastError(), "form_gdz_one launch");
}

void launch_bmat(const double* gdz, double* bmat, hipStream_t st) {
  hipLaunchKernelGGL(form_bmat, dim3(div_up(kBmatLen, 256), 1), dim3(256), 0, st, gdz, bmat);
  check_hip(hipGetLastError(), "form_bmat launch");
}

void launch_kfield(hlbl_contractor* c, hlbl_coord proc, hlbl_coord gsx, hlbl_coord yv,
                   double xunit, hlbl_geom global, hlbl_geom local, hipStream_t st) {
  const size_t V = geom_volume(local);
  c->coef.ensure(V * kCoefField);
  c->kcst.ensure(2 * qed::kNL);
  hipLaunchKernelGGL(kfield_consts64, dim3(1), dim3(64), 0, st, c->qed->table64.p, global, yv, xunit,
                     c->kcst.p);
  check_hip(hipGetLastError(), "kfield_consts64 launch");
  const int nwg = (int)std::min<size_t>(div_up(V, kKfThreads / 64), (size_t)c->tune.fp64_field_wg_per_cu * c->n_cu);
  const bool whole = global.T == local.T && global.LX == local.LX && global.LY == local.LY && global.LZ == local.LZ;
  if (!whole) {
    hipLaunchKernelGGL(kfield_fp64, dim3(nwg), dim3(kKfThreads), 0, st, c->qed->table64.p, global,
                       local, proc, gsx, yv, xunit, c->kcst.p, c->coef.p);
    check_hip(hipGetLastError(), "kfield_fp64 launch");
    return;
  }
  // one process holds the lattice: the site-only rows once per xunit, then three
  // evaluations per site
  if (c->ax64_xunit != xunit) {
    c->ax64.ensure((size_t)2 * V * qed::kNL);
    hipLaunchKernelGGL(kfield_axes64, dim3(nwg), dim3(kKfThreads), 0, st, c->qed->table64.p, global, xunit,
                       c->ax64.p);
    check_hip(hipGetLastError(), "kfield_axes64 launch");
    c->ax64_xunit = xunit;
  }
  hipLaunchKernelGGL(kfield_fp64_axes, dim3(nwg), dim3(kKfThreads), 0, st, c->qed->table64.p, global, gsx, yv,
                     xunit, c->kcst.p, c->ax64.p, c->coef.p);
  check_hip(hipGetLastError(), "kfield_fp64_axes launch");
}

void launch_pass2(hlbl_contractor* c, const double* S, const double* Y, const double* bmat,
                  hlbl_coord proc, hlbl_coord gsx, hlbl_coord yv, double xunit, hlbl_geom global,
                  hlbl_geom local, int flav_mask, double* out, int accumulate, int single_flavour,
                  hipStream_t st) {
  launch_kfield(c, proc, gsx, yv, xunit, global, local, st);
  const size_t V = geom_volume(local);
  const size_t spw = div_up(div_up(V, (size_t)c->nwg2), kP2Sites) * kP2Sites;
  const int nwg = (int)div_up(V, spw);
  c->part2.ensure((size_t)c->nwg2 * kP2Waves);
  hipLaunchKernelGGL(pass2_fp64, dim3(nwg), dim3(kP2Threads), 0, st, S, Y, bmat, c->coef.p, local,
                     flav_mask, spw, c->part2.p);
  check_hip(hipGetLastError(), "pass2_fp64 launch");
  hipLaunchKernelGGL(reduce_pass2, dim3(1), dim3(64), 0, st, c->part2.p, nwg, flav_mask, out,
                     accumulate, single_flavour);
  check_hip(hipGetLastError(), "reduce_pass2 launch");
}

}  // namespace hlbl
