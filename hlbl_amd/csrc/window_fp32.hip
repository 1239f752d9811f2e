// Window-batched mixed-precision contraction: all pairs of a window and both
// flavours per call. Pass 1 (pass1.hip) is one launch over (plane, flavour,
// pair), pass 2 one persistent launch over the pairs. Propagator products run in
// FP32 (pass-2 traces on FP32 MFMA), the lattice sums and the per-pair sums are
// FP64, combined with Kahan compensation. The coefficient field of the QED kernel depends only on
// (x - gsx, yv), so it is built once per distinct yv of the window (at gsx = 0)
// and read at the shifted site; sites near the origin (|x| or |x - y| below
// kNearRadius) evaluate the kernel in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "context.hpp"
#include "contract_kernels.hpp"
#include "kpart.hpp"
#include "pass1.hpp"
#include "pass_common.hpp"
#include "qed.hpp"
#include "window.hpp"

namespace hlbl {

__constant__ qed::TermTable c_terms32;

void upload_terms_fp32(const qed::TermTable& t) {
  check_hip(hipMemcpyToSymbol(HIP_SYMBOL(c_terms32), &t, sizeof(t)), "upload terms");
}

typedef float f4v __attribute__((ext_vector_type(4)));

namespace {
size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }
}  // namespace

// ------------------------------------------------------------ coefficients --
// The field F_yv[rel site][16][40] (FP32) for each distinct yv, gsx = 0. A site
// with |x| or |x - y| below kNearRadius is evaluated in FP64 with the FP64
// table; the others in FP32 from the FP32 table, or from the surrogate's
// channels (chbuf, see below) when given.
constexpr int kKmThreads = 256;

__global__ __launch_bounds__(kKmThreads) void kfield_mixed(
    const float* __restrict__ table32, const double* __restrict__ table64,
    const hlbl_coord* __restrict__ yvs, int n_unique, hlbl_geom global, double xunit,
    size_t sites_per_wg, const float* __restrict__ chbuf, float* __restrict__ coef) {
  __shared__ double lds64[kKfieldLds];
  __shared__ float lds32[kKfieldLds];
  const size_t V = geom_volume(global);
  const size_t begin = (size_t)blockIdx.x * sites_per_wg;
  const size_t end = begin + sites_per_wg < V ? begin + sites_per_wg : V;
  if (begin >= end) return;
  const hlbl_coord zero = {0, 0, 0, 0};
  const double near2 = kNearRadius * kNearRadius;
  for (int u = 0; u < n_unique; ++u) {
    const hlbl_coord yv = yvs[u];
    SiteGeom<double> g64;
    SiteGeom<float> g32;
    site_geom(0, global, global, zero, zero, yv, xunit, g64);
    site_geom(0, global, global, zero, zero, yv, xunit, g32);
    kfield_constants(table64, c_terms32, g64, lds64);
    // FP32 copies of the pair constants
    for (int t = threadIdx.x; t < 2 * qed::kNL; t += blockDim.x)
      lds32[8 * qed::kNCh + 8 * qed::kNT + 6 * qed::kNL + t] = (float)lds64[8 * qed::kNCh + 8 * qed::kNT + 6 * qed::kNL + t];
    __syncthreads();
    float* cu = coef + (size_t)u * V * kCoefField;
    for (size_t site = begin; site < end; ++site) {
      site_geom(site, global, global, zero, zero, yv, xunit, g64);
      const bool near = norm2(g64.xm) < near2 || norm2(g64.dm) < near2;
      if (near) {
        kfield_site<double, double, float>(table64, c_terms32, g64, ChannelRows::none(), lds64,
                                           cu + site * kCoefField);
      } else {
        site_geom(site, global, global, zero, zero, yv, xunit, g32);
        if (chbuf) {
          // surrogate channel rows: blocks 0 (x,0) and 1 (0,x) shared by every yv,
          // block 2+3u+j the yv-dependent (x,y), (y,x), (x,d); (0,d) is (0,x) at
          // the site d = x - y
          int crd[4], dsh[4];
          site_coords(site, global, crd);
          for (int r = 0; r < 4; ++r) dsh[r] = mod_pos(crd[r] - coord_dim(yv, r), geom_dim(global, r));
          const size_t sd = (((size_t)dsh[0] * global.LX + dsh[1]) * global.LY + dsh[2]) * global.LZ + dsh[3];
          ChannelRows rows;
          rows.given = true;
          rows.p[0] = chbuf + ((size_t)(2 + 3 * u + 0) * V + site) * qed::kNCh;
          rows.p[1] = chbuf + ((size_t)0 * V + site) * qed::kNCh;
          rows.p[2] = chbuf + ((size_t)(2 + 3 * u + 1) * V + site) * qed::kNCh;
          rows.p[3] = chbuf + ((size_t)1 * V + site) * qed::kNCh;
          rows.p[4] = chbuf + ((size_t)(2 + 3 * u + 2) * V + site) * qed::kNCh;
          rows.p[5] = chbuf + ((size_t)1 * V + sd) * qed::kNCh;
          kfield_site<float, float, float>(table32, c_terms32, g32, rows, lds32, cu + site * kCoefField);
        } else {
          kfield_site<float, float, float>(table32, c_terms32, g32, ChannelRows::none(), lds32,
                                           cu + site * kCoefField);
        }
      }
      __syncthreads();
    }
    __syncthreads();
  }
}

// The surrogate's judge: at sample sites, the FP64 plain coefficients against
// the field in use; err[u][k] = max |c - c_plain| / max |c_plain| for sample k.
__global__ __launch_bounds__(kKmThreads) void judge_coef(
    const double* __restrict__ table64, const hlbl_coord* __restrict__ yvs,
    hlbl_geom global, double xunit, const size_t* __restrict__ samples, int n_samples,
    const float* __restrict__ coef, double* __restrict__ err) {
  __shared__ double lds64[kKfieldLds];
  __shared__ double ref[kCoefField];
  __shared__ double red[2][kKmThreads];
  const int u = blockIdx.y, k = blockIdx.x;
  const size_t V = geom_volume(global);
  const hlbl_coord zero = {0, 0, 0, 0};
  SiteGeom<double> g;
  site_geom(0, global, global, zero, zero, yvs[u], xunit, g);
  kfield_constants(table64, c_terms32, g, lds64);
  const size_t site = samples[k];
  site_geom(site, global, global, zero, zero, yvs[u], xunit, g);
  kfield_site<double, double, double>(table64, c_terms32, g, ChannelRows::none(), lds64, ref);
  __syncthreads();
  const float* c = coef + ((size_t)u * V + site) * kCoefField;
  double dmax = 0.0, rmax = 0.0;
  for (int t = threadIdx.x; t < kCoefField; t += blockDim.x) {
    dmax = fmax(dmax, fabs((double)c[t] - ref[t]));
    rmax = fmax(rmax, fabs(ref[t]));
  }
  red[0][threadIdx.x] = dmax;
  red[1][threadIdx.x] = rmax;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + o]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) err[(size_t)u * n_samples + k] = red[1][0] > 0 ? red[0][0] / red[1][0] : red[0][0];
}

// ---------------------------------------------------------------- pass 2 --
// As pass2_fp64 on FP32 MFMA: 8 waves = (flavour, quarter of K), three
// 16-column blocks each. Per lane and site the 12 weighted products are summed
// in FP32 and added to an FP64 Kahan sum. Persistent over the pairs.
constexpr int kM2Threads = 512;
constexpr int kM2Waves = 8;
constexpr int kM2KSteps = 18;
constexpr int kM2Sites = 4;

__global__ __launch_bounds__(kM2Threads) void pass2_mixed(
    const WinPair* __restrict__ pairs, int n_pairs, const float* __restrict__ bmat,
    const float* __restrict__ coef, hlbl_geom global, size_t sites_per_wg, int nwg,
    double* __restrict__ part) {
  __shared__ float sSY[kM2Sites][2][2][12][24];
  __shared__ float sDxu[kM2Sites][2][4][12][24];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int f = wave / 4, kq = wave & 3;
  const size_t V = geom_volume(global);
  const int arow = lane & 15, amu = arow >> 2, alam = arow & 3;
  int aoff[6];
  float asg[6];
#pragma unroll
  for (int jq = 0; jq < 6; ++jq) {
    const int j = 4 * jq + (lane >> 4);
    const int pj = gamma_perm24(alam, j);
    aoff[jq] = pj;
    asg[jq] = (float)(gamma_sign24(4, pj) * gamma_sign24(alam, j));
  }
  const size_t begin = (size_t)blockIdx.x * sites_per_wg;
  const size_t end = begin + sites_per_wg < V ? begin + sites_per_wg : V;

  for (int ip = 0; ip < n_pairs; ++ip) {
    const WinPair wp = pairs[ip];
    const float* bm = bmat + (size_t)ip * 2 * 288 * kCoefCols;
    float breg[3][kM2KSteps];
#pragma unroll
    for (int cb = 0; cb < 3; ++cb)
#pragma unroll
      for (int i = 0; i < kM2KSteps; ++i)
        breg[cb][i] = bm[((size_t)f * 288 + 4 * (kq * kM2KSteps + i) + (lane >> 4)) * kCoefCols +
                         16 * cb + (lane & 15)];
    const float* cu = coef + (size_t)wp.yv_slot * V * kCoefField;
    Kahan acc;
    for (size_t z0 = begin; z0 < end; z0 += kM2Sites) {
      const int ns = (end - z0) < (size_t)kM2Sites ? (int)(end - z0) : kM2Sites;
      __syncthreads();
      for (int i = tid; i < kM2Sites * 1152; i += kM2Threads) {
        const int zi = i / 1152, rr = i % 1152;
        const int bl = rr / 576, rem = rr % 576, ff = rem / 288, a = (rem / 24) % 12, c = rem % 24;
        float v = 0.f;
        if (zi < ns) v = (float)(bl ? wp.Y : wp.S)[(((size_t)(ff * 12 + a)) * V + z0 + zi) * 24 + c];
        sSY[zi][bl][ff][a][c] = v;
      }
      __syncthreads();
      for (int t = tid; t < kM2Sites * 288; t += kM2Threads) {
        const int zi = t / 288, f2 = (t / 144) % 2, r = (t / 12) % 12, a = t % 12;
        dxu_site<float, float>(sSY[zi][0][1 - f2][a], sSY[zi][1][f2][r], &sDxu[zi][f2][0][r][2 * a], 12 * 24);
      }
      __syncthreads();
      for (int zi = 0; zi < ns; ++zi) {
        const float* dx = &sDxu[zi][f][amu][0][0];
        f4v c0 = {0.f, 0.f, 0.f, 0.f}, c1 = c0, c2 = c0;
#pragma unroll
        for (int i = 0; i < kM2KSteps; ++i) {
          const int st = kq * kM2KSteps + i;
          const int r = st / 6, jq = i % 6;
          const float a = asg[jq] * dx[r * 24 + aoff[jq]];
          c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, breg[0][i], c0, 0, 0, 0);
          c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, breg[1][i], c1, 0, 0, 0);
          c2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, breg[2][i], c2, 0, 0, 0);
        }
        // the coefficient field at the site relative to gsx
        int crd[4], rel[4];
        site_coords(z0 + zi, global, crd);
        for (int r = 0; r < 4; ++r) rel[r] = mod_pos(crd[r] - coord_dim(wp.gsx, r), geom_dim(global, r));
        const size_t rs = (((size_t)rel[0] * global.LX + rel[1]) * global.LY + rel[2]) * global.LZ + rel[3];
        const float* cs = cu + rs * kCoefField + (4 * (lane >> 4)) * kCoefUsed + (lane & 15);
        float v = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          v += c0[i] * cs[i * kCoefUsed] + c1[i] * cs[i * kCoefUsed + 16];
          if ((lane & 15) < kCoefUsed - 32) v += c2[i] * cs[i * kCoefUsed + 32];
        }
        acc.add((double)v);
      }
    }
    // wave sum (pairwise, FP64), one partial per wave
    double s = acc.s - acc.c;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) part[((size_t)ip * nwg + blockIdx.x) * kM2Waves + wave] = s;
  }
}

// out[pair][f] = Kahan sum over workgroups and waves of the pair's partials
__global__ void reduce_window2(const double* __restrict__ part, int nwg, int n_pairs,
                               double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (pair, f)
  if (i >= 2 * n_pairs) return;
  const int ip = i / 2, f = i % 2;
  Kahan k;
  for (int w = 0; w < nwg; ++w)
    for (int j = 0; j < kM2Waves / 2; ++j) k.add(part[((size_t)ip * nwg + w) * kM2Waves + f * (kM2Waves / 2) + j]);
  out[i] = k.s;
}

// ------------------------------------------------------------- launchers --

void window_record(hlbl_contractor* c, WindowPlan& w, hipStream_t st) {
  window_stage(c, w, 0, st);
  window_stage(c, w, 1, st);
  window_stage(c, w, 2, st);
}

void window_stage(hlbl_contractor* c, WindowPlan& w, int stage, hipStream_t st) {
  const hlbl_geom g = c->global;
  const size_t V = c->volume;
  const int np = w.n_pairs, nu = w.n_unique;
  // pass 1
  if (stage == 0) {
    launch_pass1_planes<float>(w.d_pairs.p, WinPair{}, np, g, g, c->proc, 3, w.part1.p, w.red1.p,
                               w.xi.p, st);
    launch_forms_batched(w.xi.p, w.dz.p, w.gdz.p, nullptr, w.bmat.p, np, st);
  }
  // coefficient fields, one per distinct yv
  if (stage == 1) {
    const int nwg = 4 * c->n_cu;
    const size_t spw = div_up(V, (size_t)nwg);
    const int n = (int)div_up(V, spw);
    hipLaunchKernelGGL(kfield_mixed, dim3(n), dim3(kKmThreads), 0, st, c->qed->table32.p,
                       c->qed->table64.p, w.d_yvs.p, nu, g, w.xunit, spw,
                       w.form == 2 ? w.chbuf.p : nullptr, w.coef.p);
    check_hip(hipGetLastError(), "kfield_mixed launch");
    if (w.form == 2) {
      hipLaunchKernelGGL(judge_coef, dim3(w.n_samples, nu), dim3(kKmThreads), 0, st,
                         c->qed->table64.p, w.d_yvs.p, g, w.xunit, w.d_samples.p, w.n_samples,
                         w.coef.p, w.judge.p);
      check_hip(hipGetLastError(), "judge_coef launch");
    }
  }
  // pass 2
  if (stage == 2) {
    const int nwg = 2 * c->n_cu;
    const size_t spw = div_up(div_up(V, (size_t)nwg), kM2Sites) * kM2Sites;
    const int n = (int)div_up(V, spw);
    hipLaunchKernelGGL(pass2_mixed, dim3(n), dim3(kM2Threads), 0, st, w.d_pairs.p, np, w.bmat.p,
                       w.coef.p, g, spw, n, w.part2.p);
    check_hip(hipGetLastError(), "pass2_mixed launch");
    hipLaunchKernelGGL(reduce_window2, dim3((unsigned)div_up(2 * np, 64)), dim3(64), 0, st,
                       w.part2.p, n, np, w.result.p);
    check_hip(hipGetLastError(), "reduce_window2 launch");
  }
}

size_t window_part2_len(const hlbl_contractor* c, int n_pairs) {
  return (size_t)n_pairs * 2 * c->n_cu * kM2Waves;
}

}  // namespace hlbl
