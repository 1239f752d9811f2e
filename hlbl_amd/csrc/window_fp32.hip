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
#include <cmath>
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
// The field F_yv[rel site][16][40] (FP32) for each distinct yv, gsx = 0, one wave
// per site. Every site is first computed in FP32 (FP32 table, or the surrogate's
// channels from chbuf); the sites near the origin (|x| or |x - y| below
// kNearRadius, a list built on the host) are then recomputed in FP64 with the
// FP64 table, where the Lambda subtraction cancels.
constexpr int kKmThreads = 256;

// pair constants of each distinct yv: cst64[u][768] and an FP32 copy; one wave per u
__global__ __launch_bounds__(64) void kfield_consts_win(const double* __restrict__ table64,
                                                        const hlbl_coord* __restrict__ yvs,
                                                        hlbl_geom global, double xunit,
                                                        double* __restrict__ cst64,
                                                        float* __restrict__ cst32) {
  __shared__ TermLds tt;
  __shared__ double w[kKwLen];
  __shared__ double c[2 * qed::kNL];
  stage_terms(c_terms32, &tt);
  __syncthreads();
  const int u = blockIdx.x;
  const hlbl_coord zero = {0, 0, 0, 0};
  SiteGeom<double> g;
  site_geom(0, global, global, zero, zero, yvs[u], xunit, g);
  kfield_wave_constants<double, double>(table64, tt, g, w, c);
  for (int i = threadIdx.x; i < 2 * qed::kNL; i += 64) {
    cst64[(size_t)u * 2 * qed::kNL + i] = c[i];
    cst32[(size_t)u * 2 * qed::kNL + i] = (float)c[i];
  }
}

__global__ __launch_bounds__(kKmThreads) void kfield_mixed(
    const float* __restrict__ table32, const hlbl_coord* __restrict__ yvs, int n_unique,
    hlbl_geom global, double xunit, const float* __restrict__ cst32,
    const float* __restrict__ chbuf, float* __restrict__ coef) {
  __shared__ TermLds tt;
  __shared__ float w[kKmThreads / 64][kKwLen];
  stage_terms(c_terms32, &tt);
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const size_t V = geom_volume(global);
  const size_t total = (size_t)n_unique * V;
  const size_t nw = (size_t)gridDim.x * (kKmThreads / 64);
  const hlbl_coord zero = {0, 0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * (kKmThreads / 64) + wave; i < total; i += nw) {
    const int u = (int)(i / V);
    const size_t site = i - (size_t)u * V;
    const hlbl_coord yv = yvs[u];
    SiteGeom<float> g;
    site_geom(site, global, global, zero, zero, yv, xunit, g);
    ChannelRows rows = ChannelRows::none();
    if (chbuf) {
      // surrogate channel rows: blocks 0 (x,0) and 1 (0,x) shared by every yv,
      // block 2+3u+j the yv-dependent (x,y), (y,x), (x,d); (0,d) is (0,x) at the
      // site d = x - y
      int crd[4], dsh[4];
      site_coords(site, global, crd);
#pragma unroll
      for (int r = 0; r < 4; ++r) dsh[r] = mod_pos(crd[r] - coord_dim(yv, r), geom_dim(global, r));
      const size_t sd = (((size_t)dsh[0] * global.LX + dsh[1]) * global.LY + dsh[2]) * global.LZ + dsh[3];
      rows.given = true;
      rows.p[0] = chbuf + ((size_t)(2 + 3 * u + 0) * V + site) * qed::kNCh;
      rows.p[1] = chbuf + ((size_t)0 * V + site) * qed::kNCh;
      rows.p[2] = chbuf + ((size_t)(2 + 3 * u + 1) * V + site) * qed::kNCh;
      rows.p[3] = chbuf + ((size_t)1 * V + site) * qed::kNCh;
      rows.p[4] = chbuf + ((size_t)(2 + 3 * u + 2) * V + site) * qed::kNCh;
      rows.p[5] = chbuf + ((size_t)1 * V + sd) * qed::kNCh;
    }
    kfield_wave<float, float, float>(table32, tt, g, rows, cst32 + (size_t)u * 2 * qed::kNL, w[wave],
                                     coef + i * kCoefField);
  }
}

// FP64 recomputation of the near sites. Candidates: for each distinct yv u, the
// sites within box[r] lattice units of the origin (which = 0) and of y (which =
// 1) in every direction; a candidate with |x| or |x - y| below kNearRadius is
// recomputed (a site in both boxes twice, with the same result).
__global__ __launch_bounds__(kKmThreads) void kfield_near(
    const double* __restrict__ table64, const hlbl_coord* __restrict__ yvs, int n_unique,
    hlbl_geom global, double xunit, const double* __restrict__ cst64, hlbl_coord box,
    float* __restrict__ coef) {
  __shared__ TermLds tt;
  __shared__ double w[kKmThreads / 64][kKwLen];
  stage_terms(c_terms32, &tt);
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const size_t V = geom_volume(global);
  const int side[4] = {2 * box.t + 1, 2 * box.x + 1, 2 * box.y + 1, 2 * box.z + 1};
  const size_t per_box = (size_t)side[0] * side[1] * side[2] * side[3];
  const size_t total = (size_t)n_unique * 2 * per_box;
  const double near2 = kNearRadius * kNearRadius;
  const hlbl_coord zero = {0, 0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * (kKmThreads / 64) + wave; i < total;
       i += (size_t)gridDim.x * (kKmThreads / 64)) {
    const int u = (int)(i / (2 * per_box));
    const int which = (int)((i / per_box) & 1);
    size_t rem = i % per_box;
    const hlbl_coord yv = yvs[u];
    int crd[4];
    for (int r = 3; r >= 0; --r) {
      const int off = (int)(rem % side[r]) - coord_dim(box, r);
      rem /= side[r];
      crd[r] = mod_pos((which ? coord_dim(yv, r) : 0) + off, geom_dim(global, r));
    }
    const size_t site = (((size_t)crd[0] * global.LX + crd[1]) * global.LY + crd[2]) * global.LZ + crd[3];
    SiteGeom<double> g;
    site_geom(site, global, global, zero, zero, yv, xunit, g);
    if (!(norm2(g.xm) < near2 || norm2(g.dm) < near2)) continue;  // wave-uniform
    kfield_wave<double, double, float>(table64, tt, g, ChannelRows::none(),
                                       cst64 + (size_t)u * 2 * qed::kNL, w[wave],
                                       coef + ((size_t)u * V + site) * kCoefField);
  }
}

// The surrogate's judge: at sample sites, the FP64 plain coefficients against
// the field in use; err[u][k] = max |c - c_plain| / max |c_plain| for sample k.
// One wave per (sample, yv).
__global__ __launch_bounds__(64) void judge_coef(
    const double* __restrict__ table64, const hlbl_coord* __restrict__ yvs, hlbl_geom global,
    double xunit, const double* __restrict__ cst64, const size_t* __restrict__ samples,
    int n_samples, const float* __restrict__ coef, double* __restrict__ err) {
  __shared__ TermLds tt;
  __shared__ double w[kKwLen];
  __shared__ double ref[kCoefField];
  stage_terms(c_terms32, &tt);
  __syncthreads();
  const int u = blockIdx.y, k = blockIdx.x, lane = threadIdx.x;
  const size_t V = geom_volume(global);
  const hlbl_coord zero = {0, 0, 0, 0};
  const size_t site = samples[k];
  SiteGeom<double> g;
  site_geom(site, global, global, zero, zero, yvs[u], xunit, g);
  kfield_wave<double, double, double>(table64, tt, g, ChannelRows::none(),
                                      cst64 + (size_t)u * 2 * qed::kNL, w, ref);
  const float* c = coef + ((size_t)u * V + site) * kCoefField;
  double dmax = 0.0, rmax = 0.0;
  for (int t = lane; t < kCoefField; t += 64) {  // the entries this lane wrote
    dmax = fmax(dmax, fabs((double)c[t] - ref[t]));
    rmax = fmax(rmax, fabs(ref[t]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, off));
    rmax = fmax(rmax, __shfl_xor(rmax, off));
  }
  if (lane == 0) err[(size_t)u * n_samples + k] = rmax > 0 ? dmax / rmax : dmax;
}

// ---------------------------------------------------------------- pass 2 --
// Per site and flavour f the traces
//   C[(mu lam)][col] = Re sum_{r,i} q^{mu lam}_r[i] X_col[i][r],  q^{mu lam}_r = gamma_lam gamma_5 dxu_mu[r],
// (cuda_lattice.cu:324-344, 403-476) as a 16 x 48 x 288 product on FP32 MFMA,
// A = q (from dxu in LDS, gamma_lam gamma_5 as a per-lane signed permutation), B =
// the pair's X (registers), then the kernel sum += sum coef * C.
// Twelve waves with two roles, over groups of 4 sites and one barrier per group:
//  - waves 0..7 (flavour, quarter of K) run the MFMAs of group g; the sum over K
//    quarters is linear in the coefficients, so every wave dots its own partial C
//    with them (read from LDS);
//  - waves 8..11 (producers) store group g+2's spinors (fetched into registers
//    during the previous step) as FP32, stage group g+1's coefficients into LDS
//    (direct loads), build dxu of group g+1 from its FP32 spinors, and fetch
//    group g+3's spinors.
constexpr int kW2Threads = 768;
constexpr int kW2Waves = 8;               // MFMA waves (partials per workgroup)
constexpr int kW2Prod = 4;                // producer waves, one per site of a group
constexpr int kW2Sites = 4;
constexpr int kW2KSteps = 18;             // k-steps (of 4) per MFMA wave: a quarter of 288
constexpr int kW2RowLen = kW2Sites * 24;  // values of one staged spinor row (4 sites)
constexpr int kW2RowPad = kW2RowLen + 4;  // FP32 row stride: rows start 36 banks apart
constexpr int kW2BufLen = 48 * kW2RowPad; // floats per staging buffer: (S/Y, f, a) rows
constexpr int kW2DxuRow = 304;            // floats per (site, f, mu) row of dxu: 288 + pad, so
                                          // that an A fragment's 16 rows x 4 k hit distinct banks
constexpr int kW2DxuLen = kW2Sites * 2 * 4 * kW2DxuRow;  // floats per dxu buffer
constexpr int kW2CoefRow = 44;            // staged coefficient row (40 used): conflict-free reads
constexpr int kW2CoefSite = 16 * kW2CoefRow;
constexpr int kW2Fetch = 48 * kW2RowLen / 2 / (64 * kW2Prod);  // double2 per producer lane: 9
static_assert(kW2Fetch == 9, "barrier_vm_fetch leaves exactly 9 loads in flight");

// barrier that waits for this wave's LDS operations and for its global loads
// (direct loads into LDS included)
__device__ inline void barrier_all() { asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n s_barrier" ::: "memory"); }
// the same, leaving the producer's last kW2Fetch (= 9) global loads in flight
__device__ inline void barrier_vm_fetch() { asm volatile("s_waitcnt vmcnt(9) lgkmcnt(0)\n s_barrier" ::: "memory"); }

// Producer lane ptid fetches its 9 double2 of the spinors of sites [z0, z0 + 4)
// of pair wp: element e2 = ptid + 256 k is double pair e2 % 48 of row e2 / 48,
// rows (S/Y, f, a). Sites at or beyond V are read from site V-1 (masked later).
__device__ inline void w2_fetch(int ptid, const WinPair& wp, size_t V, size_t z0, double2 (&r)[kW2Fetch]) {
#pragma unroll
  for (int k = 0; k < kW2Fetch; ++k) {
    const int e2 = ptid + 64 * kW2Prod * k;
    const int row = e2 / (kW2RowLen / 2), within = 2 * (e2 % (kW2RowLen / 2));
    const int bl = row / 24, ff = (row / 12) & 1, a = row % 12;
    size_t z = z0 + within / 24;
    z = z < V ? z : V - 1;
    const double* src = (bl ? wp.Y : wp.S) + (((size_t)(ff * 12 + a)) * V + z) * 24 + within % 24;
    r[k] = *reinterpret_cast<const double2*>(src);
  }
}

// ... and stores them as FP32 into buf [48][kW2RowPad]
__device__ inline void w2_store(int ptid, const double2 (&r)[kW2Fetch], float* buf) {
#pragma unroll
  for (int k = 0; k < kW2Fetch; ++k) {
    const int e2 = ptid + 64 * kW2Prod * k;
    const int row = e2 / (kW2RowLen / 2), within = 2 * (e2 % (kW2RowLen / 2));
    float2 v;
    v.x = (float)r[k].x;
    v.y = (float)r[k].y;
    *reinterpret_cast<float2*>(buf + row * kW2RowPad + within) = v;
  }
}

// Producer p stages the 16 x 40 coefficients of site z0 + p (relative to gsx)
// into dst + p * kW2CoefSite, rows kW2CoefRow apart, with direct loads: one load of
// 10 lanes x 16 bytes per row.
__device__ inline void w2_coef(int p, const WinPair& wp, hlbl_geom global, size_t V, size_t z0,
                               size_t end, const float* __restrict__ coef, float* dst) {
  const int lane = threadIdx.x & 63;
  const size_t z = z0 + p;
  if (z >= end) return;
  int crd[4];
  site_coords(z, global, crd);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    crd[r] -= coord_dim(wp.gsx, r);
    crd[r] += crd[r] < 0 ? geom_dim(global, r) : 0;
  }
  const size_t rs = (((size_t)crd[0] * global.LX + crd[1]) * global.LY + crd[2]) * global.LZ + crd[3];
  if (lane >= kCoefUsed / 4) return;
  const float* src = coef + ((size_t)wp.yv_slot * V + rs) * kCoefField + 4 * lane;
  for (int row = 0; row < 16; ++row)
    __builtin_amdgcn_global_load_lds((const void*)(src + row * kCoefUsed),
                                     (__attribute__((address_space(3))) void*)(dst + p * kW2CoefSite + row * kW2CoefRow),
                                     16, 0, 0);
}

// dxu of both flavours of a staged group by the producer lanes: task (site, f2, r, a)
__device__ inline void w2_dxu(int ptid, const float* buf, float* dxu) {
  for (int t = ptid; t < kW2Sites * 288; t += 64 * kW2Prod) {
    const int zi = t / 288, f2 = (t / 144) & 1, r = (t / 12) % 12, a = t % 12;
    const float* sa = buf + ((1 - f2) * 12 + a) * kW2RowPad + zi * 24;
    const float* ur = buf + (24 + f2 * 12 + r) * kW2RowPad + zi * 24;
    dxu_site_lean<float, float, float>(sa, ur, dxu + ((zi * 2 + f2) * 4) * kW2DxuRow + r * 24 + 2 * a,
                                       kW2DxuRow);
  }
}

__global__ __launch_bounds__(kW2Threads) void pass2_window(
    const WinPair* __restrict__ pairs, int n_pairs, const float* __restrict__ bmat,
    const float* __restrict__ coef, hlbl_geom global, size_t sites_per_wg, int nwg,
    double* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) float sF[2][kW2BufLen];
  __shared__ float sDxu[2][kW2DxuLen];
  __shared__ __attribute__((aligned(16))) float sCoef[2][kW2Sites * kW2CoefSite];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const bool producer = wave >= kW2Waves;
  const size_t V = geom_volume(global);
  const size_t begin = (size_t)blockIdx.x * sites_per_wg;
  const size_t end = begin + sites_per_wg < V ? begin + sites_per_wg : V;
  if (begin >= end) return;
  const int ng = (int)((end - begin + kW2Sites - 1) / kW2Sites);
  const int total = ng * n_pairs;
  auto group_pair = [&](int it) { return it / ng; };
  auto group_z0 = [&](int it) { return begin + (size_t)(it - (it / ng) * ng) * kW2Sites; };

  if (producer) {
    const int pw = wave - kW2Waves, ptid = tid - 64 * kW2Waves;
    double2 r[kW2Fetch];
    // prologue: spinors of groups 0 and 1 staged, coefficients of group 0 staged,
    // group 2 in flight, dxu of group 0 built
    w2_fetch(ptid, pairs[0], V, begin, r);
    w2_store(ptid, r, sF[0]);
    if (total > 1) {
      w2_fetch(ptid, pairs[group_pair(1)], V, group_z0(1), r);
      w2_store(ptid, r, sF[1]);
    }
    w2_coef(pw, pairs[0], global, V, begin, end, coef, sCoef[0]);
    bool inflight = total > 2;
    if (inflight) w2_fetch(ptid, pairs[group_pair(2)], V, group_z0(2), r);
    // the barrier waits for the coefficient loads but lets the later fetch fly
    if (inflight) barrier_vm_fetch(); else barrier_all();
    w2_dxu(ptid, sF[0], sDxu[0]);
    for (int it = 0; it < total; ++it) {
      if (inflight) barrier_vm_fetch(); else barrier_all();
      if (it + 2 < total) w2_store(ptid, r, sF[it & 1]);  // fetched during the last step
      if (it + 1 < total) {
        w2_coef(pw, pairs[group_pair(it + 1)], global, V, group_z0(it + 1), end, coef, sCoef[(it + 1) & 1]);
        w2_dxu(ptid, sF[(it + 1) & 1], sDxu[(it + 1) & 1]);
      }
      inflight = it + 3 < total;
      if (inflight) w2_fetch(ptid, pairs[group_pair(it + 3)], V, group_z0(it + 3), r);
    }
    return;
  }

  // MFMA waves
  const int f = wave >> 2, kq = wave & 3;
  // A fragments: row (mu, lam) = lane%16, K index r*24 + j, j = 4*jq + lane/16,
  // read at dxu offset perm(j) with a sign: packed 5 bits per jq (offsets) and one
  // bit per jq (negate)
  const int arow = lane & 15, amu = arow >> 2, alam = arow & 3;
  unsigned aoffp = 0, asgp = 0;
#pragma unroll
  for (int jq = 0; jq < 6; ++jq) {
    const int j = 4 * jq + (lane >> 4);
    const int pj = gamma_perm24(alam, j);
    aoffp |= (unsigned)pj << (5 * jq);
    if (gamma_sign24(4, pj) * gamma_sign24(alam, j) < 0) asgp |= 1u << jq;
  }
  const bool c2on = (lane & 15) < kCoefUsed - 32;
  // FP32 16x16x4 result: element i of lane l is row 4*(l/16) + i, column l%16
  const int coff = (4 * (lane >> 4)) * kW2CoefRow + (lane & 15);
  float breg[3][kW2KSteps];
  Kahan acc;
  barrier_all();  // the producers' prologue
  for (int it = 0; it < total; ++it) {
    const int ip = group_pair(it);
    const size_t z0 = group_z0(it);
    if (z0 == begin) {
      const float* bm = bmat + (size_t)ip * 2 * 288 * kCoefCols;
#pragma unroll
      for (int cb = 0; cb < 3; ++cb)
#pragma unroll
        for (int i = 0; i < kW2KSteps; ++i)
          breg[cb][i] = bm[((size_t)f * 288 + 4 * (kq * kW2KSteps + i) + (lane >> 4)) * kCoefCols +
                           16 * cb + (lane & 15)];
      acc = Kahan();
    }
    barrier_all();
    const float* dg = sDxu[it & 1];
    const float* cg = sCoef[it & 1];
    float v = 0.f;
#pragma unroll 1
    for (int zi = 0; zi < kW2Sites; ++zi) {
      if (z0 + zi >= end) break;
      // this site's coefficients (FP32 16x16x4 result: element i of lane l is row
      // 4*(l/16) + i, column l%16)
      const float* cs = cg + zi * kW2CoefSite + coff;
      float cv[4][3];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        cv[i][0] = cs[i * kW2CoefRow];
        cv[i][1] = cs[i * kW2CoefRow + 16];
        cv[i][2] = c2on ? cs[i * kW2CoefRow + 32] : 0.f;
      }
      // A values, read six at a time (one row r of dxu) ahead of their MFMAs
      const float* dx = dg + ((zi * 2 + f) * 4 + amu) * kW2DxuRow + (kq * kW2KSteps / 6) * 24;
      f4v c0 = {0.f, 0.f, 0.f, 0.f}, c1 = c0, c2 = c0;
#pragma unroll
      for (int rb = 0; rb < kW2KSteps / 6; ++rb) {  // kq*18 is a multiple of 6: jq = i % 6
        float av[6];
#pragma unroll
        for (int jq = 0; jq < 6; ++jq) av[jq] = dx[rb * 24 + ((aoffp >> (5 * jq)) & 31)];
#pragma unroll
        for (int jq = 0; jq < 6; ++jq) {
          const int i = 6 * rb + jq;
          const float a = __uint_as_float(__float_as_uint(av[jq]) ^ (((asgp >> jq) & 1u) << 31));
          c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, breg[0][i], c0, 0, 0, 0);
          c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, breg[1][i], c1, 0, 0, 0);
          c2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a, breg[2][i], c2, 0, 0, 0);
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) v += c0[i] * cv[i][0] + c1[i] * cv[i][1] + c2[i] * cv[i][2];
    }
    acc.add((double)v);
    if (z0 + kW2Sites >= end) {
      double s = acc.s - acc.c;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) part[((size_t)ip * nwg + blockIdx.x) * kW2Waves + wave] = s;
    }
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
    for (int j = 0; j < kW2Waves / 2; ++j) k.add(part[((size_t)ip * nwg + w) * kW2Waves + f * (kW2Waves / 2) + j]);
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
    hipLaunchKernelGGL(kfield_consts_win, dim3(nu), dim3(64), 0, st, c->qed->table64.p, w.d_yvs.p, g,
                       w.xunit, w.cst64.p, w.cst32.p);
    check_hip(hipGetLastError(), "kfield_consts_win launch");
    const int nwg = (int)std::min<size_t>(div_up((size_t)nu * V, kKmThreads / 64), (size_t)4 * c->n_cu);
    hipLaunchKernelGGL(kfield_mixed, dim3(nwg), dim3(kKmThreads), 0, st, c->qed->table32.p,
                       w.d_yvs.p, nu, g, w.xunit, w.cst32.p, w.form == 2 ? w.chbuf.p : nullptr,
                       w.coef.p);
    check_hip(hipGetLastError(), "kfield_mixed launch");
    // candidate boxes: ceil(kNearRadius / xunit) lattice units, at most the whole extent
    hlbl_coord box;
    const int hb = (int)std::ceil(kNearRadius / w.xunit);
    box.t = std::min(hb, (g.T - 1) / 2);
    box.x = std::min(hb, (g.LX - 1) / 2);
    box.y = std::min(hb, (g.LY - 1) / 2);
    box.z = std::min(hb, (g.LZ - 1) / 2);
    hipLaunchKernelGGL(kfield_near, dim3((unsigned)c->n_cu), dim3(kKmThreads), 0, st, c->qed->table64.p,
                       w.d_yvs.p, nu, g, w.xunit, w.cst64.p, box, w.coef.p);
    check_hip(hipGetLastError(), "kfield_near launch");
    if (w.form == 2) {
      hipLaunchKernelGGL(judge_coef, dim3(w.n_samples, nu), dim3(64), 0, st, c->qed->table64.p,
                         w.d_yvs.p, g, w.xunit, w.cst64.p, w.d_samples.p, w.n_samples, w.coef.p,
                         w.judge.p);
      check_hip(hipGetLastError(), "judge_coef launch");
    }
  }
  // pass 2
  if (stage == 2) {
    const int nwg = c->n_cu;
    const size_t spw = div_up(div_up(V, (size_t)nwg), kW2Sites) * kW2Sites;
    const int n = (int)div_up(V, spw);
    hipLaunchKernelGGL(pass2_window, dim3(n), dim3(kW2Threads), 0, st, w.d_pairs.p, np, w.bmat.p,
                       w.coef.p, g, spw, n, w.part2.p);
    check_hip(hipGetLastError(), "pass2_window launch");
    hipLaunchKernelGGL(reduce_window2, dim3((unsigned)div_up(2 * np, 64)), dim3(64), 0, st,
                       w.part2.p, n, np, w.result.p);
    check_hip(hipGetLastError(), "reduce_window2 launch");
  }
}

size_t window_part2_len(const hlbl_contractor* c, int n_pairs) {
  return (size_t)n_pairs * c->n_cu * kW2Waves;
}


}  // namespace hlbl
