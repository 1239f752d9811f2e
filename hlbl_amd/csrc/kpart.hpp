// Per-site QED-kernel part of pass 2, computed cooperatively by a workgroup.
//
// For the site x of a pair (gsx, yv) it forms the coefficient matrix
//   coef[(mu*4+lam)][col], col = k*4+nu (24 columns, corr_I) and
//   24 + s*4+nu (16 columns, the H traces that make corr_II), 8 zero pad columns,
// such that the site's contribution to the kernel sum of one flavour is
//   sum_{row,col} coef[row][col] * C[row][col],
//   C[(mu lam)][(k nu)] = Re Tr(q^{mu lam} G_{k nu}),  C[(mu lam)][(s nu)] = Re Tr(q^{mu lam} H_{s nu}),
// the combination of ker_4pt_contraction (kernels/cuda_lattice.cu:403-476):
//   dtmp = sum (K1[k][mu][nu][lam] + K2[k][nu][mu][lam] - K3[k][lam][nu][mu]) corr_I
//        + K3[k][lam][nu][mu] corr_II,
// with K1 = L(xm, ym), K2 = L(ym, xm), K3 = L(xm, x-y) and corr = -C.
// The QED kernel of the site is formed once for both flavours.
#pragma once

#include "lattice.hpp"
#include "qed.hpp"

namespace hlbl {

constexpr int kCoefCols = 48;  // 40 used + 8 pad (three 16-column MFMA blocks)
constexpr int kCoefLen = 16 * kCoefCols;

// Coordinates of one site relative to the pair.
template <class R>
struct SiteGeom {
  R xm[4], ym[4], dm[4];  // muon units: coord_map(x-gsx), coord_map(yv), min-image x-y
  R xz[4];                // coord_map_zerohalf(x-gsx), the odd weights
  bool in12, in3;         // K1,K2 and K3 inside the kernel radius
};

template <class R>
HLBL_HD R norm2(const R v[4]) {
  return v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
}

// x_rel = (local + proc*Llocal - gsx) mod L (cuda_lattice.cu:308-316, 409-424)
template <class R>
HLBL_HD void site_geom(size_t site, const hlbl_geom& global, const hlbl_geom& local,
                       const hlbl_coord& proc, const hlbl_coord& gsx, const hlbl_coord& yv,
                       double xunit, SiteGeom<R>& g) {
  int c[4];
  site_coords(site, local, c);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int L = geom_dim(global, r);
    const int xr = mod_pos(c[r] + coord_dim(proc, r) * geom_dim(local, r) - coord_dim(gsx, r), L);
    const int yr = coord_dim(yv, r);
    g.xm[r] = (R)((double)coord_map(xr, L) * xunit);
    g.ym[r] = (R)((double)coord_map(mod_pos(yr, L), L) * xunit);
    g.dm[r] = (R)((double)coord_map_diff(xr, yr, L) * xunit);
    g.xz[r] = (R)coord_map_zerohalf(xr, L);
  }
  const R rr = (R)(qed::kRadius * qed::kRadius);
  const R nx = norm2(g.xm), ny = norm2(g.ym), nd = norm2(g.dm);
  g.in12 = nx <= rr && ny <= rr;
  g.in3 = nx <= rr && nd <= rr;
}

// Argument pair of evaluation e of L0 at the site:
//   0: (x, y)  1: (x, 0)  2: (y, x)  3: (0, x)  4: (x, d)  5: (0, d)
// and the pair constants 6: (0, y)  7: (y, 0).
// (Selects instead of pointers into g, so that g stays in registers.)
template <class R>
HLBL_HD void eval_args(int e, const SiteGeom<R>& g, R u[4], R v[4]) {
  const bool ux = e == 0 || e == 1 || e == 4, uy = e == 2 || e == 7;
  const bool vx = e == 2 || e == 3, vy = e == 0 || e == 6, vd = e == 4 || e == 5;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    u[r] = ux ? g.xm[r] : (uy ? g.ym[r] : R(0));
    v[r] = vx ? g.xm[r] : (vy ? g.ym[r] : (vd ? g.dm[r] : R(0)));
  }
}

// Phase 1: channel ch of evaluation e (interpolated in the table).
template <class R, class TT>
__device__ inline R kp_channel(int e, int ch, const SiteGeom<R>& g, const TT* table) {
  R u[4], v[4], a, c, b;
  eval_args(e, g, u, v);
  qed::acb(u, v, a, c, b);
  const qed::Cell<R> cell = qed::locate(a, c, b);
  return qed::interp(table, cell, ch);
}

// Phase 2: component i of the T tensors of evaluation e.
template <class R>
__device__ inline R kp_tensor(int e, int i, const SiteGeom<R>& g, const R* ch_e) {
  R u[4], v[4];
  eval_args(e, g, u, v);
  return qed::t_component(i, u, v, ch_e);
}

// v[i] as a blend of the four values: the compiler cannot turn it into a load
// through a selected pointer, which would put the array in private memory.
template <class R>
__device__ inline R blend4(const R v[4], int i) {
  return v[0] * R(i == 0) + v[1] * R(i == 1) + v[2] * R(i == 2) + v[3] * R(i == 3);
}

// D_{m a}(p) = (delta - Lambda p_m p_a) exp(-Lambda p^2 / 2), entry i = m*4+a,
// into d[i], d[16+i], d[32+i] for p = xm, ym, dm
template <class R>
__device__ inline void kp_dmat(int i, const SiteGeom<R>& g, R* d) {
  const int m = i >> 2, a = i & 3;
  const R dl = (m == a) ? R(1) : R(0), lam = R(qed::kLambda);
  d[i] = (dl - lam * blend4(g.xm, m) * blend4(g.xm, a)) * exp(-lam * norm2(g.xm) * R(0.5));
  d[16 + i] = (dl - lam * blend4(g.ym, m) * blend4(g.ym, a)) * exp(-lam * norm2(g.ym) * R(0.5));
  d[32 + i] = (dl - lam * blend4(g.dm, m) * blend4(g.dm, a)) * exp(-lam * norm2(g.dm) * R(0.5));
}

// Phase 4: component o of K_m (m = 0: K1, 1: K2, 2: K3). L0 holds the eight
// evaluations [8][384] (six site evaluations, then (0,y) and (y,0)); D the
// subtraction matrices of xm, ym, dm.
template <class R>
__device__ inline R kp_kernel(int m, int o, const SiteGeom<R>& g, const R* L0, const R* D) {
  const R *l0xy, *l00y, *l0x0, *du, *dv;
  bool in;
  const R* Dx = D;
  const R* Dy = D + 16;
  const R* Dd = D + 32;
  if (m == 0) {         // L(x, y)
    l0xy = L0 + 0 * qed::kNL; l00y = L0 + 6 * qed::kNL; l0x0 = L0 + 1 * qed::kNL;
    du = Dx; dv = Dy; in = g.in12;
  } else if (m == 1) {  // L(y, x)
    l0xy = L0 + 2 * qed::kNL; l00y = L0 + 3 * qed::kNL; l0x0 = L0 + 7 * qed::kNL;
    du = Dy; dv = Dx; in = g.in12;
  } else {              // L(x, d)
    l0xy = L0 + 4 * qed::kNL; l00y = L0 + 5 * qed::kNL; l0x0 = L0 + 1 * qed::kNL;
    du = Dx; dv = Dd; in = g.in3;
  }
  if (!in) return R(0);
  return qed::lambda_subtract(o, l0xy, l00y, l0x0, du, dv);
}

// Phase 5: coefficient (row, col) from K = [3][384] (K1, K2, K3).
template <class R>
__device__ inline R kp_coef(int row, int col, const SiteGeom<R>& g, const R* K) {
  const int mu = row >> 2, lam = row & 3;
  const R* K1 = K;
  const R* K2 = K + qed::kNL;
  const R* K3 = K + 2 * qed::kNL;
  auto idx = [](int k, int a, int b, int c) { return ((k * 4 + a) * 4 + b) * 4 + c; };
  if (col < 24) {
    const int k = col >> 2, nu = col & 3;
    return -(K1[idx(k, mu, nu, lam)] + K2[idx(k, nu, mu, lam)] - K3[idx(k, lam, nu, mu)]);
  }
  if (col < 40) {
    const int s = (col - 24) >> 2, nu = (col - 24) & 3;
    R r = R(0);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int rho = comb_rho(k), sigma = comb_sigma(k);  // compile-time after unrolling
      const R b = K3[idx(k, lam, nu, mu)];
      if (sigma == s) r -= b * g.xz[rho];
      if (rho == s) r += b * g.xz[sigma];
    }
    return r;
  }
  return R(0);
}

constexpr int kCoefUsed = 40;
constexpr int kCoefField = 16 * kCoefUsed;  // coefficients per site

// LDS of one workgroup running the kfield functions (R elements)
constexpr int kKfieldLds = 8 * qed::kNCh + 8 * qed::kNT + 8 * qed::kNL + 48;

// Pair constants L0(0, y), L0(y, 0) (evaluations 6 and 7) into the L0 area of
// lds; the whole workgroup calls it and it ends with a barrier.
template <class R, class TT>
__device__ __forceinline__ void kfield_constants(const TT* __restrict__ table, const qed::TermTable& tt,
                                                 const SiteGeom<R>& g, R* lds) {
  R* sCh = lds;
  R* sT = sCh + 8 * qed::kNCh;
  R* sL0 = sT + 8 * qed::kNT;
  const int tid = threadIdx.x, nt = blockDim.x;
  constexpr int nch = qed::kNChUsed;
  for (int t = tid; t < 2 * nch; t += nt)
    sCh[(6 + t / nch) * qed::kNCh + t % nch] = kp_channel(6 + t / nch, t % nch, g, table);
  __syncthreads();
  for (int t = tid; t < 2 * qed::kNT; t += nt)
    sT[6 * qed::kNT + t] = kp_tensor(6 + t / qed::kNT, t % qed::kNT, g, sCh + (6 + t / qed::kNT) * qed::kNCh);
  __syncthreads();
  for (int t = tid; t < 2 * qed::kNL; t += nt)
    sL0[6 * qed::kNL + t] = qed::l0_component(tt, t % qed::kNL, sT + (6 + t / qed::kNL) * qed::kNT);
  __syncthreads();
}

// Where the channels of the six site evaluations come from when they are not
// interpolated: p[e] points at the 32 channels of evaluation e of this site
// (none: interpolate). Passed by value and read through selects, so that it
// stays in registers.
struct ChannelRows {
  const float* p[6];
  bool given;
  __device__ const float* row(int e) const {
    return e == 0 ? p[0] : e == 1 ? p[1] : e == 2 ? p[2] : e == 3 ? p[3] : e == 4 ? p[4] : p[5];
  }
  __device__ static ChannelRows none() {
    ChannelRows r{};
    r.given = false;
    return r;
  }
};

// The 16 x 40 coefficients of one site into out[640] (element type O). The
// channels of the six site evaluations come from the table, or from rows when
// given. Requires kfield_constants on the same lds (same pair). Whole
// workgroup; no barrier at the end (the next call writes sCh first, and sT only
// after its first barrier).
template <class R, class TT, class O>
__device__ __forceinline__ void kfield_site(const TT* __restrict__ table, const qed::TermTable& tt,
                                            const SiteGeom<R>& g, const ChannelRows rows, R* lds,
                                            O* __restrict__ out) {
  R* sCh = lds;                     // [8][32]
  R* sT = sCh + 8 * qed::kNCh;      // [8][192]; K [3][384] over the first 1152
  R* sL0 = sT + 8 * qed::kNT;       // [8][384]
  R* sD = sL0 + 8 * qed::kNL;       // [48]
  R* sK = sT;
  const int tid = threadIdx.x, nt = blockDim.x;
  constexpr int nch = qed::kNChUsed;
  if (rows.given) {
    for (int t = tid; t < 6 * nch; t += nt)
      sCh[(t / nch) * qed::kNCh + t % nch] = (R)rows.row(t / nch)[t % nch];
  } else {
    for (int t = tid; t < 6 * nch; t += nt)
      sCh[(t / nch) * qed::kNCh + t % nch] = kp_channel(t / nch, t % nch, g, table);
  }
  __syncthreads();
  for (int t = tid; t < 6 * qed::kNT; t += nt)
    sT[t] = kp_tensor(t / qed::kNT, t % qed::kNT, g, sCh + (t / qed::kNT) * qed::kNCh);
  __syncthreads();
  for (int t = tid; t < 6 * qed::kNL; t += nt)
    sL0[t] = qed::l0_component(tt, t % qed::kNL, sT + (t / qed::kNL) * qed::kNT);
  if (tid < 16) kp_dmat(tid, g, sD);
  __syncthreads();
  for (int t = tid; t < 3 * qed::kNL; t += nt) sK[t] = kp_kernel(t / qed::kNL, t % qed::kNL, g, sL0, sD);
  __syncthreads();
  for (int t = tid; t < kCoefField; t += nt) out[t] = (O)kp_coef(t / kCoefUsed, t % kCoefUsed, g, sK);
}

// The coefficient field coef[site - begin][16][40] of sites [begin, end) of a
// pair, plain form (channels interpolated in `table`), whole workgroup.
template <class R, class TT>
__device__ __forceinline__ void kfield_sites(const TT* __restrict__ table, const qed::TermTable& tt,
                                             hlbl_geom global, hlbl_geom local, hlbl_coord proc,
                                             hlbl_coord gsx, hlbl_coord yv, double xunit,
                                             size_t begin, size_t end, R* __restrict__ coef, R* lds) {
  SiteGeom<R> g;
  site_geom(0, global, local, proc, gsx, yv, xunit, g);
  kfield_constants(table, tt, g, lds);
  for (size_t site = begin; site < end; ++site) {
    site_geom(site, global, local, proc, gsx, yv, xunit, g);
    kfield_site<R, TT, R>(table, tt, g, ChannelRows::none(), lds, coef + (site - begin) * kCoefField);
  }
}

}  // namespace hlbl
