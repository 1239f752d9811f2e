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
  R u[4], v[4], a, be, b;
  eval_args(e, g, u, v);
  qed::grid_coords(u, v, a, be, b);
  const qed::Cell<R> cell = qed::locate(a, be, b);
  return qed::interp(table, cell, ch);
}

// Phase 2: the T tensors of evaluation e, components j, 64 + j and 128 + j of lane j
// (they share the frame entries) into T, from the 42 frame components ch_e.
template <class R>
__device__ inline void kp_tensors(int e, int j, const SiteGeom<R>& g, const R* ch_e, R* T) {
  static_assert(qed::kNT == 3 * 64, "three components per lane");
  R u[4], v[4], e0[4], e1[4];
  eval_args(e, g, u, v);
  qed::frame(u, v, e0, e1);
  const qed::TArgs<R> t = qed::t_args(j, e0, e1);
#pragma unroll
  for (int A = 0; A < 3; ++A) T[64 * A + j] = qed::t_value(t, ch_e + qed::kFrame * A);
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

constexpr int kCoefUsed = 40;
constexpr int kCoefField = 16 * kCoefUsed;  // coefficients per site

// Where the channels of the six site evaluations come from when they are not
// interpolated: p[e] points at the kNCh channels of evaluation e of this site
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

// ---- the coefficient field, one wave per site --------------------------------
// A wave computes the 16 x 40 coefficients of one site in its own LDS area, with
// no workgroup barrier: a wave's LDS operations complete in order, and
// wave_sync() keeps the compiler from moving accesses across phases.
//   ch[6][kNCh]  channels of the six evaluations, all fetched first (one latency)
//   T[192]  its T tensors
//   L[3][384]  L0 slots: 0 = (x, 0), 1 = the partner (0, x) / (0, d), 2 = the main
//           evaluation, turned in place into K_m (each lane updates its own entries)
//   C[640]  coefficients; the K_m scatter into the 24 G columns as they complete
//           (one coefficient per K element), the H columns gather from K3 at the end
//   D[48]   subtraction matrices of xm, ym, dm
constexpr int kChS = qed::kNChUsed;  // channel stride of one evaluation in LDS
constexpr int kKwLen = 6 * kChS + qed::kNT + 3 * qed::kNL + kCoefField + 48;
// the same for the FP32 window field (kfield_wave_axes, kfield_wave_axis_rows), which
// keeps the channels of three evaluations at stride kChS3, ch[3][42], and forms each
// lane's rows of the subtraction matrices in registers (no D area)
constexpr int kChS3 = qed::kNChUsed;
constexpr int kKwLen3 = 3 * kChS3 + qed::kNT + 3 * qed::kNL + kCoefField;

// L0 sparse terms, staged in LDS by the workgroup, one layout per precision.
// FP32: each output's terms padded to kMaxTerms with zero terms: an output is a fixed
// kMaxTerms-term sum of 32-bit words (low half: the coefficient as FP16, exact for these
// small integers; high half: the byte offset of the T entry), terms 0..3 and 4..7 in
// two 16-byte rows and term 8 beside them, read with three independent loads; a term
// costs one address add, one load and one mixed-precision multiply-add. FP64, where the
// arithmetic costs more than the latency: the compact rows, each output looping over
// its nterm terms, outputs taken in order of decreasing term count.
static_assert(qed::kMaxTerms == 9, "terms 0..3, 4..7 in 16-byte rows, term 8 beside them");
template <class R>
struct TermLds;
template <>
struct TermLds<float> {
  __attribute__((aligned(16))) unsigned term[2][qed::kNL][4];
  unsigned term8[qed::kNL];
};
template <>
struct TermLds<double> {
  unsigned short term[qed::kNL][qed::kMaxTerms];
  unsigned short order[qed::kNL];  // outputs by decreasing term count
  unsigned char nterm[qed::kNL];
};

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// copy the term table into LDS; the whole workgroup calls it, then a barrier
__device__ inline void stage_terms(const qed::TermTable& src, TermLds<float>* dst) {
  for (int i = threadIdx.x; i < qed::kNL * qed::kMaxTerms; i += blockDim.x) {
    const int o = i / qed::kMaxTerms, j = i % qed::kMaxTerms;
    unsigned w = 0u;  // padding: coefficient 0 (T entry 0)
    if (j < src.nterm[o]) {
      const unsigned short e = src.term[o][j];
      const _Float16 c = (_Float16)(float)qed::term_coef(e);
      unsigned short cb;
      __builtin_memcpy(&cb, &c, 2);
      w = (unsigned)cb | ((unsigned)(qed::term_index(e) * sizeof(float)) << 16);
    }
    if (j < 8) dst->term[j >> 2][o][j & 3] = w;
    else dst->term8[o] = w;
  }
}
__device__ inline void stage_terms(const qed::TermTable& src, TermLds<double>* dst) {
  const unsigned short* s = &src.term[0][0];
  unsigned short* d = &dst->term[0][0];
  for (int i = threadIdx.x; i < qed::kNL * qed::kMaxTerms; i += blockDim.x) d[i] = s[i];
  for (int i = threadIdx.x; i < qed::kNL; i += blockDim.x) {
    dst->nterm[i] = src.nterm[i];
    dst->order[i] = src.order[i];
  }
}

// L0 output o: FP32, the padded terms in term order, all term reads and all T
// reads independent of each other
__device__ inline float l0_lds(const TermLds<float>& tt, int o, const float* t) {
  const uint4 w0 = *reinterpret_cast<const uint4*>(&tt.term[0][o][0]);
  const uint4 w1 = *reinterpret_cast<const uint4*>(&tt.term[1][o][0]);
  const unsigned words[qed::kMaxTerms] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w, tt.term8[o]};
  const char* tb = reinterpret_cast<const char*>(t);
  float r = 0.f;
#pragma unroll
  for (int j = 0; j < qed::kMaxTerms; ++j) {
    const unsigned short cb = (unsigned short)(words[j] & 0xffffu);
    _Float16 c;
    __builtin_memcpy(&c, &cb, 2);
    r = fmaf((float)c, *reinterpret_cast<const float*>(tb + (words[j] >> 16)), r);
  }
  return r;
}
// FP64: its nterm terms
__device__ inline double l0_lds(const TermLds<double>& tt, int o, const double* t) {
  double r = 0.0;
  const int n = tt.nterm[o];
  for (int j = 0; j < n; ++j) {
    const unsigned short e = tt.term[o][j];
    r += double(qed::term_coef(e)) * t[qed::term_index(e)];
  }
  return r;
}
// the output of round slot t
__device__ inline int l0_output(const TermLds<float>&, int t) { return t; }
__device__ inline int l0_output(const TermLds<double>& tt, int t) { return tt.order[t]; }

// the channels of the six site evaluations into ch[6][kNCh]: from the rows when
// given, else interpolated in the table
template <class R, class TT>
__device__ __forceinline__ void wave_channels(const TT* __restrict__ table, const SiteGeom<R>& g,
                                              const ChannelRows& rows, R* ch) {
  const int lane = threadIdx.x & 63;
  static_assert(6 * qed::kNChUsed <= 4 * 64, "four rounds of the wave");
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int t = lane + 64 * k;
    if (t < 6 * qed::kNChUsed) {
      const int e = t / qed::kNChUsed, c = t - e * qed::kNChUsed;
      ch[e * kChS + c] = rows.given ? (R)rows.row(e)[c] : kp_channel(e, c, g, table);
    }
  }
  wave_sync();
}

// T tensors and L0 of evaluation e (0..7) into dst[384] from its channels ch[kNCh]
template <class R>
__device__ __forceinline__ void wave_l0_from(int e, const TermLds<R>& tt, const SiteGeom<R>& g,
                                             const R* ch, R* T, R* dst) {
  const int lane = threadIdx.x & 63;
  kp_tensors(e, lane, g, ch, T);
  wave_sync();
#pragma unroll 2
  for (int t = lane; t < qed::kNL; t += 64) {
    const int o = l0_output(tt, t);
    dst[o] = l0_lds(tt, o, T);
  }
  wave_sync();
}

// channels (interpolated), T tensors and L0 of evaluation e into dst[384]
template <class R, class TT>
__device__ __forceinline__ void wave_l0(int e, const TT* __restrict__ table, const TermLds<R>& tt,
                                        const SiteGeom<R>& g, R* ch, R* T, R* dst) {
  const int lane = threadIdx.x & 63;
  if (lane < qed::kNChUsed) ch[lane] = kp_channel(e, lane, g, table);
  wave_sync();
  wave_l0_from(e, tt, g, ch, T, dst);
}

// K_m (m = 0: L(x,y), 1: L(y,x), 2: L(x,d)) in place over `main`, and its G-column
// contributions: coef[(mu lam)][(k nu)] = -(K1[k mu nu lam] + K2[k nu mu lam] - K3[k lam nu mu])
// (ker_4pt_contraction, kernels/cuda_lattice.cu:403-476)
template <class R>
__device__ __forceinline__ void wave_kernel(int m, R* main, const R* l00y, const R* l0x0,
                                            const R* du, const R* dv, bool in, R* C) {
  // the lane's outputs o = lane + 64 kk share (a, b, c) = ((o >> 4) & 3, (o >> 2) & 3, o & 3),
  // so their rows of the subtraction matrices are loaded once (qed::lambda_subtract:
  // the same terms in the same order)
  const int lane = threadIdx.x & 63;
  const int c = lane & 3, b = (lane >> 2) & 3, a = (lane >> 4) & 3;  // of every o = lane + 64 kk
  R dm[4], dn[4];
#pragma unroll
  for (int al = 0; al < 4; ++al) {
    dm[al] = du[a * 4 + al];
    dn[al] = dv[b * 4 + al];
  }
#pragma unroll 2
  for (int kk = 0; kk < qed::kNL / 64; ++kk) {
    const int o = lane + 64 * kk;
    R k = main[o];
#pragma unroll
    for (int al = 0; al < 4; ++al) {
      k -= dm[al] * l00y[((kk * 4 + al) * 4 + b) * 4 + c];
      k -= dn[al] * l0x0[((kk * 4 + a) * 4 + al) * 4 + c];
    }
    k = in ? k : R(0);
    main[o] = k;
    if (m == 0) C[(a * 4 + c) * kCoefUsed + kk * 4 + b] = -k;       // K1[kk][mu=a][nu=b][lam=c]
    else if (m == 1) C[(b * 4 + c) * kCoefUsed + kk * 4 + a] -= k;  // K2[kk][nu=a][mu=b][lam=c]
    else C[(c * 4 + a) * kCoefUsed + kk * 4 + b] += k;              // K3[kk][lam=a][nu=b][mu=c]
  }
  wave_sync();
}

// FP32: L0 of evaluation e and K_m from it in one pass. With the padded term table a
// lane's L0 outputs are o = lane + 64 kk, the same outputs it turns into K_m, so the
// L0 value stays in a register (wave_l0_from followed by wave_kernel, without the
// round trip of `main` through LDS and one wave_sync).
// The lane's rows of the subtraction matrices, D_{a al}(pu) and D_{b al}(pv), are
// formed from the argument vectors (qed::subtraction_matrix, entry by entry).
__device__ __forceinline__ void wave_l0_kernel(int e, int m, const TermLds<float>& tt, const SiteGeom<float>& g,
                                               const float* ch, float* T, float* main, const float* l00y,
                                               const float* l0x0, const float pu[4], const float pv[4], bool in,
                                               float* C) {
  const int lane = threadIdx.x & 63;
  kp_tensors(e, lane, g, ch, T);
  wave_sync();
  const int c = lane & 3, b = (lane >> 2) & 3, a = (lane >> 4) & 3;
  float dm[4], dn[4];
  {
    const float lam = (float)qed::kLambda;
    const float eu = expf(-lam * norm2(pu) * 0.5f), ev = expf(-lam * norm2(pv) * 0.5f);
    const float ua = blend4(pu, a), vb = blend4(pv, b);
#pragma unroll
    for (int al = 0; al < 4; ++al) {
      dm[al] = ((a == al ? 1.f : 0.f) - lam * ua * pu[al]) * eu;
      dn[al] = ((b == al ? 1.f : 0.f) - lam * vb * pv[al]) * ev;
    }
  }
#pragma unroll 2
  for (int kk = 0; kk < qed::kNL / 64; ++kk) {
    const int o = lane + 64 * kk;
    float k = l0_lds(tt, o, T);
#pragma unroll
    for (int al = 0; al < 4; ++al) {
      k -= dm[al] * l00y[((kk * 4 + al) * 4 + b) * 4 + c];
      k -= dn[al] * l0x0[((kk * 4 + a) * 4 + al) * 4 + c];
    }
    k = in ? k : 0.f;
    main[o] = k;
    if (m == 0) C[(a * 4 + c) * kCoefUsed + kk * 4 + b] = -k;
    else if (m == 1) C[(b * 4 + c) * kCoefUsed + kk * 4 + a] -= k;
    else C[(c * 4 + a) * kCoefUsed + kk * 4 + b] += k;
  }
  wave_sync();
}

// H column (s, nu) of row (mu, lam): sum over pairs k = (rho, sigma) of
// K3[k][lam][nu][mu] (xz_sigma [rho == s] - xz_rho [sigma == s])
template <class R>
__device__ inline R coef_h(int row, int hc, const SiteGeom<R>& g, const R* K3) {
  const int mu = row >> 2, lam = row & 3, s = hc >> 2, nu = hc & 3;
  R r = R(0);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int rho = comb_rho(k), sigma = comb_sigma(k);
    const R b = K3[((k * 4 + lam) * 4 + nu) * 4 + mu];
    if (sigma == s) r -= b * g.xz[rho];
    if (rho == s) r += b * g.xz[sigma];
  }
  return r;
}

// The 640 coefficients of one site into out (element type O), computed by the
// calling wave in w (kKwLen elements of LDS). cst = [L0(0,y) | L0(y,0)] of the
// pair (768 elements, any address space readable by the wave).
template <class R, class TT, class O>
__device__ __forceinline__ void kfield_wave(const TT* __restrict__ table, const TermLds<R>& tt,
                                            const SiteGeom<R>& g, const ChannelRows rows,
                                            const R* cst, R* w, O* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  R* ch = w;
  R* T = ch + 6 * kChS;
  R* L = T + qed::kNT;
  R* C = L + 3 * qed::kNL;
  R* D = C + kCoefField;
  R* s0 = L;
  R* s1 = L + qed::kNL;
  R* s2 = L + 2 * qed::kNL;
  if (lane < 16) kp_dmat(lane, g, D);
  wave_channels(table, g, rows, ch);
  const R* Dx = D;
  const R* Dy = D + 16;
  const R* Dd = D + 32;
  wave_l0_from(1, tt, g, ch + 1 * kChS, T, s0);                  // (x, 0)
  wave_l0_from(0, tt, g, ch + 0 * kChS, T, s2);                  // (x, y)
  wave_kernel(0, s2, cst, s0, Dx, Dy, g.in12, C);              // K1 = L(x, y)
  wave_l0_from(3, tt, g, ch + 3 * kChS, T, s1);                  // (0, x)
  wave_l0_from(2, tt, g, ch + 2 * kChS, T, s2);                  // (y, x)
  wave_kernel(1, s2, s1, cst + qed::kNL, Dy, Dx, g.in12, C);   // K2 = L(y, x)
  wave_l0_from(5, tt, g, ch + 5 * kChS, T, s1);                  // (0, d)
  wave_l0_from(4, tt, g, ch + 4 * kChS, T, s2);                  // (x, d)
  wave_kernel(2, s2, s1, s0, Dx, Dd, g.in3, C);                // K3 = L(x, d)
  for (int t = lane; t < 16 * 16; t += 64) {
    const int row = t >> 4, hc = t & 15;
    C[row * kCoefUsed + 24 + hc] = coef_h(row, hc, g, s2);
  }
  wave_sync();
#pragma unroll
  for (int k = 0; k < kCoefField / 64; ++k) out[64 * k + lane] = (O)C[64 * k + lane];
  wave_sync();
}

// The evaluations of L0 that depend on the site alone, L0(x, 0) and L0(0, x), into
// x0[384] and ox[384] (one wave per site; w: kKwLen elements of LDS). The same
// functions as in kfield_wave, so the rows equal what it computes in place.
template <class R, class TT, class O>
__device__ __forceinline__ void kfield_wave_axis_rows(const TT* __restrict__ table, const TermLds<R>& tt,
                                                      const SiteGeom<R>& g, const ChannelRows rows, R* w,
                                                      O* __restrict__ x0, O* __restrict__ ox) {
  const int lane = threadIdx.x & 63;
  R* ch = w;  // [2][kChS3]: (x, 0), (0, x)
  R* T = ch + 3 * kChS3;
  R* s0 = T + qed::kNT;
  if (lane < qed::kNChUsed) {
    ch[0 * kChS3 + lane] = rows.given ? (R)rows.row(1)[lane] : kp_channel(1, lane, g, table);
    ch[1 * kChS3 + lane] = rows.given ? (R)rows.row(3)[lane] : kp_channel(3, lane, g, table);
  }
  wave_sync();
  wave_l0_from(1, tt, g, ch + 0 * kChS3, T, s0);  // (x, 0)
#pragma unroll
  for (int k = 0; k < qed::kNL / 64; ++k) x0[64 * k + lane] = (O)s0[64 * k + lane];
  wave_sync();
  wave_l0_from(3, tt, g, ch + 1 * kChS3, T, s0);  // (0, x)
#pragma unroll
  for (int k = 0; k < qed::kNL / 64; ++k) ox[64 * k + lane] = (O)s0[64 * k + lane];
  wave_sync();
}

// kfield_wave with the site-only evaluations read from rows computed by
// kfield_wave_axis_rows: x0 = L0(x, 0) and ox = L0(0, x) of this site, od =
// L0(0, d), the (0, x) row of the site d = x - y (all 16-byte aligned). Only (x, y),
// (y, x) and (x, d) are evaluated here. w holds kKwLen3 floats.
template <class TT, class O>
__device__ __forceinline__ void kfield_wave_axes(const TT* __restrict__ table, const TermLds<float>& tt,
                                                 const SiteGeom<float>& g, const ChannelRows rows,
                                                 const float* cst, const float* __restrict__ x0,
                                                 const float* __restrict__ ox, const float* __restrict__ od,
                                                 float* w, O* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  float* ch = w;  // [3][kChS3]: (x, y), (y, x), (x, d)
  float* T = ch + 3 * kChS3;
  float* L = T + qed::kNT;
  float* C = L + 3 * qed::kNL;
  float* s0 = L;
  float* s1 = L + qed::kNL;
  float* s2 = L + 2 * qed::kNL;
  // the site's rows (x, 0) and (0, x), loaded first (one latency with the channels)
  float rx0[qed::kNL / 64], rox[qed::kNL / 64];
#pragma unroll
  for (int k = 0; k < qed::kNL / 64; ++k) {
    rx0[k] = x0[64 * k + lane];
    rox[k] = ox[64 * k + lane];
  }
  // channels of the evaluations 0 (x, y), 2 (y, x), 4 (x, d)
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int t = lane + 64 * k;
    if (t < 3 * qed::kNChUsed) {
      const int j = t / qed::kNChUsed, c = t - j * qed::kNChUsed, e = 2 * j;
      ch[j * kChS3 + c] = rows.given ? rows.row(e)[c] : kp_channel(e, c, g, table);
    }
  }
#pragma unroll
  for (int k = 0; k < qed::kNL / 64; ++k) {
    s0[64 * k + lane] = rx0[k];
    s1[64 * k + lane] = rox[k];
  }
  wave_sync();
  wave_l0_kernel(0, 0, tt, g, ch + 0 * kChS3, T, s2, cst, s0, g.xm, g.ym, g.in12, C);             // K1 = L(x, y)
  wave_l0_kernel(2, 1, tt, g, ch + 1 * kChS3, T, s2, s1, cst + qed::kNL, g.ym, g.xm, g.in12, C);  // K2 = L(y, x)
  // the (0, d) row into s1, loaded now so that its registers are not live above
#pragma unroll
  for (int k = 0; k < qed::kNL / 64; ++k) s1[64 * k + lane] = od[64 * k + lane];
  wave_sync();
  wave_l0_kernel(4, 2, tt, g, ch + 2 * kChS3, T, s2, s1, s0, g.xm, g.dm, g.in3, C);               // K3 = L(x, d)
  for (int t = lane; t < 16 * 16; t += 64) {
    const int row = t >> 4, hc = t & 15;
    C[row * kCoefUsed + 24 + hc] = coef_h(row, hc, g, s2);
  }
  wave_sync();
#pragma unroll
  for (int k = 0; k < kCoefField / 64; ++k) out[64 * k + lane] = (O)C[64 * k + lane];
  wave_sync();
}

// FP64 kfield_wave with the site-only evaluations read from FP64 axis rows (x0 =
// L0(x, 0) and ox = L0(0, x) of this site, od = L0(0, d) of the site d = x - y), as
// kfield_wave_axes does in FP32: the same steps as kfield_wave in the same order, with
// only (x, y), (y, x) and (x, d) evaluated. w holds kKwLen64A doubles.
constexpr int kKwLen64A = 3 * kChS3 + qed::kNT + 3 * qed::kNL + kCoefField + 48;
template <class TT, class O>
__device__ __forceinline__ void kfield_wave_axes64(const TT* __restrict__ table, const TermLds<double>& tt,
                                                   const SiteGeom<double>& g, const double* cst,
                                                   const double* __restrict__ x0, const double* __restrict__ ox,
                                                   const double* __restrict__ od, double* w,
                                                   O* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  double* ch = w;  // [3][kChS3]: (x, y), (y, x), (x, d)
  double* T = ch + 3 * kChS3;
  double* L = T + qed::kNT;
  double* C = L + 3 * qed::kNL;
  double* D = C + kCoefField;
  double* s0 = L;
  double* s1 = L + qed::kNL;
  double* s2 = L + 2 * qed::kNL;
  if (lane < 16) kp_dmat(lane, g, D);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int t = lane + 64 * k;
    if (t < 3 * qed::kNChUsed) {
      const int j = t / qed::kNChUsed, c = t - j * qed::kNChUsed;
      ch[j * kChS3 + c] = kp_channel(2 * j, c, g, table);
    }
  }
#pragma unroll
  for (int k = 0; k < qed::kNL / 64; ++k) {
    s0[64 * k + lane] = x0[64 * k + lane];
    s1[64 * k + lane] = ox[64 * k + lane];
  }
  wave_sync();
  const double* Dx = D;
  const double* Dy = D + 16;
  const double* Dd = D + 32;
  wave_l0_from(0, tt, g, ch + 0 * kChS3, T, s2);                 // (x, y)
  wave_kernel(0, s2, cst, s0, Dx, Dy, g.in12, C);               // K1 = L(x, y)
  wave_l0_from(2, tt, g, ch + 1 * kChS3, T, s2);                 // (y, x)
  wave_kernel(1, s2, s1, cst + qed::kNL, Dy, Dx, g.in12, C);    // K2 = L(y, x)
#pragma unroll
  for (int k = 0; k < qed::kNL / 64; ++k) s1[64 * k + lane] = od[64 * k + lane];
  wave_sync();
  wave_l0_from(4, tt, g, ch + 2 * kChS3, T, s2);                 // (x, d)
  wave_kernel(2, s2, s1, s0, Dx, Dd, g.in3, C);                 // K3 = L(x, d)
  for (int t = lane; t < 16 * 16; t += 64) {
    const int row = t >> 4, hc = t & 15;
    C[row * kCoefUsed + 24 + hc] = coef_h(row, hc, g, s2);
  }
  wave_sync();
#pragma unroll
  for (int k = 0; k < kCoefField / 64; ++k) out[64 * k + lane] = (O)C[64 * k + lane];
  wave_sync();
}

// The pair constants [L0(0,y) | L0(y,0)] into cst (768 elements), by one wave
// (w: kKwLen elements of LDS).
template <class R, class TT>
__device__ __forceinline__ void kfield_wave_constants(const TT* __restrict__ table, const TermLds<R>& tt,
                                                      const SiteGeom<R>& g, R* w, R* cst) {
  R* ch = w;
  R* T = ch + 6 * kChS;
  wave_l0(6, table, tt, g, ch, T, cst);
  wave_l0(7, table, tt, g, ch, T, cst + qed::kNL);
}

}  // namespace hlbl
