// Window-batched mixed-precision contraction: all pairs of a window and both
// flavours per call. Pass 1 (pass1.hip) is one launch over (plane, flavour,
// pair), pass 2 one persistent launch over the pairs. Propagator products run in
// FP32 (pass-2 traces as scaled FP16 hi/lo products on MFMA, about 21 bits), the
// lattice sums and the per-pair sums are FP64, combined with Kahan compensation. The coefficient field of the QED kernel depends only on
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
// kfield_mixed: Waves (4 or 8) waves share the term table. With 8, two workgroups per
// CU (four waves per SIMD, which its launch bound asks for) fill the LDS exactly (term
// table 13,824 B + 8 wave areas of 8,512 B = 81,920 B); with 4, three fit (47,872 B).

// pair constants of each distinct yv: cst64[u][768] and an FP32 copy; one wave per u
__global__ __launch_bounds__(64) void kfield_consts_win(const double* __restrict__ table64,
                                                        const hlbl_coord* __restrict__ yvs,
                                                        hlbl_geom global, double xunit,
                                                        double* __restrict__ cst64,
                                                        float* __restrict__ cst32) {
  __shared__ TermLds<double> tt;
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

// The site-only evaluations L0(x, 0), L0(0, x) of every site (gsx = 0): ax[0][V][384]
// and ax[1][V][384], FP32; channels from the FP32 table, or from chbuf blocks 0
// ((x, 0)) and 1 ((0, x)) for the surrogate. One wave per site.
__global__ __launch_bounds__(kKmThreads) void kfield_axes(const float* __restrict__ table32, hlbl_geom global,
                                                          double xunit, const float* __restrict__ chbuf,
                                                          float* __restrict__ ax) {
  __shared__ TermLds<float> tt;
  __shared__ float w[kKmThreads / 64][kKwLen3];
  stage_terms(c_terms32, &tt);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: site math on the scalar unit
  const size_t V = geom_volume(global);
  const size_t nw = (size_t)gridDim.x * (kKmThreads / 64);
  const hlbl_coord zero = {0, 0, 0, 0};
  for (size_t site = (size_t)blockIdx.x * (kKmThreads / 64) + wave; site < V; site += nw) {
    SiteGeom<float> g;
    site_geom(site, global, global, zero, zero, zero, xunit, g);
    ChannelRows rows = ChannelRows::none();
    if (chbuf) {
      rows.given = true;
      rows.p[1] = chbuf + ((size_t)0 * V + site) * qed::kNCh;
      rows.p[3] = chbuf + ((size_t)1 * V + site) * qed::kNCh;
    }
    kfield_wave_axis_rows<float, float, float>(table32, tt, g, rows, w[wave], ax + site * qed::kNL,
                                               ax + (V + site) * qed::kNL);
  }
}

// The field of each distinct yv in FP32 from the axis rows: per (yv, site) only
// (x, y), (y, x) and (x, d) are evaluated; (x, 0) and (0, x) are this site's rows
// and (0, d) is the (0, x) row of the site d = x - y.
template <int Waves>
__global__ __launch_bounds__(64 * Waves, Waves == 8 ? 4 : 3) void kfield_mixed(
    const float* __restrict__ table32, const hlbl_coord* __restrict__ yvs, const int* __restrict__ slots,
    int n_unique, hlbl_geom global, double xunit, const float* __restrict__ cst32,
    const float* __restrict__ chbuf, const float* __restrict__ ax, float* __restrict__ coef) {
  static_assert(sizeof(TermLds<float>) + Waves * kKwLen3 * sizeof(float) <= 163840 / (Waves == 8 ? 2 : 3),
                "two (8 waves) or three (4 waves) workgroups per CU");
  __shared__ TermLds<float> tt;
  __shared__ float w[Waves][kKwLen3];
  stage_terms(c_terms32, &tt);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: site math on the scalar unit
  const size_t V = geom_volume(global);
  const size_t total = (size_t)n_unique * V;
  const size_t nw = (size_t)gridDim.x * Waves;
  const hlbl_coord zero = {0, 0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * Waves + wave; i < total; i += nw) {
    const int u = (int)(i / V);
    const size_t site = i - (size_t)u * V;
    const hlbl_coord yv = yvs[u];
    SiteGeom<float> g;
    site_geom(site, global, global, zero, zero, yv, xunit, g);
    // the site d = x - y (the argument of (0, d))
    int crd[4], dsh[4];
    site_coords(site, global, crd);
#pragma unroll
    for (int r = 0; r < 4; ++r) dsh[r] = mod_pos(crd[r] - coord_dim(yv, r), geom_dim(global, r));
    const size_t sd = (((size_t)dsh[0] * global.LX + dsh[1]) * global.LY + dsh[2]) * global.LZ + dsh[3];
    ChannelRows rows = ChannelRows::none();
    if (chbuf) {
      // surrogate channel rows of the yv-dependent evaluations (x,y), (y,x), (x,d)
      rows.given = true;
      rows.p[0] = chbuf + ((size_t)(2 + 3 * u + 0) * V + site) * qed::kNCh;
      rows.p[2] = chbuf + ((size_t)(2 + 3 * u + 1) * V + site) * qed::kNCh;
      rows.p[4] = chbuf + ((size_t)(2 + 3 * u + 2) * V + site) * qed::kNCh;
    }
    kfield_wave_axes<float, float>(table32, tt, g, rows, cst32 + (size_t)u * 2 * qed::kNL,
                                          ax + site * qed::kNL, ax + (V + site) * qed::kNL,
                                          ax + (V + sd) * qed::kNL, w[wave],
                                          coef + ((size_t)slots[u] * V + site) * kCoefField);
  }
}

// FP64 recomputation of the near sites. Candidates: for each distinct yv u, the
// sites within box[r] lattice units of the origin (which = 0) and of y (which =
// 1) in every direction; a candidate with |x| or |x - y| below kNearRadius is
// recomputed (a site in both boxes twice, with the same result).
__global__ __launch_bounds__(kKmThreads) void kfield_near(
    const double* __restrict__ table64, const hlbl_coord* __restrict__ yvs, const int* __restrict__ slots,
    int n_unique, hlbl_geom global, double xunit, const double* __restrict__ cst64, hlbl_coord box,
    float* __restrict__ coef) {
  __shared__ TermLds<double> tt;
  __shared__ double w[kKmThreads / 64][kKwLen];
  stage_terms(c_terms32, &tt);
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: site math on the scalar unit
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
                                       coef + ((size_t)slots[u] * V + site) * kCoefField);
  }
}

// The surrogate's judge: at sample sites, the FP64 plain coefficients against
// the field in use; err[u][k] = max |c - c_plain| / max |c_plain| for sample k.
// One wave per (sample, yv).
__global__ __launch_bounds__(64) void judge_coef(
    const double* __restrict__ table64, const hlbl_coord* __restrict__ yvs, hlbl_geom global,
    double xunit, const double* __restrict__ cst64, const size_t* __restrict__ samples,
    int n_samples, const int* __restrict__ slots, const float* __restrict__ coef, double* __restrict__ err) {
  __shared__ TermLds<double> tt;
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
  const float* c = coef + ((size_t)slots[u] * V + site) * kCoefField;
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
//   C[(mu lam)][col] = Re sum_{r,j} q^{mu lam}[r*24 + j] B_f[r*24 + j][col],
//   q^{mu lam} = gamma_lam gamma_5 dxu_mu    (cuda_lattice.cu:324-344, 403-476),
// and the kernel sum += sum coef[(mu lam)][col] C. gamma_lam gamma_5 acts on the 24
// reals of a colour-spinor as a signed permutation j -> perm_lam(j), so it moves
// from the rows of A onto the columns of B:
//   C[(zi, mu)][(lam, col)] = sum_k dxu[zi][mu][k] Bg_f[k][(lam, col)],
//   Bg_f[r*24 + perm_lam(j)][(lam, col)] = sign_lam(j) B_f[r*24 + j][col],
// a 32 x 160 x 288 product per group of 8 sites (rows (site, mu)) and flavour.
//
// Precision: A and Bg are scaled by powers of two into FP16's range (A per site,
// from the site's spinor maxima: |dxu| <= 24 max|S| max|Y|; Bg per column, from
// the column maximum), split x = hi + lo in FP16 and multiplied as
// C = hi.hi + hi.lo + lo.hi on v_mfma_f32_32x32x16_f16 (FP32 accumulation), then
// unscaled exactly: every product keeps about 21 significant bits. The sum is FP32
// per group of 8 sites and FP64 (Kahan) across groups.
//
// Work unit = (pair, flavour): the unit's Bg (hi and lo, 180 KB) lives in the
// registers of 10 MFMA waves, one (32-column block, half of K) each (72 VGPRs).
// Sixteen waves, one barrier per group:
//  - waves 0..9 run the 27 MFMAs of their block for group g and dot the result
//    with the coefficients (the sum is linear, so the K halves dot separately);
//    waves 0..7 then stage group g+1's coefficients (direct loads into LDS);
//  - waves 10..15 (producers) store group g+2's spinors (fetched into registers
//    during the previous step) as FP32 with their per-site maxima, build dxu of
//    group g+1 (scaled, split hi/lo), fetch group g+3.
// Every MFMA wave reads the whole A tile of a group from LDS; at 8 sites per
// group that is half the LDS traffic per site of a 16-row tile.
constexpr int kS2Mfma = 10;                // MFMA waves: 5 column blocks of 32 x 2 halves of K
constexpr int kS2Prod = 6;                 // producer waves
constexpr int kS2Threads = 64 * (kS2Mfma + kS2Prod);
constexpr int kS2Sites = 8;
constexpr int kS2Chunks = 18;              // K = 288 in steps of 16
constexpr int kS2HalfChunks = kS2Chunks / 2;
constexpr int kS2RowLen = kS2Sites * 24;   // values of one staged spinor row (8 sites)
constexpr int kS2RowPad = kS2RowLen + 4;   // FP32 row stride
constexpr int kS2BufLen = 24 * kS2RowPad;  // rows S[1-f][a] (0..11), Y[f][r] (12..23)
constexpr int kS2ARow = 296;               // FP16 per A row (288 + 8): 148 dwords, so that the
                                           // 16-byte reads of 8 consecutive rows hit distinct banks
constexpr int kS2ALen = 32 * kS2ARow;      // one of hi / lo
constexpr int kS2CoefSite = 656;           // floats per staged site (640 + 16): 16 banks apart
constexpr int kS2Keys = 2 * kS2Sites;      // (isy, site) keys of the spinor maxima
constexpr int kS2KeyLanes = 64 * kS2Prod / kS2Keys;  // producer lanes per key: 24
constexpr int kS2Fetch = 144 / kS2KeyLanes;          // double2 per producer lane: 6
constexpr int kS2Tasks = kS2Sites * 144;   // dxu tasks (site, r, a) per group
constexpr int kS2AMax = 14;                // |scaled value| < 2^14 (FP16 max 65504)
constexpr int kWinGdzLen = 2 * 40 * 144 * 2;
static_assert(kS2Fetch == 6, "barrier_vm6 leaves exactly 6 loads in flight");
static_assert(kS2KeyLanes % 8 == 0, "8-lane groups never straddle keys");
// Bg per (pair, flavour) in fragment order [block 5][chunk 18][hi, lo][lane 64][8]
constexpr int kBsplitLen = 5 * kS2Chunks * 2 * 64 * 8;

typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef float f16acc __attribute__((ext_vector_type(16)));

// barrier that waits for this wave's LDS operations and for its global loads
// (direct loads into LDS included)
__device__ inline void barrier_all() { asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n s_barrier" ::: "memory"); }
// the same, leaving the producer's last kS2Fetch (= 6) global loads in flight
__device__ inline void barrier_vm6() { asm volatile("s_waitcnt vmcnt(6) lgkmcnt(0)\n s_barrier" ::: "memory"); }

__device__ inline unsigned short f16_bits(float x) {
  const _Float16 h = (_Float16)x;  // nearest, ties to even
  unsigned short u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}
__device__ inline float f16_float(unsigned short u) {
  _Float16 h;
  __builtin_memcpy(&h, &u, 2);
  return (float)h;
}

// Column exponents of Bg: per (pair, flavour, col < 40), te = kS2AMax - E with
// max_k |B_f[k][col]| = m 2^E, m in [0.5, 1) (0 for an all-zero column); the
// permutation does not change a column's maximum, so the four lam share it:
// bexp[pair][f][lam*40 + col].
__global__ void bsplit_exp(const double* __restrict__ gdz, int* __restrict__ bexp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (f, col) of pair blockIdx.y
  if (i >= 2 * 40) return;
  const int f = i / 40, col = i % 40;
  gdz += (size_t)blockIdx.y * kWinGdzLen;
  double m = 0.0;
  for (int k = 0; k < 288; ++k) m = fmax(m, fabs(bmat_value(gdz, f, k, col)));
  int e = 0;
  if (m > 0.0) frexp(m, &e);
  int te = m > 0.0 ? kS2AMax - e : 0;
  te = te < -120 ? -120 : (te > 120 ? 120 : te);
  for (int lam = 0; lam < 4; ++lam) bexp[((size_t)blockIdx.y * 2 + f) * 160 + lam * 40 + col] = te;
}

// Bg of every (pair, flavour) of the window, scaled by 2^te of its column and split
// hi/lo in FP16, in fragment order: element (block cb, chunk c, hl, lane l, e) is
// Bg[k = 16c + 8(l/32) + e][g = 32cb + l%32], g = lam*40 + col (the B layout of
// v_mfma_f32_32x32x16_f16).
__global__ void form_bsplit(const double* __restrict__ gdz, const int* __restrict__ bexp,
                            unsigned short* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (f, cb, c, l, e) of pair blockIdx.y
  if (i >= kBsplitLen) return;  // 2 flavours x kBsplitLen / 2 (hi and lo per thread)
  const int e = i & 7, l = (i >> 3) & 63, c = (i >> 9) % kS2Chunks, cb = (i / (512 * kS2Chunks)) % 5,
            f = i / (512 * kS2Chunks * 5);
  gdz += (size_t)blockIdx.y * kWinGdzLen;
  const int k = 16 * c + 8 * (l >> 5) + e, g = 32 * cb + (l & 31);
  const int lam = g / 40, col = g % 40, r = k / 24, jp = k % 24;
  int j = 0;
  for (int jj = 0; jj < 24; ++jj) j = gamma_perm24(lam, jj) == jp ? jj : j;
  const int te = bexp[((size_t)blockIdx.y * 2 + f) * 160 + g];
  const double v = ldexp((gamma_sign24(4, jp) * gamma_sign24(lam, j)) * bmat_value(gdz, f, r * 24 + j, col), te);
  const unsigned short hi = f16_bits((float)v);
  const unsigned short lo = f16_bits((float)(v - (double)f16_float(hi)));
  unsigned short* o = out + ((size_t)blockIdx.y * 2 + f) * kBsplitLen + (((size_t)cb * kS2Chunks + c) * 2 * 64 + l) * 8 + e;
  o[0] = hi;
  o[64 * 8] = lo;
}

// a 64-bit value the wave holds in common, as a scalar (both halves zero-extended)
__device__ inline size_t uniform_u64(size_t v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffu));
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return ((size_t)hi << 32) | (size_t)lo;
}

typedef double d2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const d2v gd2v;

// Position of a step (unit = (pair, flavour), group) along the workgroup's walk,
// advanced one step at a time (no divisions).
struct S2Cursor {
  int u = 0, g = 0;
  __device__ inline void next(int ng) {
    if (++g == ng) {
      g = 0;
      ++u;
    }
  }
};

// Producer lane ptid serves key ptid / 24 = (isy, site) of a group [z0, z0 + 8):
// isy = 0 the rows S[1-f][a], isy = 1 the rows Y[f][r] of site z0 + site. Its k-th
// double pair is d = j + 24 k (j = ptid % 24) of the key's 144: row d / 12,
// doubles 2 (d % 12) of the site. S2Lane holds the lane's fixed byte offsets.
struct S2Lane {
  unsigned row_off[kS2Fetch];  // (row * V * 24 + 2 (d % 12)) / 2: 16-byte units (V < 2^32 / 144)
  int key, isy, site;
};

__device__ inline S2Lane s2_lane(int ptid, size_t V) {
  S2Lane L;
  L.key = ptid / kS2KeyLanes;
  L.isy = L.key >= kS2Sites;
  L.site = L.key % kS2Sites;
  const int j = ptid % kS2KeyLanes;
#pragma unroll
  for (int k = 0; k < kS2Fetch; ++k) {
    const int d = j + kS2KeyLanes * k;
    L.row_off[k] = (unsigned)((size_t)(d / 12) * V * 12 + (d % 12));
  }
  return L;
}

// fetch of flavour f's rows of the group at z0 of pair wp; sites at or beyond V
// are read from site V-1 (their coefficients are masked)
__device__ inline void s2_fetch(const S2Lane& L, const WinPair* wp, int f, size_t V, size_t z0,
                                d2v (&r)[kS2Fetch]) {
  // both block pointers as wave-uniform values in the global address space (plain
  // global loads: counted by vmcnt only)
  const size_t ps = uniform_u64((size_t)wp->S), py = uniform_u64((size_t)wp->Y);
  size_t z = z0 + L.site;
  z = z < V ? z : V - 1;
  const int ff = L.isy ? f : 1 - f;
  const size_t base = (L.isy ? py : ps) + ((size_t)ff * 12 * V + z) * 24 * sizeof(double);
#pragma unroll
  for (int k = 0; k < kS2Fetch; ++k) r[k] = *reinterpret_cast<gd2v*>(base + (size_t)L.row_off[k] * 16);
}

// maximum over each group of 8 lanes (all lanes active), by DPP
__device__ inline int max8(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false));  // row_half_mirror
  return v;
}

// ... stores them as FP32 into buf [24][kS2RowPad] and folds their largest
// magnitude into smax[key] (8-lane groups never straddle keys: one DPP maximum per
// group, one LDS atomic per group; non-negative floats order as their bits)
__device__ inline void s2_store(int ptid, const S2Lane& L, const d2v (&r)[kS2Fetch], float* buf,
                                unsigned* smax) {
  const int j = ptid % kS2KeyLanes;
  float m = 0.f;
#pragma unroll
  for (int k = 0; k < kS2Fetch; ++k) {
    const int d = j + kS2KeyLanes * k, row = d / 12, w2 = 2 * (d % 12);
    float2 v;
    v.x = (float)r[k].x;
    v.y = (float)r[k].y;
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    *reinterpret_cast<float2*>(buf + ((L.isy ? 12 : 0) + row) * kS2RowPad + L.site * 24 + w2) = v;
  }
  const int mg = max8(__float_as_int(m));
  if ((ptid & 7) == 0) atomicMax(smax + L.key, (unsigned)mg);
}

// An MFMA wave stages the 16 x 40 coefficients of site z0 + p (relative to gsx) into
// dst + p * kS2CoefSite with direct loads (lane l: 16 bytes at 16 l). p, z0 and the
// pair are wave-uniform, so the site arithmetic runs on the scalar unit (32-bit:
// the volume is below 2^31).
__device__ inline void s2_coef(int p, const WinPair& wp, hlbl_geom global, size_t V, size_t z0,
                               size_t end, const float* __restrict__ coef, float* dst) {
  const int lane = threadIdx.x & 63;
  const size_t z = z0 + p;
  if (z >= end) return;
  unsigned q = (unsigned)z;
  int crd[4];
  crd[3] = (int)(q % (unsigned)global.LZ);
  q /= (unsigned)global.LZ;
  crd[2] = (int)(q % (unsigned)global.LY);
  q /= (unsigned)global.LY;
  crd[1] = (int)(q % (unsigned)global.LX);
  crd[0] = (int)(q / (unsigned)global.LX);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    crd[r] -= coord_dim(wp.gsx, r);
    crd[r] += crd[r] < 0 ? geom_dim(global, r) : 0;
  }
  const unsigned rs = (((unsigned)crd[0] * global.LX + crd[1]) * global.LY + crd[2]) * global.LZ + crd[3];
  const float* src = coef + ((size_t)wp.yv_slot * V + rs) * kCoefField + 4 * lane;
  float* d = dst + p * kS2CoefSite;
  // 640 floats = 2.5 loads of 64 lanes x 16 bytes
  __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)d, 16, 0, 0);
  __builtin_amdgcn_global_load_lds((const void*)(src + 256), (__attribute__((address_space(3))) void*)(d + 256),
                                   16, 0, 0);
  if (lane < 32)
    __builtin_amdgcn_global_load_lds((const void*)(src + 512),
                                     (__attribute__((address_space(3))) void*)(d + 512), 16, 0, 0);
}

// Site exponent: with max|S| < 2^eS and max|Y| < 2^eY (eS, eY from the float bits of
// the maxima), |dxu| < 24 2^(eS + eY), so dxu 2^se with se = 9 - eS - eY stays
// below 24 2^9 < 2^kS2AMax.
__device__ inline int site_exp(unsigned ms, unsigned my) {
  const int es = (int)((ms >> 23) & 0xffu) - 126, ey = (int)((my >> 23) & 0xffu) - 126;
  const int se = (kS2AMax - 5) - es - ey;
  return se < -120 ? -120 : (se > 120 ? 120 : se);
}

// dxu tasks of a staged group: task (site zi, r, a) is dxu_mu[r][a] of the four mu,
// scaled by 2^se of the site and split into FP16 (re, im) pairs into ah / al
// [(zi, mu)][kS2ARow] at k = r*24 + 2a; the site exponents go to sse[zi] (from the
// task (zi, 0, 0)). Tasks are numbered t = 144 zi + 24 rp + 2 a + rb (r = 2 rp + rb),
// so that the pair task P = (zi, rp, a), both r of the row pair rp with one read of
// S row a, is the tasks 2P and 2P + 1.
// The pair's (zi, rp, a) by exact 24-bit reciprocal multiplies (P < 576: P * 911 >> 16
// = P / 72 and q * 171 >> 11 = q / 12 for q < 144), full-rate multiplies in place of
// the quarter-rate 32-bit ones a division by a constant takes.
static_assert(kS2Tasks == 1152, "the reciprocal multiplies are exact below 576 pair tasks");
struct S2Task {
  unsigned zi, rp, a;
};
__device__ inline S2Task s2_pair_task(unsigned P) {
  const unsigned zi = __umul24(P, 911u) >> 16;
  const unsigned q = P - __umul24(zi, 72u);
  const unsigned rp = __umul24(q, 171u) >> 11;
  return {zi, rp, q - __umul24(rp, 12u)};
}

// the four mu of task (zi, r, a) from dxu values, scaled and split into ah / al
__device__ inline void s2_put(unsigned zi, unsigned r, unsigned a, float sc, const float (&re)[4],
                              const float (&im)[4], unsigned* ah, unsigned* al) {
  const unsigned o0 = (__umul24(zi, 4u * kS2ARow) + __umul24(r, 24u) + 2u * a) >> 1;
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    unsigned hw, lw;
    f16_split2_scaled(re[mu], im[mu], sc, hw, lw);
    const unsigned o = o0 + mu * (kS2ARow / 2);
    ah[o] = hw;
    al[o] = lw;
  }
}

// single tasks t = t0 + 64 i + lane (i < N)
template <int N>
__device__ inline void s2_dxu(int t0, int lane, const float* buf, const unsigned* smax, unsigned* ah,
                              unsigned* al, int* sse) {
#pragma unroll 2
  for (int i = 0; i < N; ++i) {
    const unsigned t = (unsigned)(t0 + 64 * i + lane);
    const S2Task k = s2_pair_task(t >> 1);
    const unsigned r = 2u * k.rp + (t & 1u);
    const int se = site_exp(smax[k.zi], smax[kS2Sites + k.zi]);
    if (r == 0 && k.a == 0) sse[k.zi] = se;
    float re[4], im[4];
    const unsigned zo = __umul24(k.zi, 24u);
    dxu_vals_paired<float, float>(buf + __umul24(k.a, (unsigned)kS2RowPad) + zo,
                                  buf + __umul24(12u + r, (unsigned)kS2RowPad) + zo, re, im);
    s2_put(k.zi, r, k.a, pow2f(se), re, im, ah, al);
  }
}

// pair tasks P = p0 + 64 i + lane (i < N): S row a read once for both r
template <int N>
__device__ inline void s2_dxu_pairs(int p0, int lane, const float* buf, const unsigned* smax, unsigned* ah,
                                    unsigned* al, int* sse) {
#pragma unroll 1
  for (int i = 0; i < N; ++i) {
    const S2Task k = s2_pair_task((unsigned)(p0 + 64 * i + lane));
    const unsigned r = 2u * k.rp;
    const int se = site_exp(smax[k.zi], smax[kS2Sites + k.zi]);
    if (k.rp == 0 && k.a == 0) sse[k.zi] = se;
    const unsigned zo = __umul24(k.zi, 24u);
    float sa[24];
    const float4* s4 = reinterpret_cast<const float4*>(buf + __umul24(k.a, (unsigned)kS2RowPad) + zo);
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const float4 v = s4[j];
      sa[4 * j] = v.x;
      sa[4 * j + 1] = v.y;
      sa[4 * j + 2] = v.z;
      sa[4 * j + 3] = v.w;
    }
    const float sc = pow2f(se);
    const float* y0 = buf + __umul24(12u + r, (unsigned)kS2RowPad) + zo;
    float re[4], im[4];
    dxu_vals_paired<float, float>(sa, y0, re, im);
    s2_put(k.zi, r, k.a, sc, re, im, ah, al);
    dxu_vals_paired<float, float>(sa, y0 + kS2RowPad, re, im);
    s2_put(k.zi, r + 1, k.a, sc, re, im, ah, al);
  }
}

// Producer roles. Waves whose indices agree mod 4 share a SIMD (the hardware id
// register shows the SIMD numbers as a permutation of wave % 4), and on a SIMD the
// matrix and vector work of its waves add up (they hardly co-issue), so the dxu tasks
// are spread to even out MFMA + vector time per SIMD: the SIMDs of waves = 0 and 1
// (mod 4) carry three MFMA waves and producer 12 or 13 with 4 tasks per lane, as two
// rounds of pair tasks (a round of 64 pair tasks is 128 tasks; a pair reads S once for
// two tasks, fewer LDS round trips on the SIMDs that set the step: pass 2 2.6% faster); those of
// waves = 2 and 3 carry two MFMA waves and producers 10 + 14 or 11 + 15 with 3 + 2
// single tasks. Ranges: pw 2:
// pairs [0, 128) (tasks [0, 256)), pw 3: pairs [128, 256), pw 0: tasks [512, 704),
// pw 1: [704, 896), pw 4: [896, 1024), pw 5: [1024, 1152). (Pairs on all six
// producers measured the same: the SIMDs of waves = 2 and 3 do not set the step.)
__device__ inline int s2_task0(int pw) { return pw == 2 ? 0 : pw == 3 ? 128 : pw == 0 ? 512 : pw == 1 ? 704 : pw == 4 ? 896 : 1024; }
__device__ inline int s2_n_tasks(int pw) { return (pw == 2 || pw == 3) ? 4 : (pw == 0 || pw == 1) ? 3 : 2; }
static_assert(2 * 256 + 2 * 192 + 2 * 128 == kS2Tasks, "the producer task ranges cover a group");

__global__ __launch_bounds__(kS2Threads) void pass2_window(
    const WinPair* __restrict__ pairs, int n_pairs, const unsigned short* __restrict__ bsplit,
    const int* __restrict__ bexp, const float* __restrict__ coef, hlbl_geom global,
    size_t sites_per_wg, int nwg, double* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) float sF[2][kS2BufLen];
  __shared__ __attribute__((aligned(16))) unsigned short sA[2][2][kS2ALen];
  __shared__ __attribute__((aligned(16))) float sCoef[2][kS2Sites * kS2CoefSite];
  __shared__ unsigned sMax[3][kS2Keys];  // per group g % 3: spinor maxima per (isy, site)
  __shared__ int sSe[2][kS2Sites];        // per group g % 2: site exponents
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const bool producer = wave >= kS2Mfma;
  const size_t V = geom_volume(global);
  const size_t begin = (size_t)blockIdx.x * sites_per_wg;
  const size_t end = begin + sites_per_wg < V ? begin + sites_per_wg : V;
  if (begin >= end) return;
  const int ng = (int)((end - begin + kS2Sites - 1) / kS2Sites);
  const int total = ng * 2 * n_pairs;  // steps: (unit = (pair, flavour), group)
  if (tid < 3 * kS2Keys) (&sMax[0][0])[tid] = 0u;
  __syncthreads();

  if (producer) {
    const int pw = __builtin_amdgcn_readfirstlane(wave - kS2Mfma), ptid = tid - 64 * kS2Mfma;
    const S2Lane L = s2_lane(ptid, V);
    const int t0 = s2_task0(pw);
    S2Cursor fc, nc;  // the step being fetched, the step being prepared
    d2v r[kS2Fetch];
    auto fetch = [&]() {
      s2_fetch(L, pairs + (fc.u >> 1), fc.u & 1, V, begin + (size_t)fc.g * kS2Sites, r);
      fc.next(ng);
    };
    auto prepare = [&](int it) {  // dxu of step it (= nc)
      unsigned* ah = reinterpret_cast<unsigned*>(sA[it & 1][0]);
      unsigned* al = reinterpret_cast<unsigned*>(sA[it & 1][1]);
      const int nt = s2_n_tasks(pw);
      if (nt == 4) s2_dxu_pairs<2>(t0, lane, sF[it & 1], sMax[it % 3], ah, al, sSe[it & 1]);
      else if (nt == 3) s2_dxu<3>(t0, lane, sF[it & 1], sMax[it % 3], ah, al, sSe[it & 1]);
      else s2_dxu<2>(t0, lane, sF[it & 1], sMax[it % 3], ah, al, sSe[it & 1]);
      nc.next(ng);
    };
    // prologue: spinors of steps 0 and 1 staged; the dxu of step 0 built after the
    // barrier, then step 2's fetch issued (it flies through the loop's first barrier,
    // as every later fetch does)
    fetch();
    s2_store(ptid, L, r, sF[0], sMax[0]);
    if (total > 1) {
      fetch();
      s2_store(ptid, L, r, sF[1], sMax[1]);
    }
    barrier_all();
    prepare(0);
    bool inflight = total > 2;
    if (inflight) fetch();
    int m3 = 0;  // it % 3
    for (int it = 0; it < total; ++it) {
      // lets the fetch (the last 6 loads) fly
      if (inflight) barrier_vm6(); else barrier_all();
      // maxima of group it were read by its dxu (last step): clear them for group it+3
      if (ptid < kS2Keys) sMax[m3][ptid] = 0u;
      const int m3n = m3 == 2 ? 0 : m3 + 1, m3nn = m3n == 2 ? 0 : m3n + 1;
      if (it + 2 < total) s2_store(ptid, L, r, sF[it & 1], sMax[m3nn]);  // fetched last step
      if (it + 1 < total) prepare(it + 1);
      inflight = it + 3 < total;
      if (inflight) fetch();
      m3 = m3n;
    }
    return;
  }

  // MFMA waves: column block cb = wave / 2, K half kh = wave % 2 (chunks 9 kh ..
  // 9 kh + 8). A fragment of lane l: row (zi, mu) = l%32, k = 16 c + 8 (l/32) + e;
  // FP32 32x32 result: element e of lane l is row 8 (e/4) + 4 (l/32) + e%4 =
  // (zi = 2 (e/4) + l/32, mu = e%4), column g = 32 cb + l%32.
  const int cb = wave >> 1, kh = wave & 1, lrow = lane & 31, lh = lane >> 5;
  h8v bh[kS2HalfChunks], bl[kS2HalfChunks];
  int te = 0;
  Kahan acc;
  S2Cursor cur;
  // coefficients: MFMA wave w < 8 stages site w of a step with direct loads into LDS,
  // step 0's before the prologue barrier and step it + 1's after step it's products
  // (they land while the wave waits at the barrier, whose vmcnt(0) covers them; the
  // producers then issue no direct loads, which the compiler would await before their
  // LDS writes)
  auto coefs = [&](const S2Cursor& c, int it) {
    if (wave < kS2Sites)
      s2_coef(wave, pairs[c.u >> 1], global, V, begin + (size_t)c.g * kS2Sites, end, coef, sCoef[it & 1]);
  };
  coefs(cur, 0);
  barrier_all();  // the producers' prologue
  for (int it = 0; it < total; ++it, cur.next(ng)) {
    const int u = cur.u;
    const size_t z0 = begin + (size_t)cur.g * kS2Sites;
    if (cur.g == 0) {
      const h8v* bs = reinterpret_cast<const h8v*>(bsplit + (size_t)u * kBsplitLen) +
                      ((size_t)cb * kS2Chunks + kS2HalfChunks * kh) * 128 + lane;
#pragma unroll
      for (int c = 0; c < kS2HalfChunks; ++c) {
        bh[c] = bs[c * 128];
        bl[c] = bs[c * 128 + 64];
      }
      te = bexp[(size_t)u * 160 + 32 * cb + lrow];
      acc = Kahan();
    }
    barrier_all();
    const unsigned short* ah = sA[it & 1][0] + lrow * kS2ARow + 8 * lh + 16 * kS2HalfChunks * kh;
    const unsigned short* al = sA[it & 1][1] + lrow * kS2ARow + 8 * lh + 16 * kS2HalfChunks * kh;
    f16acc c;
#pragma unroll
    for (int e = 0; e < 16; ++e) c[e] = 0.f;
#pragma unroll
    for (int i = 0; i < kS2HalfChunks; ++i) {
      const h8v xh = *reinterpret_cast<const h8v*>(ah + 16 * i);
      const h8v xl = *reinterpret_cast<const h8v*>(al + 16 * i);
      c = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh, bh[i], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh, bl[i], c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_32x32x16_f16(xl, bh[i], c, 0, 0, 0);
    }
    // coefficient (mu, lam)(col) of column g = lam*40 + col: index mu*160 + g; the
    // four sites zi = 2q + l/32 of this lane unscaled by 2^-(se + te) each
    const float* cs = sCoef[it & 1] + lh * kS2CoefSite + 32 * cb + lrow;
    // sites past the end (only in a workgroup's last group): stale coefficients
    // (in bounds), selected away
    const int nlive = z0 + kS2Sites <= end ? kS2Sites : (int)(end - z0);
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int zi = 2 * q + lh;
      float vq = 0.f;
#pragma unroll
      for (int m = 0; m < 4; ++m) vq += c[4 * q + m] * cs[2 * q * kS2CoefSite + m * 160];
      const double d = (double)ldexpf(vq, -(sSe[it & 1][zi] + te));
      v += zi < nlive ? d : 0.0;
    }
    acc.add(v);
    if (cur.g == ng - 1) {
      double s = acc.s - acc.c;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) part[((size_t)u * nwg + blockIdx.x) * kS2Mfma + wave] = s;
    }
    if (it + 1 < total) {
      S2Cursor nx = cur;
      nx.next(ng);
      coefs(nx, it + 1);
    }
  }
}

// out[pair][f] = Kahan sum over workgroups and MFMA waves of the unit's partials
__global__ void reduce_window2(const double* __restrict__ part, int nwg, int n_pairs,
                               double* __restrict__ out) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;  // (pair, f)
  if (u >= 2 * n_pairs) return;
  Kahan k;
  for (int w = 0; w < nwg; ++w)
    for (int j = 0; j < kS2Mfma; ++j) k.add(part[((size_t)u * nwg + w) * kS2Mfma + j]);
  out[u] = k.s;
}

// ------------------------------------------------------------- launchers --

void window_stage(hlbl_contractor* c, WindowPlan& w, int stage, hipStream_t st) {
  const hlbl_geom g = c->global;
  const size_t V = c->volume;
  const int np = w.n_pairs, nu = w.n_unique;
  // pass 1
  if (stage == 0) {
    launch_pass1_planes<float>(w.d_pairs.p, WinPair{}, np, g, g, c->proc, 3, w.part1.p, w.red1.p,
                               w.xi.p, st, w.d_order.p);
    launch_forms_batched(w.xi.p, w.dz.p, w.gdz.p, np, st);
    hipLaunchKernelGGL(bsplit_exp, dim3(1, np), dim3(128), 0, st, w.gdz.p, w.bexp.p);
    hipLaunchKernelGGL(form_bsplit, dim3((unsigned)div_up(kBsplitLen, 256), np), dim3(256), 0, st, w.gdz.p,
                       w.bexp.p, w.bsplit.p);
    check_hip(hipGetLastError(), "form_bsplit launch");
  }
  // the coefficient fields to build, one per listed yv
  if (stage == 1 && nu > 0) {
    hipLaunchKernelGGL(kfield_consts_win, dim3(nu), dim3(64), 0, st, c->qed->table64.p, w.d_yvs.p, g,
                       w.xunit, w.cst64.p, w.cst32.p);
    check_hip(hipGetLastError(), "kfield_consts_win launch");
    const hlbl_tuning& tu = c->tune;
    const int nwa = (int)std::min<size_t>(div_up(V, kKmThreads / 64), (size_t)tu.axes_wg_per_cu * c->n_cu);
    hipLaunchKernelGGL(kfield_axes, dim3(nwa), dim3(kKmThreads), 0, st, c->qed->table32.p, g, w.xunit,
                       w.form == 2 ? w.chbuf.p : nullptr, w.ax.p);
    check_hip(hipGetLastError(), "kfield_axes launch");
    const int fw = tu.field_waves;
    const int nwg = (int)std::min<size_t>(div_up((size_t)nu * V, (size_t)fw), (size_t)tu.field_wg_per_cu * c->n_cu);
    if (fw == 8)
      hipLaunchKernelGGL(kfield_mixed<8>, dim3(nwg), dim3(64 * 8), 0, st, c->qed->table32.p,
                         w.d_yvs.p, w.d_slots.p, nu, g, w.xunit, w.cst32.p, w.form == 2 ? w.chbuf.p : nullptr,
                         w.ax.p, w.coef);
    else
      hipLaunchKernelGGL(kfield_mixed<4>, dim3(nwg), dim3(64 * 4), 0, st, c->qed->table32.p,
                         w.d_yvs.p, w.d_slots.p, nu, g, w.xunit, w.cst32.p, w.form == 2 ? w.chbuf.p : nullptr,
                         w.ax.p, w.coef);
    check_hip(hipGetLastError(), "kfield_mixed launch");
    // candidate boxes: ceil(kNearRadius / xunit) lattice units, at most the whole extent
    hlbl_coord box;
    const int hb = (int)std::ceil(kNearRadius / w.xunit);
    box.t = std::min(hb, (g.T - 1) / 2);
    box.x = std::min(hb, (g.LX - 1) / 2);
    box.y = std::min(hb, (g.LY - 1) / 2);
    box.z = std::min(hb, (g.LZ - 1) / 2);
    hipLaunchKernelGGL(kfield_near, dim3((unsigned)(tu.near_wg_per_cu * c->n_cu)), dim3(kKmThreads), 0, st,
                       c->qed->table64.p,
                       w.d_yvs.p, w.d_slots.p, nu, g, w.xunit, w.cst64.p, box, w.coef);
    check_hip(hipGetLastError(), "kfield_near launch");
    if (w.form == 2) {
      hipLaunchKernelGGL(judge_coef, dim3(w.n_samples, nu), dim3(64), 0, st, c->qed->table64.p,
                         w.d_yvs.p, g, w.xunit, w.cst64.p, w.d_samples.p, w.n_samples, w.d_slots.p,
                         w.coef, w.judge.p);
      check_hip(hipGetLastError(), "judge_coef launch");
    }
  }
  // pass 2
  if (stage == 2) {
    // the producers address spinor rows in 16-byte units with 32-bit offsets
    if (V * 144 >= ((size_t)1 << 32)) throw Error("window form: lattice too large for one block's offsets");
    const int nwg = c->tune.pass2_wg_per_cu * c->n_cu;
    const size_t spw = div_up(div_up(V, (size_t)nwg), kS2Sites) * kS2Sites;
    const int n = (int)div_up(V, spw);
    hipLaunchKernelGGL(pass2_window, dim3(n), dim3(kS2Threads), 0, st, w.d_pairs.p, np, w.bsplit.p, w.bexp.p,
                       w.coef, g, spw, n, w.part2.p);
    check_hip(hipGetLastError(), "pass2_window launch");
    hipLaunchKernelGGL(reduce_window2, dim3((unsigned)div_up(2 * np, 64)), dim3(64), 0, st,
                       w.part2.p, n, np, w.result.p);
    check_hip(hipGetLastError(), "reduce_window2 launch");
  }
}

size_t window_part2_len(const hlbl_contractor* c, int n_pairs) {
  return (size_t)n_pairs * 2 * c->tune.pass2_wg_per_cu * c->n_cu * kS2Mfma;
}

size_t window_bsplit_len(int n_pairs) { return (size_t)n_pairs * 2 * kBsplitLen; }
size_t window_bexp_len(int n_pairs) { return (size_t)n_pairs * 2 * 160; }


}  // namespace hlbl
