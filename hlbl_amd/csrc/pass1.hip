// Pass 1 of the contraction: the weighted site sums behind dzu and dzsu
// (cuda_lattice.cu:160-262), shared by the FP64 per-pair path (T = double) and
// the window-batched mixed-precision path (T = float: FP32 products and plane
// sums, FP64 Kahan reductions).
//
//   Xi^W[f][p][a][b] = sum_z W(z) sum_colour S_f[a](z)[(s', colour)] conj(Y_{1-f}[b](z)[(s, colour)])
//
// for the five weights W = 1, xz_0 .. xz_3 (coord_map_zerohalf of x - gsx, global
// coordinates) and the eight spin pairs p = cls*4 + s, s' = s^2 (cls 0: gamma_0,
// gamma_3) or 3 - s (cls 1: gamma_1, gamma_2).
//
// One workgroup per (t, x) plane of the local lattice, per flavour and per pair.
// Within a plane xz_0 and xz_1 are constant, so a workgroup keeps three sums per
// output, P = sum X(z), Py = sum xz_2 X(z), Pz = sum xz_3 X(z), and writes them as
// the plane's partial; the reductions apply the plane's xz_0 and xz_1 and sum the
// planes in a fixed order (Kahan, FP64), so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "pass1.hpp"

#include "context.hpp"
#include "lattice.hpp"
#include "pass_common.hpp"

namespace hlbl {

namespace {

constexpr int kPlThreads = 256;
constexpr int kPlSites = 16;              // sites staged per step
constexpr int kXiOut = 8 * 144;           // complex outputs per flavour
constexpr int kXiReal = 2 * kXiOut;       // 2304
constexpr int kPlPart = 3 * kXiReal;      // values per (pair, flavour, plane) partial
constexpr int kPlChunk = 32;              // planes per first-level reduction chunk
constexpr int kXiLen = 2 * 8 * 5 * 144 * 2;

size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }

HLBL_HD int n_planes(const hlbl_geom& local) { return local.T * local.LX; }
int n_chunks(const hlbl_geom& local) { return (int)div_up(n_planes(local), kPlChunk); }

}  // namespace

// Thread (half, p, blk): the 3x3 (a, b) block blk of spin pair p; the two
// halves of the workgroup take alternate sites of each staged step.
template <class T>
__global__ __launch_bounds__(kPlThreads, 2) void pass1_planes(
    const WinPair* __restrict__ pairs, WinPair single, hlbl_geom global, hlbl_geom local,
    hlbl_coord proc, int flav_mask, T* __restrict__ part) {
  __shared__ T sh[kPlSites][24][24];  // [site][12 rows S_f[a], 12 rows Y_{1-f}[b]][24]
  __shared__ T sw[kPlSites][2];       // the staged sites' weights xz_2, xz_3
  // blockIdx.x = pair * 2 + flavour (fastest, so that the pairs of a window that
  // share a block read its plane close together in time), blockIdx.y = plane
  const int ip = blockIdx.x >> 1, f = blockIdx.x & 1;
  if (!((flav_mask >> f) & 1)) return;
  const WinPair wp = pairs ? pairs[ip] : single;
  const int plane = blockIdx.y;
  const int LYZ = local.LY * local.LZ;
  const size_t V = geom_volume(local);
  const size_t base = (size_t)plane * LYZ;
  const int tid = threadIdx.x, half = tid >> 7, t7 = tid & 127;
  const int p = t7 >> 4, blk = t7 & 15;
  const int a0 = 3 * (blk >> 2), b0 = 3 * (blk & 3);
  const int s = p & 3, sp = (p < 4) ? (s ^ 2) : (3 - s);
  const double* Sf = wp.S + (size_t)f * 12 * V * 24;
  const double* Yg = wp.Y + (size_t)(1 - f) * 12 * V * 24;
  const int oy = proc.y * local.LY - wp.gsx.y, oz = proc.z * local.LZ - wp.gsx.z;

  T acc[3][9][2];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[k][i][0] = acc[k][i][1] = T(0);

  for (int z0 = 0; z0 < LYZ; z0 += kPlSites) {
    const int ns = LYZ - z0 < kPlSites ? LYZ - z0 : kPlSites;
    __syncthreads();
    // 24 rows x kPlSites sites x 12 chunks of 16 bytes; a row's sites are contiguous
    for (int i = tid; i < 24 * kPlSites * 12; i += kPlThreads) {
      const int k = i % 12, zi = (i / 12) % kPlSites, row = i / (12 * kPlSites);
      double2 v = {0.0, 0.0};
      if (zi < ns) {
        const double* src = row < 12 ? Sf + ((size_t)row * V + base + z0 + zi) * 24
                                     : Yg + ((size_t)(row - 12) * V + base + z0 + zi) * 24;
        v = reinterpret_cast<const double2*>(src)[k];
      }
      sh[zi][row][2 * k] = (T)v.x;
      sh[zi][row][2 * k + 1] = (T)v.y;
    }
    if (tid < kPlSites) {
      const int yz = z0 + tid;
      const int yl = yz / local.LZ, zl = yz - yl * local.LZ;
      sw[tid][0] = (T)coord_map_zerohalf(mod_pos(yl + oy, global.LY), global.LY);
      sw[tid][1] = (T)coord_map_zerohalf(mod_pos(zl + oz, global.LZ), global.LZ);
    }
    __syncthreads();
    for (int zi = half; zi < ns; zi += 2) {
      const T wy = sw[zi][0], wz = sw[zi][1];
      T sr[3][3], si[3][3], yr[3][3], yi[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int col = 0; col < 3; ++col) {
          sr[i][col] = sh[zi][a0 + i][2 * (3 * sp + col)];
          si[i][col] = sh[zi][a0 + i][2 * (3 * sp + col) + 1];
          yr[i][col] = sh[zi][12 + b0 + i][2 * (3 * s + col)];
          yi[i][col] = sh[zi][12 + b0 + i][2 * (3 * s + col) + 1];
        }
#pragma unroll
      for (int ia = 0; ia < 3; ++ia)
#pragma unroll
        for (int ib = 0; ib < 3; ++ib) {
          // S conj(Y) over the three colours as fused multiply-add chains
          T re = sr[ia][0] * yr[ib][0];
          T im = si[ia][0] * yr[ib][0];
          re = fma(si[ia][0], yi[ib][0], re);
          im = fma(-sr[ia][0], yi[ib][0], im);
#pragma unroll
          for (int col = 1; col < 3; ++col) {
            re = fma(sr[ia][col], yr[ib][col], re);
            re = fma(si[ia][col], yi[ib][col], re);
            im = fma(si[ia][col], yr[ib][col], im);
            im = fma(-sr[ia][col], yi[ib][col], im);
          }
          const int o = ia * 3 + ib;
          acc[0][o][0] += re;
          acc[0][o][1] += im;
          acc[1][o][0] = fma(wy, re, acc[1][o][0]);
          acc[1][o][1] = fma(wy, im, acc[1][o][1]);
          acc[2][o][0] = fma(wz, re, acc[2][o][0]);
          acc[2][o][1] = fma(wz, im, acc[2][o][1]);
        }
    }
  }
  // the halves' sums: half 1 through LDS into half 0, which stores the partial
  __syncthreads();
  T* red = &sh[0][0][0];  // 54 x 128 values
  if (half) {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < 9; ++i) {
        red[((k * 9 + i) * 2 + 0) * 128 + t7] = acc[k][i][0];
        red[((k * 9 + i) * 2 + 1) * 128 + t7] = acc[k][i][1];
      }
  }
  __syncthreads();
  if (!half) {
    T* out = part + (((size_t)ip * 2 + f) * gridDim.y + plane) * kPlPart;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int ia = 0; ia < 3; ++ia)
#pragma unroll
        for (int ib = 0; ib < 3; ++ib) {
          const int i = ia * 3 + ib;
          const int o = k * kXiReal + (p * 144 + (a0 + ia) * 12 + b0 + ib) * 2;
          out[o] = acc[k][i][0] + red[((k * 9 + i) * 2 + 0) * 128 + t7];
          out[o + 1] = acc[k][i][1] + red[((k * 9 + i) * 2 + 1) * 128 + t7];
        }
  }
}

// ---------------------------------------------------------------------------
// The window form on the matrix cores. Every spin pair p = (s', s) has s' and s in
// opposite halves ({0, 1} and {2, 3}), so for spin half g (g = 0: s in {0, 1},
// s' in {2, 3}; g = 1: s in {2, 3}, s' in {0, 1}) the four pairs of the half are
// one real 48 x 48 product, summed over K = (site, colour, re/im):
//   row (ri, a, si), s' = si + 2 (1 - g):  ri = 0: (Re S, Im S),  ri = 1: (Im S, -Re S)
//   column (w, b, sj), s = sj + 2 g:      (Re Y, Im Y) times 1 (w = 0) or xz_3(z) (w = 1)
// so that element (ri, a, si; w, b, sj) is part ri of sum_z w(z) S[a][(s', c)] conj(Y[b][(s, c)]).
// xz_2 is constant along a (t, x, y) line: with the running sum R_y after line y,
//   sum_y xz_2(y) L_y = sum_y (xz_2(y) - xz_2(y+1)) R_y,   xz_2(LY) = 0,
// one multiply-add per accumulator at the end of each line.
//
// Precision: S and Y are rounded to FP32, scaled by a power of two per chunk and
// operand so that the chunk's largest magnitude stays below 2^14, split into FP16
// hi + lo and multiplied as hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_f16 (FP32
// accumulation; about 21 significant bits per product); each chunk's sums are
// unscaled exactly into FP32 running sums.
//
// One workgroup of 12 waves per (pair, flavour, plane), walking the plane in chunks
// of 8 sites of a line. K of a chunk: site 4 ks + kb (MFMA step ks = 0, 1; lane group
// kb = l / 16) supplies the lane's 8 FP16 values, (Re, Im) of colours 0..2 and two
// zeros. MFMA waves 0..3 = (g, ks) hold all nine 16 x 16 tiles of their half and
// step. A chunk passes three stages, one per iteration: loader waves 8..11 load it
// into registers (two chunks ahead) and stage it in LDS as FP32 with its largest |S|
// and |Y| (LDS atomics); splitter waves 4..7 write the scaled hi/lo operands in MFMA
// fragment order; the MFMA waves multiply. v_mfma_f32_16x16x32_f16: A[i][k] is element k % 8 of lane
// i + 16 (k / 8), B[k][j] of lane j + 16 (k / 8); result r of lane l is
// (4 (l / 16) + r, l % 16).
//
// Placement: the planes are dealt to the XCDs in turn, and an XCD runs every unit of
// its plane, in the host's order (lists of units that share blocks, one after another).
// Each block's plane is then fetched by one XCD only, and the units running together
// on it (one per CU) read their shared blocks from its L2.
namespace {

constexpr unsigned kXcds = 8;  // MI355X: 8 XCDs of 32 CUs
constexpr int kH1Mfma = 4;
constexpr int kH1Threads = 64 * (kH1Mfma + 8);  // + 4 splitters + 4 loaders
constexpr int kH1Sites = 8;                       // sites per chunk
constexpr int kH1Units = 2 * 12 * kH1Sites * 12;  // 16-byte units of a chunk (S and Y rows)
constexpr int kH1Fetch = kH1Units / 256;          // per producer lane
static_assert(kH1Fetch * 256 == kH1Units, "the producers cover a chunk");
constexpr int kH1Row = kH1Sites * 24 + 4;         // FP32 staging floats per row a (padded)
constexpr int kH1F = 2 * 12 * kH1Row;             // [isy][a][site][24]
constexpr int kH1Frag = 64 * 8;                   // FP16 of one fragment (a ds_read_b128 per lane)
constexpr int kH1H = 2 * 2 * 2 * 3 * 2 * kH1Frag;  // [A, B][g][ks][tile][hi, lo][lane][8]
constexpr int kH1Items = 2 * 2 * 2 * 4 * 12 * 2;  // (A, B) x g x ks x kb x row a (b) x si (sj)
static_assert(kH1Items == 3 * 256, "three split items per producer lane");
constexpr int kH1Vals = 72;                       // running sums per MFMA lane (9 tiles x 4, x 2)
constexpr int kH1Max = 14;                        // scaled magnitudes stay below 2^kH1Max
static_assert(2 * kH1Vals * 64 * 4 <= 2 * kH1H * 2, "the end reduction fits the fragments");
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(1))) const d2v gd2v;

// a 64-bit value the wave holds in common, as a scalar (both halves zero-extended)
__device__ inline size_t uniform_u64(size_t v) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v & 0xffffffffu));
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return ((size_t)hi << 32) | (size_t)lo;
}

// LDS barrier that leaves global loads (the producers' register prefetch) in flight
__device__ inline void h1_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n s_barrier" ::: "memory"); }

__device__ inline int h1_frag(int op, int g, int ks, int tile, int hl) {
  return ((((op * 2 + g) * 2 + ks) * 3 + tile) * 2 + hl) * kH1Frag;
}

// exponent e with max < 2^e, from the bits of a non-negative float maximum
__device__ inline int h1_exp(unsigned m) { return (int)((m >> 23) & 0xffu) - 126; }
// scale exponents within [-63, 63]: the unscale 2^-(tS + tY) is a normal float (the
// upscale of a chunk whose largest magnitude is below 2^-49 stops at 2^63)
__device__ inline int h1_clamp(int t) { return t < -63 ? -63 : (t > 63 ? 63 : t); }

// maximum over each group of 8 lanes (all lanes active), by DPP
__device__ inline unsigned h1_max8(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false));
  return v;
}

// maximum over the wave (all lanes active), as a wave-uniform value
__device__ inline unsigned h1_max64(unsigned v) {
  v = h1_max8(v);
  unsigned m = (unsigned)__builtin_amdgcn_readlane((int)v, 0);
#pragma unroll
  for (int gidx = 1; gidx < 8; ++gidx) m = max(m, (unsigned)__builtin_amdgcn_readlane((int)v, 8 * gidx));
  return m;
}

// (x0..x5) * sc as FP16 hi and lo (elements 0..5 of a lane's eight; 6 and 7 stay zero)
__device__ inline void h1_split6(const float (&x)[6], float sc, uint3& hi, uint3& lo) {
  f16_split2(x[0] * sc, x[1] * sc, hi.x, lo.x);
  f16_split2(x[2] * sc, x[3] * sc, hi.y, lo.y);
  f16_split2(x[4] * sc, x[5] * sc, hi.z, lo.z);
}
// the im row of packed (Re, Im) pairs: (Im, -Re)
__device__ inline unsigned h1_imrow(unsigned d) { return __builtin_amdgcn_alignbit(d, d, 16) ^ 0x80000000u; }

}  // namespace

__global__ __launch_bounds__(kH1Threads, 1) void pass1_mfma(
    const WinPair* __restrict__ pairs, WinPair single, hlbl_geom global, hlbl_geom local,
    hlbl_coord proc, int flav_mask, const int* __restrict__ order, int n_units, int n_planes,
    float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) float sF[2][kH1F];
  __shared__ __attribute__((aligned(16))) _Float16 sH[2][kH1H];
  __shared__ unsigned sMax[4][2];  // per chunk c % 4: largest |S|, |Y| (float bits)
  // workgroup j of XCD k (workgroups are dealt to the XCDs in turn): plane
  // kXcds (j / n_units) + k, unit order[j % n_units]
  const unsigned xcd = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
  const int plane = kXcds * (int)(j / (unsigned)n_units) + (int)xcd;
  if (plane >= n_planes) return;
  const int unit = order ? order[j % (unsigned)n_units] : (int)(j % (unsigned)n_units);
  const int ip = unit >> 1, f = unit & 1;
  if (!((flav_mask >> f) & 1)) return;
  const WinPair wp = pairs ? pairs[ip] : single;
  const int LY = local.LY, LZ = local.LZ;
  const size_t V = geom_volume(local);
  const size_t pbase = (size_t)plane * LY * LZ;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int oy = proc.y * local.LY - wp.gsx.y, oz = proc.z * local.LZ - wp.gsx.z;
  const int cpl = (LZ + kH1Sites - 1) / kH1Sites, nch = LY * cpl;
  const int nit = (nch + 3) & ~1;  // iterations -2 .. nit - 3 (an even count, one barrier each)
  auto chunk_z0 = [&](int c) { return (c % cpl) * kH1Sites; };
  auto chunk_site0 = [&](int c) { return pbase + (size_t)(c / cpl) * LZ + chunk_z0(c); };
  auto chunk_nz = [&](int c) { return LZ - chunk_z0(c) < kH1Sites ? LZ - chunk_z0(c) : kH1Sites; };
  if (threadIdx.x < 8) (&sMax[0][0])[threadIdx.x] = 0u;
  // elements 6 and 7 of every lane's operand are the zero padding of K: written once
  for (int i = threadIdx.x; i < 2 * kH1H / 8; i += kH1Threads)
    *reinterpret_cast<unsigned*>(&sH[0][0] + 8 * i + 6) = 0u;  // halves 6 and 7

  // xz_3 headroom: |xz_3| <= LZ / 2 < 2^hz
  int hz = 0;
  while ((1 << hz) <= global.LZ / 2) ++hz;

  if (wave >= kH1Mfma + 4) {
    // ---- loaders (waves 8..11): iteration it stages chunk it + 2 (FP32 and its
    // maxima), clears the maxima slot of chunk it + 3 and loads chunk it + 4 into the
    // registers chunk it + 2 came from. Always kH1Fetch loads per lane (a chunk past
    // the end loads the last one again, a site past a short chunk's end its first
    // site; staging zeroes those), so the later chunk's loads stay in flight.
    const int ql = threadIdx.x - 64 * (kH1Mfma + 4);
    const size_t pS = (size_t)(wp.S + (size_t)f * 12 * V * 24);
    const size_t pY = (size_t)(wp.Y + (size_t)(1 - f) * 12 * V * 24);
    // unit k of this lane: u = ql + 256 k = ((isy * 12 + a) * 8 + site) * 12 + j;
    // uoff[k] the address of its row and 16 bytes at site 0 (a site adds 192 bytes)
    size_t uoff[kH1Fetch];
    int usite[kH1Fetch], uF[kH1Fetch];
    bool uisy = false;  // unit 4 is a row of Y
#pragma unroll
    for (int k = 0; k < kH1Fetch; ++k) {
      const int u = ql + 256 * k, isy = u >= kH1Units / 2, r = u - isy * (kH1Units / 2);
      const int a = r / (kH1Sites * 12), site = (r / 12) % kH1Sites, j = r % 12;
      uoff[k] = (isy ? pY : pS) + ((size_t)a * V * 24 + 2 * j) * sizeof(double);
      usite[k] = site;
      uF[k] = (isy * 12 + a) * kH1Row + site * 24 + 2 * j;
      if (k == 4) uisy = isy;
    }
    // one code path for every chunk (a site past a short chunk's end reads the chunk's
    // first site), so that the compiler's load counting lets the later chunk fly
    auto issue_any = [&](int c, d2v (&r)[kH1Fetch]) {
      c = c < nch ? c : nch - 1;
      const unsigned s0 = (unsigned)chunk_site0(c);  // the volume is below 2^32
      const int nz = chunk_nz(c);
#pragma unroll
      for (int k = 0; k < kH1Fetch; ++k) {
        const unsigned z = s0 + (unsigned)(usite[k] < nz ? usite[k] : 0);
        r[k] = *reinterpret_cast<gd2v*>(uoff[k] + (size_t)z * (24 * sizeof(double)));
      }
    };
    // units k < 4 are rows of S, k > 4 rows of Y, k = 4 either (u = ql + 1024)
    static_assert(4 * 256 <= kH1Units / 2 && 5 * 256 >= kH1Units / 2, "one mixed unit per lane");
    auto stage = [&](auto kFull, int c, const d2v (&r)[kH1Fetch]) {
      float* F = sF[c & 1];
      const int nz = chunk_nz(c);
      unsigned mS = 0u, mY = 0u;
#pragma unroll
      for (int k = 0; k < kH1Fetch; ++k) {
        float2 v;
        v.x = (float)r[k].x;
        v.y = (float)r[k].y;
        if (!decltype(kFull)::value && usite[k] >= nz) v = make_float2(0.f, 0.f);
        const unsigned m = max(__float_as_uint(fabsf(v.x)), __float_as_uint(fabsf(v.y)));
        if (k < 4) mS = max(mS, m);
        else if (k > 4) mY = max(mY, m);
        else if (uisy) mY = max(mY, m);
        else mS = max(mS, m);
        *reinterpret_cast<float2*>(F + uF[k]) = v;
      }
      mS = h1_max64(mS);
      mY = h1_max64(mY);
      if (lane == 0) {
        atomicMax(&sMax[c & 3][0], mS);
        atomicMax(&sMax[c & 3][1], mY);
      }
    };
    auto step = [&](int it, d2v (&r)[kH1Fetch]) {
      if (it + 2 < nch) {
        if (chunk_nz(it + 2) == kH1Sites) stage(std::true_type{}, it + 2, r);
        else stage(std::false_type{}, it + 2, r);
      }
      if (ql < 2) sMax[(it + 7) & 3][ql] = 0u;
      issue_any(it + 4, r);
      h1_barrier();
    };
    d2v r0[kH1Fetch], r1[kH1Fetch];
    issue_any(0, r0);
    issue_any(1, r1);
    h1_barrier();
    for (int it = -2; it < nit - 2; it += 2) {
      step(it, r0);
      step(it + 1, r1);
    }
    return;
  }

  if (wave >= kH1Mfma) {
    // ---- splitters (waves 4..7): iteration it writes the scaled hi/lo fragments of
    // chunk it + 1. Item i = pl + 256 k: op = A (S rows) or B (Y columns), then
    // ((g * 2 + ks) * 4 + kb) * 12 + a, * 2 + si; an A item writes its re and im rows,
    // a B item its w = 0 and w = 1 columns.
    const int pl = threadIdx.x - 64 * kH1Mfma;
    int iF[3], iH0[3], iH1[3], isite[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int i = pl + 256 * k, op = i >= kH1Items / 2, jj = i - op * (kH1Items / 2);
      const int si = jj & 1, a = (jj >> 1) % 12, kb = (jj / 24) & 3, gks = jj / 96;
      const int g = gks >> 1, ks = gks & 1, site = 4 * ks + kb;
      const int spin = op ? (g ? 2 : 0) + si : (g ? 0 : 2) + si;
      iF[k] = (op * 12 + a) * kH1Row + site * 24 + spin * 6;
      const int r0 = 2 * a + si, r1 = r0 + 24;
      iH0[k] = h1_frag(op, g, ks, r0 >> 4, 0) + 8 * ((r0 & 15) + 16 * kb);
      iH1[k] = h1_frag(op, g, ks, r1 >> 4, 0) + 8 * ((r1 & 15) + 16 * kb);
      isite[k] = site;
    }
    auto split = [&](int c) {
      const float* F = sF[c & 1];
      _Float16* H = sH[c & 1];
      const int tS = h1_clamp(kH1Max - h1_exp(sMax[c & 3][0]));
      const int tY = h1_clamp(kH1Max - h1_exp(sMax[c & 3][1]) - hz);
      const float scS = ldexpf(1.f, tS), scY = ldexpf(1.f, tY);
      const int zb = mod_pos(chunk_z0(c) + oz, global.LZ);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const bool opB = pl + 256 * k >= kH1Items / 2;  // uniform per wave and k
        float x[6];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          const float2 v = *reinterpret_cast<const float2*>(F + iF[k] + 2 * e);
          x[2 * e] = v.x;
          x[2 * e + 1] = v.y;
        }
        uint3 hi, lo;
        h1_split6(x, opB ? scY : scS, hi, lo);
        *reinterpret_cast<uint3*>(H + iH0[k]) = hi;
        *reinterpret_cast<uint3*>(H + iH0[k] + kH1Frag) = lo;
        if (!opB) {
          hi = uint3{h1_imrow(hi.x), h1_imrow(hi.y), h1_imrow(hi.z)};
          lo = uint3{h1_imrow(lo.x), h1_imrow(lo.y), h1_imrow(lo.z)};
        } else {
          int zs = zb + isite[k];
          zs = zs >= global.LZ ? zs - global.LZ : zs;
          zs = zs >= global.LZ ? zs % global.LZ : zs;
          h1_split6(x, scY * (float)coord_map_zerohalf(zs, global.LZ), hi, lo);
        }
        *reinterpret_cast<uint3*>(H + iH1[k]) = hi;
        *reinterpret_cast<uint3*>(H + iH1[k] + kH1Frag) = lo;
      }
    };
    h1_barrier();
    for (int it = -2; it < nit - 2; ++it) {
      if (it + 1 >= 0 && it + 1 < nch) split(it + 1);
      h1_barrier();
    }
    return;
  }

  // ---- MFMA waves: (g, ks)
  const int g = wave >> 1, ks = wave & 1;
  f4v R[3][3], py[3][3];
#pragma unroll
  for (int rt = 0; rt < 3; ++rt)
#pragma unroll
    for (int ct = 0; ct < 3; ++ct) R[rt][ct] = py[rt][ct] = f4v{0.f, 0.f, 0.f, 0.f};
  auto wy = [&](int y) {
    return y < LY ? (float)coord_map_zerohalf(mod_pos(y + oy, global.LY), global.LY) : 0.f;
  };
  h1_barrier();
  for (int it = -2; it < nit - 2; ++it) {
    if (it >= 0 && it < nch) {
      const _Float16* H = sH[it & 1];
      h8v A[3][2], B[3][2];
#pragma unroll
      for (int t = 0; t < 3; ++t)
#pragma unroll
        for (int hl = 0; hl < 2; ++hl) {
          A[t][hl] = *reinterpret_cast<const h8v*>(H + h1_frag(0, g, ks, t, hl) + 8 * lane);
          B[t][hl] = *reinterpret_cast<const h8v*>(H + h1_frag(1, g, ks, t, hl) + 8 * lane);
        }
      const int tS = h1_clamp(kH1Max - h1_exp(sMax[it & 3][0]));
      const int tY = h1_clamp(kH1Max - h1_exp(sMax[it & 3][1]) - hz);
      const float u = ldexpf(1.f, -(tS + tY));
      f4v c[3][3];
#pragma unroll
      for (int rt = 0; rt < 3; ++rt)
#pragma unroll
        for (int ct = 0; ct < 3; ++ct) {
          c[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[rt][0], B[ct][0], f4v{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
          c[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[rt][0], B[ct][1], c[rt][ct], 0, 0, 0);
          c[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[rt][1], B[ct][0], c[rt][ct], 0, 0, 0);
        }
#pragma unroll
      for (int rt = 0; rt < 3; ++rt)
#pragma unroll
        for (int ct = 0; ct < 3; ++ct)
#pragma unroll
          for (int i = 0; i < 4; ++i) R[rt][ct][i] = fmaf(c[rt][ct][i], u, R[rt][ct][i]);
      if ((it + 1) % cpl == 0) {  // end of line y: the xz_2 sums by partial summation
        const int y = it / cpl;
        const float d = wy(y) - wy(y + 1);
#pragma unroll
        for (int rt = 0; rt < 3; ++rt)
#pragma unroll
          for (int ct = 0; ct < 3; ++ct)
#pragma unroll
            for (int i = 0; i < 4; ++i) py[rt][ct][i] = fmaf(d, R[rt][ct][i], py[rt][ct][i]);
      }
    }
    h1_barrier();
  }

  // the two K steps' sums (the producers have left): step 1 through LDS into step 0
  float* red = reinterpret_cast<float*>(&sH[0][0]) + (size_t)g * kH1Vals * 64 + lane;
  if (ks) {
#pragma unroll
    for (int rt = 0; rt < 3; ++rt)
#pragma unroll
      for (int ct = 0; ct < 3; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          red[((rt * 3 + ct) * 4 + i) * 64] = R[rt][ct][i];
          red[(36 + (rt * 3 + ct) * 4 + i) * 64] = py[rt][ct][i];
        }
  }
  __syncthreads();
  if (ks) return;
  float* out = part + (((size_t)ip * 2 + f) * n_planes + plane) * kPlPart;
  const int kb = lane >> 4, r16 = lane & 15;
#pragma unroll
  for (int rt = 0; rt < 3; ++rt)
#pragma unroll
    for (int ct = 0; ct < 3; ++ct)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float vR = R[rt][ct][i] + red[((rt * 3 + ct) * 4 + i) * 64];
        const float vP = py[rt][ct][i] + red[(36 + (rt * 3 + ct) * 4 + i) * 64];
        const int r = rt * 16 + 4 * kb + i, im = r >= 24, r24 = r - 24 * im;
        const int cl = ct * 16 + r16, w = cl >= 24, c24 = cl - 24 * w;
        const int si = r24 & 1, sj = c24 & 1;
        const int p = (si != sj) * 4 + 2 * g + sj;
        const int e = (p * 144 + (r24 >> 1) * 12 + (c24 >> 1)) * 2 + im;
        if (w) {
          out[2 * kXiReal + e] = vR;
        } else {
          out[e] = vR;
          out[kXiReal + e] = vP;
        }
      }
}

// First level: for chunk c of kPlChunk planes, red1[pair][f][c][w][e] = the
// five weighted sums (W = 1, xz_0, xz_1, xz_2, xz_3) over the chunk's planes in
// plane order, e = (p*144 + a*12 + b)*2 + re/im. A flavour outside flav_mask
// gets zeros.
template <class T>
__global__ void reduce_planes1(const T* __restrict__ part, const WinPair* __restrict__ pairs,
                               WinPair single, hlbl_geom global, hlbl_geom local, hlbl_coord proc,
                               int flav_mask, double* __restrict__ red1) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kXiReal) return;
  const int chunk = blockIdx.y, pf = blockIdx.z, f = pf & 1;
  const int np = n_planes(local), nc = gridDim.y;
  double* out = red1 + ((size_t)pf * nc + chunk) * 5 * kXiReal + e;
  if (!((flav_mask >> f) & 1)) {
    for (int w = 0; w < 5; ++w) out[w * kXiReal] = 0.0;
    return;
  }
  const WinPair wp = pairs ? pairs[pf >> 1] : single;
  const int ot = proc.t * local.T - wp.gsx.t, ox = proc.x * local.LX - wp.gsx.x;
  Kahan k1, kt, kx, ky, kz;
  const int pl_end = (chunk + 1) * kPlChunk < np ? (chunk + 1) * kPlChunk : np;
  for (int pl = chunk * kPlChunk; pl < pl_end; ++pl) {
    const T* src = part + ((size_t)pf * np + pl) * kPlPart + e;
    const int tl = pl / local.LX, xl = pl - tl * local.LX;
    const double wt = coord_map_zerohalf(mod_pos(tl + ot, global.T), global.T);
    const double wx = coord_map_zerohalf(mod_pos(xl + ox, global.LX), global.LX);
    const double v0 = (double)src[0];
    k1.add(v0);
    kt.add(wt * v0);
    kx.add(wx * v0);
    ky.add((double)src[kXiReal]);
    kz.add((double)src[2 * kXiReal]);
  }
  out[0] = k1.s;
  out[kXiReal] = kt.s;
  out[2 * kXiReal] = kx.s;
  out[3 * kXiReal] = ky.s;
  out[4 * kXiReal] = kz.s;
}

// Second level: xi[pair][f][p][w][a*12+b][re/im] = sum of the chunks in order.
__global__ void reduce_planes2(const double* __restrict__ red1, int nc, double* __restrict__ xi) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;  // w * kXiReal + e
  if (j >= 5 * kXiReal) return;
  const int pf = blockIdx.z;
  Kahan k;
  for (int c = 0; c < nc; ++c) k.add(red1[((size_t)pf * nc + c) * 5 * kXiReal + j]);
  const int w = j / kXiReal, e = j - w * kXiReal;
  const int ri = e & 1, pab = e >> 1, p = pab / 144, ab = pab - p * 144;
  xi[(size_t)(pf >> 1) * kXiLen + ((((pf & 1) * 8 + p) * 5 + w) * 144 + ab) * 2 + ri] = k.s;
}

size_t pass1_part_len(const hlbl_geom& local, int n_pairs) {
  return (size_t)n_pairs * 2 * n_planes(local) * kPlPart;
}

size_t pass1_red_len(const hlbl_geom& local, int n_pairs) {
  return (size_t)n_pairs * 2 * n_chunks(local) * 5 * kXiReal;
}

template <class T>
void launch_pass1_planes(const WinPair* d_pairs, const WinPair& single, int n_pairs,
                         const hlbl_geom& global, const hlbl_geom& local, const hlbl_coord& proc,
                         int flav_mask, T* part, double* red1, double* xi, hipStream_t st,
                         const int* d_order) {
  if (n_pairs <= 0) return;
  const int np = n_planes(local), nc = n_chunks(local);
  if constexpr (std::is_same<T, float>::value) {
    // every XCD gets every unit of its planes: kXcds * ceil(np / kXcds) * units workgroups
    const size_t nwg = (size_t)kXcds * div_up(np, kXcds) * 2 * n_pairs;
    if (nwg > 0x7fffffffu) throw Error("pass 1: too many workgroups for one launch");
    hipLaunchKernelGGL(pass1_mfma, dim3((unsigned)nwg), dim3(kH1Threads), 0, st, d_pairs, single, global,
                       local, proc, flav_mask, d_order, 2 * n_pairs, np, part);
  } else
    hipLaunchKernelGGL(pass1_planes<T>, dim3(2 * n_pairs, np), dim3(kPlThreads), 0, st, d_pairs,
                       single, global, local, proc, flav_mask, part);
  check_hip(hipGetLastError(), "pass1 plane sums launch");
  hipLaunchKernelGGL(reduce_planes1<T>, dim3((unsigned)div_up(kXiReal, 256), nc, 2 * n_pairs),
                     dim3(256), 0, st, part, d_pairs, single, global, local, proc, flav_mask, red1);
  check_hip(hipGetLastError(), "reduce_planes1 launch");
  hipLaunchKernelGGL(reduce_planes2, dim3((unsigned)div_up(5 * kXiReal, 256), 1, 2 * n_pairs),
                     dim3(256), 0, st, red1, nc, xi);
  check_hip(hipGetLastError(), "reduce_planes2 launch");
}

template void launch_pass1_planes<double>(const WinPair*, const WinPair&, int, const hlbl_geom&,
                                          const hlbl_geom&, const hlbl_coord&, int, double*,
                                          double*, double*, hipStream_t, const int*);
template void launch_pass1_planes<float>(const WinPair*, const WinPair&, int, const hlbl_geom&,
                                         const hlbl_geom&, const hlbl_coord&, int, float*, double*,
                                         double*, hipStream_t, const int*);

}  // namespace hlbl
