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
// The FP32 form on the matrix cores. Every spin pair p = (s', s) has s' and s in
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
// One workgroup of 12 waves per (pair, flavour, plane): wave = (g, K slice kq).
// The plane is walked in chunks of 8 sites of a line; a chunk's FP64 rows (12 of
// S, 12 of Y, 1,536 bytes each) are staged with direct global-to-LDS loads into
// one of three buffers, two chunks ahead. A chunk is 48 K values = 12 MFMA steps
// of 4 (lane group kb = l / 16 takes k = 12 kb + step); slice kq takes steps
// 2 kq and 2 kq + 1, which are the (re, im) of one complex element: colour
// kq % 3 of site 2 kb + kq / 3. v_mfma_f32_16x16x4f32 operands: A[i][k] and
// B[k][j] in lane i + 16 k (resp. j + 16 k); result element r of lane l is
// (4 (l / 16) + r, l % 16).
namespace {
constexpr int kM1Waves = 12;
constexpr int kM1Threads = 64 * kM1Waves;
constexpr int kM1Sites = 8;                     // sites per chunk
constexpr int kM1Row = kM1Sites * 24 * 8 + 96;  // bytes per row a of a chunk (padded: the
                                                // rows of a 16-lane read fall in distinct banks)
constexpr int kM1Half = 12 * kM1Row;            // the 12 rows of S (or of Y)
constexpr int kM1Buf = 2 * kM1Half;
constexpr int kM1NBuf = 3;                      // chunks staged: one read, two in flight
constexpr int kM1Vals = 72;                     // accumulator values per lane (9 tiles x 4, x 2)
static_assert(6 * kM1Vals * 64 * 4 <= kM1NBuf * kM1Buf, "the end reduction fits the buffers");
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

// waits for this wave's loads of the chunk about to be read, then the workgroup
// barrier; keep: the next chunk is a full one (four loads), which may stay in flight
__device__ inline void m1_barrier(bool keep) {
  static_assert(kM1NBuf == 3, "vmcnt(4): one later chunk in flight");
  if (keep) asm volatile("s_waitcnt vmcnt(4) lgkmcnt(0)\n s_barrier" ::: "memory");
  else asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n s_barrier" ::: "memory");
}

// Rows 2 wave and 2 wave + 1 (row R: R < 12 S[f][R], else Y[1-f][R-12]) of the
// chunk of nz sites at site0 into buf: 12 nz 16-byte units per row, lane l unit l
// and l + 64 (four loads per wave for a full chunk). The units of a short chunk's
// missing sites are zeroed (LDS stores), so every chunk is read as a full one.
__device__ inline void m1_issue(const double* S, const double* Y, size_t V, size_t site0, int nz,
                                int wave, int lane, char* buf) {
  const int units = 12 * nz;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int R = 2 * wave + q, isy = R >= 12, a = R - 12 * isy;
    const char* src = reinterpret_cast<const char*>((isy ? Y : S) + ((size_t)a * V + site0) * 24);
    char* dst = buf + isy * kM1Half + a * kM1Row;
    if (lane < units)
      __builtin_amdgcn_global_load_lds((const void*)(src + 16 * lane),
                                       (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
    if (lane + 64 < units)
      __builtin_amdgcn_global_load_lds((const void*)(src + 16 * (lane + 64)),
                                       (__attribute__((address_space(3))) void*)(dst + 1024), 16,
                                       0, 0);
    if (units < 12 * kM1Sites) {
      for (int u = units + lane; u < 12 * kM1Sites; u += 64)
        *reinterpret_cast<d2v*>(dst + 16 * u) = d2v{0.0, 0.0};
    }
  }
}

// one chunk of one wave: two MFMA steps per tile
__device__ inline void m1_chunk(const char* buf, const int (&offA)[3], const int (&offB)[3],
                                float w3, int r16, f4v (&acc)[3][3]) {
  float a0[3], a1[3], b0[3], b1[3];
#pragma unroll
  for (int rt = 0; rt < 3; ++rt) {
    const d2v x = *reinterpret_cast<const d2v*>(buf + offA[rt]);
    const float re = (float)x.x, im = (float)x.y;
    const bool imrow = rt * 16 + r16 >= 24;
    a0[rt] = imrow ? im : re;
    a1[rt] = imrow ? -re : im;
  }
#pragma unroll
  for (int ct = 0; ct < 3; ++ct) {
    const d2v v = *reinterpret_cast<const d2v*>(buf + offB[ct]);
    const float re = (float)v.x, im = (float)v.y;
    const float m = ct * 16 + r16 >= 24 ? w3 : 1.f;
    b0[ct] = re * m;
    b1[ct] = im * m;
  }
#pragma unroll
  for (int rt = 0; rt < 3; ++rt)
#pragma unroll
    for (int ct = 0; ct < 3; ++ct) {
      acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[rt], b0[ct], acc[rt][ct], 0, 0, 0);
      acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[rt], b1[ct], acc[rt][ct], 0, 0, 0);
    }
}

}  // namespace

__global__ __launch_bounds__(kM1Threads, 1) void pass1_mfma(
    const WinPair* __restrict__ pairs, WinPair single, hlbl_geom global, hlbl_geom local,
    hlbl_coord proc, int flav_mask, const int* __restrict__ units, int unit_slots,
    float* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) char lds[kM1NBuf * kM1Buf];
  int unit = blockIdx.x;
  if (units) {
    unit = units[(blockIdx.x % kPass1Groups) * unit_slots + blockIdx.x / kPass1Groups];
    if (unit < 0) return;
  }
  const int ip = unit >> 1, f = unit & 1;
  if (!((flav_mask >> f) & 1)) return;
  const WinPair wp = pairs ? pairs[ip] : single;
  const int plane = blockIdx.y;
  const int LY = local.LY, LZ = local.LZ;
  const size_t V = geom_volume(local);
  const size_t pbase = (size_t)plane * LY * LZ;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int g = wave >= 6, kq = wave - 6 * g;
  const int kb = lane >> 4, r16 = lane & 15;
  const double* Sf = wp.S + (size_t)f * 12 * V * 24;
  const double* Yg = wp.Y + (size_t)(1 - f) * 12 * V * 24;
  const int oy = proc.y * local.LY - wp.gsx.y, oz = proc.z * local.LZ - wp.gsx.z;

  // this lane's element of the chunk: site 2 kb + kq / 3, colour kq % 3
  const int site = 2 * kb + (kq >= 3), cc = kq % 3;
  int offA[3], offB[3];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const int r = t * 16 + r16, r24 = r >= 24 ? r - 24 : r;
    offA[t] = (r24 >> 1) * kM1Row + site * 192 + ((g ? 0 : 2) + (r24 & 1)) * 48 + cc * 16;
    offB[t] = kM1Half + (r24 >> 1) * kM1Row + site * 192 + ((g ? 2 : 0) + (r24 & 1)) * 48 + cc * 16;
  }

  f4v acc[3][3], py[3][3];
#pragma unroll
  for (int rt = 0; rt < 3; ++rt)
#pragma unroll
    for (int ct = 0; ct < 3; ++ct) acc[rt][ct] = py[rt][ct] = f4v{0.f, 0.f, 0.f, 0.f};

  const int cpl = (LZ + kM1Sites - 1) / kM1Sites, nch = LY * cpl;
  auto chunk_nz = [&](int c) {
    const int z0 = (c % cpl) * kM1Sites;
    return LZ - z0 < kM1Sites ? LZ - z0 : kM1Sites;
  };
  auto issue = [&](int c) {
    const int y = c / cpl, z0 = (c - y * cpl) * kM1Sites;
    m1_issue(Sf, Yg, V, pbase + (size_t)y * LZ + z0, chunk_nz(c), wave, lane,
             lds + (c % kM1NBuf) * kM1Buf);
  };
  auto wy = [&](int y) {
    return y < LY ? (float)coord_map_zerohalf(mod_pos(y + oy, global.LY), global.LY) : 0.f;
  };
  for (int c = 0; c < kM1NBuf - 1 && c < nch; ++c) issue(c);
  int y = 0, zp = 0;  // line and part of chunk c
  for (int c = 0; c < nch; ++c) {
    m1_barrier(c + 1 < nch && chunk_nz(c + 1) == kM1Sites);
    if (c + kM1NBuf - 1 < nch) issue(c + kM1NBuf - 1);
    // xz_3 of this lane's site z0 + site: the chunk's base on the scalar unit
    const int zb = mod_pos(zp * kM1Sites + oz, global.LZ);
    int zs = zb + site;
    zs = global.LZ >= kM1Sites ? (zs >= global.LZ ? zs - global.LZ : zs) : zs % global.LZ;
    const float w3 = (float)coord_map_zerohalf(zs, global.LZ);
    m1_chunk(lds + (c % kM1NBuf) * kM1Buf, offA, offB, w3, r16, acc);
    if (++zp == cpl) {  // end of line y: the xz_2 sums by partial summation
      const float d = wy(y) - wy(y + 1);
#pragma unroll
      for (int rt = 0; rt < 3; ++rt)
#pragma unroll
        for (int ct = 0; ct < 3; ++ct)
#pragma unroll
          for (int i = 0; i < 4; ++i) py[rt][ct][i] = fmaf(d, acc[rt][ct][i], py[rt][ct][i]);
      zp = 0;
      ++y;
    }
  }

  // sum of the six K slices in a fixed order: (3 + 0, 4 + 1, 5 + 2), then 0 + 1 + 2
  float* red = reinterpret_cast<float*>(lds);
  auto put = [&](int slot) {
    float* d = red + (size_t)slot * kM1Vals * 64 + lane;
#pragma unroll
    for (int rt = 0; rt < 3; ++rt)
#pragma unroll
      for (int ct = 0; ct < 3; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          d[((rt * 3 + ct) * 4 + i) * 64] = acc[rt][ct][i];
          d[(36 + (rt * 3 + ct) * 4 + i) * 64] = py[rt][ct][i];
        }
  };
  auto take = [&](int slot) {
    const float* d = red + (size_t)slot * kM1Vals * 64 + lane;
#pragma unroll
    for (int rt = 0; rt < 3; ++rt)
#pragma unroll
      for (int ct = 0; ct < 3; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          acc[rt][ct][i] += d[((rt * 3 + ct) * 4 + i) * 64];
          py[rt][ct][i] += d[(36 + (rt * 3 + ct) * 4 + i) * 64];
        }
  };
  __syncthreads();
  if (kq >= 3) put(g * 3 + kq - 3);
  __syncthreads();
  if (kq < 3) take(g * 3 + kq);
  __syncthreads();
  if (kq == 1 || kq == 2) put(g * 2 + kq - 1);
  __syncthreads();
  if (kq != 0) return;
  take(g * 2);
  take(g * 2 + 1);
  float* out = part + (((size_t)ip * 2 + f) * gridDim.y + plane) * kPlPart;
#pragma unroll
  for (int rt = 0; rt < 3; ++rt)
#pragma unroll
    for (int ct = 0; ct < 3; ++ct)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = rt * 16 + 4 * kb + i, im = r >= 24, r24 = r - 24 * im;
        const int cl = ct * 16 + r16, w = cl >= 24, c24 = cl - 24 * w;
        const int si = r24 & 1, sj = c24 & 1;
        const int p = (si != sj) * 4 + 2 * g + sj;
        const int e = (p * 144 + (r24 >> 1) * 12 + (c24 >> 1)) * 2 + im;
        if (w) {
          out[2 * kXiReal + e] = acc[rt][ct][i];
        } else {
          out[e] = acc[rt][ct][i];
          out[kXiReal + e] = py[rt][ct][i];
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
                         const int* d_units, int unit_slots) {
  if (n_pairs <= 0) return;
  const int np = n_planes(local), nc = n_chunks(local);
  if constexpr (std::is_same<T, float>::value)
    hipLaunchKernelGGL(pass1_mfma, dim3(d_units ? kPass1Groups * unit_slots : 2 * n_pairs, np),
                       dim3(kM1Threads), 0, st, d_pairs, single, global, local, proc, flav_mask,
                       d_units, unit_slots, part);
  else
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
                                          double*, double*, hipStream_t, const int*, int);
template void launch_pass1_planes<float>(const WinPair*, const WinPair&, int, const hlbl_geom&,
                                         const hlbl_geom&, const hlbl_coord&, int, float*, double*,
                                         double*, hipStream_t, const int*, int);

}  // namespace hlbl
