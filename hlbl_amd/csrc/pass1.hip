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
                         int flav_mask, T* part, double* red1, double* xi, hipStream_t st) {
  if (n_pairs <= 0) return;
  const int np = n_planes(local), nc = n_chunks(local);
  hipLaunchKernelGGL(pass1_planes<T>, dim3(2 * n_pairs, np), dim3(kPlThreads), 0, st, d_pairs,
                     single, global, local, proc, flav_mask, part);
  check_hip(hipGetLastError(), "pass1_planes launch");
  hipLaunchKernelGGL(reduce_planes1<T>, dim3((unsigned)div_up(kXiReal, 256), nc, 2 * n_pairs),
                     dim3(256), 0, st, part, d_pairs, single, global, local, proc, flav_mask, red1);
  check_hip(hipGetLastError(), "reduce_planes1 launch");
  hipLaunchKernelGGL(reduce_planes2, dim3((unsigned)div_up(5 * kXiReal, 256), 1, 2 * n_pairs),
                     dim3(256), 0, st, red1, nc, xi);
  check_hip(hipGetLastError(), "reduce_planes2 launch");
}

template void launch_pass1_planes<double>(const WinPair*, const WinPair&, int, const hlbl_geom&,
                                          const hlbl_geom&, const hlbl_coord&, int, double*,
                                          double*, double*, hipStream_t);
template void launch_pass1_planes<float>(const WinPair*, const WinPair&, int, const hlbl_geom&,
                                         const hlbl_geom&, const hlbl_coord&, int, float*, double*,
                                         double*, hipStream_t);

}  // namespace hlbl
