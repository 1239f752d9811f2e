// The P1 piece of the 2+2 contribution of one source block (compute_2p2_gpu's
// kernel_pi and kernel_p1, kernels/hlbl_kernels/kernels.cu:241-340), FP64, in a fixed
// summation order (no atomics):
//   Pi_{mu nu}(z) = sum_a Re[(gamma_nu gamma_5 w_a)[a]],
//   w_a[b] = sum_c conj(D_b[c]) (gamma_5 gamma_mu U_a)[c],  U = block[f], D = block[1-f];
//   P1[rho][mu][nu][zeta] = sum of Pi_{mu nu}(z) over the local sites with
//   (z - w)_rho mod L_rho = zeta (zeta < lmax; bins this process does not hold are 0).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "contract_kernels.hpp"
#include "lattice.hpp"
#include "pass_common.hpp"

namespace hlbl {

namespace {

constexpr int kPiWaves = 4;
constexpr int kP1Threads = 256;

// With (gamma_g v)[s] = i^{phase(g,s)} v[col(g,s)] on the spin s of a spin-colour
// index 3 s + k, the term of source index a = (s_a, k_a) is
//   Re[i^{phase(nu,s_a) + phase(5,col(nu,s_a))} w_a[(col(nu,s_a), k_a)]],
//   w_a[b] = sum_{s,k} conj(D_b[s,k]) i^{phase(5,s) + phase(mu,s)} U_a[col(mu,s), k].
// One wave per site: lane = (mu, nu, s_a), the three colours k_a in the lane, the four
// spins s_a summed across lanes.
__global__ __launch_bounds__(64 * kPiWaves) void pi_2p2(const double* __restrict__ blk, int f, size_t V,
                                                         double* __restrict__ pi) {
  __shared__ double sh[kPiWaves][2][12 * 24];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mu = lane >> 4, nu = (lane >> 2) & 3, sa = lane & 3;
  double* u = sh[wave][0];
  double* d = sh[wave][1];
  const int sb = gamma_col(nu, sa);
  const int q_out = gamma_phase(nu, sa) + gamma_phase(4, sb);
  // the loop count is the same for every wave of the workgroup (barriers inside)
  for (size_t z0 = (size_t)blockIdx.x * kPiWaves; z0 < V; z0 += (size_t)gridDim.x * kPiWaves) {
    const size_t z = z0 + wave;
    if (z < V)
      for (int i = lane; i < 12 * 24; i += 64) {
        const int a = i / 24, c = i % 24;
        u[i] = blk[((size_t)(f * 12 + a) * V + z) * 24 + c];
        d[i] = blk[((size_t)((1 - f) * 12 + a) * V + z) * 24 + c];
      }
    __syncthreads();
    if (z < V) {
      double acc = 0.0;
      for (int ka = 0; ka < 3; ++ka) {
        const double* ua = u + (3 * sa + ka) * 24;
        const double* db = d + (3 * sb + ka) * 24;
        double wr = 0.0, wi = 0.0;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int q = gamma_phase(4, s) + gamma_phase(mu, s), cs = gamma_col(mu, s);
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            double gr = ua[2 * (3 * cs + k)], gi = ua[2 * (3 * cs + k) + 1];
            mul_ipow(gr, gi, q);
            const double dr = db[2 * (3 * s + k)], di = db[2 * (3 * s + k) + 1];
            // conj(d) g
            wr = fma(dr, gr, fma(di, gi, wr));
            wi = fma(dr, gi, fma(-di, gr, wi));
          }
        }
        mul_ipow(wr, wi, q_out);
        acc += wr;
      }
      acc += __shfl_xor(acc, 1);
      acc += __shfl_xor(acc, 2);
      if (sa == 0) pi[z * 16 + mu * 4 + nu] = acc;
    }
    __syncthreads();
  }
}

__global__ void fill_zero(double* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// One workgroup per (local coordinate zr along rho, rho): the sum of Pi over the slice
// z_rho = zr, per thread over a fixed stride, then a fixed tree; written to the bin of
// (zr + offset - w_rho) mod L_rho.
__global__ __launch_bounds__(kP1Threads) void p1_2p2(const double* __restrict__ pi, hlbl_geom local,
                                                     hlbl_geom global, hlbl_coord proc, hlbl_coord gsw,
                                                     int lmax, double* __restrict__ p1) {
  __shared__ double red[kP1Threads][17];
  const int rho = blockIdx.y, zr = blockIdx.x;
  const int L = geom_dim(local, rho);
  if (zr >= L) return;  // the whole workgroup
  const int dims[4] = {local.T, local.LX, local.LY, local.LZ};
  int other[3], n = 0;
  for (int r = 0; r < 4; ++r)
    if (r != rho) other[n++] = r;
  const size_t nslice = geom_volume(local) / (size_t)L;
  double s[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) s[k] = 0.0;
  for (size_t i = threadIdx.x; i < nslice; i += kP1Threads) {
    int c[4];
    c[rho] = zr;
    size_t q = i;
    c[other[2]] = (int)(q % (size_t)dims[other[2]]);
    q /= (size_t)dims[other[2]];
    c[other[1]] = (int)(q % (size_t)dims[other[1]]);
    c[other[0]] = (int)(q / (size_t)dims[other[1]]);
    const size_t zl = (((size_t)c[0] * dims[1] + c[1]) * dims[2] + c[2]) * dims[3] + c[3];
#pragma unroll
    for (int k = 0; k < 16; ++k) s[k] += pi[zl * 16 + k];
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) red[threadIdx.x][k] = s[k];
  __syncthreads();
  for (int h = kP1Threads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 16; ++k) red[threadIdx.x][k] += red[threadIdx.x + h][k];
    __syncthreads();
  }
  if (threadIdx.x < 16) {
    const int Lg = geom_dim(global, rho);
    const int zeta = mod_pos(zr + L * coord_dim(proc, rho) - coord_dim(gsw, rho), Lg);
    p1[((size_t)rho * 16 + threadIdx.x) * lmax + zeta] = red[0][threadIdx.x];
  }
}

}  // namespace

void launch_p1_2p2(hlbl_contractor* c, const double* blk, int f, hlbl_coord gsw, double* d_p1,
                   hipStream_t st) {
  const size_t V = c->volume;
  c->pi2p2.ensure(16 * V);
  const int lmax = std::max(std::max(c->global.T, c->global.LX), std::max(c->global.LY, c->global.LZ));
  const int nwg = (int)std::min<size_t>((V + kPiWaves - 1) / kPiWaves, (size_t)8 * c->n_cu);
  hipLaunchKernelGGL(pi_2p2, dim3(nwg), dim3(64 * kPiWaves), 0, st, blk, f, V, c->pi2p2.p);
  check_hip(hipGetLastError(), "pi_2p2 launch");
  hipLaunchKernelGGL(fill_zero, dim3(1), dim3(256), 0, st, d_p1, (size_t)64 * lmax);
  check_hip(hipGetLastError(), "fill_zero launch");
  const int lloc = std::max(std::max(c->local.T, c->local.LX), std::max(c->local.LY, c->local.LZ));
  hipLaunchKernelGGL(p1_2p2, dim3(lloc, 4), dim3(kP1Threads), 0, st, c->pi2p2.p, c->local, c->global,
                     c->proc, gsw, lmax, d_p1);
  check_hip(hipGetLastError(), "p1_2p2 launch");
}

}  // namespace hlbl
