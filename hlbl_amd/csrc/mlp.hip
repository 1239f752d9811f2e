// MLP surrogate of the tabulated QED kernel: 3 inputs (|x|/R, |y|/R, beta/pi),
// 4 hidden layers of 128 GELU, 42 outputs (the frame components of the T tensors,
// standardised per channel). Weights are stored in
// BF16 and used as BF16 values in every forward pass (the FP32 master copy is
// what the optimiser updates: quantisation-aware training); products are
// accumulated in FP32. Training (forward, backward, Adam) and inference run on
// this library's own kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "context.hpp"
#include "contract_kernels.hpp"
#include "kpart.hpp"
#include "mlp.hpp"
#include "provider.hpp"
#include "qed.hpp"
#include "qed_host.hpp"
#include "window.hpp"

namespace hlbl {

namespace {
size_t div_up(size_t a, size_t b) { return (a + b - 1) / b; }
}  // namespace

// round to the nearest BF16 value (ties to even), returned as float
__host__ __device__ inline float bf16_round(float x) {
  unsigned u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// ---------------------------------------------------------- generic GEMM --
// C[i][j] (ldc) = sum_k A(i,k) B(k,j) (+ bias[j]) (+ C if accumulate),
// A(i,k) = A[i*sai + k*sak], B(k,j) = B[k*sbk + j*sbj]. 64x64 tiles, 4x4 per thread.
__global__ __launch_bounds__(256) void gemm_f32(int M, int N, int K, const float* __restrict__ A,
                                                long sai, long sak, const float* __restrict__ B,
                                                long sbk, long sbj, float* __restrict__ Cm, int ldc,
                                                const float* __restrict__ bias, int accumulate) {
  __shared__ float As[16][68];
  __shared__ float Bs[16][68];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    for (int e = tid; e < 64 * 16; e += 256) {
      const int ii = e >> 4, kk = e & 15;
      const int gi = i0 + ii, gk = k0 + kk;
      As[kk][ii] = (gi < M && gk < K) ? A[gi * sai + gk * sak] : 0.f;
      const int jj = e >> 4, kb = e & 15;
      const int gj = j0 + jj, gk2 = k0 + kb;
      Bs[kb][jj] = (gj < N && gk2 < K) ? B[gk2 * sbk + gj * sbj] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        a[q] = As[kk][ty * 4 + q];
        b[q] = Bs[kk][tx * 4 + q];
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * b[q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = i0 + ty * 4 + p, j = j0 + tx * 4 + q;
      if (i < M && j < N) {
        float v = acc[p][q] + (bias ? bias[j] : 0.f);
        if (accumulate) v += Cm[(size_t)i * ldc + j];
        Cm[(size_t)i * ldc + j] = v;
      }
    }
}

void gemm(int M, int N, int K, const float* A, long sai, long sak, const float* B, long sbk,
          long sbj, float* Cm, int ldc, const float* bias, int accumulate, hipStream_t st) {
  hipLaunchKernelGGL(gemm_f32, dim3((unsigned)div_up(N, 64), (unsigned)div_up(M, 64)), dim3(256), 0,
                     st, M, N, K, A, sai, sak, B, sbk, sbj, Cm, ldc, bias, accumulate);
}

__global__ void k_round_bf16(const float* __restrict__ w, float* __restrict__ wq,
                             unsigned short* __restrict__ w16, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float r = bf16_round(w[i]);
  wq[i] = r;
  unsigned u;
  memcpy(&u, &r, 4);
  if (w16) w16[i] = (unsigned short)(u >> 16);
}

// a = gelu(z), rounded to BF16 when round_act (the activations of the BF16
// inference; the backward pass treats the rounding as identity)
__global__ void k_gelu(const float* __restrict__ z, float* __restrict__ a, size_t n, int round_act) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = round_act ? bf16_round(gelu(z[i])) : gelu(z[i]);
}
// dz = da * gelu'(z)
__global__ void k_gelu_back(const float* __restrict__ z, const float* __restrict__ da,
                            float* __restrict__ dz, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dz[i] = da[i] * gelu_grad(z[i]);
}
// db[j] = sum_b dz[b][j] in a fixed order: a workgroup takes 32 columns, its 8
// row groups sum rows g, g+8, ... (coalesced along j), then the 8 partials are
// added in order.
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dz, int B, int N,
                                                float* __restrict__ db) {
  __shared__ float part[8][32];
  const int c = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int j = blockIdx.x * 32 + c;
  float s = 0.f;
  if (j < N)
    for (int b = g; b < B; b += 8) s += dz[(size_t)b * N + j];
  part[g][c] = s;
  __syncthreads();
  if (g == 0 && j < N) {
    float t = 0.f;
    for (int q = 0; q < 8; ++q) t += part[q][c];
    db[j] = t;
  }
}
// dout = 2 (out - t) / (B*C); per-row loss partials (FP64)
__global__ void k_mse(const float* __restrict__ out, const float* __restrict__ t, int B, int Cn,
                      float* __restrict__ dout, double* __restrict__ row_loss) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double l = 0.0;
  const float inv = 2.0f / (float)(B * Cn);
  for (int c = 0; c < Cn; ++c) {
    const float d = out[(size_t)b * Cn + c] - t[(size_t)b * Cn + c];
    l += (double)d * d;
    dout[(size_t)b * Cn + c] = inv * d;
  }
  row_loss[b] = l;
}
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}
__global__ void k_gather_rows(const float* __restrict__ src, const int* __restrict__ idx, int B,
                              int width, float* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * width) return;
  const int b = i / width, c = i % width;
  dst[i] = src[(size_t)idx[b] * width + c];
}

// ------------------------------------------------------------ the network --

struct MlpWork {
  int B = 0;
  DevBuf<float> z[5], a[5], dz, da, dtmp;  // pre-activations and activations per layer
  DevBuf<double> row_loss;
};

// forward of B rows (x: [B][3]) with weights wq (BF16-rounded floats); keeps z/a.
// round_act: hidden activations rounded to BF16, as the inference kernel does.
void mlp_forward_train(const float* wq, const float* x, int B, MlpWork& wk, hipStream_t st,
                       bool round_act) {
  const MlpLayout L = mlp_layout();
  const float* in = x;
  for (int l = 0; l < 5; ++l) {
    wk.z[l].ensure((size_t)B * L.out[l]);
    // z = in W^T + b: A(i,k) = in[i][k], B(k,j) = W[j][k]
    gemm(B, L.out[l], L.in[l], in, L.in[l], 1, wq + L.w[l], 1, L.in[l], wk.z[l].p, L.out[l],
         wq + L.b[l], 0, st);
    if (l < 4) {
      wk.a[l].ensure((size_t)B * L.out[l]);
      const size_t n = (size_t)B * L.out[l];
      hipLaunchKernelGGL(k_gelu, dim3((unsigned)div_up(n, 256)), dim3(256), 0, st, wk.z[l].p, wk.a[l].p, n,
                         round_act ? 1 : 0);
      in = wk.a[l].p;
    }
  }
}

// loss and gradient (into grad, all parameters) of a batch
double mlp_backward(const float* wq, const float* x, const float* target, int B, MlpWork& wk,
                    float* grad, hipStream_t st, bool want_loss, bool round_act) {
  const MlpLayout L = mlp_layout();
  mlp_forward_train(wq, x, B, wk, st, round_act);
  wk.da.ensure((size_t)B * kMlpHidden);
  wk.dz.ensure((size_t)B * kMlpHidden);
  wk.dtmp.ensure((size_t)B * kMlpHidden);
  wk.row_loss.ensure(B);
  // output layer: dz4 = dL/dz4 (no activation)
  hipLaunchKernelGGL(k_mse, dim3((unsigned)div_up(B, 256)), dim3(256), 0, st, wk.z[4].p, target, B,
                     kMlpOut, wk.dz.p, wk.row_loss.p);
  for (int l = 4; l >= 0; --l) {
    const float* a_in = l == 0 ? x : wk.a[l - 1].p;
    // dW[j][k] = sum_b dz[b][j] a_in[b][k]
    gemm(L.out[l], L.in[l], B, wk.dz.p, 1, L.out[l], a_in, L.in[l], 1, grad + L.w[l], L.in[l],
         nullptr, 0, st);
    hipLaunchKernelGGL(k_colsum, dim3((unsigned)div_up(L.out[l], 32)), dim3(256), 0, st, wk.dz.p, B,
                       L.out[l], grad + L.b[l]);
    if (l > 0) {
      // da_in[b][k] = sum_j dz[b][j] W[j][k]
      gemm(B, L.in[l], L.out[l], wk.dz.p, L.out[l], 1, wq + L.w[l], L.in[l], 1, wk.dtmp.p, L.in[l],
           nullptr, 0, st);
      const size_t n = (size_t)B * L.in[l];
      hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)div_up(n, 256)), dim3(256), 0, st,
                         wk.z[l - 1].p, wk.dtmp.p, wk.dz.p, n);
    }
  }
  check_hip(hipGetLastError(), "mlp backward launch");
  if (!want_loss) return 0.0;
  std::vector<double> rl(B);
  check_hip(hipMemcpyAsync(rl.data(), wk.row_loss.p, B * sizeof(double), hipMemcpyDeviceToHost, st), "copy loss");
  check_hip(hipStreamSynchronize(st), "mlp loss");
  double s = 0.0;
  for (double v : rl) s += v;
  return s / ((double)B * kMlpOut);
}

void mlp_forward_fp32(hlbl_qed* q, const float* x, int n, float* y, hipStream_t st) {
  MlpWork wk;
  mlp_forward_train(q->mlp_w.p, x, n, wk, st, false);
  check_hip(hipMemcpyAsync(y, wk.z[4].p, (size_t)n * kMlpOut * sizeof(float), hipMemcpyDeviceToDevice, st), "copy");
  check_hip(hipStreamSynchronize(st), "mlp forward");
}

double mlp_loss_grad(hlbl_qed* q, const float* x, const float* t, int n, float* grad, hipStream_t st) {
  MlpWork wk;
  mlp_requantise(q, st);
  return mlp_backward(q->mlp_wq.p, x, t, n, wk, grad, st, true, false);
}

// ---------------------------------------------------- training-set and stats --

// counter-based uniform in [0, 1)
__host__ __device__ inline double mlp_uniform(uint64_t seed, uint64_t i, int k) {
  return (double)(hash64(seed * 0x9E3779B97F4A7C15ULL + i * 3 + k) >> 11) * (1.0 / 9007199254740992.0);
}

// |x|, |y| uniform in [inner, radius], cos beta uniform in [-1, 1]; beta = acos(cos beta)
void mlp_sample_inputs(uint64_t seed, size_t i, double inner, double radius, double& a, double& be,
                       double& b) {
  a = inner + (radius - inner) * mlp_uniform(seed, i, 0);
  b = inner + (radius - inner) * mlp_uniform(seed, i, 1);
  be = std::acos(2.0 * mlp_uniform(seed, i, 2) - 1.0);
}

// Training set of n samples (host threads): the channels of the plain evaluation
// (the table, interpolated as the plain kernel does) at |x|, |y| uniform in
// [inner, radius], cos beta uniform in [-1, 1]; x [n][3] normalised inputs
// ((|x| - inner) / (radius - inner), the same for |y|, beta/pi), t [n][42] channels.
void mlp_build_set(uint64_t seed, int n, double inner, double radius, const std::vector<double>& table,
                   std::vector<float>& x, std::vector<double>& t) {
  x.resize((size_t)n * 3);
  t.resize((size_t)n * kMlpOut);
  int nt = (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 16) nt = 16;
  std::vector<std::thread> pool;
  for (int th = 0; th < nt; ++th)
    pool.emplace_back([&, th] {
      double ch[qed::kNCh];
      for (int i = th; i < n; i += nt) {
        double a, be, b;
        mlp_sample_inputs(seed, (size_t)i, inner, radius, a, be, b);
        const qed::Cell<double> cell = qed::locate(a, be, b);
        for (int k = 0; k < kMlpOut; ++k) ch[k] = qed::interp(table.data(), cell, k);
        x[(size_t)i * 3 + 0] = (float)((a - inner) / (radius - inner));
        x[(size_t)i * 3 + 1] = (float)((b - inner) / (radius - inner));
        x[(size_t)i * 3 + 2] = (float)(be / M_PI);
        for (int k = 0; k < kMlpOut; ++k) t[(size_t)i * kMlpOut + k] = ch[k];
      }
    });
  for (auto& p : pool) p.join();
}

// ---------------------------------------------------------------- inference --
// BF16 forward on MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation): weights and
// hidden activations in BF16, the 3-input first layer in FP32 on the vector ALU.
// A wave owns 64 rows (four 16-row tiles) and keeps their activations in LDS as
// BF16 rows of kActStride (padded against bank conflicts); the weights of a layer
// are read from global memory (L2-resident, 108 KB) into registers once per 64
// rows as B fragments: 8 BF16 of W[col][k..k+7] per lane, col = lane%16 of the
// column block, k = 8*(lane/16) of the k-step. A fragments are the same 8-wide
// slices of an activation row, so A and B pair the same k.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4m __attribute__((ext_vector_type(4)));

constexpr int kInfThreads = 256;
constexpr int kInfWaveRows = 64;
constexpr int kInfRows = (kInfThreads / 64) * kInfWaveRows;  // rows per workgroup step
constexpr int kActStride = kMlpHidden + 8;                   // BF16 per activation row

__device__ inline unsigned short f2bf(float x) {
  unsigned u;
  memcpy(&u, &x, 4);
  u += 0x7fffu + ((u >> 16) & 1u);  // nearest, ties to even (inputs are finite)
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) {
  const unsigned u = (unsigned)h << 16;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

// One MFMA layer (K = 128) over the wave's 64 rows: NB column blocks of 16; a
// hidden layer writes gelu(z) back into act as BF16, the output layer hands its
// N columns to out_fn.
template <int NB, bool LAST, class OutFn>
__device__ __forceinline__ void mlp_layer(const unsigned short* __restrict__ w, const unsigned short* __restrict__ b,
                                 int N, unsigned short* act, OutFn& out_fn) {
  const int lane = threadIdx.x & 63;
  bf16x8 breg[4][NB];
  float bias[NB];
#pragma unroll
  for (int cb = 0; cb < NB; ++cb) {
    const int col = 16 * cb + (lane & 15);
    const int cc = col < N ? col : N - 1;  // clamped read; the column is dropped
    bias[cb] = bf2f(b[cc]);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      breg[ks][cb] = *reinterpret_cast<const bf16x8*>(w + cc * kMlpHidden + 32 * ks + 8 * (lane >> 4));
  }
#pragma unroll 1
  for (int t = 0; t < kInfWaveRows / 16; ++t) {
    f4m c[NB];
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) c[cb] = f4m{bias[cb], bias[cb], bias[cb], bias[cb]};
    const unsigned short* arow = act + (16 * t + (lane & 15)) * kActStride + 8 * (lane >> 4);
    bf16x8 a[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) a[ks] = *reinterpret_cast<const bf16x8*>(arow + 32 * ks);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
#pragma unroll
      for (int cb = 0; cb < NB; ++cb)
        c[cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], breg[ks][cb], c[cb], 0, 0, 0);
    // result: element i of lane l is row 4*(l/16) + i of the tile, column l%16 of the block
#pragma unroll
    for (int cb = 0; cb < NB; ++cb) {
      const int col = 16 * cb + (lane & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = 16 * t + 4 * (lane >> 4) + i;
        if (!LAST) act[row * kActStride + col] = f2bf(gelu(c[cb][i]));
        else if (col < N) out_fn(row, col, c[cb][i]);
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);  // keep the next layer's weight loads after this layer
}

// The wave's 64 rows through the network. Lane r supplies the inputs (x0, x1, x2)
// of row r; out_fn(row, col, value) receives the 42 outputs. act: the wave's
// [64][kActStride] BF16 in LDS. Rows are private to the wave, and a wave's LDS
// operations complete in order, so no barrier is needed.
template <class OutFn>
__device__ __forceinline__ void mlp_wave64(const unsigned short* __restrict__ w16, float x0, float x1, float x2,
                           unsigned short* act, OutFn out_fn) {
  const MlpLayout L = mlp_layout();
  const int lane = threadIdx.x & 63;
  // layer 0 (3 -> 128), lane = row; weights wave-uniform
  {
    unsigned* arow = reinterpret_cast<unsigned*>(act + lane * kActStride);
#pragma unroll 2
    for (int j = 0; j < kMlpHidden; j += 2) {
      float z[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int o = j + q;
        z[q] = bf2f(w16[L.b[0] + o]) + bf2f(w16[L.w[0] + 3 * o]) * x0 +
               bf2f(w16[L.w[0] + 3 * o + 1]) * x1 + bf2f(w16[L.w[0] + 3 * o + 2]) * x2;
      }
      arow[j / 2] = (unsigned)f2bf(gelu(z[0])) | ((unsigned)f2bf(gelu(z[1])) << 16);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  mlp_layer<8, false>(w16 + L.w[1], w16 + L.b[1], kMlpHidden, act, out_fn);
  mlp_layer<8, false>(w16 + L.w[2], w16 + L.b[2], kMlpHidden, act, out_fn);
  mlp_layer<8, false>(w16 + L.w[3], w16 + L.b[3], kMlpHidden, act, out_fn);
  static_assert(kMlpOut <= 48, "three 16-column output blocks");
  mlp_layer<3, true>(w16 + L.w[4], w16 + L.b[4], kMlpOut, act, out_fn);
}

// plain forward of n rows x [n][3] -> y [n][42] (normalised), BF16
__global__ __launch_bounds__(kInfThreads, 2) void mlp_infer_rows(const unsigned short* __restrict__ w16,
                                                              const float* __restrict__ x, int n,
                                                              float* __restrict__ y) {
  __shared__ __attribute__((aligned(16))) unsigned short act[kInfThreads / 64][kInfWaveRows * kActStride];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // one workgroup per kInfRows rows (a persistent loop would let the compiler hoist
  // every layer's weight fragments out of it)
  const size_t r0 = (size_t)blockIdx.x * kInfRows + (size_t)wave * kInfWaveRows;
  if (r0 >= (size_t)n) return;
  const size_t row = r0 + lane;
  float x0 = 0.f, x1 = 0.f, x2 = 0.f;
  if (row < (size_t)n) {
    x0 = x[row * 3];
    x1 = x[row * 3 + 1];
    x2 = x[row * 3 + 2];
  }
  mlp_wave64(w16, x0, x1, x2, act[wave], [=](int r, int c, float v) {
    if (r0 + r < (size_t)n) y[(r0 + r) * kMlpOut + c] = v;
  });
}

// Surrogate channels of a window (gsx = 0), chbuf[block][site][kNCh] =
// mean + sd * net, for the blocks 0: (x, 0) and 1: (0, x), shared
// by every yv, and 2 + 3u + j: (x, y), (y, x), (x, x - y) of the distinct yv u.
// Arguments outside the trained domain (|x| or |y| above radius or below inner) take
// the channels from the table instead: with an inner radius, the rows (x, 0) and (0, x)
// always do.
__global__ __launch_bounds__(kInfThreads, 2) void mlp_surrogate_channels(
    const unsigned short* __restrict__ w16, const float* __restrict__ scale,
    const float* __restrict__ table32, float inner, float radius, const hlbl_coord* __restrict__ yvs,
    int n_unique, hlbl_geom global, double xunit, float* __restrict__ chbuf) {
  __shared__ __attribute__((aligned(16))) unsigned short act[kInfThreads / 64][kInfWaveRows * kActStride];
  __shared__ float s_in[kInfThreads / 64][kInfWaveRows];  // 1: inside the trained domain
  __shared__ float s_acb[kInfThreads / 64][kInfWaveRows][3];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const size_t V = geom_volume(global);
  const size_t rows = (size_t)(2 + 3 * n_unique) * V;
  const hlbl_coord zero = {0, 0, 0, 0};
  const float inv_r = 1.0f / (radius - inner);
  {
    const size_t r0 = (size_t)blockIdx.x * kInfRows + (size_t)wave * kInfWaveRows;
    if (r0 >= rows) return;
    const size_t row = r0 + lane;
    float a = 0.f, be = 0.f, b = 0.f, in = 0.f;
    if (row < rows) {
      const int blk = (int)(row / V);
      const size_t site = row - (size_t)blk * V;
      const int u = blk < 2 ? 0 : (blk - 2) / 3;
      const int e = blk == 0 ? 1 : blk == 1 ? 3 : 2 * ((blk - 2) % 3);
      SiteGeom<float> g;
      site_geom(site, global, global, zero, zero, yvs[u], xunit, g);
      float pu[4], pv[4];
      eval_args(e, g, pu, pv);
      qed::grid_coords(pu, pv, a, be, b);
      // outside the trained domain: table
      in = (a > radius || b > radius || a < inner || b < inner) ? 0.f : 1.f;
    }
    s_in[wave][lane] = in;
    s_acb[wave][lane][0] = a;
    s_acb[wave][lane][1] = be;
    s_acb[wave][lane][2] = b;
    const float* inw = s_in[wave];
    const float(*acbw)[3] = s_acb[wave];
    mlp_wave64(w16, fminf(fmaxf((a - inner) * inv_r, 0.f), 1.f), fminf(fmaxf((b - inner) * inv_r, 0.f), 1.f),
               be * (float)M_1_PI, act[wave],
               [=](int r, int col, float v) {
      if (r0 + r >= rows) return;
      float ch;
      if (inw[r] > 0.f) {
        ch = scale[col] + scale[kMlpOut + col] * v;
      } else {
        const qed::Cell<float> cell = qed::locate(acbw[r][0], acbw[r][1], acbw[r][2]);
        ch = qed::interp(table32, cell, col);
      }
      chbuf[(r0 + r) * qed::kNCh + col] = ch;
    });
  }
}

void launch_surrogate_channels(hlbl_contractor* c, WindowPlan& w, hipStream_t st) {
  hlbl_qed* q = c->qed;
  const size_t rows = (size_t)(2 + 3 * w.n_unique) * c->volume;
  const unsigned nwg = (unsigned)div_up(rows, kInfRows);
  hipLaunchKernelGGL(mlp_surrogate_channels, dim3(nwg), dim3(kInfThreads), 0, st, q->mlp_w_bf16.p,
                     q->mlp_scale.p, q->table32.p, (float)q->mlp_inner, (float)q->mlp_radius, w.d_yvs.p,
                     w.n_unique,
                     c->global, w.xunit, w.chbuf.p);
  check_hip(hipGetLastError(), "mlp_surrogate_channels launch");
}

// surrogate channels at points (a, beta, b): inputs, network and scale; points
// outside the trained domain take the table's channels
__global__ void k_surrogate_inputs(const double* __restrict__ acb, int n, double inner, double radius,
                                   float* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double inv = 1.0 / (radius - inner);
  x[3 * i + 0] = (float)fmin(fmax((acb[3 * i + 0] - inner) * inv, 0.0), 1.0);
  x[3 * i + 1] = (float)fmin(fmax((acb[3 * i + 2] - inner) * inv, 0.0), 1.0);
  x[3 * i + 2] = (float)(acb[3 * i + 1] * M_1_PI);
}
__global__ void k_surrogate_outputs(const double* __restrict__ acb, const float* __restrict__ y,
                                    const float* __restrict__ scale, const double* __restrict__ table,
                                    double inner, double radius, int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * kMlpOut) return;
  const int r = i / kMlpOut, c = i % kMlpOut;
  const double a = acb[3 * r], cb = acb[3 * r + 1], b = acb[3 * r + 2];
  if (a > radius || b > radius || a < inner || b < inner) {
    out[i] = qed::interp(table, qed::locate(a, cb, b), c);
    return;
  }
  out[i] = (double)scale[c] + (double)scale[kMlpOut + c] * (double)y[i];
}

void launch_surrogate_points(hlbl_qed* q, const double* acb, int n, double* out, hipStream_t st) {
  if (!q->mlp_ready) throw Error("hlbl_qed_channels: mode 1 needs a trained surrogate");
  float *x = nullptr, *y = nullptr;
  check_hip(hipMallocAsync((void**)&x, (size_t)n * 3 * sizeof(float), st), "hipMallocAsync");
  check_hip(hipMallocAsync((void**)&y, (size_t)n * kMlpOut * sizeof(float), st), "hipMallocAsync");
  hipLaunchKernelGGL(k_surrogate_inputs, dim3((unsigned)div_up(n, 256)), dim3(256), 0, st, acb, n,
                     q->mlp_inner, q->mlp_radius, x);
  mlp_infer(q, x, n, y, st);
  hipLaunchKernelGGL(k_surrogate_outputs, dim3((unsigned)div_up((size_t)n * kMlpOut, 256)), dim3(256), 0,
                     st, acb, y, q->mlp_scale.p, q->table64.p, q->mlp_inner, q->mlp_radius, n, out);
  hipError_t e = hipGetLastError();
  (void)hipFreeAsync(x, st);
  (void)hipFreeAsync(y, st);
  check_hip(e, "surrogate channels launch");
}

void mlp_infer(hlbl_qed* q, const float* x, int n, float* y, hipStream_t st) {
  if (n <= 0) return;
  const unsigned nwg = (unsigned)div_up(n, kInfRows);
  hipLaunchKernelGGL(mlp_infer_rows, dim3(nwg), dim3(kInfThreads), 0, st, q->mlp_w_bf16.p, x, n, y);
  check_hip(hipGetLastError(), "mlp_infer_rows launch");
}

// ------------------------------------------------------------------ training --

void mlp_init_params(hlbl_qed* q, uint64_t seed) {
  const MlpLayout L = mlp_layout();
  std::vector<float> w(L.total, 0.f);
  for (int l = 0; l < 5; ++l) {
    // uniform(-s, s), s = sqrt(6 / (fan_in + fan_out)) (Glorot), biases 0
    const double s = std::sqrt(6.0 / (L.in[l] + L.out[l]));
    for (int i = 0; i < L.in[l] * L.out[l]; ++i)
      w[L.w[l] + i] = (float)(s * (2.0 * mlp_uniform(seed + 17 * l, (uint64_t)i, 0) - 1.0));
  }
  q->mlp_w.ensure(L.total);
  check_hip(hipMemcpy(q->mlp_w.p, w.data(), L.total * sizeof(float), hipMemcpyHostToDevice), "copy params");
  q->mlp_wq.ensure(L.total);
  q->mlp_w_bf16.ensure(L.total);
  hipLaunchKernelGGL(k_round_bf16, dim3((unsigned)div_up(L.total, 256)), dim3(256), 0, 0, q->mlp_w.p,
                     q->mlp_wq.p, q->mlp_w_bf16.p, L.total);
  check_hip(hipDeviceSynchronize(), "mlp init");
  if (!q->mlp_scale.p) {
    std::vector<float> sc(2 * kMlpOut);
    for (int k = 0; k < kMlpOut; ++k) {
      sc[k] = 0.f;
      sc[kMlpOut + k] = 1.f;
    }
    q->mlp_scale.ensure(2 * kMlpOut);
    check_hip(hipMemcpy(q->mlp_scale.p, sc.data(), sc.size() * sizeof(float), hipMemcpyHostToDevice), "copy scale");
  }
}

void mlp_requantise(hlbl_qed* q, hipStream_t st) {
  const int n = mlp_layout().total;
  hipLaunchKernelGGL(k_round_bf16, dim3((unsigned)div_up(n, 256)), dim3(256), 0, st, q->mlp_w.p,
                     q->mlp_wq.p, q->mlp_w_bf16.p, n);
}

void mlp_train(hlbl_qed* q, int n_samples, int epochs, int batch, float lr, uint64_t seed,
               float* loss_hist, hipStream_t st) {
  const MlpLayout L = mlp_layout();
  if (!q->mlp_w.p) mlp_init_params(q, seed ^ 0x5eedULL);
  std::vector<float> x;
  std::vector<double> t;
  mlp_build_set(seed, n_samples, q->mlp_inner, q->mlp_radius, q->host_table, x, t);
  // per-channel standardisation of the targets
  std::vector<float> sc(2 * kMlpOut);
  std::vector<float> tn((size_t)n_samples * kMlpOut);
  for (int k = 0; k < kMlpOut; ++k) {
    double m = 0, v = 0;
    for (int i = 0; i < n_samples; ++i) m += t[(size_t)i * kMlpOut + k];
    m /= n_samples;
    for (int i = 0; i < n_samples; ++i) {
      const double d = t[(size_t)i * kMlpOut + k] - m;
      v += d * d;
    }
    const double sd = std::sqrt(v / n_samples) + 1e-30;
    sc[k] = (float)m;
    sc[kMlpOut + k] = (float)sd;
    for (int i = 0; i < n_samples; ++i) tn[(size_t)i * kMlpOut + k] = (float)((t[(size_t)i * kMlpOut + k] - m) / sd);
  }
  q->mlp_scale.ensure(2 * kMlpOut);
  check_hip(hipMemcpy(q->mlp_scale.p, sc.data(), sc.size() * sizeof(float), hipMemcpyHostToDevice), "copy scale");
  DevBuf<float> dx, dt, bx, bt, grad, m, v;
  dx.ensure(x.size());
  dt.ensure(tn.size());
  check_hip(hipMemcpy(dx.p, x.data(), x.size() * 4, hipMemcpyHostToDevice), "copy x");
  check_hip(hipMemcpy(dt.p, tn.data(), tn.size() * 4, hipMemcpyHostToDevice), "copy t");
  bx.ensure((size_t)batch * 3);
  bt.ensure((size_t)batch * kMlpOut);
  grad.ensure(L.total);
  m.ensure(L.total);
  v.ensure(L.total);
  check_hip(hipMemsetAsync(m.p, 0, L.total * 4, st), "memset");
  check_hip(hipMemsetAsync(v.p, 0, L.total * 4, st), "memset");
  DevBuf<int> didx;
  didx.ensure(batch);
  std::vector<int> idx(n_samples), bidx(batch);
  for (int i = 0; i < n_samples; ++i) idx[i] = i;
  MlpWork wk;
  long step = 0;
  const float b1 = 0.9f, b2 = 0.999f;
  const int steps_per_epoch = n_samples / batch;
  for (int ep = 0; ep < epochs; ++ep) {
    // deterministic shuffle
    for (int i = n_samples - 1; i > 0; --i) {
      const int j = (int)(hash64(seed + 1000003ULL * ep + i) % (uint64_t)(i + 1));
      std::swap(idx[i], idx[j]);
    }
    // cosine learning-rate decay over the run
    double eloss = 0.0;
    for (int s = 0; s < steps_per_epoch; ++s) {
      for (int b = 0; b < batch; ++b) bidx[b] = idx[(size_t)s * batch + b];
      check_hip(hipMemcpyAsync(didx.p, bidx.data(), batch * 4, hipMemcpyHostToDevice, st), "copy idx");
      hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)div_up(batch * 3, 256)), dim3(256), 0, st, dx.p, didx.p, batch, 3, bx.p);
      hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)div_up(batch * kMlpOut, 256)), dim3(256), 0, st, dt.p, didx.p, batch, kMlpOut, bt.p);
      mlp_requantise(q, st);
      const bool want = (s == steps_per_epoch - 1);
      const double l = mlp_backward(q->mlp_wq.p, bx.p, bt.p, batch, wk, grad.p, st, want, true);
      if (want) eloss = l;
      ++step;
      const double frac = (double)step / ((double)epochs * steps_per_epoch);
      const float lr_t = (float)(lr * 0.5 * (1.0 + std::cos(3.14159265358979 * frac)));
      const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
      hipLaunchKernelGGL(k_adam, dim3((unsigned)div_up(L.total, 256)), dim3(256), 0, st, q->mlp_w.p,
                         grad.p, m.p, v.p, L.total, lr_t, b1, b2, 1e-8f, c1, c2);
      // the staging copy of bidx is synchronous with respect to the host vector reuse
      check_hip(hipStreamSynchronize(st), "mlp step");
    }
    if (loss_hist) loss_hist[ep] = (float)eloss;
  }
  mlp_requantise(q, st);
  check_hip(hipStreamSynchronize(st), "mlp train");
  q->mlp_ready = true;
  ++q->mlp_gen;
}

}  // namespace hlbl
