// mlp_kernels.hpp - BF16 MFMA building blocks of the geodesic surrogate (gfx950).
//
// The surrogate is a plain MLP (4 inputs -> 6 x 256 SiLU -> 16 outputs).
// Training runs entirely on these kernels:
//   * k_gemm_nt: C[M x N] = A[M x K] . B[N x K]^T (+ bias) with BF16 operands
//     and FP32 accumulation on v_mfma_f32_16x16x32_bf16; both operands are
//     K-contiguous ("NT"), so every lane loads its 8-element K slice of a
//     row as one 16-byte load.  Epilogues: FP32 store, or bias + SiLU with
//     the pre-activation kept in FP32 and the activation stored in BF16.
//   * transpose, SiLU backward, bias-gradient (column sum), loss and Adam.
// Activation layout: row-major [rows][features], features padded to a
// multiple of 32 (the MFMA K step).  Weight layout: [out][in] (K = in is
// contiguous), FP32 master copy + BF16 copy + BF16 transpose [in][out].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bf16.hpp"

namespace bhk {
namespace mlp {

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

enum Epilogue { EPI_F32 = 0, EPI_BIAS_SILU = 1 };

// Block tile 64 x 64, four waves in 2 x 2, each wave 32 x 32 = 2 x 2 MFMA
// tiles of 16 x 16; K advances by 32.  MFMA operand layout (16x16x32):
// lane l holds A[row = l % 16][k = 8 (l / 16) .. +7] and
// B[col = l % 16][k = 8 (l / 16) .. +7]; the result lane l holds
// C[row = 4 (l / 16) + i][col = l % 16], i = 0..3.
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_nt(int M, int N, int K, const uint16_t* __restrict__ A,
                                                 int lda, const uint16_t* __restrict__ B, int ldb,
                                                 const float* __restrict__ bias, float* __restrict__ C,
                                                 int ldc, uint16_t* __restrict__ H, int ldh) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 64 + (wave >> 1) * 32;
  const int n0 = blockIdx.x * 64 + (wave & 1) * 32;
  const int lr = lane & 15;
  const int lk = (lane >> 4) * 8;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int row = m0 + i * 16 + lr;
      const uint4 v = (row < M) ? *reinterpret_cast<const uint4*>(A + static_cast<size_t>(row) * lda + k0 + lk) : zero;
      a[i] = __builtin_bit_cast(bf16x8, v);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + j * 16 + lr;
      const uint4 v = (col < N) ? *reinterpret_cast<const uint4*>(B + static_cast<size_t>(col) * ldb + k0 + lk) : zero;
      b[j] = __builtin_bit_cast(bf16x8, v);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int col = n0 + j * 16 + lr;
      if (col >= N) continue;
      const float bb = bias ? bias[col] : 0.0f;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = m0 + i * 16 + (lane >> 4) * 4 + e;
        if (row >= M) continue;
        const float z = acc[i][j][e] + bb;
        if (C) C[static_cast<size_t>(row) * ldc + col] = z;
        if (EPI == EPI_BIAS_SILU) H[static_cast<size_t>(row) * ldh + col] = f2bf(silu(z));
      }
    }
  }
}

// out[c][r] = in[r][c] (BF16), tiled through LDS.
__global__ void k_transpose_bf16(int rows, int cols, const uint16_t* __restrict__ in, int ldi,
                                 uint16_t* __restrict__ out, int ldo) {
  __shared__ uint16_t t[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int r = by + j, c = bx + tx;
    t[j][tx] = (r < rows && c < cols) ? in[static_cast<size_t>(r) * ldi + c] : 0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = bx + j, r = by + tx;
    if (c < cols && r < rows) out[static_cast<size_t>(c) * ldo + r] = t[tx][j];
  }
}

// dZ = dH * SiLU'(Z) -> BF16, elementwise over [rows][cols].
__global__ void k_silu_bwd(long long n, int cols, int ld_g, const float* __restrict__ dH_f32,
                           const float* __restrict__ Z, int ld_z, uint16_t* __restrict__ dZ, int ld_d) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long r = i / cols;
  const int c = static_cast<int>(i - r * cols);
  const float g = dH_f32[r * ld_g + c] * silu_grad(Z[r * ld_z + c]);
  dZ[r * ld_d + c] = f2bf(g);
}

// db[c] = sum_r dZ[r][c] (BF16 input, FP32 output): one block per column group.
__global__ void k_colsum_bf16(int rows, int cols, const uint16_t* __restrict__ dZ, int ld,
                              float* __restrict__ db) {
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int part = threadIdx.x >> 6;  // 4 partial sums per column
  __shared__ float s[4][64];
  float acc = 0.0f;
  if (c < cols)
    for (int r = part; r < rows; r += 4) acc += bf2f(dZ[static_cast<size_t>(r) * ld + c]);
  s[part][threadIdx.x & 63] = acc;
  __syncthreads();
  if (part == 0 && c < cols) db[c] = s[0][threadIdx.x] + s[1][threadIdx.x] + s[2][threadIdx.x] + s[3][threadIdx.x];
}

__global__ void k_f32_to_bf16(long long n, const float* __restrict__ in, uint16_t* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(in[i]);
}

// Weighted squared-error loss over [rows][n_out]; writes dOut (BF16, padded
// to ld_d columns, zero beyond n_out) and accumulates sum w (y - t)^2 and
// sum w into loss[0], loss[1].  Output column conf_col (if >= 0) is trained
// toward log10 of the row's weighted squared error of the other outputs
// (detached), so it predicts the surrogate's own error.
__global__ void k_loss(int rows, int n_out, const float* __restrict__ Y, int ldy,
                       const float* __restrict__ T, const float* __restrict__ W, int ldt,
                       int conf_col, float conf_weight, float scale, uint16_t* __restrict__ dOut,
                       int ld_d, float* __restrict__ loss) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float err = 0.0f, wsum = 0.0f, l = 0.0f;
  for (int c = 0; c < n_out; ++c) {
    if (c == conf_col) continue;
    const float d = Y[static_cast<size_t>(r) * ldy + c] - T[static_cast<size_t>(r) * ldt + c];
    const float w = W[static_cast<size_t>(r) * ldt + c];
    err += w * d * d;
    wsum += w;
    dOut[static_cast<size_t>(r) * ld_d + c] = f2bf(2.0f * w * d * scale);
  }
  l = err;
  if (conf_col >= 0) {
    const float tgt = log10f(err + 1e-12f);
    const float d = Y[static_cast<size_t>(r) * ldy + conf_col] - tgt;
    dOut[static_cast<size_t>(r) * ld_d + conf_col] = f2bf(2.0f * conf_weight * d * scale);
    l += conf_weight * d * d;
  }
  for (int c = n_out; c < ld_d; ++c) dOut[static_cast<size_t>(r) * ld_d + c] = 0;
  atomicAdd(&loss[0], l);
  atomicAdd(&loss[1], wsum);
}

// Adam (Kingma & Ba 2015) on FP32 master parameters; refreshes the BF16
// copy.  bc1 = 1 - beta1^t, bc2 = 1 - beta2^t.
__global__ void k_adam(long long n, float* __restrict__ p, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, uint16_t* __restrict__ p_bf,
                       float lr, float b1, float b2, float eps, float bc1, float bc2, float wd) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * p[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float upd = lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  p[i] -= upd;
  if (p_bf) p_bf[i] = f2bf(p[i]);
}

// First layer in FP32 (the inputs are not quantised to BF16): Z0 = X W0^T + b0
// over the n_feat live features, H1 = BF16(SiLU(Z0)).  One thread per
// (row, output).  X [n][ldx], W0 [out][ldw], Z0 (optional) / H1 [n][out].
__global__ void k_layer0_fwd(int n, int n_out, int n_feat, const float* __restrict__ X, int ldx,
                             const float* __restrict__ W0, int ldw, const float* __restrict__ b0,
                             float* __restrict__ Z0, uint16_t* __restrict__ H1) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<long long>(n) * n_out) return;
  const int r = static_cast<int>(i / n_out), o = static_cast<int>(i - static_cast<long long>(r) * n_out);
  const float* x = X + static_cast<size_t>(r) * ldx;
  const float* w = W0 + static_cast<size_t>(o) * ldw;
  float z = b0[o];
  for (int k = 0; k < n_feat; ++k) z = fmaf(w[k], x[k], z);
  if (Z0) Z0[i] = z;
  H1[i] = f2bf(silu(z));
}

// dW0[o][k] = sum_r dZ0[r][o] X[r][k] for k < n_feat (FP32; pad columns 0).
// One block per output o; 256 threads stride over rows.
__global__ void k_layer0_wgrad(int n, int n_out, int n_feat, const uint16_t* __restrict__ dZ0, int ldz,
                               const float* __restrict__ X, int ldx, float* __restrict__ dW0, int ldw) {
  const int o = blockIdx.x;
  float acc[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) acc[k] = 0.0f;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const float g = bf2f(dZ0[static_cast<size_t>(r) * ldz + o]);
    const float* x = X + static_cast<size_t>(r) * ldx;
#pragma unroll
    for (int k = 0; k < 32; ++k)
      if (k < n_feat) acc[k] = fmaf(g, x[k], acc[k]);
  }
  __shared__ float red[256];
  for (int k = 0; k < ldw; ++k) {
    red[threadIdx.x] = (k < 32) ? acc[k < 32 ? k : 0] : 0.0f;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) dW0[static_cast<size_t>(o) * ldw + k] = (k < n_feat) ? red[0] : 0.0f;
    __syncthreads();
  }
}

}  // namespace mlp
}  // namespace bhk

