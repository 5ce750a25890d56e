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
#include "../../include/bhk_surrogate.h"

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
//
// Split K (gridDim.z > 1, EPI_F32 only): block z accumulates K-columns
// [z K, (z + 1) K) of A and B (K is then the slice length) into the partial
// C + z M ldc; k_sum_partials adds the partials in a fixed order.
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_nt(int M, int N, int K, const uint16_t* __restrict__ A,
                                                 int lda, const uint16_t* __restrict__ B, int ldb,
                                                 const float* __restrict__ bias, float* __restrict__ C,
                                                 int ldc, uint16_t* __restrict__ H, int ldh) {
  if (gridDim.z > 1) {
    A += static_cast<size_t>(blockIdx.z) * K;
    B += static_cast<size_t>(blockIdx.z) * K;
    C += static_cast<size_t>(blockIdx.z) * M * ldc;
  }
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

// Row chunk of the split reductions (weight gradients, layer-0 gradient).
constexpr int kColRows = 512;

// db[o] = sum_i dZt[o][i] for i < n (BF16 rows of the transposed dZ, row
// stride ld, n and ld multiples of 8): one 256-thread block per row, 16-byte
// loads, then a fixed-order tree in LDS.
__global__ void k_rowsum_bf16(int n, const uint16_t* __restrict__ dZt, int ld, float* __restrict__ db) {
  const uint16_t* row = dZt + static_cast<size_t>(blockIdx.x) * ld;
  float acc = 0.0f;
  for (int i = threadIdx.x * 8; i < n; i += 256 * 8) {
    const uint4 v = *reinterpret_cast<const uint4*>(row + i);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) acc += bf2f(static_cast<uint16_t>(w[k] & 0xffffu)) + bf2f(static_cast<uint16_t>(w[k] >> 16));
  }
  __shared__ float s[256];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (static_cast<int>(threadIdx.x) < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[blockIdx.x] = s[0];
}

// out[i] = sum_z part[z][i] for z = 0 .. nz - 1, in order.
__global__ void k_sum_partials(long long n, int nz, const float* __restrict__ part, float* __restrict__ out) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc = 0.0f;
  for (int z = 0; z < nz; ++z) acc += part[static_cast<size_t>(z) * n + i];
  out[i] = acc;
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
// over the n_feat live features, H1 = BF16(SiLU(Z0)).  Block (x, y) holds
// the weights of outputs 256 y .. 256 y + 255 in LDS and runs rows
// kL0Rows x .. + kL0Rows - 1, one output per thread (coalesced stores of Z0
// and H1 rows).  X [n][ldx], W0 [out][ldw], Z0 (optional) / H1 [n][n_out];
// grid (ceil(n / kL0Rows), ceil(n_out / 256)).
constexpr int kL0Rows = 16;
__global__ void __launch_bounds__(256) k_layer0_fwd(int n, int n_out, int n_feat, const float* __restrict__ X,
                                                    int ldx, const float* __restrict__ W0, int ldw,
                                                    const float* __restrict__ b0, float* __restrict__ Z0,
                                                    uint16_t* __restrict__ H1) {
  __shared__ float w[256][BH_SUR_FEATURES + 1];
  __shared__ float x[kL0Rows][BH_SUR_FEATURES];
  const int o0 = blockIdx.y * 256, o = o0 + threadIdx.x;
  for (int i = threadIdx.x; i < 256 * BH_SUR_FEATURES; i += 256) {
    const int oo = i / BH_SUR_FEATURES, k = i - oo * BH_SUR_FEATURES;
    w[oo][k] = (o0 + oo < n_out && k < n_feat) ? W0[static_cast<size_t>(o0 + oo) * ldw + k] : 0.0f;
  }
  const int r0 = blockIdx.x * kL0Rows;
  for (int i = threadIdx.x; i < kL0Rows * BH_SUR_FEATURES; i += 256) {
    const int rr = i / BH_SUR_FEATURES, k = i - rr * BH_SUR_FEATURES;
    x[rr][k] = (r0 + rr < n) ? X[static_cast<size_t>(r0 + rr) * ldx + k] : 0.0f;
  }
  __syncthreads();
  if (o >= n_out) return;
  const float b = b0[o];
  for (int rr = 0; rr < kL0Rows && r0 + rr < n; ++rr) {
    float z = b;
    for (int k = 0; k < n_feat && k < BH_SUR_FEATURES; ++k) z = fmaf(w[threadIdx.x][k], x[rr][k], z);
    const size_t i = static_cast<size_t>(r0 + rr) * n_out + o;
    if (Z0) Z0[i] = z;
    H1[i] = f2bf(silu(z));
  }
}

// dW0[o][k] = sum_r dZ0[r][o] X[r][k] for k < n_feat (FP32; pad columns 0),
// in row chunks: block (x, y) covers outputs 64 x .. 64 x + 63 (one per lane;
// dZ0 reads are coalesced, the X row is the same for every lane) and rows
// [y kColRows, (y + 1) kColRows), four waves interleaving the rows, and
// writes part[y][o][k] for k < ldw; k_sum_partials adds the chunks.
__global__ void k_layer0_wgrad_part(int n, int n_out, int n_feat, const uint16_t* __restrict__ dZ0, int ldz,
                                    const float* __restrict__ X, int ldx, float* __restrict__ part, int ldw) {
  const int o = blockIdx.x * 64 + (threadIdx.x & 63);
  const int q = threadIdx.x >> 6;
  const int r0 = blockIdx.y * kColRows;
  const int r1 = min(n, r0 + kColRows);
  float acc[BH_SUR_FEATURES];
#pragma unroll
  for (int k = 0; k < BH_SUR_FEATURES; ++k) acc[k] = 0.0f;
  if (o < n_out) {
    for (int r = r0 + q; r < r1; r += 4) {
      const float g = bf2f(dZ0[static_cast<size_t>(r) * ldz + o]);
      const float* x = X + static_cast<size_t>(r) * ldx;
#pragma unroll
      for (int k = 0; k < BH_SUR_FEATURES; ++k) acc[k] = fmaf(g, x[k], acc[k]);
    }
  }
  __shared__ float s[4][64][BH_SUR_FEATURES + 1];
#pragma unroll
  for (int k = 0; k < BH_SUR_FEATURES; ++k) s[q][threadIdx.x & 63][k] = acc[k];
  __syncthreads();
  if (q == 0 && o < n_out) {
    float* dst = part + (static_cast<size_t>(blockIdx.y) * n_out + o) * ldw;
    for (int k = 0; k < ldw; ++k) {
      const int l = threadIdx.x & 63;
      dst[k] = (k < n_feat && k < BH_SUR_FEATURES)
                   ? (s[0][l][k] + s[1][l][k]) + (s[2][l][k] + s[3][l][k])
                   : 0.0f;
    }
  }
}

}  // namespace mlp
}  // namespace bhk

