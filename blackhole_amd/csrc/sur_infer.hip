// sur_infer.hip - image rendering with the geodesic surrogate (gfx950).
//
// k_sur_render: one workgroup (4 waves) takes a tile of 128 consecutive
// pixels through the whole network without leaving the CU.  The layers are
// computed transposed, C^T[out][ray] = W[out][k] H[ray][k]^T, so that the
// weights (BF16, streamed from L2) are the MFMA A operand and the
// activations (BF16 [ray][feature], resident in LDS) the B operand.  Wave w
// owns output features [64 w, 64 w + 64): 2 feature blocks x 4 ray blocks =
// 8 tiles of v_mfma_f32_32x32x16_bf16 (128 accumulator registers per lane),
// so every weight matrix is read once per 128 rays.  The 32x32x16 form
// keeps the matrix pipe near full from a single wave (one 16x16x32 stream
// per wave leaves half of it idle), which matters because the two waves of
// a SIMD alternate between MFMA work and their SiLU epilogues.  A lane's
// accumulator holds, for one ray, four groups of four consecutive features;
// each group goes back to LDS as one 8-byte store.  The hidden weights are
// read from a copy swizzled into fragment order (k_swizzle), so each wave's
// weight fragment is one contiguous 1 KiB load.  LDS per workgroup is 74 KiB,
// so two workgroups share a CU.
//
// Layer 0 (16 features -> 256) keeps the inputs' FP32 resolution: features
// x and weights W are split into BF16 hi/lo parts and the layer is three
// MFMA K-steps, W_hi x_hi + W_hi x_lo + W_lo x_hi (error ~2^-16 relative).
//
// The network runs rescaled by L = log2(e): with W_0 and every hidden bias
// multiplied by L, each accumulator holds z' = L z, SiLU becomes
// z' / (1 + 2^-z') = L silu(z) (exp2 with a negated operand, add, reciprocal,
// multiply), the stored activations are L h, the next layer's z' = W (L h) +
// L b = L z again, and the outputs are (W_out (L h)) / L + b_out.  Each
// accumulator starts from its bias, so the epilogue needs no bias add.
//
// Decode (thread t < 128 = ray t of the tile): ln r_k = output + 3 x
// feature 14 (the light-bending estimate), g from the exact Keplerian
// formula at the predicted radius with the ray's exact xi, and the intensity
// of the first crossing inside [r_ISCO, r_out] from the camera's disk LUT
// exactly as the integrator computes it.  A ray is listed for the
// integrator (fallback) when its confidence output is low, when it lies
// within ring_band of the critical curve (per-camera polar table), when a
// predicted crossing lies within edge_tol of the disk's inner or outer edge,
// or when the crossing-count / capture outputs are ambiguous (with the
// Mino-time solve to follow, only the ring band; with the band solved too,
// none: band rays are written with no crossings for the solver to start
// afresh).  A deterministic 1-in-audit_every sample of the other rays is
// listed for the audit.  Appends use one atomic per wave (ballot + popcount).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bhk_surrogate.h"
#include "bf16.hpp"
#include "sur_features.hpp"
#include "trace_common.hpp"

namespace bhk {
int fail(const std::string& msg);

// View of a surrogate's device parameters (surrogate.hip).
struct SurView {
  const float* p32;        // FP32 parameter vector
  const uint16_t* pbf;     // its BF16 mirror
  const long long* w_off;  // per layer offsets of W_l and b_l
  const long long* b_off;
  int width, L;            // hidden width, number of weight layers (depth + 1)
};
SurView surrogate_view(BH_Surrogate* s);
// Device scratch owned by the surrogate, one buffer per (device, stream).
void* surrogate_render_scratch(BH_Surrogate* s, void* stream, size_t bytes);

namespace sinf {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(16))) float f32x16;

constexpr int kTile = 128;    // rays per workgroup
constexpr int kW = 256;       // hidden width of the fused kernel
constexpr int kLdH = kW + 8;  // LDS row stride of the activations (BF16): 528 B, conflict-free
constexpr int kMaxLayers = 8;
constexpr float kLog2e = 1.4426950408889634f;
constexpr float kInvLog2e = 0.6931471805599453f;

struct InferParams {
  const CamDev* cams;
  const double* luts;
  const float* ring;
  const uint16_t* w0cat;  // [256][64]: [W_hi | W_hi | W_lo | 0] of L W_0
  const float* bs;        // L b_l for l = 0..n_hidden, [n_hidden + 1][256]
  const uint16_t* w[kMaxLayers];
  const uint16_t* wsw;    // hidden layers 1..n_hidden in fragment order (k_swizzle), 128 KiB each
  const float* b[kMaxLayers];
  BH_PixelRecord* out;
  long long* fb_ids;
  unsigned long long* fb_count;
  long long* audit_ids;
  unsigned long long* audit_count;
  long long n_total, fb_cap, audit_cap;
  double disk_r_out, nu, T_max, r_escape;
  int n_cams, lut_n, n_hidden, audit_every;  // weight layers 1..n_hidden are 256 x 256
  float ring_band, edge_tol, conf_max, amb;
  int refine;      // the Mino-time solver follows: only the ring band falls back here
  int band_solve;  // ... and solves the ring band too (no crossings written, nothing listed)
  double keep_r;   // with the solve to follow: rays escaped or captured with every predicted
                   // crossing at r >= keep_r keep the network's record (0: none)
};

__device__ __forceinline__ uint32_t hash32(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return static_cast<uint32_t>(x);
}

__device__ __forceinline__ bf16x8 ld16(const uint16_t* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

// Two FP32 values to packed BF16 (round to nearest even) with the gfx950
// conversion instruction (v_cvt_pk_bf16_f32); same rounding as mlp::f2bf.
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
__device__ __forceinline__ uint32_t pack_bf16x2(float lo, float hi) {
  const bf16x2 v = {static_cast<__bf16>(lo), static_cast<__bf16>(hi)};
  return __builtin_bit_cast(uint32_t, v);
}

// L silu(z) from z' = L z (L = log2 e): z' / (1 + 2^-z').
__device__ __forceinline__ float silu_scaled(float zs) {
  return zs * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-zs));
}

// v_mfma_f32_32x32x16_bf16 operand layout (lane l): A[m = l % 32][k = 8 (l / 32) + i],
// B[k = 8 (l / 32) + i][n = l % 32], accumulator element e at
// C[m = 8 (e / 4) + 4 (l / 32) + e % 4][n = l % 32].
// A wave's accumulators: acc[ob][rb] holds (rescaled) features
// o0 + 32 ob + 8 (e / 4) + 4 (lane / 32) + e % 4 of ray 32 rb + lane % 32.
typedef f32x16 WaveAcc[2][4];

// The wave's bias slice in accumulator order.
__device__ __forceinline__ void load_bias(f32x16 (&b16)[2], const float* __restrict__ bias, int o0, int lh) {
#pragma unroll
  for (int ob = 0; ob < 2; ++ob)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const float4 bb = *reinterpret_cast<const float4*>(bias + o0 + 32 * ob + 8 * g + 4 * lh);
      b16[ob][4 * g] = bb.x;
      b16[ob][4 * g + 1] = bb.y;
      b16[ob][4 * g + 2] = bb.z;
      b16[ob][4 * g + 3] = bb.w;
    }
}

// H[ray][f .. f + 3] = BF16(L silu) of every tile of the wave.
__device__ __forceinline__ void store_act(const WaveAcc& acc, uint16_t* __restrict__ H, int o0, int l32, int lh) {
#pragma unroll
  for (int ob = 0; ob < 2; ++ob)
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      uint16_t* row = H + (32 * rb + l32) * kLdH + o0 + 32 * ob + 4 * lh;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        uint2 v;
        v.x = pack_bf16x2(silu_scaled(acc[ob][rb][4 * g]), silu_scaled(acc[ob][rb][4 * g + 1]));
        v.y = pack_bf16x2(silu_scaled(acc[ob][rb][4 * g + 2]), silu_scaled(acc[ob][rb][4 * g + 3]));
        *reinterpret_cast<uint2*>(row + 8 * g) = v;
      }
    }
}

// Weight fragments of one K-step for the wave's two feature blocks.  Wsw is
// a layer in fragment order: for feature block fb = o0 / 32 + ob and K-step
// ks, lane l's 8 weights sit at Wsw[((fb * 16 + ks) * 64 + l) * 8].
struct AFrag {
  bf16x8 a0, a1;
};
__device__ __forceinline__ AFrag load_afrag(const uint16_t* __restrict__ Wsw, int o0, int lane, int ks) {
  const uint16_t* wp0 = Wsw + ((o0 / 32) * 16 * 64 + lane) * 8 + ks * 512;
  return AFrag{ld16(wp0), ld16(wp0 + 16 * 64 * 8)};
}

// Weight fragments stream from L2 kPrefetch K-steps ahead of their MFMAs
// (a ring of kPrefetch AFrags; kPrefetch divides the 16 K-steps), and each
// K-step's activation fragments are read from LDS one step ahead.  The
// compiler's scheduler would otherwise sink every load next to its first use
// and expose the L2 and LDS latencies on every K-step; a scheduling barrier
// after each step's loads keeps them where they are issued.
constexpr int kPrefetch = 4;
static_assert((kW / 16) % kPrefetch == 0, "the prefetch ring must divide the K-steps");

// One hidden layer's accumulation: acc = bias + W[o0 + ..][:] . H[ray][:]^T,
// 16 K-steps of 16, with the first kPrefetch K-steps' weights already in pa;
// the ring is refilled kPrefetch steps ahead and left empty.  The first
// K-step reads the bias as the MFMA's C operand.
__device__ __forceinline__ void hidden_mfma(WaveAcc& acc, const uint16_t* __restrict__ Wsw,
                                            const uint16_t* __restrict__ H, const f32x16 (&b16)[2],
                                            AFrag (&pa)[kPrefetch], int o0, int lane, int l32, int lh) {
  const uint16_t* hp = H + l32 * kLdH + 8 * lh;
  bf16x8 bn[4];
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) bn[rb] = ld16(hp + 32 * rb * kLdH);
#pragma unroll
  for (int ks = 0; ks < kW / 16; ++ks) {
    const AFrag a = pa[ks % kPrefetch];
    if (ks + kPrefetch < kW / 16) pa[ks % kPrefetch] = load_afrag(Wsw, o0, lane, ks + kPrefetch);
    bf16x8 bv[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) bv[rb] = bn[rb];
    if (ks + 1 < kW / 16) {
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) bn[rb] = ld16(hp + 32 * rb * kLdH + 16 * (ks + 1));
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      acc[0][rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.a0, bv[rb], ks == 0 ? b16[0] : acc[0][rb], 0, 0, 0);
      acc[1][rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.a1, bv[rb], ks == 0 ? b16[1] : acc[1][rb], 0, 0, 0);
    }
  }
}

__global__ void __launch_bounds__(256, 2) k_sur_render(const InferParams P) {
  __shared__ __attribute__((aligned(16))) uint16_t H[kTile * kLdH];  // activations, 66 KiB
  __shared__ __attribute__((aligned(16))) uint16_t X[kTile * 32];    // [x_hi | x_lo]; later FP32 outputs [128][16]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lq = lane >> 4;    // 16x16x32 layout (output layer)
  const int l32 = lane & 31, lh = lane >> 5;   // 32x32x16 layout (layer 0, hidden layers)
  const long long base = static_cast<long long>(blockIdx.x) * kTile;

  // ---- ray setup and features (threads 0..127) -----------------------------
  int cam = 0;
  double alpha = 0.0, beta = 0.0;
  bool valid = false;
  float lnbase = 0.0f;
  if (tid < kTile) {
    const long long gid = base + tid;
    if (gid < P.n_total) {
      cam = find_camera(P.cams, P.n_cams, gid);
      const CamDev& k = P.cams[cam];
      const long long local = gid - k.offset;
      if (local >= 0 && local < k.w * k.h) {
        const long long py = local / k.w, px = local - py * k.w;
        alpha = k.alpha_min + (static_cast<double>(px) + 0.5) * k.dalpha;
        beta = k.beta_max - (static_cast<double>(py) + 0.5) * k.dbeta;
        valid = true;
      }
    }
    float f[BH_SUR_FEATURES];
    if (valid) {
      surrogate_features(static_cast<float>(alpha), static_cast<float>(beta), static_cast<float>(P.cams[cam].a),
                         static_cast<float>(P.cams[cam].cos_o), f);
    } else {
#pragma unroll
      for (int k = 0; k < BH_SUR_FEATURES; ++k) f[k] = 0.0f;
    }
    lnbase = 3.0f * f[kSurBaseFeature];
    uint32_t row[16];
#pragma unroll
    for (int k = 0; k < BH_SUR_FEATURES; k += 2) {
      const uint16_t h0 = mlp::f2bf(f[k]), h1 = mlp::f2bf(f[k + 1]);
      row[k / 2] = static_cast<uint32_t>(h0) | (static_cast<uint32_t>(h1) << 16);
      row[8 + k / 2] = pack_bf16x2(f[k] - mlp::bf2f(h0), f[k + 1] - mlp::bf2f(h1));
    }
    uint4* dst = reinterpret_cast<uint4*>(X + tid * 32);
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[v] = make_uint4(row[4 * v], row[4 * v + 1], row[4 * v + 2], row[4 * v + 3]);
  }
  __syncthreads();

  const int o0 = wave * 64;
  WaveAcc acc;
  f32x16 b16[2];

  // ---- layer 0: W_hi x_hi + W_hi x_lo + W_lo x_hi over the [x_hi | x_lo] rows ---
  load_bias(b16, P.bs, o0, lh);
  {
    bf16x8 a[2][3];
#pragma unroll
    for (int ob = 0; ob < 2; ++ob)
#pragma unroll
      for (int c = 0; c < 3; ++c) a[ob][c] = ld16(P.w0cat + (o0 + 32 * ob + l32) * 64 + 16 * c + 8 * lh);
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      const bf16x8 xh = ld16(X + (32 * rb + l32) * 32 + 8 * lh);
      const bf16x8 xl = ld16(X + (32 * rb + l32) * 32 + 16 + 8 * lh);
#pragma unroll
      for (int ob = 0; ob < 2; ++ob) {
        acc[ob][rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ob][0], xh, b16[ob], 0, 0, 0);
        acc[ob][rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ob][1], xl, acc[ob][rb], 0, 0, 0);
        acc[ob][rb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[ob][2], xh, acc[ob][rb], 0, 0, 0);
      }
    }
  }
  AFrag pa[kPrefetch];  // the next layer's first K-steps, in flight across the epilogue
  if (P.n_hidden > 0) {
#pragma unroll
    for (int d = 0; d < kPrefetch; ++d) pa[d] = load_afrag(P.wsw, o0, lane, d);
  }
  store_act(acc, H, o0, l32, lh);

  // ---- hidden layers 1..n_hidden: 256 x 256 ----------------------------------------
  for (int l = 1; l <= P.n_hidden; ++l) {
    const uint16_t* Wl = P.wsw + static_cast<size_t>(l - 1) * kW * kW;
    load_bias(b16, P.bs + l * kW, o0, lh);
    __syncthreads();  // H complete
    hidden_mfma(acc, Wl, H, b16, pa, o0, lane, l32, lh);
    if (l < P.n_hidden) {
#pragma unroll
      for (int d = 0; d < kPrefetch; ++d) pa[d] = load_afrag(Wl + kW * kW, o0, lane, d);
    }
    __syncthreads();  // every wave has read H
    store_act(acc, H, o0, l32, lh);
  }
  __syncthreads();

  // ---- output layer: 16 x 128, wave w owns ray blocks 2w, 2w + 1 --------------
  {
    const int lo = P.n_hidden + 1;
    const uint16_t* __restrict__ W = P.w[lo];
    f32x4 o[2] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int kc = 0; kc < kW; kc += 32) {
      const bf16x8 a = ld16(W + lr * kW + kc + lq * 8);
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        const bf16x8 bv = ld16(H + (16 * (2 * wave + jj) + lr) * kLdH + kc + lq * 8);
        o[jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, o[jj], 0, 0, 0);
      }
    }
    // X is free since layer 0; it now holds the FP32 outputs [128][16].
    float* O = reinterpret_cast<float*>(X);
    const float4 bo = *reinterpret_cast<const float4*>(P.b[lo] + 4 * lq);
#pragma unroll
    for (int jj = 0; jj < 2; ++jj) {
      const int ray = 16 * (2 * wave + jj) + lr;
      *reinterpret_cast<float4*>(O + ray * 16 + 4 * lq) =
          make_float4(fmaf(o[jj][0], kInvLog2e, bo.x), fmaf(o[jj][1], kInvLog2e, bo.y),
                      fmaf(o[jj][2], kInvLog2e, bo.z), fmaf(o[jj][3], kInvLog2e, bo.w));
    }
  }
  __syncthreads();
  if (tid >= kTile) return;  // waves 2, 3 are done; waves 0, 1 decode whole

  // ---- decode --------------------------------------------------------------------
  const float* O = reinterpret_cast<const float*>(X) + tid * 16;
  const long long gid = base + tid;
  bool fb = false, aud = false;
  if (valid) {
    const CamDev& k = P.cams[cam];
    const double a = k.a;
    const double xi = -(alpha * k.sin_o);
    const float ncf = O[BH_SUR_NCROSS];
    const float cap = O[BH_SUR_CAPTURE];
    const float conf = O[BH_SUR_CONF];
    const int nc = static_cast<int>(fminf(fmaxf(rintf(ncf), 0.0f), 3.0f));
    bool band = false;
    fb = !P.refine && (!(conf <= P.conf_max) || fabsf(ncf - rintf(ncf)) > P.amb);
    // critical-curve band: radial distance from the curve's [lo, hi] range in
    // the pixel's polar bin, against the band scaled by the bin's obliquity
    {
      const float* rg = P.ring + static_cast<size_t>(cam) * BH_SUR_RING_FLOATS;
      const float da = static_cast<float>(alpha) - rg[0], db = static_cast<float>(beta) - rg[1];
      float psi = atan2f(db, da);
      if (psi < 0.0f) psi += 6.2831853071795865f;
      const int bin = min(static_cast<int>(psi * (BH_SUR_RING_BINS / 6.2831853071795865f)), BH_SUR_RING_BINS - 1);
      const float rho = sqrtf(da * da + db * db);
      const float lo = rg[2 + bin], hi = rg[2 + BH_SUR_RING_BINS + bin], obl = rg[2 + 2 * BH_SUR_RING_BINS + bin];
      const float d = fmaxf(fmaxf(lo - rho, rho - hi), 0.0f);
      band = d < P.ring_band * obl;
      if (band && !P.band_solve) fb = true;
    }
    // with the solve to follow, a ray the network places escaped or captured
    // with an unambiguous count and capture and every predicted crossing
    // beyond keep_r keeps the network's record (its intensity is 0 whatever
    // the radii); the margin makes the class safe
    bool keep = P.refine && P.keep_r > 0.0 && !band && fabsf(ncf - rintf(ncf)) <= P.amb &&
                fabsf(cap - rintf(cap)) <= P.amb;
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (c < nc) keep = keep && exp(static_cast<double>(O[2 * c] + lnbase)) >= P.keep_r;
    double cr0 = 0.0, cr1 = 0.0, cr2 = 0.0, cg0 = 0.0, cg1 = 0.0, cg2 = 0.0, I = 0.0;
    int hit = -1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (c < nc && hit < 0) {
        const double r = exp(static_cast<double>(O[2 * c] + lnbase));
        // with the solve to follow, g and the intensity are written by the solve
        // (or the integrator) for every other ray, so only the radii are decoded
        const double g = ((!P.refine || keep) && r >= k.r_isco) ? disk_g(r, a, xi) : 0.0;
        if (c == 0) { cr0 = r; cg0 = g; }
        if (c == 1) { cr1 = r; cg1 = g; }
        if (c == 2) { cr2 = r; cg2 = g; }
        if (!P.refine &&
            (fabs(r - k.r_isco) < P.edge_tol * k.r_isco || fabs(r - P.disk_r_out) < P.edge_tol * P.disk_r_out))
          fb = true;
        if (r >= k.r_isco && r <= P.disk_r_out) {
          hit = c;
          if (!P.refine) {
            const double* lut = P.luts + static_cast<long long>(k.lut_index) * P.lut_n;
            const double F = lut_linear(lut, P.lut_n, k.r_isco, P.disk_r_out, r);
            I = disk_intensity(F, g, P.nu, P.T_max);
          }
        }
      }
    }
    if (!P.refine && hit < 0 && fabsf(cap - rintf(cap)) > P.amb) fb = true;
    const int cls = hit >= 0 ? BH_HIT_DISK : (cap > 0.5f ? BH_HIT_HORIZON : BH_HIT_ESCAPED);
    const int ncr = (band && P.band_solve) ? 0 : (hit >= 0 ? hit + 1 : nc);
    double r_exit = 0.0, u_exit = 0.0, phi_exit = 0.0;
    if (cls == BH_HIT_DISK) {
      r_exit = hit == 0 ? cr0 : (hit == 1 ? cr1 : cr2);
    } else if (cls == BH_HIT_HORIZON) {
      r_exit = k.r_stop;
    } else {
      r_exit = P.r_escape;
    }
    // the escape direction: for the network's escaped rays, and with the solve
    // to follow for every ray (the solve keeps it where it finds the ray
    // escaping; it computes crossings and the class, not the path's end)
    if (cls == BH_HIT_ESCAPED || P.refine) {
      u_exit = O[BH_SUR_ESC_U];
      phi_exit = atan2f(O[BH_SUR_ESC_SINPHI], O[BH_SUR_ESC_COSPHI]);
    }
    BH_PixelRecord* rec = P.out + gid;
    double2* d2 = reinterpret_cast<double2*>(rec);
    d2[0] = make_double2(I, r_exit);
    d2[1] = make_double2(u_exit, phi_exit);
    d2[2] = make_double2(cr0, cr1);
    d2[3] = make_double2(cr2, cg0);
    d2[4] = make_double2(cg1, cg2);
    d2[5] = make_double2(0.0, 0.0);
    int* i4 = reinterpret_cast<int*>(rec);
    *reinterpret_cast<int4*>(i4 + 24) = make_int4(0, 0, 0, 0);  // drift_carter, steps, attempts
    const int flags = BH_FLAG_SURROGATE | (fb ? BH_FLAG_FALLBACK : 0) |
                      (keep ? BH_FLAG_NETWORK_KEPT | (cls == BH_HIT_ESCAPED ? BH_FLAG_EXIT_NETWORK : 0) : 0);
    *reinterpret_cast<int4*>(i4 + 28) = make_int4(cls | (ncr << 8) | (flags << 16), __float_as_int(conf), 0, 0);
    aud = !fb && P.audit_every > 0 &&
          (hash32(static_cast<uint64_t>(gid)) % static_cast<uint32_t>(P.audit_every)) == 0u;
  }
  wave_append(fb, gid, P.fb_ids, P.fb_count, P.fb_cap);
  wave_append(aud, gid, P.audit_ids, P.audit_count, P.audit_cap);
}

// Rescaled parameters of the fused kernel: [W_hi | W_hi | W_lo | 0] of
// L W_0 (256 outputs x 16 live inputs at a row stride of ld) and L b_l for
// the n_bias = n_hidden + 1 SiLU layers.
struct BiasPtrs {
  const float* p[kMaxLayers];
};

// Hidden layers 1..n_hidden (row-major [out][in] BF16) in fragment order:
// layer j's fragment of feature block fb and K-step ks for lane l is the 8
// weights W[32 fb + l % 32][16 ks + 8 (l / 32) .. + 7], stored at
// out[j kW^2 + ((fb 16 + ks) 64 + l) 8].  Grid n_hidden x 8 x 16, 64 lanes.
struct LayerPtrs {
  const uint16_t* p[kMaxLayers];
};

__global__ void k_swizzle(const LayerPtrs W, uint16_t* __restrict__ out) {
  const int b = blockIdx.x, l = threadIdx.x;
  const int j = b / 128, fb = (b / 16) % 8, ks = b % 16;
  const uint16_t* src = W.p[j] + (32 * fb + (l & 31)) * kW + 16 * ks + 8 * (l >> 5);
  *reinterpret_cast<uint4*>(out + static_cast<size_t>(j) * kW * kW + ((fb * 16 + ks) * 64 + l) * 8) =
      *reinterpret_cast<const uint4*>(src);
}

__global__ void k_prep(const float* __restrict__ W0, int ld, const BiasPtrs biases, int n_bias,
                       uint16_t* __restrict__ w0cat, float* __restrict__ bs) {
  const int o = blockIdx.x, k = threadIdx.x;  // 256 x 64
  uint16_t v = 0;
  if (k < 32) {
    v = mlp::f2bf(kLog2e * W0[o * ld + (k & 15)]);
  } else if (k < 48) {
    const float w = kLog2e * W0[o * ld + (k - 32)];
    v = mlp::f2bf(w - mlp::bf2f(mlp::f2bf(w)));
  }
  w0cat[o * 64 + k] = v;
  if (k < n_bias) bs[k * kW + o] = kLog2e * biases.p[k][o];
}

__global__ void k_zero_u64(unsigned long long* a, unsigned long long* b) {
  if (threadIdx.x == 0) {
    *a = 0ull;
    *b = 0ull;
  }
}

__global__ void k_zero_f64(double* p, int n) {
  if (static_cast<int>(threadIdx.x) < n) p[threadIdx.x] = 0.0;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off));
  return v;
}

__device__ __forceinline__ double hit_radius(const BH_PixelRecord& r) {
  const int n = r.n_cross < 3 ? r.n_cross : 3;
  return n > 0 ? r.cross_r[n - 1] : 0.0;
}

__global__ void k_audit(const BH_PixelRecord* __restrict__ out, const long long* __restrict__ ids,
                        const unsigned long long* __restrict__ count, long long cap,
                        const BH_PixelRecord* __restrict__ traced, double floor_I, double* __restrict__ stats) {
  const long long n = min(static_cast<long long>(*count), cap);
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long long q = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; q < n;
       q += static_cast<long long>(gridDim.x) * blockDim.x) {
    const BH_PixelRecord& a = out[ids[q]];
    const BH_PixelRecord& b = traced[q];
    const double dI = fabs(a.intensity - b.intensity);
    s[0] += 1.0;
    if (b.intensity > floor_I) {
      s[1] += 1.0;
      const double rel = dI / b.intensity;
      s[2] += rel;
      s[3] = fmax(s[3], rel);
    }
    s[4] += (a.hit_class != b.hit_class) ? 1.0 : 0.0;
    s[5] += dI;
    s[6] += b.intensity;
    if (a.hit_class == BH_HIT_DISK && b.hit_class == BH_HIT_DISK) s[7] += fabs(log(hit_radius(a) / hit_radius(b)));
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = (k == 3) ? wave_max(s[k]) : wave_sum(s[k]);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k == 3) {
        atomicMax(reinterpret_cast<unsigned long long*>(stats + 3),
                  static_cast<unsigned long long>(__double_as_longlong(s[3])));
      } else if (s[k] != 0.0) {
        atomicAdd(stats + k, s[k]);
      }
    }
  }
}

}  // namespace sinf
}  // namespace bhk

using namespace bhk;
using namespace bhk::sinf;

namespace {

// Critical curve of a distant observer, ordered around the curve: upper
// branch (beta >= 0) for r from r_pro to r_retro, lower branch back
// (criticalImpactParams, src/physics/analytic_kerr_geodesic.h:410-448).
void critical_curve(double a, double th, int n, std::vector<double>* xa, std::vector<double>* xb) {
  xa->clear();
  xb->clear();
  if (a < 1e-4) {
    const double b = 3.0 * std::sqrt(3.0);
    for (int i = 0; i < 2 * n; ++i) {
      const double t = kTwoPi * i / (2.0 * n);
      xa->push_back(b * std::cos(t));
      xb->push_back(b * std::sin(t));
    }
    return;
  }
  const double rp = 2.0 * (1.0 + std::cos((2.0 / 3.0) * std::acos(-a)));
  const double rm = 2.0 * (1.0 + std::cos((2.0 / 3.0) * std::acos(a)));
  const double s = std::sin(th), c = std::cos(th);
  std::vector<double> ua, ub;
  for (int i = 0; i < n; ++i) {
    const double r = rp + (rm - rp) * 0.5 * (1.0 - std::cos(kPi * i / (n - 1.0)));
    const double delta = r * r - 2.0 * r + a * a;
    const double dp = 2.0 * r - 2.0;
    const double xi = (r * r + a * a) / a - 4.0 * r * delta / (a * dp);
    const double eta = 16.0 * r * r * delta / (dp * dp) - (xi - a) * (xi - a);
    const double b2 = eta + a * a * c * c - xi * xi * (c / s) * (c / s);
    if (b2 < 0.0) continue;
    ua.push_back(-xi / s);
    ub.push_back(std::sqrt(b2));
  }
  for (size_t i = 0; i < ua.size(); ++i) {
    xa->push_back(ua[i]);
    xb->push_back(ub[i]);
  }
  for (size_t i = ua.size(); i-- > 0;) {
    xa->push_back(ua[i]);
    xb->push_back(-ub[i]);
  }
}

}  // namespace

extern "C" {

int bh_surrogate_ring_tables(const BH_Camera* cams, int n_cams, float* host_out) {
  if (!cams || n_cams <= 0 || !host_out) return fail("bh_surrogate_ring_tables: invalid arguments");
  constexpr int nb = BH_SUR_RING_BINS;
  for (int ci = 0; ci < n_cams; ++ci) {
    const BH_Camera& cm = cams[ci];
    if (!(cm.spin >= 0.0 && cm.spin < 1.0) || !(cm.theta_obs > 0.0 && cm.theta_obs < kPi))
      return fail("bh_surrogate_ring_tables: invalid camera");
    std::vector<double> xa, xb;
    critical_curve(cm.spin, cm.theta_obs, 8192, &xa, &xb);
    const size_t m = xa.size();
    if (m < 16) return fail("bh_surrogate_ring_tables: degenerate critical curve");
    const double amin = *std::min_element(xa.begin(), xa.end());
    const double amax = *std::max_element(xa.begin(), xa.end());
    const double ac = 0.5 * (amin + amax), bc = 0.0;
    std::vector<double> lo(nb, 1e30), hi(nb, -1e30), cmin(nb, 1.0);
    for (size_t i = 0; i < m; ++i) {
      const double da = xa[i] - ac, db = xb[i] - bc;
      const double rho = std::sqrt(da * da + db * db);
      double psi = std::atan2(db, da);
      if (psi < 0.0) psi += kTwoPi;
      const int bin = std::min(static_cast<int>(psi * (nb / kTwoPi)), nb - 1);
      const size_t ip = (i + 1) % m, im = (i + m - 1) % m;
      const double ta = xa[ip] - xa[im], tb = xb[ip] - xb[im];
      const double tn = std::sqrt(ta * ta + tb * tb);
      // |cos| of the angle between the radial direction and the normal = |r^ x t^|
      const double cs = (tn > 0.0 && rho > 0.0) ? std::fabs(da * tb - db * ta) / (rho * tn) : 1.0;
      lo[bin] = std::min(lo[bin], rho);
      hi[bin] = std::max(hi[bin], rho);
      cmin[bin] = std::min(cmin[bin], cs);
    }
    // bins the sampling missed take their neighbours' range
    for (int pass = 0; pass < nb; ++pass) {
      bool any = false;
      for (int b = 0; b < nb; ++b) {
        if (hi[b] >= lo[b]) continue;
        const int l = (b + nb - 1) % nb, r = (b + 1) % nb;
        const bool okl = hi[l] >= lo[l], okr = hi[r] >= lo[r];
        if (okl || okr) {
          lo[b] = std::min(okl ? lo[l] : 1e30, okr ? lo[r] : 1e30);
          hi[b] = std::max(okl ? hi[l] : -1e30, okr ? hi[r] : -1e30);
          cmin[b] = std::min(cmin[l], cmin[r]);
        } else {
          any = true;
        }
      }
      if (!any) break;
    }
    float* o = host_out + static_cast<size_t>(ci) * BH_SUR_RING_FLOATS;
    o[0] = static_cast<float>(ac);
    o[1] = static_cast<float>(bc);
    for (int b = 0; b < nb; ++b) {
      o[2 + b] = static_cast<float>(lo[b]);
      o[2 + nb + b] = static_cast<float>(hi[b]);
      o[2 + 2 * nb + b] = static_cast<float>(1.0 / std::max(cmin[b], 0.2));
    }
  }
  return 0;
}

int bh_surrogate_render(BH_Surrogate* s, const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                        const double* disk_luts, int disk_lut_n, const float* ring_tables,
                        const BH_SurrogateRenderParams* rp, BH_PixelRecord* out, int64_t* fallback_ids,
                        uint64_t* fallback_count, int64_t fallback_cap, int64_t* audit_ids,
                        uint64_t* audit_count, int64_t audit_cap, void* stream) {
  if (!s || !cams || n_cams <= 0 || !opts || !disk_luts || disk_lut_n < 2 || !ring_tables || !rp || !out ||
      !fallback_ids || !fallback_count || !audit_ids || !audit_count)
    return fail("bh_surrogate_render: NULL argument");
  if (fallback_cap <= 0 || audit_cap <= 0) return fail("bh_surrogate_render: list capacities must be > 0");
  if (!(rp->ring_band >= 0.0f && rp->edge_tol >= 0.0f && rp->ambiguity >= 0.0f) || rp->audit_every < 0 ||
      rp->refine < 0 || rp->refine > 3 || !(rp->keep_factor == 0.0f || rp->keep_factor >= 1.0f))
    return fail("bh_surrogate_render: invalid render parameters");
  if (opts->mino_exit < 0 || opts->mino_exit > 2) return fail("bh_surrogate_render: mino_exit must be 0, 1 or 2");
  const SurView v = surrogate_view(s);
  if (v.width != kW || v.L < 2 || v.L > kMaxLayers)
    return fail("bh_surrogate_render: the fused kernel needs width 256 and 0..6 hidden 256 x 256 layers");
  for (int i = 0; i < n_cams; ++i)
    if (std::fabs(cams[i].r_obs - static_cast<double>(kSurRobs)) > 1e-9)
      return fail("bh_surrogate_render: cameras must have r_obs = 100 M (the surrogate's training distance)");
  std::vector<CamDev> hc;
  long long total = 0;
  if (make_cam_devs(cams, n_cams, opts->horizon_factor, "bh_surrogate_render", &hc, &total) != 0) return -1;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int n_bias = v.L - 1;  // SiLU layers 0..n_hidden
  const size_t w0_bytes = 256 * 64 * sizeof(uint16_t);
  const size_t bs_bytes = static_cast<size_t>(kMaxLayers) * kW * sizeof(float);
  const int n_hidden = v.L - 2;
  const size_t wsw_bytes = static_cast<size_t>(n_hidden) * kW * kW * sizeof(uint16_t);
  const size_t bytes = w0_bytes + bs_bytes + wsw_bytes + hc.size() * sizeof(CamDev);
  const size_t cap = std::max<size_t>(bytes, w0_bytes + bs_bytes + wsw_bytes + 64 * sizeof(CamDev));
  char* scratch = static_cast<char*>(surrogate_render_scratch(s, stream, cap));
  if (!scratch) return fail("bh_surrogate_render: hipMalloc failed");
  uint16_t* d_w0 = reinterpret_cast<uint16_t*>(scratch);
  float* d_bs = reinterpret_cast<float*>(scratch + w0_bytes);
  uint16_t* d_wsw = reinterpret_cast<uint16_t*>(scratch + w0_bytes + bs_bytes);
  CamDev* d_cams = reinterpret_cast<CamDev*>(scratch + w0_bytes + bs_bytes + wsw_bytes);
  // Pageable source: the runtime stages it before returning.
  if (hipMemcpyAsync(d_cams, hc.data(), hc.size() * sizeof(CamDev), hipMemcpyHostToDevice, st) != hipSuccess)
    return fail("bh_surrogate_render: camera upload failed");
  BiasPtrs bp;
  std::memset(&bp, 0, sizeof(bp));
  for (int l = 0; l < n_bias; ++l) bp.p[l] = v.p32 + v.b_off[l];
  hipLaunchKernelGGL(k_prep, dim3(256), dim3(64), 0, st, v.p32 + v.w_off[0], BH_SUR_IN_PAD, bp, n_bias, d_w0, d_bs);
  if (n_hidden > 0) {
    LayerPtrs lp;
    std::memset(&lp, 0, sizeof(lp));
    for (int l = 1; l <= n_hidden; ++l) lp.p[l - 1] = v.pbf + v.w_off[l];
    hipLaunchKernelGGL(k_swizzle, dim3(n_hidden * 8 * 16), dim3(64), 0, st, lp, d_wsw);
  }
  InferParams P;
  std::memset(&P, 0, sizeof(P));
  P.cams = d_cams;
  P.luts = disk_luts;
  P.ring = ring_tables;
  P.w0cat = d_w0;
  P.bs = d_bs;
  for (int l = 0; l < v.L; ++l) {
    P.w[l] = v.pbf + v.w_off[l];
    P.b[l] = v.p32 + v.b_off[l];
  }
  P.wsw = d_wsw;
  P.out = out;
  P.fb_ids = reinterpret_cast<long long*>(fallback_ids);
  P.fb_count = reinterpret_cast<unsigned long long*>(fallback_count);
  P.audit_ids = reinterpret_cast<long long*>(audit_ids);
  P.audit_count = reinterpret_cast<unsigned long long*>(audit_count);
  P.n_total = total;
  P.fb_cap = fallback_cap;
  P.audit_cap = audit_cap;
  P.disk_r_out = opts->disk_r_out;
  P.nu = opts->nu_obs;
  P.T_max = opts->disk_T_max;
  P.r_escape = opts->r_escape;
  P.n_cams = n_cams;
  P.lut_n = disk_lut_n;
  P.n_hidden = v.L - 2;
  P.audit_every = rp->audit_every;
  P.ring_band = rp->ring_band;
  P.edge_tol = rp->edge_tol;
  P.conf_max = rp->conf_max;
  P.amb = rp->ambiguity;
  P.refine = rp->refine != 0;
  P.band_solve = rp->refine == 3;
  P.keep_r = (P.refine && opts->mino_exit != 1) ? static_cast<double>(rp->keep_factor) * opts->disk_r_out : 0.0;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, P.fb_count, P.audit_count);
  const long long nblk = (total + kTile - 1) / kTile;
  if (nblk > 0x7fffffffll) return fail("bh_surrogate_render: too many pixels");
  hipLaunchKernelGGL(k_sur_render, dim3(static_cast<unsigned>(nblk)), dim3(256), 0, st, P);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && rp->refine == 1) {
    MinoParams M;
    std::memset(&M, 0, sizeof(M));
    M.cams = d_cams;
    M.luts = disk_luts;
    M.out = out;
    M.fb_ids = P.fb_ids;
    M.fb_count = P.fb_count;
    M.fb_cap = fallback_cap;
    M.n_total = total;
    M.disk_r_out = opts->disk_r_out;
    M.nu = opts->nu_obs;
    M.T_max = opts->disk_T_max;
    M.r_escape = opts->r_escape;
    M.tol = kMinoTol;
    M.n_cams = n_cams;
    M.lut_n = disk_lut_n;
    M.refine = 1;
    M.exit_dir = opts->mino_exit == 1;
    e = launch_mino_disk(M, st);
  }
  return e == hipSuccess ? 0 : fail(std::string("bh_surrogate_render: launch failed: ") + hipGetErrorString(e));
}

int bh_surrogate_audit(const BH_PixelRecord* out, const int64_t* ids, const uint64_t* count, int64_t cap,
                       const BH_PixelRecord* traced, double floor_I, double* stats, void* stream) {
  if (!out || !ids || !count || cap <= 0 || !traced || !stats) return fail("bh_surrogate_audit: invalid arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st, stats, 8);
  hipLaunchKernelGGL(k_audit, dim3(256), dim3(256), 0, st, out, reinterpret_cast<const long long*>(ids),
                     reinterpret_cast<const unsigned long long*>(count), static_cast<long long>(cap), traced,
                     floor_I, stats);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(std::string("bh_surrogate_audit: launch failed: ") + hipGetErrorString(e));
}

}  // extern "C"
