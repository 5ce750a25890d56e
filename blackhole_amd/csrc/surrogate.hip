// surrogate.hip - native training and inference driver of the geodesic
// surrogate MLP (host orchestration of the kernels in mlp_kernels.hpp, plus the
// analytic feature embedding).  Everything stays on the device: FP32 master
// weights, BF16 working copies (row-major [out][in] and transposed
// [in][out]), Adam moments, activations and gradients.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bhk_surrogate.h"

#include "mlp_kernels.hpp"
#include "sur_features.hpp"

namespace bhk {
int fail(const std::string& msg);

__global__ void k_features(const float* __restrict__ rays4, long long n, float* __restrict__ feat) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float f[BH_SUR_FEATURES];
  surrogate_features(rays4[4 * i], rays4[4 * i + 1], rays4[4 * i + 2], rays4[4 * i + 3], f);
  float* o = feat + i * BH_SUR_IN_PAD;
#pragma unroll
  for (int k = 0; k < BH_SUR_FEATURES; ++k) o[k] = f[k];
#pragma unroll
  for (int k = BH_SUR_FEATURES; k < BH_SUR_IN_PAD; ++k) o[k] = 0.0f;
}

// Training targets and loss weights from traced records (disk disabled, so
// every equatorial crossing is recorded; cross_g = 0 inside the ISCO).
__global__ void k_targets(const BH_PixelRecord* __restrict__ rec, const float* __restrict__ feat, long long n,
                          float* __restrict__ T, float* __restrict__ W, float w_r, float w_g, float w_n,
                          float w_c, float w_e) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const BH_PixelRecord r = rec[i];
  float* t = T + i * BH_SUR_OUT;
  float* w = W + i * BH_SUR_OUT;
#pragma unroll
  for (int c = 0; c < BH_SUR_OUT; ++c) {
    t[c] = 0.0f;
    w[c] = 0.0f;
  }
  const int nc = r.n_cross;
  const double lnbase = 3.0 * static_cast<double>(feat[i * BH_SUR_IN_PAD + kSurBaseFeature]);
  for (int k = 0; k < 3; ++k) {
    if (nc > k && r.cross_r[k] > 0.0) {
      t[2 * k] = static_cast<float>(log(r.cross_r[k]) - lnbase);
      w[2 * k] = w_r;
      if (r.cross_g[k] > 0.0) {
        t[2 * k + 1] = static_cast<float>(r.cross_g[k]);
        w[2 * k + 1] = w_g;
      }
    }
  }
  t[BH_SUR_NCROSS] = static_cast<float>(nc < 3 ? nc : 3);
  w[BH_SUR_NCROSS] = w_n;
  const bool cap = r.hit_class == BH_HIT_HORIZON;
  t[BH_SUR_CAPTURE] = cap ? 1.0f : 0.0f;
  w[BH_SUR_CAPTURE] = (r.hit_class == BH_HIT_HORIZON || r.hit_class == BH_HIT_ESCAPED) ? w_c : 0.0f;
  if (r.hit_class == BH_HIT_ESCAPED) {
    double sp, cp;
    sincos(r.phi_exit, &sp, &cp);
    t[BH_SUR_ESC_U] = static_cast<float>(r.u_exit);
    t[BH_SUR_ESC_SINPHI] = static_cast<float>(sp);
    t[BH_SUR_ESC_COSPHI] = static_cast<float>(cp);
    w[BH_SUR_ESC_U] = w[BH_SUR_ESC_SINPHI] = w[BH_SUR_ESC_COSPHI] = w_e;
  }
}

}  // namespace bhk

using namespace bhk;
using namespace bhk::mlp;

struct BH_Surrogate {
  int width = 256, depth = 6, max_batch = 0, L = 7;
  std::vector<int> in_pad, out_dim, out_pad;
  std::vector<long long> w_off, b_off;
  long long n_params = 0;
  float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;
  uint16_t* d_pbf = nullptr;                 // BF16 mirror of the parameter vector
  float* d_x0 = nullptr;                     // FP32 input features [B][32]
  std::vector<uint16_t*> d_wt;               // per layer BF16 W^T [in_pad][out_pad]
  std::vector<uint16_t*> d_x;                // activations X_l BF16 [B][in_pad_l], l = 0..L
  std::vector<float*> d_z;                   // pre-activations Z_l FP32 [B][width], l < L
  float* d_y = nullptr;                      // outputs FP32 [B][16]
  float* d_dx = nullptr;                     // FP32 [B][width]
  uint16_t* d_dz = nullptr;                  // BF16 [B][width]
  uint16_t* d_dout = nullptr;                // BF16 [B][32]
  uint16_t* d_t1 = nullptr;                  // transpose scratch [width][B]
  uint16_t* d_t2 = nullptr;                  // transpose scratch [width][B]
  float* d_loss = nullptr;                   // 2 floats
  float* d_part = nullptr;                   // partial sums of the split reductions
  long long t = 0;                           // Adam step
  // Rendering scratch (sur_infer.hip), one buffer per (device, stream):
  // renders of one surrogate on different streams never share it.
  struct RenderBuf {
    void* d = nullptr;
    size_t cap = 0;
  };
  std::mutex render_mu;
  std::map<std::pair<int, void*>, RenderBuf> render_bufs;
};

namespace bhk {

// Device parameters of a surrogate for the fused rendering kernel.
struct SurView {
  const float* p32;
  const uint16_t* pbf;
  const long long* w_off;
  const long long* b_off;
  int width, L;
};

SurView surrogate_view(BH_Surrogate* s) {
  return SurView{s->d_p, s->d_pbf, s->w_off.data(), s->b_off.data(), s->width, s->L};
}

// The surrogate's rendering scratch for (current device, stream), at least
// `bytes` long; nullptr if it cannot be allocated.  Growing a buffer frees
// the old one (hipFree waits for the work that still reads it).
void* surrogate_render_scratch(BH_Surrogate* s, void* stream, size_t bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(s->render_mu);
  BH_Surrogate::RenderBuf& b = s->render_bufs[{dev, stream}];
  if (b.cap < bytes) {
    if (b.d) (void)hipFree(b.d);
    b.d = nullptr;
    b.cap = 0;
    if (hipMalloc(&b.d, bytes) != hipSuccess) return nullptr;
    b.cap = bytes;
  }
  return b.d;
}

}  // namespace bhk

namespace {

template <class T>
bool dalloc(T** p, size_t n) {
  return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T)) == hipSuccess;
}

inline dim3 grid1(long long n, int b = 256) { return dim3(static_cast<unsigned>((n + b - 1) / b)); }

void refresh_bf16(BH_Surrogate* s, hipStream_t st) {
  hipLaunchKernelGGL(k_f32_to_bf16, grid1(s->n_params), dim3(256), 0, st, s->n_params, s->d_p, s->d_pbf);
  for (int l = 0; l < s->L; ++l) {
    const int rows = s->out_dim[l], cols = s->in_pad[l];
    hipLaunchKernelGGL(k_transpose_bf16, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(256), 0, st, rows,
                       cols, s->d_pbf + s->w_off[l], cols, s->d_wt[l], s->out_pad[l]);
  }
}

// Forward through all layers for n rows (FP32 features in d_x0).  Layer 0
// runs in FP32 on the VALU (inputs unquantised), layers 1..L on BF16 MFMA.
// keep_z: store the pre-activations (training).
void forward(BH_Surrogate* s, int n, bool keep_z, hipStream_t st) {
  hipLaunchKernelGGL(k_layer0_fwd, dim3((n + kL0Rows - 1) / kL0Rows, (s->out_dim[0] + 255) / 256), dim3(256), 0,
                     st, n, s->out_dim[0], BH_SUR_FEATURES, s->d_x0, BH_SUR_IN_PAD, s->d_p + s->w_off[0],
                     BH_SUR_IN_PAD, s->d_p + s->b_off[0], keep_z ? s->d_z[0] : nullptr, s->d_x[1]);
  for (int l = 1; l < s->L; ++l) {
    const int N = s->out_dim[l], K = s->in_pad[l];
    const dim3 g((N + 63) / 64, (n + 63) / 64);
    const uint16_t* W = s->d_pbf + s->w_off[l];
    const float* b = s->d_p + s->b_off[l];
    if (l < s->L - 1)
      hipLaunchKernelGGL(k_gemm_nt<EPI_BIAS_SILU>, g, dim3(256), 0, st, n, N, K, s->d_x[l], K, W, K, b,
                         keep_z ? s->d_z[l] : nullptr, N, s->d_x[l + 1], s->width);
    else
      hipLaunchKernelGGL(k_gemm_nt<EPI_F32>, g, dim3(256), 0, st, n, N, K, s->d_x[l], K, W, K, b, s->d_y,
                         BH_SUR_OUT, static_cast<uint16_t*>(nullptr), 0);
  }
}


// Loss + backward into d_g.
void backward(BH_Surrogate* s, int n, const float* T, const float* Wt, hipStream_t st) {
  (void)hipMemsetAsync(s->d_loss, 0, 2 * sizeof(float), st);
  hipLaunchKernelGGL(k_loss, grid1(n), dim3(256), 0, st, n, BH_SUR_OUT, s->d_y, BH_SUR_OUT, T, Wt,
                     BH_SUR_OUT, static_cast<int>(BH_SUR_CONF), 0.01f, 1.0f / static_cast<float>(n),
                     s->d_dout, 32, s->d_loss);
  const int W = s->width;
  for (int l = s->L - 1; l >= 0; --l) {
    const int N_out = s->out_dim[l], K_in = s->in_pad[l];
    const uint16_t* dZ;  // BF16 [n][ld]
    int ldz;
    if (l == s->L - 1) {
      dZ = s->d_dout;
      ldz = 32;
    } else {
      hipLaunchKernelGGL(k_silu_bwd, grid1(static_cast<long long>(n) * W), dim3(256), 0, st,
                         static_cast<long long>(n) * W, W, W, s->d_dx, s->d_z[l], W, s->d_dz, W);
      dZ = s->d_dz;
      ldz = W;
    }
    const int n_chunks = (n + kColRows - 1) / kColRows;
    if (l == 0) {
      // FP32 first layer: dW0 from the FP32 features
      hipLaunchKernelGGL(k_layer0_wgrad_part, dim3((N_out + 63) / 64, n_chunks), dim3(256), 0, st, n, N_out,
                         BH_SUR_FEATURES, dZ, ldz, s->d_x0, BH_SUR_IN_PAD, s->d_part, K_in);
      hipLaunchKernelGGL(k_sum_partials, grid1(static_cast<long long>(N_out) * K_in), dim3(256), 0, st,
                         static_cast<long long>(N_out) * K_in, n_chunks, s->d_part, s->d_g + s->w_off[0]);
      // bias gradient: row sums of dZ0^T
      hipLaunchKernelGGL(k_transpose_bf16, dim3((N_out + 31) / 32, (n + 31) / 32), dim3(256), 0, st, n, N_out,
                         dZ, ldz, s->d_t1, n);
      hipLaunchKernelGGL(k_rowsum_bf16, dim3(N_out), dim3(256), 0, st, n, s->d_t1, n, s->d_g + s->b_off[0]);
      continue;
    }
    // grad W_l [N_out][K_in] = dZ^T X_l : A = dZ^T [N_out][n], B = X_l^T [K_in][n]
    hipLaunchKernelGGL(k_transpose_bf16, dim3((N_out + 31) / 32, (n + 31) / 32), dim3(256), 0, st, n, N_out,
                       dZ, ldz, s->d_t1, n);
    hipLaunchKernelGGL(k_transpose_bf16, dim3((K_in + 31) / 32, (n + 31) / 32), dim3(256), 0, st, n, K_in,
                       s->d_x[l], K_in, s->d_t2, n);
    if (n % kColRows == 0 && n_chunks > 1) {
      // split K = n over row chunks (the 256 x 256 result alone is only 16 tiles)
      hipLaunchKernelGGL(k_gemm_nt<EPI_F32>, dim3((K_in + 63) / 64, (N_out + 63) / 64, n_chunks), dim3(256), 0, st,
                         N_out, K_in, kColRows, s->d_t1, n, s->d_t2, n, static_cast<const float*>(nullptr),
                         s->d_part, K_in, static_cast<uint16_t*>(nullptr), 0);
      hipLaunchKernelGGL(k_sum_partials, grid1(static_cast<long long>(N_out) * K_in), dim3(256), 0, st,
                         static_cast<long long>(N_out) * K_in, n_chunks, s->d_part, s->d_g + s->w_off[l]);
    } else {
      hipLaunchKernelGGL(k_gemm_nt<EPI_F32>, dim3((K_in + 63) / 64, (N_out + 63) / 64), dim3(256), 0, st, N_out,
                         K_in, n, s->d_t1, n, s->d_t2, n, static_cast<const float*>(nullptr),
                         s->d_g + s->w_off[l], K_in, static_cast<uint16_t*>(nullptr), 0);
    }
    hipLaunchKernelGGL(k_rowsum_bf16, dim3(N_out), dim3(256), 0, st, n, s->d_t1, n, s->d_g + s->b_off[l]);
    if (l > 0) {
      // dX_l [n][K_in] = dZ W_l : A = dZ [n][out_pad], B = W_l^T [K_in][out_pad]
      hipLaunchKernelGGL(k_gemm_nt<EPI_F32>, dim3((K_in + 63) / 64, (n + 63) / 64), dim3(256), 0, st, n,
                         K_in, s->out_pad[l], dZ, ldz, s->d_wt[l], s->out_pad[l],
                         static_cast<const float*>(nullptr), s->d_dx, W, static_cast<uint16_t*>(nullptr), 0);
    }
  }
}

int load_features(BH_Surrogate* s, const float* features, int n, hipStream_t st) {
  if (!s || !features || n <= 0 || n > s->max_batch) return fail("surrogate: invalid batch");
  if (hipMemcpyAsync(s->d_x0, features, static_cast<size_t>(n) * BH_SUR_IN_PAD * sizeof(float),
                     hipMemcpyDeviceToDevice, st) != hipSuccess)
    return fail("surrogate: feature copy failed");
  return 0;
}

}  // namespace

extern "C" {

BH_Surrogate* bh_surrogate_create(const BH_SurrogateConfig* cfg) {
  if (!cfg || cfg->width <= 0 || cfg->width % 64 != 0 || cfg->depth < 1 || cfg->max_batch <= 0 ||
      cfg->max_batch % 64 != 0) {
    fail("bh_surrogate_create: width and max_batch must be positive multiples of 64, depth >= 1");
    return nullptr;
  }
  BH_Surrogate* s = new BH_Surrogate();
  s->width = cfg->width;
  s->depth = cfg->depth;
  s->max_batch = cfg->max_batch;
  s->L = cfg->depth + 1;
  const int W = s->width, B = s->max_batch;
  for (int l = 0; l < s->L; ++l) {
    s->in_pad.push_back(l == 0 ? BH_SUR_IN_PAD : W);
    s->out_dim.push_back(l == s->L - 1 ? BH_SUR_OUT : W);
    s->out_pad.push_back(l == s->L - 1 ? 32 : W);
  }
  long long off = 0;
  for (int l = 0; l < s->L; ++l) {
    s->w_off.push_back(off);
    off += static_cast<long long>(s->out_dim[l]) * s->in_pad[l];
    s->b_off.push_back(off);
    off += s->out_dim[l];
  }
  s->n_params = off;
  bool ok = dalloc(&s->d_p, off) && dalloc(&s->d_g, off) && dalloc(&s->d_m, off) && dalloc(&s->d_v, off) &&
            dalloc(&s->d_pbf, off);
  s->d_wt.resize(s->L, nullptr);
  for (int l = 0; l < s->L && ok; ++l) ok = dalloc(&s->d_wt[l], static_cast<size_t>(s->in_pad[l]) * s->out_pad[l]);
  s->d_x.resize(s->L + 1, nullptr);
  for (int l = 0; l <= s->L && ok; ++l)
    ok = dalloc(&s->d_x[l], static_cast<size_t>(B) * (l == 0 ? BH_SUR_IN_PAD : W));
  s->d_z.resize(s->L - 1, nullptr);
  for (int l = 0; l < s->L - 1 && ok; ++l) ok = dalloc(&s->d_z[l], static_cast<size_t>(B) * W);
  ok = ok && dalloc(&s->d_x0, static_cast<size_t>(B) * BH_SUR_IN_PAD) && dalloc(&s->d_y, static_cast<size_t>(B) * BH_SUR_OUT) && dalloc(&s->d_dx, static_cast<size_t>(B) * W) &&
       dalloc(&s->d_dz, static_cast<size_t>(B) * W) && dalloc(&s->d_dout, static_cast<size_t>(B) * 32) &&
       dalloc(&s->d_t1, static_cast<size_t>(B) * W) && dalloc(&s->d_t2, static_cast<size_t>(B) * W) &&
       dalloc(&s->d_loss, 2) &&
       dalloc(&s->d_part, static_cast<size_t>((B + kColRows - 1) / kColRows) *
                              std::max<size_t>(static_cast<size_t>(W) * W, static_cast<size_t>(W) * BH_SUR_IN_PAD));
  if (!ok) {
    fail("bh_surrogate_create: hipMalloc failed");
    bh_surrogate_destroy(s);
    return nullptr;
  }
  for (int l = 0; l < s->L; ++l)
    (void)hipMemset(s->d_wt[l], 0, static_cast<size_t>(s->in_pad[l]) * s->out_pad[l] * sizeof(uint16_t));
  // He-style initialisation: W ~ N(0, 2 / fan_in) (SplitMix64 + Box-Muller), b = 0.
  std::vector<float> p(off, 0.0f);
  uint64_t st = cfg->seed ? cfg->seed : 0x5EED5EEDull;
  auto next = [&st]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  for (int l = 0; l < s->L; ++l) {
    const int fan_in = (l == 0) ? BH_SUR_FEATURES : s->in_pad[l];
    const double sd = std::sqrt((l == s->L - 1 ? 1.0 : 2.0) / fan_in);
    for (int o = 0; o < s->out_dim[l]; ++o)
      for (int i = 0; i < s->in_pad[l]; ++i) {
        const double u1 = (static_cast<double>(next() >> 11) + 0.5) / 9007199254740992.0;
        const double u2 = (static_cast<double>(next() >> 11) + 0.5) / 9007199254740992.0;
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        p[s->w_off[l] + static_cast<long long>(o) * s->in_pad[l] + i] =
            (l == 0 && i >= BH_SUR_FEATURES) ? 0.0f : static_cast<float>(sd * z);
      }
  }
  if (bh_surrogate_set_params(s, p.data()) != 0) {
    bh_surrogate_destroy(s);
    return nullptr;
  }
  (void)hipMemset(s->d_m, 0, off * sizeof(float));
  (void)hipMemset(s->d_v, 0, off * sizeof(float));
  (void)hipDeviceSynchronize();
  return s;
}

void bh_surrogate_destroy(BH_Surrogate* s) {
  if (!s) return;
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  f(s->d_p); f(s->d_g); f(s->d_m); f(s->d_v); f(s->d_pbf);
  for (auto p : s->d_wt) f(p);
  for (auto p : s->d_x) f(p);
  for (auto p : s->d_z) f(p);
  f(s->d_x0); f(s->d_y); f(s->d_dx); f(s->d_dz); f(s->d_dout); f(s->d_t1); f(s->d_t2); f(s->d_loss);
  f(s->d_part);
  for (auto& kv : s->render_bufs) f(kv.second.d);
  delete s;
}

int64_t bh_surrogate_param_count(const BH_Surrogate* s) { return s ? s->n_params : -1; }

int bh_surrogate_get_params(const BH_Surrogate* s, float* host_out) {
  if (!s || !host_out) return fail("bh_surrogate_get_params: NULL argument");
  if (hipMemcpy(host_out, s->d_p, s->n_params * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return fail("bh_surrogate_get_params: copy failed");
  return 0;
}

int bh_surrogate_set_params(BH_Surrogate* s, const float* host_in) {
  if (!s || !host_in) return fail("bh_surrogate_set_params: NULL argument");
  if (hipMemcpy(s->d_p, host_in, s->n_params * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
    return fail("bh_surrogate_set_params: copy failed");
  refresh_bf16(s, nullptr);
  return hipDeviceSynchronize() == hipSuccess ? 0 : fail("bh_surrogate_set_params: refresh failed");
}

int bh_surrogate_features(const float* rays4, int64_t n, float* features, void* stream) {
  if (!rays4 || !features || n <= 0) return fail("bh_surrogate_features: invalid arguments");
  hipLaunchKernelGGL(k_features, grid1(n), dim3(256), 0, static_cast<hipStream_t>(stream), rays4,
                     static_cast<long long>(n), features);
  return hipGetLastError() == hipSuccess ? 0 : fail("bh_surrogate_features: launch failed");
}

int bh_surrogate_targets(const BH_PixelRecord* records, const float* features, int64_t n, float* targets,
                         float* weights, const float* w5, void* stream) {
  if (!records || !features || !targets || !weights || !w5 || n <= 0)
    return fail("bh_surrogate_targets: invalid arguments");
  hipLaunchKernelGGL(k_targets, grid1(n), dim3(256), 0, static_cast<hipStream_t>(stream), records, features,
                     static_cast<long long>(n), targets, weights, w5[0], w5[1], w5[2], w5[3], w5[4]);
  return hipGetLastError() == hipSuccess ? 0 : fail("bh_surrogate_targets: launch failed");
}

int bh_surrogate_forward(BH_Surrogate* s, const float* features, int n, float* out, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (load_features(s, features, n, st) != 0) return -1;
  if (!out) return fail("bh_surrogate_forward: NULL output");
  forward(s, n, false, st);
  if (hipMemcpyAsync(out, s->d_y, static_cast<size_t>(n) * BH_SUR_OUT * sizeof(float), hipMemcpyDeviceToDevice,
                     st) != hipSuccess)
    return fail("bh_surrogate_forward: copy failed");
  return hipGetLastError() == hipSuccess ? 0 : fail("bh_surrogate_forward: launch failed");
}

int bh_surrogate_gradients(BH_Surrogate* s, const float* features, const float* targets, const float* weights,
                           int n, float* host_grads, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (n % 32 != 0) return fail("surrogate: batch must be a multiple of 32");
  if (load_features(s, features, n, st) != 0) return -1;
  if (!targets || !weights || !host_grads) return fail("bh_surrogate_gradients: NULL argument");
  forward(s, n, true, st);
  backward(s, n, targets, weights, st);
  if (hipStreamSynchronize(st) != hipSuccess) return fail("bh_surrogate_gradients: kernel failure");
  if (hipMemcpy(host_grads, s->d_g, s->n_params * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return fail("bh_surrogate_gradients: copy failed");
  return 0;
}

int bh_surrogate_train_step(BH_Surrogate* s, const float* features, const float* targets, const float* weights,
                            int n, float lr, float* loss_out, void* stream) {
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (n % 32 != 0) return fail("surrogate: batch must be a multiple of 32");
  if (load_features(s, features, n, st) != 0) return -1;
  if (!targets || !weights) return fail("bh_surrogate_train_step: NULL argument");
  forward(s, n, true, st);
  backward(s, n, targets, weights, st);
  s->t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = 1.0f - std::pow(b1, static_cast<float>(s->t));
  const float bc2 = 1.0f - std::pow(b2, static_cast<float>(s->t));
  hipLaunchKernelGGL(k_adam, grid1(s->n_params), dim3(256), 0, st, s->n_params, s->d_p, s->d_g, s->d_m, s->d_v,
                     s->d_pbf, lr, b1, b2, 1e-8f, bc1, bc2, 0.0f);
  for (int l = 0; l < s->L; ++l) {
    const int rows = s->out_dim[l], cols = s->in_pad[l];
    hipLaunchKernelGGL(k_transpose_bf16, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(256), 0, st, rows, cols,
                       s->d_pbf + s->w_off[l], cols, s->d_wt[l], s->out_pad[l]);
  }
  if (loss_out) {
    if (hipMemcpyAsync(loss_out, s->d_loss, 2 * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess)
      return fail("bh_surrogate_train_step: loss copy failed");
  }
  if (hipStreamSynchronize(st) != hipSuccess) return fail("bh_surrogate_train_step: kernel failure");
  return 0;
}

int bh_mlp_gemm_nt(int M, int N, int K, const void* A, int lda, const void* Bt, int ldb,
                   const float* bias, float* C, int ldc, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0 || (K % 32) != 0 || (lda % 8) != 0 || (ldb % 8) != 0 ||
      lda < K || ldb < K || ldc < N || !A || !Bt || !C)
    return fail("bh_mlp_gemm_nt: invalid shape (K must be a multiple of 32, lda/ldb multiples of 8)");
  const dim3 grid((N + 63) / 64, (M + 63) / 64);
  hipLaunchKernelGGL(k_gemm_nt<EPI_F32>, grid, dim3(256), 0, static_cast<hipStream_t>(stream), M, N, K,
                     static_cast<const uint16_t*>(A), lda, static_cast<const uint16_t*>(Bt), ldb, bias, C,
                     ldc, static_cast<uint16_t*>(nullptr), 0);
  return hipGetLastError() == hipSuccess ? 0 : fail("bh_mlp_gemm_nt: launch failed");
}

}  // extern "C"
