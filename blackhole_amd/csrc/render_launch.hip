// render_launch.hip - the render-launch ABI (include/bh_kernel_launch.h):
// BH_LaunchParams -> a pinhole camera traced by the library's tracer, then
// a shading pass from the per-pixel records into the caller's float4
// framebuffer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bh_kernel_launch.h"
#include "grmhd_device.hpp"
#include "trace_common.hpp"

namespace bhk {
int fail(const std::string& msg);

namespace rl {

constexpr int kLutN = 256;
constexpr double kDiskOuterRs = 30.0;  // disk outer edge in units of rs

struct ShadeParams {
  const BH_PixelRecord* rec;
  float4* fb;
  long long n;
  const double* lut;
  int lut_n;
  double r_isco, r_out;
  float lit, bg;
  int redshift, bg_on;
  // GRMHD modulation (use_luts and a volume in the grmhd slot): disk colour x rho u
  unsigned long long grmhd, grmhd_right;
  float grmhd_alpha, grmhd_r_min, grmhd_r_max;
  int grmhd_on;
  double len;  // launch length unit per M (rs / 2)
};

// Framebuffer colour of one record: disk -> blackbody tint of the observed
// temperature g F^(1/4) times the bolometric brightness 2 F g^4 (g^4 only
// with redshift on), scaled by max(rho, 0) max(u, 0) of the GRMHD volume at
// the hit point (r at the crossing, theta = pi/2, phi at the end of the
// crossing step) when GRMHD is on (d_shade_hit, device_physics.cuh:1842-1868);
// horizon -> opaque black, otherwise the background.
__global__ void k_shade(const ShadeParams p) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const BH_PixelRecord& r = p.rec[i];
  float4 c;
  if (r.hit_class == BH_HIT_DISK) {
    const int k = (r.n_cross < 3 ? r.n_cross : 3) - 1;
    const double rc = k >= 0 ? r.cross_r[k] : p.r_isco;
    const double g = k >= 0 ? r.cross_g[k] : 0.0;
    const double F = lut_linear(p.lut, p.lut_n, p.r_isco, p.r_out, rc);
    const float T = static_cast<float>(g * sqrt(sqrt(F > 0.0 ? F : 0.0)));
    const double g4 = p.redshift ? (g * g) * (g * g) : 1.0;
    const float lum = p.lit * static_cast<float>(2.0 * F * g4);
    const float tg = fminf(fmaxf(0.25f + 0.75f * T, 0.0f), 1.0f);
    const float tb = fminf(T * T, 1.0f);
    c = make_float4(lum, lum * tg, lum * tb, 1.0f);
    if (p.grmhd_on) {
      const float rr = static_cast<float>(rc * p.len);
      const float ph = static_cast<float>(r.phi_exit);
      const float4 q = grmhd_sample(p.grmhd, p.grmhd_right, p.grmhd_alpha, p.grmhd_r_min, p.grmhd_r_max,
                                    rr * cosf(ph), rr * sinf(ph), 0.0f);
      const float scale = fmaxf(q.x, 0.0f) * fmaxf(q.y, 0.0f);
      c.x *= scale;
      c.y *= scale;
      c.z *= scale;
    }
  } else if (r.hit_class == BH_HIT_HORIZON) {
    c = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  } else {
    const float b = p.bg_on ? p.bg : 0.0f;
    c = make_float4(b, b, b, 0.0f);
  }
  p.fb[i] = c;
}

// Per (device, stream): the records of the last launch and the disk LUT.
struct RenderCtx {
  BH_PixelRecord* rec = nullptr;
  size_t rec_cap = 0;
  double* lut = nullptr;
  double lut_key[3] = {-1.0, -1.0, -1.0};  // spin, r_isco, r_out of the uploaded LUT
};
std::mutex g_mu;
std::map<std::pair<int, void*>, RenderCtx> g_ctx;
std::mutex g_tex_mu;
unsigned long long g_tex[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

// The launch as the library's tracer sees it.
struct RenderSetup {
  BH_Camera cam;
  BH_TraceOptions opts;
  PinholeDev pin;
  double r_isco, r_out;
};

int make_setup(const BH_LaunchParams* lp, int variant, RenderSetup* S, const char* who) {
  const std::string w(who);
  if (!lp) return fail(w + ": NULL params");
  const BH_LaunchParams& p = *lp;
  if (p.width <= 0 || p.height <= 0) return fail(w + ": empty framebuffer");
  if (!(p.rs > 0.0f) || !(p.spin >= 0.0f && p.spin < 1.0f)) return fail(w + ": rs must be > 0 and spin in [0, 1)");
  if (!(p.fov_scale > 0.0f) || p.max_steps <= 0 || !(p.max_dist > 0.0f))
    return fail(w + ": fov_scale, max_steps and max_dist must be > 0");
  const double M = 0.5 * static_cast<double>(p.rs);
  const double x = static_cast<double>(p.cam_pos[0]) / M;
  const double y = static_cast<double>(p.cam_pos[1]) / M;
  const double z = static_cast<double>(p.cam_pos[2]) / M;
  const double r = std::sqrt(x * x + y * y + z * z);
  const double rho = std::sqrt(x * x + y * y);
  if (!(rho > 0.0)) return fail(w + ": the camera must be off the spin axis");
  if (!(r > 2.0)) return fail(w + ": the camera must be outside r = rs (outside every ergosphere)");
  PinholeDev& pin = S->pin;
  pin.r = r;
  pin.cos_t = z / r;
  pin.sin_t = rho / r;
  pin.phi = std::atan2(y, x);
  const double cp = x / rho, sp = y / rho;
  pin.e_r[0] = x / r;
  pin.e_r[1] = y / r;
  pin.e_r[2] = z / r;
  pin.e_t[0] = pin.cos_t * cp;
  pin.e_t[1] = pin.cos_t * sp;
  pin.e_t[2] = -pin.sin_t;
  pin.e_p[0] = -sp;
  pin.e_p[1] = cp;
  pin.e_p[2] = 0.0;
  for (int i = 0; i < 9; ++i) pin.basis[i] = static_cast<double>(p.cam_basis[i]);
  pin.fov = static_cast<double>(p.fov_scale);
  pin.aspect = static_cast<double>(p.width) / static_cast<double>(p.height);
  pin.shift_x = static_cast<double>(p.frame_shift_x);
  pin.shift_y = static_cast<double>(p.frame_shift_y);
  pin.inv_w = 1.0 / static_cast<double>(p.width);
  pin.inv_h = 1.0 / static_cast<double>(p.height);
  // camera table entry: pixel (px, py) -> (alpha, beta) = (px + 0.5, py + 0.5)
  BH_Camera& c = S->cam;
  std::memset(&c, 0, sizeof(c));
  c.spin = static_cast<double>(p.spin);
  c.theta_obs = std::acos(pin.cos_t);
  c.r_obs = r;
  c.alpha_min = 0.0;
  c.alpha_max = static_cast<double>(p.width);
  c.beta_min = static_cast<double>(p.height);
  c.beta_max = 0.0;
  c.width = p.width;
  c.height = p.height;
  c.record_offset = 0;
  BH_TraceOptions& o = S->opts;
  bh_trace_options_default(&o);
  o.integrator = BH_INTEGRATOR_DP45;
  o.precision = (variant == BH_KERNEL_FP32_BASELINE || variant == BH_KERNEL_FP32_COARSENED) ? BH_PRECISION_FP32
                                                                                          : BH_PRECISION_FP64;
  o.max_steps = p.max_steps;
  o.max_attempts = 4 * p.max_steps + 100;
  o.disk_enabled = p.adisk_enabled ? 1 : 0;
  o.volume_enabled = 0;
  o.r_escape = static_cast<double>(p.max_dist) / M;
  S->r_out = kDiskOuterRs * static_cast<double>(p.rs) / M;
  o.disk_r_out = S->r_out;
  S->r_isco = p.isco > 0.0f ? static_cast<double>(p.isco) / M : bh_isco_prograde(c.spin);
  if (!(S->r_isco < S->r_out)) return fail(w + ": isco must lie inside the disk's outer edge (30 rs)");
  return 0;
}

int resolve_variant(int variant) {
  if (variant < 0) variant = bh_select_kernel_variant();
  if (variant < 0 || variant >= BH_KERNEL_COUNT) variant = BH_KERNEL_FP32_BASELINE;
  return variant;
}

// Traces the launch into `rec` (w x h records) and returns the context's LUT.
int trace_params(const BH_LaunchParams* lp, int variant, BH_PixelRecord* rec, void* stream, RenderSetup* S,
                 RenderCtx** ctx_out, const char* who) {
  if (make_setup(lp, variant, S, who) != 0) return -1;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(std::string(who) + ": no HIP device");
  RenderCtx* ctx;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ctx = &g_ctx[{dev, stream}];
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (!ctx->lut && hipMalloc(&ctx->lut, kLutN * sizeof(double)) != hipSuccess)
    return fail(std::string(who) + ": hipMalloc of the disk LUT failed");
  if (ctx->lut_key[0] != S->cam.spin || ctx->lut_key[1] != S->r_isco || ctx->lut_key[2] != S->r_out) {
    std::vector<double> lut(kLutN);
    if (bh_disk_lut_build(S->cam.spin, S->r_isco, S->r_out, kLutN, lut.data()) != 0) return -1;
    // the table is small; a synchronous copy keeps the host vector's lifetime simple
    if (hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(ctx->lut, lut.data(), kLutN * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(std::string(who) + ": disk LUT upload failed");
    ctx->lut_key[0] = S->cam.spin;
    ctx->lut_key[1] = S->r_isco;
    ctx->lut_key[2] = S->r_out;
  }
  if (ctx_out) *ctx_out = ctx;
  return trace_pinhole(S->cam, S->opts, ctx->lut, kLutN, S->pin, S->r_isco, rec, stream, who);
}

}  // namespace rl
}  // namespace bhk

using namespace bhk;
using namespace bhk::rl;

extern "C" {

int bh_select_kernel_variant(void) {
  // On gfx950 FP64 issues at the rate of (unpacked) FP32, and the FP32 form's
  // FP64 islands diverge within waves, so the FP64 form is the faster one.
  return BH_KERNEL_FP16_STORAGE;
}

int bh_trace_launch_params(const struct BH_LaunchParams* params, int variant, BH_PixelRecord* out, void* stream) {
  if (!out) return fail("bh_trace_launch_params: NULL output");
  RenderSetup S;
  return trace_params(params, resolve_variant(variant), out, stream, &S, nullptr, "bh_trace_launch_params");
}

int bh_launch_geodesic_kernel(void* framebuffer, const struct BH_LaunchParams* params, int variant, void* stream) {
  const char* who = "bh_launch_geodesic_kernel";
  if (!framebuffer || !params) return fail(std::string(who) + ": NULL framebuffer or params");
  const long long n = static_cast<long long>(params->width) * params->height;
  if (params->width <= 0 || params->height <= 0) return fail(std::string(who) + ": empty framebuffer");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(std::string(who) + ": no HIP device");
  RenderCtx* ctx;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ctx = &g_ctx[{dev, stream}];
    if (ctx->rec_cap < static_cast<size_t>(n)) {
      if (ctx->rec) (void)hipFree(ctx->rec);
      ctx->rec = nullptr;
      ctx->rec_cap = 0;
      if (hipMalloc(&ctx->rec, static_cast<size_t>(n) * sizeof(BH_PixelRecord)) != hipSuccess)
        return fail(std::string(who) + ": hipMalloc of the record buffer failed");
      ctx->rec_cap = static_cast<size_t>(n);
    }
  }
  RenderSetup S;
  if (trace_params(params, resolve_variant(variant), ctx->rec, stream, &S, &ctx, who) != 0) return -1;
  ShadeParams sp;
  sp.rec = ctx->rec;
  sp.fb = static_cast<float4*>(framebuffer);
  sp.n = n;
  sp.lut = ctx->lut;
  sp.lut_n = kLutN;
  sp.r_isco = S.r_isco;
  sp.r_out = S.r_out;
  sp.lit = params->adisk_lit > 0.0f ? params->adisk_lit : 1.0f;
  sp.bg = params->background_intensity;
  sp.redshift = params->redshift_enabled;
  sp.bg_on = params->background_enabled;
  {
    std::lock_guard<std::mutex> lk(g_tex_mu);
    sp.grmhd = grmhd_volume_known(g_tex[5]) ? g_tex[5] : 0ull;
    sp.grmhd_right = grmhd_volume_known(g_tex[7]) ? g_tex[7] : 0ull;
  }
  sp.grmhd_on = params->use_luts != 0 && sp.grmhd != 0ull;
  sp.grmhd_alpha = params->grmhd_alpha;
  sp.grmhd_r_min = params->grmhd_r_min;
  sp.grmhd_r_max = params->grmhd_r_max;
  sp.len = 0.5 * static_cast<double>(params->rs);
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(k_shade, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), sp);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : fail(std::string(who) + ": shading launch failed: " + hipGetErrorString(e));
}

void bh_upload_lut_textures(unsigned long long emissivity, unsigned long long redshift, unsigned long long spectral,
                            unsigned long long grb, unsigned long long galaxy, unsigned long long grmhd,
                            unsigned long long synch_g, unsigned long long grmhd_right) {
  std::lock_guard<std::mutex> lk(g_tex_mu);
  const unsigned long long v[8] = {emissivity, redshift, spectral, grb, galaxy, grmhd, synch_g, grmhd_right};
  for (int i = 0; i < 8; ++i) g_tex[i] = v[i];
}

void bh_upload_bridge_background_texture(unsigned long long background_equirect) {
  std::lock_guard<std::mutex> lk(g_tex_mu);
  g_tex[8] = background_equirect;
}

unsigned long long bh_lut_texture_handle(int slot) {
  if (slot < 0 || slot > 8) return 0ull;
  std::lock_guard<std::mutex> lk(g_tex_mu);
  return g_tex[slot];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Renderer backend (include/bh_backend.h)
// ---------------------------------------------------------------------------

#include "../../include/bh_backend.h"

struct BhCudaBackend {
  float4* fb = nullptr;
  int width = 0, height = 0;
  int variant = BH_KERNEL_FP16_STORAGE;
  hipStream_t stream = nullptr;
  unsigned int gl_fb = 0;
  unsigned int gl_lut[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
};

extern "C" {

BH_CudaBackend* bhCudaInit(unsigned int glTexture, int width, int height) {
  if (width <= 0 || height <= 0) {
    fail("bhCudaInit: empty framebuffer");
    return nullptr;
  }
  BH_CudaBackend* b = new BH_CudaBackend();
  b->gl_fb = glTexture;
  b->variant = bh_select_kernel_variant();
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&b->fb, static_cast<size_t>(width) * height * sizeof(float4)) != hipSuccess) {
    fail("bhCudaInit: device allocation failed");
    bhCudaShutdown(b);
    return nullptr;
  }
  b->width = width;
  b->height = height;
  return b;
}

int bhCudaRenderFrame(BH_CudaBackend* b, const struct BH_LaunchParams* params) {
  if (!b || !params) return fail("bhCudaRenderFrame: NULL argument");
  if (params->width != b->width || params->height != b->height)
    return fail("bhCudaRenderFrame: params size differs from the framebuffer (bhCudaResize first)");
  return bh_launch_geodesic_kernel(b->fb, params, b->variant, b->stream);
}

int bhCudaResize(BH_CudaBackend* b, unsigned int glTexture, int width, int height) {
  if (!b || width <= 0 || height <= 0) return fail("bhCudaResize: invalid arguments");
  if (hipStreamSynchronize(b->stream) != hipSuccess) return fail("bhCudaResize: device error");
  if (b->fb) (void)hipFree(b->fb);
  b->fb = nullptr;
  if (hipMalloc(&b->fb, static_cast<size_t>(width) * height * sizeof(float4)) != hipSuccess)
    return fail("bhCudaResize: device allocation failed");
  b->width = width;
  b->height = height;
  b->gl_fb = glTexture;
  return 0;
}

int bhCudaRegisterLut(BH_CudaBackend* b, int slot, unsigned int glTexture, unsigned int /*target*/) {
  if (!b || slot < 0 || slot > 7) return fail("bhCudaRegisterLut: invalid arguments");
  b->gl_lut[slot] = glTexture;
  return 0;
}

int bhCudaRegisterBackgroundTexture(BH_CudaBackend* b, unsigned int glTexture, unsigned int /*target*/) {
  if (!b) return fail("bhCudaRegisterBackgroundTexture: NULL backend");
  b->gl_lut[8] = glTexture;
  return 0;
}

int bhCudaSetVariant(BH_CudaBackend* b, int variant) {
  if (!b) return -1;
  b->variant = resolve_variant(variant);
  return b->variant;
}

int bhCudaGetVariant(BH_CudaBackend* b) { return b ? b->variant : -1; }

void bhCudaShutdown(BH_CudaBackend* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->fb) (void)hipFree(b->fb);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

void* bhCudaFramebuffer(BH_CudaBackend* b, int* width, int* height) {
  if (!b) return nullptr;
  if (width) *width = b->width;
  if (height) *height = b->height;
  return b->fb;
}

int bhCudaSynchronize(BH_CudaBackend* b) {
  if (!b) return fail("bhCudaSynchronize: NULL backend");
  return hipStreamSynchronize(b->stream) == hipSuccess ? 0 : fail("bhCudaSynchronize: device error");
}

unsigned int bhCudaGlTexture(BH_CudaBackend* b, int slot) {
  if (!b || slot < -1 || slot > 8) return 0u;
  return slot < 0 ? b->gl_fb : b->gl_lut[slot];
}

}  // extern "C"
