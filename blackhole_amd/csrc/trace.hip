// trace.hip - persistent per-pixel Kerr tracer kernels for gfx950 and the
// bh_trace_image_plane entry point.
//
// Replaces the reference's one-thread-per-pixel launches
// (geodesicTraceFp32, src/cuda/kernels_fp32.cu:32-111, launched by
// bh_launch_geodesic_kernel, src/cuda/kernel_launch.cu:220-269) for the
// image-plane tracer.
//
// Execution model (wave64, one ray per lane):
//   * The grid is sized to exactly fill the device (occupancy x CUs) and
//     every wave loops until the ray queue is empty.
//   * Lane refill: every `batch` steps the lanes whose ray has ended are
//     counted with a wave ballot; one lane performs a single atomicAdd on the
//     queue head for the whole wave and each idle lane takes base + (number
//     of idle lanes below it).  Consecutive ray ids (neighbouring pixels) land
//     in the same wave, so their volume fetches stay cache-coherent, and no
//     lane idles for longer than `batch` steps while rays remain -- the
//     divergent trip counts (tens to thousands of steps) never hold a wave.
//   * Ray state lives in registers; the per-ray record (128 B) is written
//     once, when the ray ends.
//   * The thermal-synchrotron K_2 table is staged in LDS once per workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "trace_common.hpp"

namespace bhk {

struct Lane {
  RayC c;
  double y[kN], k1[kN];
  double h;
  RtAcc rt;
  double cr0, cr1, cr2, cg0, cg1, cg2;
  double r_stop, r_isco;
  long long gid;
  int cam;
  int steps, attempts, ncross, nsamp;
  bool last_rej;
};

// Sets up the ray of queue entry q; false for an id in a gap between camera
// record ranges (no record).
template <bool VOL>
__device__ __forceinline__ bool lane_setup(Lane& L, const DevParams& P, long long q) {
  CamDev k;
  double alpha, beta;
  if (!resolve_ray(P, q, &L.gid, &L.cam, &k, &alpha, &beta)) return false;
  L.c = setup_ray_ab(k, alpha, beta, L.y);
  rhs(L.c, L.y, L.k1);
  L.h = P.h_init;
  L.rt.I = 0.0;
  L.rt.Tr = 1.0;
  L.rt.tau = 0.0;
  L.rt.a0 = 0.0;
  L.rt.e0 = 0.0;
  L.rt.s0 = VOL ? affine_rate(L.c, L.y[0], L.y[1]) : 0.0;
  L.cr0 = L.cr1 = L.cr2 = 0.0;
  L.cg0 = L.cg1 = L.cg2 = 0.0;
  L.r_stop = k.r_stop;
  L.r_isco = k.r_isco;
  L.steps = L.attempts = L.ncross = L.nsamp = 0;
  L.last_rej = false;
  return true;
}

// Volume transfer over the accepted step of length h ending at y.
__device__ __forceinline__ void volume_segment(Lane& L, const DevParams& P, const double* y,
                                               double h) {
  double a1, e1;
  if (volume_point(P.vol, L.c, y[0], y[1], y[2], P.nu_obs, P.length_unit_cm, &a1, &e1)) L.nsamp++;
  rt_segment(L.rt, h, a1, e1, affine_rate(L.c, y[0], y[1]));
}

// Equatorial crossing between (y0, f0) and (y1, f1); true if the disk ends the ray.
__device__ __forceinline__ bool crossing(Lane& L, const DevParams& P, const double* y0,
                                         const double* f0, const double* y1, const double* f1,
                                         double h) {
  if ((y0[1] < 0.0) == (y1[1] < 0.0)) return false;
  const double rc = crossing_radius(y0, f0, y1, f1, h);
  const double g = (rc >= L.r_isco) ? disk_g(rc, L.c.a, L.c.xi) : 0.0;
  // Branch-free slot update keeps the six slots in registers (an if/else
  // chain on ncross is lowered to a dynamically indexed scratch array).
  const bool rec_ok = L.ncross < P.max_crossings;
  const bool s0 = rec_ok && L.ncross == 0, s1 = rec_ok && L.ncross == 1, s2 = rec_ok && L.ncross == 2;
  L.cr0 = s0 ? rc : L.cr0;
  L.cg0 = s0 ? g : L.cg0;
  L.cr1 = s1 ? rc : L.cr1;
  L.cg1 = s1 ? g : L.cg1;
  L.cr2 = s2 ? rc : L.cr2;
  L.cg2 = s2 ? g : L.cg2;
  if (L.ncross < 255) L.ncross++;
  if (P.disk_enabled && rc >= L.r_isco && rc <= P.disk_r_out) {
    const double* lut = P.luts + static_cast<long long>(P.cams[L.cam].lut_index) * P.lut_n;
    const double F = lut_linear(lut, P.lut_n, L.r_isco, P.disk_r_out, rc);
    L.rt.I = __builtin_fma(L.rt.Tr, disk_intensity(F, g, P.nu_obs, P.disk_T_max), L.rt.I);
    return true;
  }
  return false;
}

__device__ __forceinline__ int terminate_check(const Lane& L, const DevParams& P) {
  if (L.y[0] < L.r_stop) return BH_HIT_HORIZON;
  if (L.y[0] > P.r_escape) return BH_HIT_ESCAPED;
  if (P.transmit_min > 0.0 && L.rt.Tr < P.transmit_min) return BH_HIT_OPAQUE;
  return 0;
}

// Field-by-field stores straight into the caller's record (a local record
// struct would be staged through scratch).
__device__ __forceinline__ void write_record(const Lane& L, const DevParams& P, int status) {
  BH_PixelRecord* __restrict__ rec = P.out + L.gid;
  const double R = radial_potential(L.c, L.y[0]);
  const double U = polar_potential(L.c, L.y[1]);
  const double rs = L.y[0] * L.y[0] + L.c.a2;
  double2* d2 = reinterpret_cast<double2*>(rec);
  d2[0] = make_double2(L.rt.I, L.y[0]);
  d2[1] = make_double2(L.y[1], L.y[2]);
  d2[2] = make_double2(L.cr0, L.cr1);
  d2[3] = make_double2(L.cr2, L.cg0);
  d2[4] = make_double2(L.cg1, L.cg2);
  d2[5] = make_double2(L.rt.tau, __builtin_fabs(__builtin_fma(L.y[3], L.y[3], -R)) / (rs * rs));
  const double dc = __builtin_fabs(__builtin_fma(L.y[4], L.y[4], -U)) /
                    (__builtin_fabs(L.c.eta) + L.c.a2 + L.c.xi * L.c.xi);
  int* i4 = reinterpret_cast<int*>(rec);
  reinterpret_cast<double*>(rec)[12] = dc;
  i4[26] = L.steps;
  i4[27] = L.attempts;
  i4[28] = status | (L.ncross << 8);  // hit_class, n_cross, flags = 0, pad0 = 0
  i4[29] = __float_as_int(1.0f);      // confidence
  i4[30] = L.steps;                   // fp64_steps
  i4[31] = L.nsamp;
}

// One RK4 step; returns the terminal status (0 = continue).
template <bool VOL>
__device__ __forceinline__ int step_rk4(Lane& L, const DevParams& P) {
  if (L.steps >= P.max_steps) return BH_HIT_MAX_STEPS;
  const double h = P.h_fixed;
  double yn[kN], fn[kN];
  rk4_step(L.c, L.y, L.k1, h, yn);
  rhs(L.c, yn, fn);
  project_radial(L.c, yn, fn);
  L.steps++;
  L.attempts++;
  if (VOL) volume_segment(L, P, yn, h);
  const bool hit = crossing(L, P, L.y, L.k1, yn, fn, h);
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    L.y[i] = yn[i];
    L.k1[i] = fn[i];
  }
  if (hit) return BH_HIT_DISK;
  return terminate_check(L, P);
}

// One DP45 attempt; returns the terminal status (0 = continue).
template <bool VOL>
__device__ __forceinline__ int step_dp45(Lane& L, const DevParams& P) {
  if (L.steps >= P.max_steps || L.attempts >= P.max_attempts) return BH_HIT_MAX_STEPS;
  double cap = P.h_max;
  if (VOL && in_support(P.vol, L.y[0], L.y[1]) && P.h_vol < cap) cap = P.h_vol;
  if (L.h > cap) L.h = cap;
  L.attempts++;
  double y5[kN], k7[kN];
  const double err = dp45_attempt(L.c, L.y, L.k1, L.h, P.rtol, P.atol, y5, k7);
  if (err <= 1.0) {
    const double fac = dp45_factor(err, true, L.last_rej);
    project_radial(L.c, y5, k7);
    L.steps++;
    if (VOL) volume_segment(L, P, y5, L.h);
    const bool hit = crossing(L, P, L.y, L.k1, y5, k7, L.h);
#pragma unroll
    for (int i = 0; i < kN; ++i) {
      L.y[i] = y5[i];
      L.k1[i] = k7[i];
    }
    L.h = L.h * fac;
    L.last_rej = false;
    if (hit) return BH_HIT_DISK;
    return terminate_check(L, P);
  }
  L.h = L.h * dp45_factor(err, false, L.last_rej);
  L.last_rej = true;
  return 0;
}

template <int INTEG, bool VOL>
__global__ void __launch_bounds__(kBlock) k_trace_fp64(const DevParams* __restrict__ Pp) {
  const DevParams& P = *Pp;
  Lane L;
  L.gid = -1;
  bool done = false;
  const long long n_queue = queue_length(P);
  for (;;) {
    const bool need = (L.gid < 0) && !done;
    const long long id = refill_id(need, P.head);
    if (need) {
      if (id < n_queue) {
        if (!lane_setup<VOL>(L, P, id)) L.gid = -1;
      } else {
        done = true;
      }
    }
    if (__ballot(!done) == 0ull) break;
    if (L.gid >= 0) {
      for (int b = 0; b < P.batch; ++b) {
        const int st = (INTEG == BH_INTEGRATOR_DP45) ? step_dp45<VOL>(L, P) : step_rk4<VOL>(L, P);
        if (st) {
          write_record(L, P, st);
          L.gid = -1;
          break;
        }
      }
    }
  }
}

__global__ void k_reset_head(unsigned long long* head) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *head = 0ull;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

thread_local std::string g_err;
thread_local hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
thread_local bool g_have_events = false;

int fail(const std::string& msg) {
  g_err = msg;
  std::fprintf(stderr, "[bhk] %s\n", msg.c_str());
  return -1;
}

// Per (device, stream) launch context: one device block holding the
// DevParams followed by the camera table, and the ray-queue head.  Reused
// across launches on the same stream (stream order serialises them).
struct StreamCtx {
  unsigned char* d_block = nullptr;
  size_t cap = 0;
  unsigned long long* d_head = nullptr;
};
std::mutex g_mu;
std::map<std::pair<int, void*>, StreamCtx> g_ctx;

template <int I, bool V>
int occupancy_grid() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -1;
  if (dev < 0 || dev >= 64) return -1;
  if (cached[dev] > 0) return cached[dev];
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace_fp64<I, V>, kBlock, 0) != hipSuccess)
    return -1;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
  cached[dev] = std::max(1, nb) * prop.multiProcessorCount;
  return cached[dev];
}

template <int I, bool V>
hipError_t launch(const DevParams* d_params, hipStream_t s, int grid_override) {
  const int grid = grid_override > 0 ? grid_override : occupancy_grid<I, V>();
  if (grid <= 0) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((k_trace_fp64<I, V>), dim3(grid), dim3(kBlock), 0, s, d_params);
  return hipGetLastError();
}

}  // namespace bhk

using namespace bhk;

extern "C" {

int bh_version(void) { return BHK_VERSION_MAJOR * 100 + BHK_VERSION_MINOR; }

const char* bh_last_error(void) { return g_err.c_str(); }

int bh_trace_options_default(BH_TraceOptions* o) {
  if (!o) return fail("bh_trace_options_default: NULL options");
  std::memset(o, 0, sizeof(*o));
  o->integrator = BH_INTEGRATOR_RK4;
  o->precision = BH_PRECISION_FP64;
  o->max_steps = 2000;
  o->max_attempts = 100000;
  o->disk_enabled = 1;
  o->volume_enabled = 0;
  o->max_crossings = 3;
  o->refill_batch = 0;
  o->h_fixed = 0.01;
  o->h_init = 0.01;
  o->h_max = 0.1;
  o->h_vol = 0.01;
  o->rtol = 1e-8;
  o->atol = 1e-10;
  o->horizon_factor = 1.01;
  o->r_escape = 200.0;
  o->disk_r_out = 20.0;
  o->nu_obs = 230e9;
  o->disk_T_max = 1e7;
  o->length_unit_cm = 6.6743e-8 * 4.0e6 * 1.98841e33 / (2.99792458e10 * 2.99792458e10);
  o->transmit_min = 0.0;
  o->island_r_factor = 1.2;
  o->island_turn_dist = 0.05;
  return 0;
}

double bh_outer_horizon(double a) { return 1.0 + std::sqrt(1.0 - a * a); }

double bh_isco_prograde(double a) {
  const double z1 = 1.0 + std::cbrt(1.0 - a * a) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
  const double z2 = std::sqrt(3.0 * a * a + z1 * z1);
  return 3.0 + z2 - std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

}  // extern "C"

namespace bhk {

// Shared launcher: every pixel of the cameras (rays == ids == NULL), a list
// of pixel ids, or an explicit ray list.
// Per-camera device constants, with the checks every camera must pass
// (spin in [0, 1), theta_obs in (0, pi), non-empty, increasing disjoint
// record ranges, observer outside the horizon stop).  Returns -1 with the
// error set (prefixed by who) on the first invalid camera.
int make_cam_devs(const BH_Camera* cams, int n_cams, double horizon_factor, const char* who,
                  std::vector<CamDev>* out, long long* total_out) {
  const std::string w(who);
  out->assign(n_cams, CamDev{});
  long long total = 0;
  for (int i = 0; i < n_cams; ++i) {
    const BH_Camera& c = cams[i];
    if (!(c.spin >= 0.0 && c.spin < 1.0)) return fail(w + ": spin outside [0, 1)");
    if (!(c.theta_obs > 0.0 && c.theta_obs < kPi)) return fail(w + ": theta_obs outside (0, pi)");
    if (c.width <= 0 || c.height <= 0) return fail(w + ": empty camera");
    if (c.record_offset < 0) return fail(w + ": negative record_offset");
    if (i > 0 && c.record_offset < cams[i - 1].record_offset +
                                       static_cast<long long>(cams[i - 1].width) * cams[i - 1].height)
      return fail(w + ": camera record ranges must be increasing and disjoint");
    const double rh = bh_outer_horizon(c.spin);
    if (!(c.r_obs > horizon_factor * rh)) return fail(w + ": r_obs inside the horizon stop");
    CamDev& k = (*out)[i];
    k.a = c.spin;
    k.sin_o = std::sin(c.theta_obs);
    k.cos_o = std::cos(c.theta_obs);
    k.r_obs = c.r_obs;
    k.r_stop = horizon_factor * rh;
    k.r_isco = bh_isco_prograde(c.spin);
    k.r_plus = rh;
    k.alpha_min = c.alpha_min;
    k.dalpha = (c.alpha_max - c.alpha_min) / static_cast<double>(c.width);
    k.beta_max = c.beta_max;
    k.dbeta = (c.beta_max - c.beta_min) / static_cast<double>(c.height);
    k.w = c.width;
    k.h = c.height;
    k.offset = c.record_offset;
    k.lut_index = i;
    total = std::max(total, c.record_offset + static_cast<long long>(c.width) * c.height);
  }
  *total_out = total;
  return 0;
}

int trace_launch(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                 const double* disk_luts, int disk_lut_n, const BH_Volume* vol,
                 const BH_RaySpec* rays, long long n_rays, const int64_t* ids,
                 const uint64_t* n_ids_dev, long long max_ids, BH_PixelRecord* out, void* stream,
                 bool compact) {
  if (!rays && (!cams || n_cams <= 0)) return fail("bh_trace_image_plane: no cameras");
  if (rays && n_rays <= 0) return fail("bh_trace_rays: empty ray list");
  if (ids && (!n_ids_dev || max_ids <= 0)) return fail("bh_trace_pixel_list: invalid id list");
  if (rays && opts && opts->disk_enabled) return fail("bh_trace_rays: the thin disk must be disabled");
  if (!opts) return fail("bh_trace_image_plane: NULL options");
  if (!out) return fail("bh_trace_image_plane: NULL output");
  const BH_TraceOptions& o = *opts;
  if (o.integrator != BH_INTEGRATOR_RK4 && o.integrator != BH_INTEGRATOR_DP45)
    return fail("bh_trace_image_plane: unknown integrator");
  if (o.precision != BH_PRECISION_FP64 && o.precision != BH_PRECISION_FP32)
    return fail("bh_trace_image_plane: unknown precision");
  if (o.precision == BH_PRECISION_FP32 &&
      !(o.island_r_factor >= 1.0 && o.island_turn_dist >= 0.0))
    return fail("bh_trace_image_plane: FP32 form needs island_r_factor >= 1, island_turn_dist >= 0");
  if (o.integrator == BH_INTEGRATOR_RK4 && !(o.h_fixed > 0.0))
    return fail("bh_trace_image_plane: h_fixed must be > 0");
  if (o.integrator == BH_INTEGRATOR_DP45 &&
      !(o.h_init > 0.0 && o.h_max > 0.0 && o.rtol > 0.0 && o.atol > 0.0))
    return fail("bh_trace_image_plane: DP45 needs h_init, h_max, rtol, atol > 0");
  if (o.max_crossings < 0 || o.max_crossings > 3)
    return fail("bh_trace_image_plane: max_crossings must be in [0, 3]");
  if (o.max_steps < 0 || o.max_attempts < 0) return fail("bh_trace_image_plane: negative budget");
  if (o.disk_enabled && (!disk_luts || disk_lut_n < 2))
    return fail("bh_trace_image_plane: disk enabled without disk LUTs");
  const bool vol_on = o.volume_enabled != 0;
  if (vol_on) {
    if (!vol || !vol->cells || vol->n_r < 1 || vol->n_th < 1 || vol->n_ph < 1 ||
        !(vol->r_max > vol->r_min))
      return fail("bh_trace_image_plane: volume enabled with an invalid BH_Volume");
    if (static_cast<long long>(vol->n_r) * vol->n_th * vol->n_ph >= (1ll << 31))
      return fail("bh_trace_image_plane: volume grid too large (>= 2^31 cells)");
    if (o.integrator == BH_INTEGRATOR_DP45 && !(o.h_vol > 0.0))
      return fail("bh_trace_image_plane: h_vol must be > 0");
  }
  std::vector<CamDev> hc;
  long long total = 0;
  if (!rays && make_cam_devs(cams, n_cams, o.horizon_factor, "bh_trace_image_plane", &hc, &total) != 0) return -1;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail("bh_trace_image_plane: no HIP device");
  const size_t bytes = sizeof(DevParams) + hc.size() * sizeof(CamDev);
  StreamCtx* ctx;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ctx = &g_ctx[{dev, stream}];
    if (ctx->cap < bytes) {
      if (ctx->d_block) (void)hipFree(ctx->d_block);
      ctx->d_block = nullptr;
      ctx->cap = 0;
      const size_t cap = std::max<size_t>(bytes, sizeof(DevParams) + 64 * sizeof(CamDev));
      if (hipMalloc(&ctx->d_block, cap) != hipSuccess)
        return fail("bh_trace_image_plane: hipMalloc of the launch block failed");
      ctx->cap = cap;
    }
    if (!ctx->d_head && hipMalloc(&ctx->d_head, 256) != hipSuccess)
      return fail("bh_trace_image_plane: hipMalloc of the queue head failed");
  }
  std::vector<unsigned char> blk(bytes);
  DevParams P;
  std::memset(&P, 0, sizeof(P));
  P.h_fixed = o.h_fixed;
  P.h_init = o.h_init;
  P.h_max = o.h_max;
  P.h_vol = o.h_vol;
  P.rtol = o.rtol;
  P.atol = o.atol;
  P.r_escape = o.r_escape;
  P.disk_r_out = o.disk_r_out;
  P.nu_obs = o.nu_obs;
  P.disk_T_max = o.disk_T_max;
  P.length_unit_cm = o.length_unit_cm;
  P.transmit_min = o.transmit_min;
  if (vol_on) {
    P.vol.cells = static_cast<const float4*>(vol->cells);
    P.vol.n_r = vol->n_r;
    P.vol.n_th = vol->n_th;
    P.vol.n_ph = vol->n_ph;
    P.vol.r_min = vol->r_min;
    P.vol.r_max = vol->r_max;
    P.vol.r_span = vol->r_max - vol->r_min;
    P.vol.sup_r_lo = vol->sup_r_lo;
    P.vol.sup_r_hi = vol->sup_r_hi;
    P.vol.sup_u_max = vol->sup_u_max;
    P.vol.ell = vol->ell;
  }
  P.cams = reinterpret_cast<const CamDev*>(ctx->d_block + sizeof(DevParams));
  P.luts = disk_luts;
  P.out = out;
  P.head = ctx->d_head;
  P.n_total = rays ? n_rays : (ids ? max_ids : total);
  P.rays = rays;
  P.ids = reinterpret_cast<const long long*>(ids);
  P.count_dev = reinterpret_cast<const unsigned long long*>(ids ? n_ids_dev : nullptr);
  P.horizon_factor = o.horizon_factor;
  P.n_cams = rays ? 0 : n_cams;
  P.lut_n = disk_lut_n;
  P.batch = o.refill_batch > 0 ? o.refill_batch : (o.integrator == BH_INTEGRATOR_DP45 ? 8 : 16);
  P.max_steps = o.max_steps;
  P.max_attempts = o.max_attempts;
  P.max_crossings = o.max_crossings;
  P.disk_enabled = o.disk_enabled;
  P.compact_ids = (ids && compact) ? 1 : 0;
  P.island_r_factor = o.island_r_factor;
  P.island_turn_dist = o.island_turn_dist;
  P.rtol32 = static_cast<float>(std::max(o.rtol, 1e-6));
  P.atol32 = static_cast<float>(std::max(o.atol, 1e-8));
  {
    const double nu = o.nu_obs;
    const double me_c2 = kMElectron * kC * kC;
    const double nuB_per_G = kECharge / (kTwoPi * kMElectron * kC);
    P.CT = static_cast<float>(kH * nu / me_c2);
    P.SB = static_cast<float>(2.0 * kH * nu * nu * nu / (kC * kC));
    P.Cx = static_cast<float>(nu / (1.5 * nuB_per_G));
    P.CjL = static_cast<float>(1.7320508075688772 * kECharge * kECharge * kECharge / me_c2 * nuB_per_G *
                               0.5 * o.length_unit_cm);
    P.Cd = static_cast<float>(o.disk_T_max > 0.0 ? kH * nu / (kKB * o.disk_T_max) : 1e30);
  }
  std::memcpy(blk.data(), &P, sizeof(P));
  std::memcpy(blk.data() + sizeof(DevParams), hc.data(), hc.size() * sizeof(CamDev));
  // Pageable source: the runtime stages it before returning, so blk may die.
  if (hipMemcpyAsync(ctx->d_block, blk.data(), bytes, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail("bh_trace_image_plane: parameter upload failed");
  if (!g_have_events) {
    if (hipEventCreate(&g_ev0) != hipSuccess || hipEventCreate(&g_ev1) != hipSuccess)
      return fail("bh_trace_image_plane: hipEventCreate failed");
    g_have_events = true;
  }
  hipLaunchKernelGGL(k_reset_head, dim3(1), dim3(64), 0, s, ctx->d_head);
  if (hipGetLastError() != hipSuccess) return fail("bh_trace_image_plane: reset launch failed");
  (void)hipEventRecord(g_ev0, s);
  const DevParams* dp = reinterpret_cast<const DevParams*>(ctx->d_block);
  hipError_t e;
  if (o.precision == BH_PRECISION_FP32)
    e = launch_fp32(o.integrator, vol_on, dp, s, 0);
  else if (o.integrator == BH_INTEGRATOR_DP45)
    e = vol_on ? launch<BH_INTEGRATOR_DP45, true>(dp, s, 0) : launch<BH_INTEGRATOR_DP45, false>(dp, s, 0);
  else
    e = vol_on ? launch<BH_INTEGRATOR_RK4, true>(dp, s, 0) : launch<BH_INTEGRATOR_RK4, false>(dp, s, 0);
  if (e != hipSuccess)
    return fail(std::string("bh_trace_image_plane: launch failed: ") + hipGetErrorString(e));
  (void)hipEventRecord(g_ev1, s);
  return 0;
}

}  // namespace bhk

extern "C" {

int bh_trace_image_plane(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                         const double* disk_luts, int disk_lut_n, const BH_Volume* vol,
                         BH_PixelRecord* out, void* stream) {
  return trace_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, nullptr, 0, nullptr, nullptr, 0, out,
                      stream, false);
}

int bh_trace_pixel_list(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                        const double* disk_luts, int disk_lut_n, const BH_Volume* vol, const int64_t* ids,
                        const uint64_t* n_ids_dev, int64_t max_ids, BH_PixelRecord* out, void* stream) {
  if (!ids) return fail("bh_trace_pixel_list: NULL id list");
  return trace_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, nullptr, 0, ids, n_ids_dev, max_ids, out,
                      stream, false);
}

int bh_trace_pixel_list_compact(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                                const double* disk_luts, int disk_lut_n, const BH_Volume* vol,
                                const int64_t* ids, const uint64_t* n_ids_dev, int64_t max_ids,
                                BH_PixelRecord* out, void* stream) {
  if (!ids) return fail("bh_trace_pixel_list_compact: NULL id list");
  return trace_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, nullptr, 0, ids, n_ids_dev, max_ids, out,
                      stream, true);
}

int bh_trace_rays(const BH_RaySpec* rays, int64_t n, const BH_TraceOptions* opts, const BH_Volume* vol,
                  BH_PixelRecord* out, void* stream) {
  if (!rays) return fail("bh_trace_rays: NULL ray list");
  return trace_launch(nullptr, 0, opts, nullptr, 0, vol, rays, n, nullptr, nullptr, 0, out, stream, false);
}

float bh_last_kernel_ms(void) {
  if (!g_have_events) return -1.0f;
  float ms = -1.0f;
  if (hipEventElapsedTime(&ms, g_ev0, g_ev1) != hipSuccess) return -1.0f;
  return ms;
}

}  // extern "C"
