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

#include "kerr_device.hpp"

namespace bhk {

constexpr int kBlock = 256;
constexpr int kMaxK2 = 1024;  // LDS K2 table capacity (doubles)

struct TraceArgs {
  const CamDev* cams;
  int n_cams;
  int batch;
  long long n_total;
  unsigned long long* head;  // ray-queue head
  BH_TraceOptions o;
  const double* luts;
  int lut_n;
  BH_Volume vol;
  const double* k2tab;
  int k2_n;
  double k2_lo, k2_hi;
  BH_PixelRecord* out;
};

struct Lane {
  RayC c;
  double y[kN], k1[kN];
  double h;
  double I, Tr, tau;
  double cr0, cr1, cr2, cg0, cg1, cg2;
  double r_stop, r_isco;
  long long gid;
  int cam;
  int steps, attempts, ncross, nsamp;
  bool last_rej;
};

__device__ __forceinline__ int find_camera(const CamDev* cams, int n, long long gid) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cams[mid].offset <= gid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Returns false for ids in a gap between camera record ranges (no record).
__device__ __forceinline__ bool lane_setup(Lane& L, const TraceArgs& A) {
  L.cam = find_camera(A.cams, A.n_cams, L.gid);
  const CamDev& k = A.cams[L.cam];
  const long long local = L.gid - k.offset;
  if (local >= k.w * k.h) return false;
  const long long py = local / k.w, px = local - py * k.w;
  L.c = setup_ray(k, px, py, L.y);
  rhs(L.c, L.y, L.k1);
  L.h = A.o.h_init;
  L.I = 0.0;
  L.Tr = 1.0;
  L.tau = 0.0;
  L.cr0 = L.cr1 = L.cr2 = 0.0;
  L.cg0 = L.cg1 = L.cg2 = 0.0;
  L.r_stop = k.r_stop;
  L.r_isco = k.r_isco;
  L.steps = L.attempts = L.ncross = L.nsamp = 0;
  L.last_rej = false;
  return true;
}

__device__ __forceinline__ bool in_support(const BH_Volume& v, double r, double u) {
  return r >= v.sup_r_lo && r <= v.sup_r_hi && __builtin_fabs(u) <= v.sup_u_max;
}

__device__ __forceinline__ double k2_ratio(const double* k2s, const TraceArgs& A, double th) {
  const double lt = log(th);
  return exp(lut_linear(k2s, A.k2_n, A.k2_lo, A.k2_hi, lt));
}

// One formal-solution segment ending at y (front-to-back), as
// d_rte_step / d_trace_geodesic_rte (src/cuda/device_physics.cuh:1988-2173)
// with the thermal-synchrotron coefficients of rte_integrator.h:291-357.
__device__ __forceinline__ void volume_segment(Lane& L, const TraceArgs& A, const double* k2s,
                                               const double* y, double h) {
  const BH_Volume& v = A.vol;
  const double r = y[0], u = y[1];
  if (!(r >= v.r_min && r <= v.r_max && in_support(v, r, u))) return;
  double f3[3];
  volume_fetch(static_cast<const float4*>(v.cells), v, r, u, y[2], f3);
  L.nsamp++;
  const double ne = f3[0], th = f3[1], B = f3[2];
  if (!(ne > 0.0 && th > 0.0 && B > 0.0)) return;
  const double g = fluid_g(r, u, L.c.a, L.c.xi, v.ell);
  if (!(g > 0.0)) return;
  const double nu_f = A.o.nu_obs / g;
  const double tempK = th * (kMElectron * kC * kC / kKB);
  const double Bnu = planck(nu_f, tempK);
  if (!(Bnu > 0.0)) return;
  const double j = synch_j(nu_f, B, ne, th, k2_ratio(k2s, A, th));
  const double alpha_nu = j / Bnu;
  const double sig = __builtin_fma(L.c.a2, u * u, r * r);
  const double dl = h * sig / r;
  const double dtau = alpha_nu * (dl / g) * A.o.length_unit_cm;
  const double om = -expm1(-dtau);
  const double g3 = g * g * g;
  L.I = __builtin_fma(L.Tr * g3 * Bnu, om, L.I);
  L.Tr = L.Tr * (1.0 - om);
  L.tau = L.tau + dtau;
}

// Equatorial crossing between (y0, f0) and (y1, f1); true if the disk ends the ray.
__device__ __forceinline__ bool crossing(Lane& L, const TraceArgs& A, const double* y0,
                                         const double* f0, const double* y1, const double* f1,
                                         double h) {
  if ((y0[1] < 0.0) == (y1[1] < 0.0)) return false;
  const double rc = crossing_radius(y0, f0, y1, f1, h);
  const double g = (rc >= L.r_isco) ? disk_g(rc, L.c.a, L.c.xi) : 0.0;
  if (L.ncross < A.o.max_crossings) {
    if (L.ncross == 0) { L.cr0 = rc; L.cg0 = g; }
    else if (L.ncross == 1) { L.cr1 = rc; L.cg1 = g; }
    else if (L.ncross == 2) { L.cr2 = rc; L.cg2 = g; }
  }
  if (L.ncross < 255) L.ncross++;
  if (A.o.disk_enabled && rc >= L.r_isco && rc <= A.o.disk_r_out) {
    const double* lut = A.luts + static_cast<long long>(A.cams[L.cam].lut_index) * A.lut_n;
    const double F = lut_linear(lut, A.lut_n, L.r_isco, A.o.disk_r_out, rc);
    L.I = __builtin_fma(L.Tr, disk_intensity(F, g, A.o.nu_obs, A.o.disk_T_max), L.I);
    return true;
  }
  return false;
}

__device__ __forceinline__ int terminate_check(const Lane& L, const TraceArgs& A) {
  if (L.y[0] < L.r_stop) return BH_HIT_HORIZON;
  if (L.y[0] > A.o.r_escape) return BH_HIT_ESCAPED;
  if (A.o.transmit_min > 0.0 && L.Tr < A.o.transmit_min) return BH_HIT_OPAQUE;
  return 0;
}

__device__ __forceinline__ void write_record(const Lane& L, const TraceArgs& A, int status) {
  BH_PixelRecord rec;
  rec.intensity = L.I;
  rec.r_exit = L.y[0];
  rec.u_exit = L.y[1];
  rec.phi_exit = L.y[2];
  rec.cross_r[0] = L.cr0;
  rec.cross_r[1] = L.cr1;
  rec.cross_r[2] = L.cr2;
  rec.cross_g[0] = L.cg0;
  rec.cross_g[1] = L.cg1;
  rec.cross_g[2] = L.cg2;
  rec.tau = L.tau;
  const double R = radial_potential(L.c, L.y[0]);
  const double U = polar_potential(L.c, L.y[1]);
  const double rs = L.y[0] * L.y[0] + L.c.a2;
  rec.drift_null = __builtin_fabs(__builtin_fma(L.y[3], L.y[3], -R)) / (rs * rs);
  rec.drift_carter = __builtin_fabs(__builtin_fma(L.y[4], L.y[4], -U)) /
                     (__builtin_fabs(L.c.eta) + L.c.a2 + L.c.xi * L.c.xi);
  rec.steps = L.steps;
  rec.attempts = L.attempts;
  rec.hit_class = static_cast<uint8_t>(status);
  rec.n_cross = static_cast<uint8_t>(L.ncross);
  rec.flags = 0;
  rec.pad0 = 0;
  rec.confidence = 1.0f;
  rec.fp64_steps = L.steps;
  rec.n_samples = L.nsamp;
  A.out[L.gid] = rec;
}

// One RK4 step; returns the terminal status (0 = continue).
template <bool VOL>
__device__ __forceinline__ int step_rk4(Lane& L, const TraceArgs& A, const double* k2s) {
  if (L.steps >= A.o.max_steps) return BH_HIT_MAX_STEPS;
  const double h = A.o.h_fixed;
  double yn[kN], fn[kN];
  rk4_step(L.c, L.y, L.k1, h, yn);
  rhs(L.c, yn, fn);
  project_radial(L.c, yn, fn);
  L.steps++;
  L.attempts++;
  if (VOL) volume_segment(L, A, k2s, yn, h);
  const bool hit = crossing(L, A, L.y, L.k1, yn, fn, h);
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    L.y[i] = yn[i];
    L.k1[i] = fn[i];
  }
  if (hit) return BH_HIT_DISK;
  return terminate_check(L, A);
}

// One DP45 attempt; returns the terminal status (0 = continue).
template <bool VOL>
__device__ __forceinline__ int step_dp45(Lane& L, const TraceArgs& A, const double* k2s) {
  if (L.steps >= A.o.max_steps || L.attempts >= A.o.max_attempts) return BH_HIT_MAX_STEPS;
  double cap = A.o.h_max;
  if (VOL && in_support(A.vol, L.y[0], L.y[1]) && A.o.h_vol < cap) cap = A.o.h_vol;
  if (L.h > cap) L.h = cap;
  L.attempts++;
  double y5[kN], k7[kN];
  const double err = dp45_attempt(L.c, L.y, L.k1, L.h, A.o.rtol, A.o.atol, y5, k7);
  if (err <= 1.0) {
    const double fac = dp45_factor(err, true, L.last_rej);
    project_radial(L.c, y5, k7);
    L.steps++;
    if (VOL) volume_segment(L, A, k2s, y5, L.h);
    const bool hit = crossing(L, A, L.y, L.k1, y5, k7, L.h);
#pragma unroll
    for (int i = 0; i < kN; ++i) {
      L.y[i] = y5[i];
      L.k1[i] = k7[i];
    }
    L.h = L.h * fac;
    L.last_rej = false;
    if (hit) return BH_HIT_DISK;
    return terminate_check(L, A);
  }
  L.h = L.h * dp45_factor(err, false, L.last_rej);
  L.last_rej = true;
  return 0;
}

template <int INTEG, bool VOL>
__global__ void __launch_bounds__(kBlock) k_trace_fp64(TraceArgs A) {
  __shared__ double k2s[kMaxK2];
  if (VOL) {
    for (int i = threadIdx.x; i < A.k2_n; i += kBlock) k2s[i] = A.k2tab[i];
    __syncthreads();
  }
  const int lane = static_cast<int>(__lane_id());
  const unsigned long long below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  Lane L;
  L.gid = -1;
  bool done = false;
  for (;;) {
    const bool need = (L.gid < 0) && !done;
    const unsigned long long m = __ballot(need);
    if (m) {
      const int leader = __ffsll(static_cast<long long>(m)) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(A.head, static_cast<unsigned long long>(__popcll(m)));
      const unsigned int lo = __shfl(static_cast<unsigned int>(base & 0xffffffffull), leader);
      const unsigned int hi = __shfl(static_cast<unsigned int>(base >> 32), leader);
      base = (static_cast<unsigned long long>(hi) << 32) | lo;
      if (need) {
        const long long id = static_cast<long long>(base) + __popcll(m & below);
        if (id < A.n_total) {
          L.gid = id;
          if (!lane_setup(L, A)) L.gid = -1;
        } else {
          done = true;
        }
      }
    }
    if (__ballot(!done) == 0ull) break;
    if (L.gid >= 0) {
      for (int b = 0; b < A.batch; ++b) {
        const int st = (INTEG == BH_INTEGRATOR_DP45) ? step_dp45<VOL>(L, A, k2s)
                                                      : step_rk4<VOL>(L, A, k2s);
        if (st) {
          write_record(L, A, st);
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

struct StreamCtx {
  CamDev* d_cams = nullptr;
  size_t cam_cap = 0;
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
hipError_t launch(const TraceArgs& A, hipStream_t s, int grid_override) {
  const int grid = grid_override > 0 ? grid_override : occupancy_grid<I, V>();
  if (grid <= 0) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((k_trace_fp64<I, V>), dim3(grid), dim3(kBlock), 0, s, A);
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

int bh_trace_image_plane(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                         const double* disk_luts, int disk_lut_n, const BH_Volume* vol,
                         const BH_K2Table* k2, BH_PixelRecord* out, void* stream) {
  if (!cams || n_cams <= 0) return fail("bh_trace_image_plane: no cameras");
  if (!opts) return fail("bh_trace_image_plane: NULL options");
  if (!out) return fail("bh_trace_image_plane: NULL output");
  const BH_TraceOptions& o = *opts;
  if (o.integrator != BH_INTEGRATOR_RK4 && o.integrator != BH_INTEGRATOR_DP45)
    return fail("bh_trace_image_plane: unknown integrator");
  if (o.precision != BH_PRECISION_FP64)
    return fail("bh_trace_image_plane: precision form not available in this build");
  if (o.integrator == BH_INTEGRATOR_RK4 && !(o.h_fixed > 0.0))
    return fail("bh_trace_image_plane: h_fixed must be > 0");
  if (o.integrator == BH_INTEGRATOR_DP45 &&
      !(o.h_init > 0.0 && o.h_max > 0.0 && o.rtol > 0.0 && o.atol > 0.0))
    return fail("bh_trace_image_plane: DP45 needs h_init, h_max, rtol, atol > 0");
  if (o.max_crossings < 0 || o.max_crossings > 3)
    return fail("bh_trace_image_plane: max_crossings must be in [0, 3]");
  if (o.disk_enabled && (!disk_luts || disk_lut_n < 2))
    return fail("bh_trace_image_plane: disk enabled without disk LUTs");
  const bool vol_on = o.volume_enabled != 0;
  if (vol_on) {
    if (!vol || !vol->cells || vol->n_r < 1 || vol->n_th < 1 || vol->n_ph < 1 ||
        !(vol->r_max > vol->r_min))
      return fail("bh_trace_image_plane: volume enabled with an invalid BH_Volume");
    if (o.integrator == BH_INTEGRATOR_DP45 && !(o.h_vol > 0.0))
      return fail("bh_trace_image_plane: h_vol must be > 0");
    if (!k2 || !k2->tab || k2->n < 2 || k2->n > kMaxK2 || !(k2->lnth_max > k2->lnth_min))
      return fail("bh_trace_image_plane: volume enabled with an invalid BH_K2Table");
  }
  std::vector<CamDev> hc(n_cams);
  long long total = 0;
  for (int i = 0; i < n_cams; ++i) {
    const BH_Camera& c = cams[i];
    if (!(c.spin >= 0.0 && c.spin < 1.0)) return fail("bh_trace_image_plane: spin outside [0, 1)");
    if (!(c.theta_obs > 0.0 && c.theta_obs < kPi))
      return fail("bh_trace_image_plane: theta_obs outside (0, pi)");
    if (c.width <= 0 || c.height <= 0) return fail("bh_trace_image_plane: empty camera");
    if (i > 0 && c.record_offset < cams[i - 1].record_offset + static_cast<long long>(cams[i - 1].width) * cams[i - 1].height)
      return fail("bh_trace_image_plane: camera record ranges must be increasing and disjoint");
    if (c.record_offset < 0) return fail("bh_trace_image_plane: negative record_offset");
    const double rh = bh_outer_horizon(c.spin);
    if (!(c.r_obs > o.horizon_factor * rh)) return fail("bh_trace_image_plane: r_obs inside the horizon stop");
    CamDev& k = hc[i];
    k.a = c.spin;
    k.sin_o = std::sin(c.theta_obs);
    k.cos_o = std::cos(c.theta_obs);
    k.r_obs = c.r_obs;
    k.r_stop = o.horizon_factor * rh;
    k.r_isco = bh_isco_prograde(c.spin);
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
  hipStream_t s = static_cast<hipStream_t>(stream);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail("bh_trace_image_plane: no HIP device");
  StreamCtx* ctx;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ctx = &g_ctx[{dev, stream}];
    if (ctx->cam_cap < hc.size()) {
      if (ctx->d_cams) (void)hipFree(ctx->d_cams);
      ctx->d_cams = nullptr;
      ctx->cam_cap = 0;
      const size_t cap = std::max<size_t>(hc.size(), 64);
      if (hipMalloc(&ctx->d_cams, cap * sizeof(CamDev)) != hipSuccess)
        return fail("bh_trace_image_plane: hipMalloc cameras failed");
      ctx->cam_cap = cap;
    }
    if (!ctx->d_head && hipMalloc(&ctx->d_head, 256) != hipSuccess)
      return fail("bh_trace_image_plane: hipMalloc queue head failed");
  }
  if (hipMemcpyAsync(ctx->d_cams, hc.data(), hc.size() * sizeof(CamDev), hipMemcpyHostToDevice, s) !=
      hipSuccess)
    return fail("bh_trace_image_plane: camera upload failed");
  TraceArgs A;
  A.cams = ctx->d_cams;
  A.n_cams = n_cams;
  A.batch = o.refill_batch > 0 ? o.refill_batch : (o.integrator == BH_INTEGRATOR_DP45 ? 8 : 16);
  A.n_total = total;
  A.head = ctx->d_head;
  A.o = o;
  A.luts = disk_luts;
  A.lut_n = disk_lut_n;
  std::memset(&A.vol, 0, sizeof(A.vol));
  A.k2tab = nullptr;
  A.k2_n = 0;
  A.k2_lo = A.k2_hi = 0.0;
  if (vol_on) {
    A.vol = *vol;
    A.k2tab = k2->tab;
    A.k2_n = k2->n;
    A.k2_lo = k2->lnth_min;
    A.k2_hi = k2->lnth_max;
  }
  A.out = out;
  if (!g_have_events) {
    if (hipEventCreate(&g_ev0) != hipSuccess || hipEventCreate(&g_ev1) != hipSuccess)
      return fail("bh_trace_image_plane: hipEventCreate failed");
    g_have_events = true;
  }
  hipLaunchKernelGGL(k_reset_head, dim3(1), dim3(64), 0, s, ctx->d_head);
  if (hipGetLastError() != hipSuccess) return fail("bh_trace_image_plane: reset launch failed");
  (void)hipEventRecord(g_ev0, s);
  hipError_t e;
  if (o.integrator == BH_INTEGRATOR_DP45)
    e = vol_on ? launch<BH_INTEGRATOR_DP45, true>(A, s, 0) : launch<BH_INTEGRATOR_DP45, false>(A, s, 0);
  else
    e = vol_on ? launch<BH_INTEGRATOR_RK4, true>(A, s, 0) : launch<BH_INTEGRATOR_RK4, false>(A, s, 0);
  if (e != hipSuccess) return fail(std::string("bh_trace_image_plane: launch failed: ") + hipGetErrorString(e));
  (void)hipEventRecord(g_ev1, s);
  return 0;
}

float bh_last_kernel_ms(void) {
  if (!g_have_events) return -1.0f;
  float ms = -1.0f;
  if (hipEventElapsedTime(&ms, g_ev0, g_ev1) != hipSuccess) return -1.0f;
  return ms;
}

}  // extern "C"
