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
  PolAcc pol;  // polarised transfer (POL kernels)
  double r_stop, r_isco;
  long long gid;
  int cam;
  int vb;  // volume of a volume set (multi-volume launches)
  int npts;  // path points written (path recording)
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
  // a pinhole camera's table entry maps pixels to their centres (px + 0.5, py + 0.5)
  L.c = P.pin ? setup_pinhole(*P.pin, k.a, alpha, beta, L.y) : setup_ray_ab(k, alpha, beta, L.y);
  rhs(L.c, L.y, L.k1);
  L.h = P.h_init;
  L.rt.I = 0.0;
  L.rt.Tr = 1.0;
  L.rt.tau = 0.0;
  L.rt.a0 = 0.0;
  L.rt.e0 = 0.0;
  L.rt.s0 = VOL ? affine_rate(L.c, L.y[0], L.y[1]) : 0.0;
  clear_crossing_slots(P.out + L.gid);
  L.r_stop = k.r_stop;
  L.r_isco = k.r_isco;
  L.vb = k.vol_index;
  L.steps = L.attempts = L.ncross = L.nsamp = 0;
  L.last_rej = false;
  return true;
}

// Volume transfer over the accepted step of length h ending at y.  MV: the
// lane's camera samples its own volume of a volume set (same geometry and
// support as P.vol, own cells and fluid angular momentum).  POL: the linear
// polarisation is carried beside I (pol_segment).
template <bool MV, bool POL>
__device__ __forceinline__ void volume_segment(Lane& L, const DevParams& P, const double* y,
                                               double h) {
  double a1, e1, r1 = 0.0;
  VolDev v = P.vol;
  if (MV) {
    v.cells = P.buckets[L.vb].cells;
    v.ell = P.buckets[L.vb].ell;
  }
  if (volume_point(v, L.c, y[0], y[1], y[2], P.rad, &a1, &e1, POL ? &r1 : nullptr))
    L.nsamp++;
  const double s1 = affine_rate(L.c, y[0], y[1]);
  if (POL) pol_segment(L.pol, L.rt, h, a1, e1, r1, s1, P.pol_c2, P.pol_s2, P.pol_frac);
  rt_segment(L.rt, h, a1, e1, s1);
}

// Equatorial crossing between (y0, f0) and (y1, f1); true if the disk ends the ray.
__device__ __forceinline__ bool crossing(Lane& L, const DevParams& P, const double* y0,
                                         const double* f0, const double* y1, const double* f1,
                                         double h) {
  if ((y0[1] < 0.0) == (y1[1] < 0.0)) return false;
  const double rc = crossing_radius(y0, f0, y1, f1, h);
  const double g = (rc >= L.r_isco) ? disk_g(rc, L.c.a, L.c.xi) : 0.0;
  if (L.ncross < P.max_crossings) store_crossing(P.out + L.gid, L.ncross, rc, g);
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
  // outbound only: a camera beyond r_escape traces its rays in first (inside
  // r_escape the test equals r > r_escape, no inbound ray can exceed it)
  if (L.y[0] > P.r_escape && L.y[3] > 0.0) return BH_HIT_ESCAPED;
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

// Appends the lane's current position to its polyline (path recording).
__device__ __forceinline__ void path_point(const DevParams& P, Lane& L) {
  if (L.npts >= P.path_max) return;
  const double r = L.y[0], u = L.y[1];
  const double st = __builtin_sqrt(fmax(__builtin_fma(-u, u, 1.0), 0.0));
  double sp, cp;
  sincos(L.y[2], &sp, &cp);
  float* q = P.path + (L.gid * static_cast<long long>(P.path_max) + L.npts) * 3;
  q[0] = static_cast<float>(r * st * cp);
  q[1] = static_cast<float>(r * st * sp);
  q[2] = static_cast<float>(r * u);
  L.npts++;
}

// One RK4 step; returns the terminal status (0 = continue).
template <bool VOL, bool MV, bool PATH, bool POL>
__device__ __forceinline__ int step_rk4(Lane& L, const DevParams& P) {
  if (L.steps >= P.max_steps) return BH_HIT_MAX_STEPS;
  const double h = P.h_fixed;
  double yn[kN], fn[kN];
  rk4_step(L.c, L.y, L.k1, h, yn);
  rhs(L.c, yn, fn);
  project_radial(L.c, yn, fn);
  L.steps++;
  L.attempts++;
  if (VOL) volume_segment<MV, POL>(L, P, yn, h);
  const bool hit = crossing(L, P, L.y, L.k1, yn, fn, h);
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    L.y[i] = yn[i];
    L.k1[i] = fn[i];
  }
  if (PATH) path_point(P, L);
  if (hit) return BH_HIT_DISK;
  return terminate_check(L, P);
}

// One DP45 attempt; returns the terminal status (0 = continue).
template <bool VOL, bool MV, bool PATH, bool POL>
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
    if (VOL) volume_segment<MV, POL>(L, P, y5, L.h);
    const bool hit = crossing(L, P, L.y, L.k1, y5, k7, L.h);
#pragma unroll
    for (int i = 0; i < kN; ++i) {
      L.y[i] = y5[i];
      L.k1[i] = k7[i];
    }
    L.h = L.h * fac;
    L.last_rej = false;
    if (PATH) path_point(P, L);
    if (hit) return BH_HIT_DISK;
    return terminate_check(L, P);
  }
  L.h = L.h * dp45_factor(err, false, L.last_rej);
  L.last_rej = true;
  return 0;
}

// Workgroups per CU the register allocation targets.  Volume variants: two
// (they would otherwise take a few registers past the 256 that two waves per
// SIMD allow; the handful of spilled values cost less than the lost latency
// hiding, and a third wave per SIMD measured slower: c2 13.8 against 10.7
// ms).  Thin-disk variants: three waves per SIMD in 168 registers
// (c5 plain DP45 101 against 110 ms, records identical).
template <bool VOL, bool POL>
struct Fp64Occupancy {
  static constexpr int kBlocks = (VOL || POL) ? 2 : 3;
};
template <int INTEG, bool VOL, bool MV, bool PATH, bool POL = false>
__global__ void __launch_bounds__(kBlock, (Fp64Occupancy<VOL, POL>::kBlocks)) k_trace_fp64(
    const DevParams* __restrict__ Pp) {
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
        if (!lane_setup<VOL>(L, P, id)) {
          L.gid = -1;
        } else {
          if (PATH) {
            L.npts = 0;
            path_point(P, L);
          }
          if (POL) L.pol = PolAcc{0.0, 0.0, 1.0, 0.0, 0.0};
        }
      } else {
        done = true;
      }
    }
    if (__ballot(!done) == 0ull) break;
    if (L.gid >= 0) {
      for (int b = 0; b < P.batch; ++b) {
        const int st = (INTEG == BH_INTEGRATOR_DP45) ? step_dp45<VOL, MV, PATH, POL>(L, P)
                                                      : step_rk4<VOL, MV, PATH, POL>(L, P);
        if (st) {
          write_record(L, P, st);
          if (POL) P.stokes[L.gid] = make_double4(L.rt.I, L.pol.q, L.pol.u, 0.0);
          if (PATH) P.path_count[L.gid] = L.npts;
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

// Kernel selection: integrator x precision x volume x volume set.
struct KernelKey {
  int integ, prec;
  bool vol, multi;
  bool path = false;  // polyline recording (DP45, FP64, no volume only)
  bool pol = false;   // polarised transfer (FP64, one volume)
};

template <int I, bool V, bool M, bool PT = false, bool PO = false>
int max_blocks_per_cu() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
  if (cached[dev] > 0) return cached[dev];
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace_fp64<I, V, M, PT, PO>, kBlock, 0) != hipSuccess)
    return -1;
  cached[dev] = std::max(1, nb);
  return cached[dev];
}

int cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
  if (cached[dev] > 0) return cached[dev];
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return -1;
  cached[dev] = prop.multiProcessorCount;
  return cached[dev];
}

int occupancy_limit(const KernelKey& k) {
  if (k.prec == BH_PRECISION_FP32) return occupancy_limit_fp32(k.integ, k.vol);
  if (k.path) return max_blocks_per_cu<BH_INTEGRATOR_DP45, false, false, true>();
  if (k.pol)
    return k.integ == BH_INTEGRATOR_DP45 ? max_blocks_per_cu<BH_INTEGRATOR_DP45, true, false, false, true>()
                                         : max_blocks_per_cu<BH_INTEGRATOR_RK4, true, false, false, true>();
  if (k.integ == BH_INTEGRATOR_DP45) {
    if (k.multi) return max_blocks_per_cu<BH_INTEGRATOR_DP45, true, true>();
    return k.vol ? max_blocks_per_cu<BH_INTEGRATOR_DP45, true, false>()
                 : max_blocks_per_cu<BH_INTEGRATOR_DP45, false, false>();
  }
  if (k.multi) return max_blocks_per_cu<BH_INTEGRATOR_RK4, true, true>();
  return k.vol ? max_blocks_per_cu<BH_INTEGRATOR_RK4, true, false>() : max_blocks_per_cu<BH_INTEGRATOR_RK4, false, false>();
}

hipError_t launch_kernel(const KernelKey& k, const DevParams* dp, hipStream_t s, int grid) {
  if (k.prec == BH_PRECISION_FP32) return launch_fp32(k.integ, k.vol, dp, s, grid);
#define BHK_LAUNCH(I, V, M) \
  hipLaunchKernelGGL((k_trace_fp64<I, V, M, false>), dim3(grid), dim3(kBlock), 0, s, dp)
  if (k.path) {
    hipLaunchKernelGGL((k_trace_fp64<BH_INTEGRATOR_DP45, false, false, true>), dim3(grid), dim3(kBlock), 0, s, dp);
  } else if (k.pol) {
    if (k.integ == BH_INTEGRATOR_DP45)
      hipLaunchKernelGGL((k_trace_fp64<BH_INTEGRATOR_DP45, true, false, false, true>), dim3(grid), dim3(kBlock), 0, s,
                         dp);
    else
      hipLaunchKernelGGL((k_trace_fp64<BH_INTEGRATOR_RK4, true, false, false, true>), dim3(grid), dim3(kBlock), 0, s,
                         dp);
  } else if (k.integ == BH_INTEGRATOR_DP45) {
    if (k.multi) BHK_LAUNCH(BH_INTEGRATOR_DP45, true, true);
    else if (k.vol) BHK_LAUNCH(BH_INTEGRATOR_DP45, true, false);
    else BHK_LAUNCH(BH_INTEGRATOR_DP45, false, false);
  } else {
    if (k.multi) BHK_LAUNCH(BH_INTEGRATOR_RK4, true, true);
    else if (k.vol) BHK_LAUNCH(BH_INTEGRATOR_RK4, true, false);
    else BHK_LAUNCH(BH_INTEGRATOR_RK4, false, false);
  }
#undef BHK_LAUNCH
  return hipGetLastError();
}

// Autotuned launch shapes, per (device, kernel): read-mostly, mutex-guarded.
struct TuneEntry {
  int blocks_per_cu, refill_batch;
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, bool, bool>, TuneEntry> g_tune;

bool tuned_shape(const KernelKey& k, TuneEntry* e) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find({dev, k.integ, k.prec, k.vol, k.multi});
  if (it == g_tune.end()) return false;
  *e = it->second;
  return true;
}

void store_tuned_shape(const KernelKey& k, const TuneEntry& e) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[{dev, k.integ, k.prec, k.vol, k.multi}] = e;
}

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

// Host image of one launch: the parameter block and the tables it points to
// (camera constants, volume-set entries), laid out as they go to the device:
// [DevParams][CamDev x n_cams][VolBucket x n_buckets].
struct LaunchImage {
  DevParams P;
  std::vector<CamDev> cams;
  std::vector<VolBucket> buckets;
  std::vector<PinholeDev> pin;  // 0 or 1 entries
  KernelKey key;
  int blocks_per_cu;  // 0: tuned or occupancy limit

  size_t bytes() const {
    return sizeof(DevParams) + cams.size() * sizeof(CamDev) + buckets.size() * sizeof(VolBucket) +
           pin.size() * sizeof(PinholeDev);
  }
  // Serialises into host memory dst, with the table pointers of the block
  // fixed up for its device address d_block and the ray-queue head the
  // kernel's refills count on (d_head, device memory, must not be NULL).
  void serialize(unsigned char* dst, unsigned char* d_block, unsigned long long* d_head) const {
    DevParams q = P;
    q.head = d_head;
    q.cams = reinterpret_cast<const CamDev*>(d_block + sizeof(DevParams));
    q.buckets = buckets.empty() ? nullptr
                                : reinterpret_cast<const VolBucket*>(d_block + sizeof(DevParams) +
                                                                     cams.size() * sizeof(CamDev));
    const size_t pin_off = sizeof(DevParams) + cams.size() * sizeof(CamDev) + buckets.size() * sizeof(VolBucket);
    q.pin = pin.empty() ? nullptr : reinterpret_cast<const PinholeDev*>(d_block + pin_off);
    std::memcpy(dst, &q, sizeof(q));
    if (!cams.empty()) std::memcpy(dst + sizeof(DevParams), cams.data(), cams.size() * sizeof(CamDev));
    if (!buckets.empty())
      std::memcpy(dst + sizeof(DevParams) + cams.size() * sizeof(CamDev), buckets.data(),
                  buckets.size() * sizeof(VolBucket));
    if (!pin.empty()) std::memcpy(dst + pin_off, pin.data(), sizeof(PinholeDev));
  }
  // Grid of the persistent kernel.
  int grid() const {
    int bpc = blocks_per_cu;
    if (bpc <= 0) {
      TuneEntry e;
      if (tuned_shape(key, &e)) bpc = e.blocks_per_cu;
    }
    const int lim = occupancy_limit(key);
    const int n_cu = cu_count();
    if (lim <= 0 || n_cu <= 0) return -1;
    if (bpc <= 0 || bpc > lim) bpc = lim;
    return bpc * n_cu;
  }
};

VolDev vol_dev(const BH_Volume& v) {
  VolDev d;
  d.cells = static_cast<const float4*>(v.cells);
  d.n_r = v.n_r;
  d.n_th = v.n_th;
  d.n_ph = v.n_ph;
  d.r_min = v.r_min;
  d.r_max = v.r_max;
  d.r_span = v.r_max - v.r_min;
  d.sup_r_lo = v.sup_r_lo;
  d.sup_r_hi = v.sup_r_hi;
  d.sup_u_max = v.sup_u_max;
  d.ell = v.ell;
  d.s_r = static_cast<double>(v.n_r) / d.r_span;
  d.s_th = static_cast<double>(v.n_th) / kPi;
  d.s_ph = static_cast<double>(v.n_ph) / kTwoPi;
  d.inv_nph = 1.0 / static_cast<double>(v.n_ph);
  return d;
}

int check_volume(const BH_Volume* vol, const char* who) {
  const std::string w(who);
  if (!vol || !vol->cells || vol->n_r < 1 || vol->n_th < 1 || vol->n_ph < 1 || !(vol->r_max > vol->r_min))
    return fail(w + ": volume enabled with an invalid BH_Volume");
  if (static_cast<long long>(vol->n_r) * vol->n_th * vol->n_ph >= (1ll << 31))
    return fail(w + ": volume grid too large (>= 2^31 cells)");
  return 0;
}

// Validates a launch and builds its image: every pixel of the cameras
// (rays == ids == NULL), a list of pixel ids, or an explicit ray list; one
// volume (vol) or a volume set (vs).
int build_launch(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts, const double* disk_luts,
                 int disk_lut_n, const BH_Volume* vol, const BH_VolumeSet* vs, const BH_RaySpec* rays,
                 long long n_rays, const int64_t* ids, const uint64_t* n_ids_dev, long long max_ids,
                 BH_PixelRecord* out, bool compact, const char* who, LaunchImage* im) {
  const std::string w(who);
  if (!rays && (!cams || n_cams <= 0)) return fail(w + ": no cameras");
  if (rays && n_rays <= 0) return fail(w + ": empty ray list");
  if (ids && (!n_ids_dev || max_ids <= 0)) return fail(w + ": invalid id list");
  if (rays && opts && opts->disk_enabled) return fail(w + ": the thin disk must be disabled for explicit rays");
  if (!opts) return fail(w + ": NULL options");
  if (!out) return fail(w + ": NULL output");
  const BH_TraceOptions& o = *opts;
  if (o.integrator != BH_INTEGRATOR_RK4 && o.integrator != BH_INTEGRATOR_DP45) return fail(w + ": unknown integrator");
  if (o.precision != BH_PRECISION_FP64 && o.precision != BH_PRECISION_FP32) return fail(w + ": unknown precision");
  if (o.precision == BH_PRECISION_FP32 && !(o.island_r_factor >= 1.0 && o.island_turn_dist >= 0.0))
    return fail(w + ": FP32 form needs island_r_factor >= 1, island_turn_dist >= 0");
  if (o.integrator == BH_INTEGRATOR_RK4 && !(o.h_fixed > 0.0)) return fail(w + ": h_fixed must be > 0");
  if (o.integrator == BH_INTEGRATOR_DP45 && !(o.h_init > 0.0 && o.h_max > 0.0 && o.rtol > 0.0 && o.atol > 0.0))
    return fail(w + ": DP45 needs h_init, h_max, rtol, atol > 0");
  if (o.max_crossings < 0 || o.max_crossings > 3) return fail(w + ": max_crossings must be in [0, 3]");
  if (o.max_steps < 0 || o.max_attempts < 0) return fail(w + ": negative budget");
  if (o.blocks_per_cu < 0 || o.refill_batch < 0) return fail(w + ": negative launch shape");
  if (o.disk_enabled && (!disk_luts || disk_lut_n < 2)) return fail(w + ": disk enabled without disk LUTs");
  const bool vol_on = o.volume_enabled != 0;
  const bool multi = vol_on && vs != nullptr;
  if (vol_on) {
    if (multi) {
      if (!vs->vols || vs->n_vols < 1 || !vs->cam_vol) return fail(w + ": invalid BH_VolumeSet");
      if (rays) return fail(w + ": a volume set needs cameras");
      if (o.precision == BH_PRECISION_FP32) return fail(w + ": the FP32 form takes a single volume");
      for (int i = 0; i < vs->n_vols; ++i) {
        const BH_Volume& a = vs->vols[i];
        const BH_Volume& b = vs->vols[0];
        if (check_volume(&a, who) != 0) return -1;
        if (a.n_r != b.n_r || a.n_th != b.n_th || a.n_ph != b.n_ph || a.r_min != b.r_min || a.r_max != b.r_max ||
            a.sup_r_lo != b.sup_r_lo || a.sup_r_hi != b.sup_r_hi || a.sup_u_max != b.sup_u_max)
          return fail(w + ": the volumes of a set must share grid geometry and support");
      }
    } else if (check_volume(vol, who) != 0) {
      return -1;
    }
    if (o.integrator == BH_INTEGRATOR_DP45 && !(o.h_vol > 0.0)) return fail(w + ": h_vol must be > 0");
  }
  im->cams.clear();
  im->buckets.clear();
  long long total = 0;
  if (!rays && make_cam_devs(cams, n_cams, o.horizon_factor, who, &im->cams, &total) != 0) return -1;
  if (multi) {
    for (int i = 0; i < n_cams; ++i) {
      const int v = vs->cam_vol[i];
      if (v < 0 || v >= vs->n_vols) return fail(w + ": camera volume index out of range");
      im->cams[i].vol_index = v;
    }
    for (int i = 0; i < vs->n_vols; ++i)
      im->buckets.push_back(VolBucket{static_cast<const float4*>(vs->vols[i].cells), vs->vols[i].ell});
  }
  DevParams& P = im->P;
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
  if (vol_on) P.vol = vol_dev(multi ? vs->vols[0] : *vol);
  P.luts = disk_luts;
  P.out = out;
  P.n_total = rays ? n_rays : (ids ? max_ids : total);
  P.rays = rays;
  P.ids = reinterpret_cast<const long long*>(ids);
  P.count_dev = reinterpret_cast<const unsigned long long*>(ids ? n_ids_dev : nullptr);
  P.horizon_factor = o.horizon_factor;
  P.n_cams = rays ? 0 : n_cams;
  P.lut_n = disk_lut_n;
  KernelKey key{o.integrator, o.precision, vol_on, multi};
  int batch = o.refill_batch;
  if (batch <= 0) {
    TuneEntry e;
    batch = tuned_shape(key, &e) ? e.refill_batch : (o.integrator == BH_INTEGRATOR_DP45 ? 8 : 16);
  }
  P.batch = batch;
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
    P.CjL = static_cast<float>(1.7320508075688772 * kECharge * kECharge * kECharge / me_c2 * nuB_per_G * 0.5 *
                               o.length_unit_cm);
    P.Cd = static_cast<float>(o.disk_T_max > 0.0 ? kH * nu / (kKB * o.disk_T_max) : 1e30);
    P.rad.CT = kH * nu / me_c2;
    P.rad.Cx = nu / (1.5 * nuB_per_G);
    P.rad.CjL = 1.7320508075688772 * kECharge * kECharge * kECharge / me_c2 * nuB_per_G * 0.5 * o.length_unit_cm;
    P.rad.SB = 2.0 * kH * nu * nu * nu / (kC * kC);
    P.rad.iSB = 1.0 / P.rad.SB;
  }
  im->key = key;
  im->blocks_per_cu = o.blocks_per_cu;
  return 0;
}

// Queue reset and the persistent kernel: no host work, so it can be captured.
int launch_image(const LaunchImage& im, const DevParams* d_params, unsigned long long* d_head, hipStream_t s,
                 const char* who) {
  if (!d_params || !d_head) return fail(std::string(who) + ": launch block or queue head missing");
  const int grid = im.grid();
  if (grid <= 0) return fail(std::string(who) + ": no launch configuration for this device");
  hipLaunchKernelGGL(k_reset_head, dim3(1), dim3(64), 0, s, d_head);
  const hipError_t e = launch_kernel(im.key, d_params, s, grid);
  if (e != hipSuccess) return fail(std::string(who) + ": launch failed: " + hipGetErrorString(e));
  return 0;
}

// Per (device, stream) context of the one-shot entry points: one device
// block for the launch image and the ray-queue head, reused across launches
// on the same stream (stream order serialises them).
struct StreamCtx {
  unsigned char* d_block = nullptr;
  size_t cap = 0;
  unsigned long long* d_head = nullptr;
};
std::mutex g_mu;
std::map<std::pair<int, void*>, StreamCtx> g_ctx;

// Uploads a launch image into the stream's context block and launches it.
int launch_on_stream(const LaunchImage& im, void* stream, const char* who) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(std::string(who) + ": no HIP device");
  const size_t bytes = im.bytes();
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
        return fail(std::string(who) + ": hipMalloc of the launch block failed");
      ctx->cap = cap;
    }
    if (!ctx->d_head && hipMalloc(&ctx->d_head, 256) != hipSuccess)
      return fail(std::string(who) + ": hipMalloc of the queue head failed");
  }
  std::vector<unsigned char> blk(bytes);
  im.serialize(blk.data(), ctx->d_block, ctx->d_head);
  // Pageable source: the runtime stages it before returning, so blk may die.
  if (hipMemcpyAsync(ctx->d_block, blk.data(), bytes, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(std::string(who) + ": parameter upload failed");
  if (!g_have_events) {
    if (hipEventCreate(&g_ev0) != hipSuccess || hipEventCreate(&g_ev1) != hipSuccess)
      return fail(std::string(who) + ": hipEventCreate failed");
    g_have_events = true;
  }
  (void)hipEventRecord(g_ev0, s);
  if (launch_image(im, reinterpret_cast<const DevParams*>(ctx->d_block), ctx->d_head, s, who) != 0) return -1;
  (void)hipEventRecord(g_ev1, s);
  return 0;
}

// Allocation-free launches (bh_trace_image_plane_noalloc): the queue head
// and the launch block live in this file-scope device array, one copy per
// GPU, held until the process ends.  Layout: [0, 256) the ray-queue head,
// [256, kNoAllocBytes) the launch block, sized for kNoAllocMaxCams cameras
// without a volume set (about 64 KiB).  Every call on a GPU shares its
// copy, so calls that overlap on one GPU (different streams or threads)
// race; calls on one stream are ordered by the stream.
constexpr size_t kNoAllocMaxCams = 540;
constexpr size_t kNoAllocHead = 256;
constexpr size_t kNoAllocBytes =
    (kNoAllocHead + sizeof(DevParams) + kNoAllocMaxCams * sizeof(CamDev) + 255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_noalloc_mu;
unsigned char* g_noalloc_ptr[64] = {};  // device address of scratch, per device

// Device address of scratch on the current device, looked up once per device
// with that device current.  The array's name is a host-side stand-in and is
// never passed to the device.
unsigned char* noalloc_scratch(const char* who) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    fail(std::string(who) + ": no HIP device");
    return nullptr;
  }
  if (dev < 0 || dev >= 64) {
    fail(std::string(who) + ": device index beyond 63");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fail(std::string(who) + ": hipGetSymbolAddress(scratch) failed: " + hipGetErrorString(e));
      return nullptr;
    }
    g_noalloc_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return g_noalloc_ptr[dev];
}

// Uploads a launch image into the device's scratch array and launches it:
// no device allocation and no event.
int launch_noalloc(const LaunchImage& im, void* stream, const char* who) {
  const size_t bytes = im.bytes();
  if (bytes > kNoAllocBytes - kNoAllocHead || im.cams.size() > kNoAllocMaxCams)
    return fail(std::string(who) + ": launch block of " + std::to_string(bytes) + " bytes (" +
                std::to_string(im.cams.size()) + " cameras) exceeds the " +
                std::to_string(kNoAllocBytes - kNoAllocHead) + "-byte scratch (at most " +
                std::to_string(kNoAllocMaxCams) + " cameras without a volume set)");
  unsigned char* base = noalloc_scratch(who);
  if (!base) return -1;
  unsigned long long* d_head = reinterpret_cast<unsigned long long*>(base);
  unsigned char* d_block = base + kNoAllocHead;
  thread_local unsigned char host_img[kNoAllocBytes];
  im.serialize(host_img, d_block, d_head);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // Pageable source: the runtime stages it before returning, so host_img may
  // be reused by the next call on this thread.
  if (hipMemcpyAsync(d_block, host_img, bytes, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(std::string(who) + ": parameter upload failed");
  return launch_image(im, reinterpret_cast<const DevParams*>(d_block), d_head, s, who);
}

int trace_launch(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts, const double* disk_luts,
                 int disk_lut_n, const BH_Volume* vol, const BH_VolumeSet* vs, const BH_RaySpec* rays,
                 long long n_rays, const int64_t* ids, const uint64_t* n_ids_dev, long long max_ids,
                 BH_PixelRecord* out, void* stream, bool compact, const char* who) {
  LaunchImage im;
  if (build_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, vs, rays, n_rays, ids, n_ids_dev, max_ids, out,
                   compact, who, &im) != 0)
    return -1;
  return launch_on_stream(im, stream, who);
}

int trace_paths(const BH_Camera& cam, const BH_TraceOptions& opts, float* d_path, int* d_count, int path_max,
                BH_PixelRecord* out, void* stream, const char* who) {
  if (opts.integrator != BH_INTEGRATOR_DP45 || opts.precision != BH_PRECISION_FP64 || opts.volume_enabled ||
      opts.disk_enabled)
    return fail(std::string(who) + ": path recording runs DP45 FP64 without disk or volume");
  if (!d_path || !d_count || path_max < 1) return fail(std::string(who) + ": invalid path buffers");
  LaunchImage im;
  if (build_launch(&cam, 1, &opts, nullptr, 0, nullptr, nullptr, nullptr, 0, nullptr, nullptr, 0, out, false, who,
                   &im) != 0)
    return -1;
  im.P.path = d_path;
  im.P.path_count = d_count;
  im.P.path_max = path_max;
  im.key.path = true;
  return launch_on_stream(im, stream, who);
}

int trace_pinhole(const BH_Camera& cam, const BH_TraceOptions& opts, const double* disk_lut, int lut_n,
                  const PinholeDev& pin, double r_isco, BH_PixelRecord* out, void* stream, const char* who) {
  LaunchImage im;
  if (build_launch(&cam, 1, &opts, disk_lut, lut_n, nullptr, nullptr, nullptr, 0, nullptr, nullptr, 0, out, false,
                   who, &im) != 0)
    return -1;
  if (opts.volume_enabled) return fail(std::string(who) + ": the pinhole camera takes no volume");
  if (r_isco > 0.0) im.cams[0].r_isco = r_isco;
  im.pin.push_back(pin);
  return launch_on_stream(im, stream, who);
}

}  // namespace bhk

// A prepared frame: its launch image lives in device memory owned by the
// frame, so launching it needs no host work (and can be captured in a graph).
struct BH_Frame {
  bhk::LaunchImage im;
  unsigned char* d_block = nullptr;
  unsigned long long* d_head = nullptr;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int device = 0;
  // id sample for autotuning: every k-th pixel, records to a compact buffer
  long long* d_tune_ids = nullptr;
  unsigned long long* d_tune_count = nullptr;
  BH_PixelRecord* d_tune_out = nullptr;
  unsigned char* d_tune_block = nullptr;
  bhk::LaunchImage tune_im;
};

using namespace bhk;

namespace {

void frame_drop_graph(BH_Frame* f) {
  if (f->exec) (void)hipGraphExecDestroy(f->exec);
  if (f->graph) (void)hipGraphDestroy(f->graph);
  f->exec = nullptr;
  f->graph = nullptr;
}

int frame_upload(BH_Frame* f) {
  if (!f->d_block || !f->d_head) return fail("bh_frame: frame without device memory");
  std::vector<unsigned char> blk(f->im.bytes());
  f->im.serialize(blk.data(), f->d_block, f->d_head);
  if (hipMemcpy(f->d_block, blk.data(), blk.size(), hipMemcpyHostToDevice) != hipSuccess)
    return fail("bh_frame: parameter upload failed");
  return 0;
}

}  // namespace

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

extern "C" {

int bh_trace_image_plane(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                         const double* disk_luts, int disk_lut_n, const BH_Volume* vol,
                         BH_PixelRecord* out, void* stream) {
  return trace_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, nullptr, nullptr, 0, nullptr, nullptr, 0, out,
                      stream, false, "bh_trace_image_plane");
}

int bh_trace_image_plane_noalloc(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                                 const double* disk_luts, int disk_lut_n, const BH_Volume* vol,
                                 BH_PixelRecord* out, void* stream) {
  const char* who = "bh_trace_image_plane_noalloc";
  LaunchImage im;
  if (build_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, nullptr, nullptr, 0, nullptr, nullptr, 0, out,
                   false, who, &im) != 0)
    return -1;
  return launch_noalloc(im, stream, who);
}

int bh_trace_image_plane_pol(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts, const double* disk_luts,
                             int disk_lut_n, const BH_Volume* vol, const BH_PolParams* pp, BH_PixelRecord* out,
                             double* stokes, void* stream) {
  const char* who = "bh_trace_image_plane_pol";
  if (!pp || !stokes) return fail(std::string(who) + ": NULL polarisation parameters or Stokes buffer");
  if (!opts || !opts->volume_enabled || !vol)
    return fail(std::string(who) + ": polarised transfer needs volume_enabled and a volume");
  if (opts->precision != BH_PRECISION_FP64) return fail(std::string(who) + ": the polarised form runs in FP64");
  if (!(pp->lin_pol_frac >= 0.0 && pp->lin_pol_frac <= 1.0) || !std::isfinite(pp->evpa_b) ||
      !std::isfinite(pp->faraday_per_ne))
    return fail(std::string(who) + ": lin_pol_frac must be in [0, 1], evpa_b and faraday_per_ne finite");
  LaunchImage im;
  if (build_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, nullptr, nullptr, 0, nullptr, nullptr, 0, out,
                   false, who, &im) != 0)
    return -1;
  im.P.stokes = reinterpret_cast<double4*>(stokes);
  im.P.pol_c2 = std::cos(2.0 * pp->evpa_b);
  im.P.pol_s2 = std::sin(2.0 * pp->evpa_b);
  im.P.pol_frac = pp->lin_pol_frac;
  im.P.faraday_per_ne = pp->faraday_per_ne;
  im.P.rad.FL = pp->faraday_per_ne * im.P.length_unit_cm;
  im.key.pol = true;
  return launch_on_stream(im, stream, who);
}

int bh_trace_image_plane_multi(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                               const double* disk_luts, int disk_lut_n, const BH_VolumeSet* vs,
                               BH_PixelRecord* out, void* stream) {
  if (!vs) return fail("bh_trace_image_plane_multi: NULL volume set");
  if (!opts || !opts->volume_enabled) return fail("bh_trace_image_plane_multi: volume_enabled must be set");
  return trace_launch(cams, n_cams, opts, disk_luts, disk_lut_n, nullptr, vs, nullptr, 0, nullptr, nullptr, 0, out,
                      stream, false, "bh_trace_image_plane_multi");
}

int bh_trace_pixel_list(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                        const double* disk_luts, int disk_lut_n, const BH_Volume* vol, const int64_t* ids,
                        const uint64_t* n_ids_dev, int64_t max_ids, BH_PixelRecord* out, void* stream) {
  if (!ids) return fail("bh_trace_pixel_list: NULL id list");
  return trace_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, nullptr, nullptr, 0, ids, n_ids_dev, max_ids,
                      out, stream, false, "bh_trace_pixel_list");
}

int bh_trace_pixel_list_compact(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                                const double* disk_luts, int disk_lut_n, const BH_Volume* vol,
                                const int64_t* ids, const uint64_t* n_ids_dev, int64_t max_ids,
                                BH_PixelRecord* out, void* stream) {
  if (!ids) return fail("bh_trace_pixel_list_compact: NULL id list");
  return trace_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, nullptr, nullptr, 0, ids, n_ids_dev, max_ids,
                      out, stream, true, "bh_trace_pixel_list_compact");
}

int bh_trace_rays(const BH_RaySpec* rays, int64_t n, const BH_TraceOptions* opts, const BH_Volume* vol,
                  BH_PixelRecord* out, void* stream) {
  if (!rays) return fail("bh_trace_rays: NULL ray list");
  return trace_launch(nullptr, 0, opts, nullptr, 0, vol, nullptr, rays, n, nullptr, nullptr, 0, out, stream, false,
                      "bh_trace_rays");
}

int bh_trace_rays_host(const BH_RaySpec* rays, int64_t n, const BH_TraceOptions* opts, BH_PixelRecord* out) {
  if (!rays || !out || n <= 0) return fail("bh_trace_rays_host: invalid arguments");
  BH_RaySpec* d_rays = nullptr;
  BH_PixelRecord* d_out = nullptr;
  int rc = 0;
  if (hipMalloc(&d_rays, n * sizeof(BH_RaySpec)) != hipSuccess || hipMalloc(&d_out, n * sizeof(BH_PixelRecord)) != hipSuccess) {
    rc = fail("bh_trace_rays_host: device allocation failed");
  } else if (hipMemcpy(d_rays, rays, n * sizeof(BH_RaySpec), hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail("bh_trace_rays_host: upload failed");
  } else if (bh_trace_rays(d_rays, n, opts, nullptr, d_out, nullptr) != 0) {
    rc = -1;
  } else if (hipStreamSynchronize(nullptr) != hipSuccess ||
             hipMemcpy(out, d_out, n * sizeof(BH_PixelRecord), hipMemcpyDeviceToHost) != hipSuccess) {
    rc = fail("bh_trace_rays_host: trace or copy failed");
  }
  if (d_rays) (void)hipFree(d_rays);
  if (d_out) (void)hipFree(d_out);
  return rc;
}

float bh_last_kernel_ms(void) {
  if (!g_have_events) return -1.0f;
  float ms = -1.0f;
  if (hipEventElapsedTime(&ms, g_ev0, g_ev1) != hipSuccess) return -1.0f;
  return ms;
}

BH_Frame* bh_frame_create(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts, const double* disk_luts,
                          int disk_lut_n, const BH_Volume* vol, const BH_VolumeSet* vs, BH_PixelRecord* out) {
  BH_Frame* f = new BH_Frame();
  if (build_launch(cams, n_cams, opts, disk_luts, disk_lut_n, vol, vs, nullptr, 0, nullptr, nullptr, 0, out, false,
                   "bh_frame_create", &f->im) != 0) {
    delete f;
    return nullptr;
  }
  if (hipGetDevice(&f->device) != hipSuccess || hipMalloc(&f->d_block, f->im.bytes()) != hipSuccess ||
      hipMalloc(&f->d_head, 256) != hipSuccess) {
    fail("bh_frame_create: device allocation failed");
    bh_frame_destroy(f);
    return nullptr;
  }
  if (frame_upload(f) != 0) {
    bh_frame_destroy(f);
    return nullptr;
  }
  return f;
}

void bh_frame_destroy(BH_Frame* f) {
  if (!f) return;
  frame_drop_graph(f);
  if (f->capture_stream) (void)hipStreamDestroy(f->capture_stream);
  auto fr = [](void* p) { if (p) (void)hipFree(p); };
  fr(f->d_block);
  fr(f->d_head);
  fr(f->d_tune_ids);
  fr(f->d_tune_count);
  fr(f->d_tune_out);
  fr(f->d_tune_block);
  delete f;
}

int bh_frame_launch(BH_Frame* f, void* stream) {
  if (!f) return fail("bh_frame_launch: NULL frame");
  return launch_image(f->im, reinterpret_cast<const DevParams*>(f->d_block), f->d_head,
                      static_cast<hipStream_t>(stream), "bh_frame_launch");
}

int bh_frame_graph_launch(BH_Frame* f, void* stream) {
  if (!f) return fail("bh_frame_graph_launch: NULL frame");
  if (!f->exec) {
    if (!f->capture_stream && hipStreamCreateWithFlags(&f->capture_stream, hipStreamNonBlocking) != hipSuccess)
      return fail("bh_frame_graph_launch: hipStreamCreate failed");
    if (hipStreamBeginCapture(f->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail("bh_frame_graph_launch: begin capture failed");
    const int rc = launch_image(f->im, reinterpret_cast<const DevParams*>(f->d_block), f->d_head, f->capture_stream,
                                "bh_frame_graph_launch");
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(f->capture_stream, &g);
    if (rc != 0 || e != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      return rc != 0 ? -1 : fail("bh_frame_graph_launch: end capture failed");
    }
    f->graph = g;
    if (hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0) != hipSuccess) {
      frame_drop_graph(f);
      return fail("bh_frame_graph_launch: graph instantiation failed");
    }
  }
  const hipError_t e = hipGraphLaunch(f->exec, static_cast<hipStream_t>(stream));
  return e == hipSuccess ? 0 : fail(std::string("bh_frame_graph_launch: ") + hipGetErrorString(e));
}

int bh_frame_set_launch_shape(BH_Frame* f, int blocks_per_cu, int refill_batch) {
  if (!f || blocks_per_cu < 0 || refill_batch < 0) return fail("bh_frame_set_launch_shape: invalid arguments");
  f->im.blocks_per_cu = blocks_per_cu;
  if (refill_batch > 0) f->im.P.batch = refill_batch;
  frame_drop_graph(f);  // the captured launch has the old shape
  return frame_upload(f);
}

int bh_frame_autotune(BH_Frame* f, int sample_stride, BH_TuneResult* best, void* stream) {
  if (!f || sample_stride < 1 || !best) return fail("bh_frame_autotune: invalid arguments");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long total = f->im.P.n_total;
  const long long n_s = (total + sample_stride - 1) / sample_stride;
  if (!f->d_tune_ids) {
    std::vector<long long> ids(n_s);
    for (long long i = 0; i < n_s; ++i) ids[i] = i * sample_stride;
    const unsigned long long cnt = static_cast<unsigned long long>(n_s);
    if (hipMalloc(&f->d_tune_ids, n_s * sizeof(long long)) != hipSuccess ||
        hipMalloc(&f->d_tune_count, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&f->d_tune_out, n_s * sizeof(BH_PixelRecord)) != hipSuccess ||
        hipMalloc(&f->d_tune_block, f->im.bytes()) != hipSuccess)
      return fail("bh_frame_autotune: device allocation failed");
    if (hipMemcpy(f->d_tune_ids, ids.data(), n_s * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->d_tune_count, &cnt, sizeof(cnt), hipMemcpyHostToDevice) != hipSuccess)
      return fail("bh_frame_autotune: sample upload failed");
  }
  // the sample launch: the frame's image on an id list into the compact buffer
  f->tune_im = f->im;
  f->tune_im.P.ids = f->d_tune_ids;
  f->tune_im.P.count_dev = f->d_tune_count;
  f->tune_im.P.n_total = n_s;
  f->tune_im.P.compact_ids = 1;
  f->tune_im.P.out = f->d_tune_out;
  const int lim = occupancy_limit(f->im.key);
  if (lim <= 0) return fail("bh_frame_autotune: no occupancy for this kernel");
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail("bh_frame_autotune: hipEventCreate failed");
  const int batches[] = {1, 2, 4, 8, 16, 32};
  float best_ms = 1e30f;
  int best_bpc = lim, best_batch = f->im.P.batch;
  int rc = 0;
  for (int bpc = 1; bpc <= lim && rc == 0; ++bpc) {
    for (int batch : batches) {
      f->tune_im.blocks_per_cu = bpc;
      f->tune_im.P.batch = batch;
      std::vector<unsigned char> blk(f->tune_im.bytes());
      f->tune_im.serialize(blk.data(), f->d_tune_block, f->d_head);
      if (hipMemcpyAsync(f->d_tune_block, blk.data(), blk.size(), hipMemcpyHostToDevice, s) != hipSuccess) {
        rc = fail("bh_frame_autotune: upload failed");
        break;
      }
      float ms = 0.0f;
      for (int rep = 0; rep < 2 && rc == 0; ++rep) {  // first run warms caches
        (void)hipEventRecord(e0, s);
        rc = launch_image(f->tune_im, reinterpret_cast<const DevParams*>(f->d_tune_block), f->d_head, s,
                          "bh_frame_autotune");
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess) rc = fail("bh_frame_autotune: trial failed");
        if (rc == 0) (void)hipEventElapsedTime(&ms, e0, e1);
      }
      if (rc != 0) break;
      if (ms < best_ms) {
        best_ms = ms;
        best_bpc = bpc;
        best_batch = batch;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != 0) return -1;
  store_tuned_shape(f->im.key, TuneEntry{best_bpc, best_batch});
  best->blocks_per_cu = best_bpc;
  best->refill_batch = best_batch;
  best->occupancy_limit = lim;
  best->sample_rays = static_cast<int32_t>(std::min<long long>(n_s, 0x7fffffff));
  best->sample_ms = best_ms;
  best->pad0 = 0;
  return bh_frame_set_launch_shape(f, best_bpc, best_batch);
}

}  // extern "C"
