// bridge.hip - the Blender bridge's GPU entry points (include/bh_bridge.h) on
// the library's tracer and render launch.  Synchronous, host buffers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bh_bridge.h"
#include "../../include/bh_kernel_launch.h"
#include "trace_common.hpp"

namespace bhk {
int fail(const std::string& msg);

namespace {

// Device allocation that frees itself.
template <class T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  bool alloc(size_t n) { return hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T)) == hipSuccess; }
};

int check_observer(double a, double r_obs, double inc, const char* who) {
  const std::string w(who);
  if (!(a >= 0.0 && a < 1.0)) return fail(w + ": spin outside [0, 1)");
  if (!(inc > 0.0 && inc < kPi)) return fail(w + ": inclination outside (0, pi)");
  if (!(r_obs > 1.01 * bh_outer_horizon(a))) return fail(w + ": observer inside the horizon stop");
  return 0;
}

// Camera whose pixel centres fall on the inclusive grid lo + (hi - lo) i / (n - 1).
void inclusive_grid(double lo, double hi, int n, double* start, double* step) {
  if (n > 1) {
    *step = (hi - lo) / static_cast<double>(n - 1);
    *start = lo;
  } else {
    *step = 1.0;
    *start = 0.5 * (lo + hi);
  }
}

int trace_polylines(double a, double r_obs, double inc, double amin, double amax, int na, double bmin, double bmax,
                    int nb, int max_points, float* out_xyz, int* out_counts, const char* who) {
  if (!out_xyz || !out_counts || na < 1 || nb < 1 || max_points < 1) return fail(std::string(who) + ": invalid arguments");
  if (check_observer(a, r_obs, inc, who) != 0) return -1;
  double a0, da, b0, db;
  inclusive_grid(amin, amax, na, &a0, &da);
  inclusive_grid(bmin, bmax, nb, &b0, &db);
  // pixel (i, j): alpha = alpha_min + (i + 0.5) dalpha = a0 + i da; row j from
  // the top: beta = beta_max - (j + 0.5) dbeta = b0 + j db
  BH_Camera cam;
  std::memset(&cam, 0, sizeof(cam));
  cam.spin = a;
  cam.theta_obs = inc;
  cam.r_obs = r_obs;
  cam.alpha_min = a0 - 0.5 * da;
  cam.alpha_max = cam.alpha_min + na * da;
  cam.beta_max = b0 - 0.5 * db;
  cam.beta_min = cam.beta_max + nb * db;
  cam.width = na;
  cam.height = nb;
  BH_TraceOptions o;
  bh_trace_options_default(&o);
  o.integrator = BH_INTEGRATOR_DP45;
  o.precision = BH_PRECISION_FP64;
  o.disk_enabled = 0;
  o.max_steps = max_points - 1;
  o.max_attempts = 20 * max_points + 1000;
  o.r_escape = 3.0 * r_obs;
  const long long n = static_cast<long long>(na) * nb;
  DevBuf<float> path;
  DevBuf<int> cnt;
  DevBuf<BH_PixelRecord> rec;
  if (!path.alloc(static_cast<size_t>(n) * max_points * 3) || !cnt.alloc(n) || !rec.alloc(n))
    return fail(std::string(who) + ": device allocation failed");
  if (trace_paths(cam, o, path.p, cnt.p, max_points, rec.p, nullptr, who) != 0) return -1;
  if (hipStreamSynchronize(nullptr) != hipSuccess ||
      hipMemcpy(out_xyz, path.p, static_cast<size_t>(n) * max_points * 3 * sizeof(float), hipMemcpyDeviceToHost) !=
          hipSuccess ||
      hipMemcpy(out_counts, cnt.p, n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(std::string(who) + ": trace or copy failed");
  return 0;
}

int lensing_map(double a, double r_obs, double inc, int w, int h, float* out, const char* who) {
  if (!out || w < 1 || h < 1) return fail(std::string(who) + ": invalid arguments");
  if (check_observer(a, r_obs, inc, who) != 0) return -1;
  // the reference's grid: alpha = 30 (px / w - 0.5), beta = 30 (0.5 - py / h)
  const double da = 30.0 / w, db = 30.0 / h;
  BH_Camera cam;
  std::memset(&cam, 0, sizeof(cam));
  cam.spin = a;
  cam.theta_obs = inc;
  cam.r_obs = r_obs;
  cam.alpha_min = -15.0 - 0.5 * da;
  cam.alpha_max = cam.alpha_min + w * da;
  cam.beta_max = 15.0 + 0.5 * db;
  cam.beta_min = cam.beta_max - h * db;
  cam.width = w;
  cam.height = h;
  BH_TraceOptions o;
  bh_trace_options_default(&o);
  o.integrator = BH_INTEGRATOR_DP45;
  o.max_steps = 20000;
  o.disk_enabled = 1;
  o.disk_r_out = 100.0;
  o.r_escape = 3.0 * r_obs;
  const double r_isco = bh_isco_prograde(a);
  if (!(r_isco < o.disk_r_out)) return fail(std::string(who) + ": ISCO beyond the disk");
  std::vector<double> lut(256);
  if (bh_disk_lut_build(a, r_isco, o.disk_r_out, 256, lut.data()) != 0) return -1;
  const long long n = static_cast<long long>(w) * h;
  DevBuf<double> d_lut;
  DevBuf<BH_PixelRecord> rec;
  if (!d_lut.alloc(256) || !rec.alloc(n)) return fail(std::string(who) + ": device allocation failed");
  if (hipMemcpy(d_lut.p, lut.data(), 256 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return fail(std::string(who) + ": LUT upload failed");
  if (bh_trace_image_plane(&cam, 1, &o, d_lut.p, 256, nullptr, rec.p, nullptr) != 0) return -1;
  std::vector<BH_PixelRecord> host(n);
  if (hipStreamSynchronize(nullptr) != hipSuccess ||
      hipMemcpy(host.data(), rec.p, n * sizeof(BH_PixelRecord), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(std::string(who) + ": trace or copy failed");
  for (long long i = 0; i < n; ++i) {
    const BH_PixelRecord& r = host[i];
    float* q = out + 4 * i;
    if (r.hit_class == BH_HIT_DISK) {
      const int k = (r.n_cross < 3 ? r.n_cross : 3) - 1;
      const double rc = r.cross_r[k], g = r.cross_g[k];
      const double grav = std::sqrt(1.0 - 2.0 / rc);
      q[0] = static_cast<float>(grav);
      q[1] = static_cast<float>(g / grav);
      q[2] = 1.0f;
    } else if (r.hit_class == BH_HIT_HORIZON) {
      q[0] = 0.0f;
      q[1] = 1.0f;
      q[2] = 0.0f;
    } else {
      q[0] = static_cast<float>(r.r_exit > 2.0 ? std::sqrt(1.0 - 2.0 / r.r_exit) : 0.0);
      q[1] = 1.0f;
      q[2] = 0.0f;
    }
    q[3] = 1.0f;
  }
  return 0;
}

// The reference bridge's launch defaults (rs = 1, lengths in r_s).
void bridge_defaults(BH_LaunchParams* p, float spin, float fov, int w, int h) {
  std::memset(p, 0, sizeof(*p));
  p->rs = 1.0f;
  p->spin = spin;
  p->isco = static_cast<float>(0.5 * bh_isco_prograde(spin));
  p->step_size = 0.08f;
  p->fov_scale = fov > 0.05f ? fov : 0.05f;
  p->max_dist = 500.0f;
  p->max_steps = 600;
  p->width = w;
  p->height = h;
  p->adisk_enabled = 1;
  p->redshift_enabled = 1;
  p->kerr_enabled = spin > 1e-6f ? 1 : 0;
  p->doppler_strength = 1.0f;
  p->background_intensity = 0.8f;
  p->background_enabled = 1;
  p->photon_glow_strength = 1.0f;
  p->adisk_lit = 0.35f;
}

int render_to_host(const BH_LaunchParams& p, float* out, const char* who) {
  if (!out || p.width < 1 || p.height < 1) return fail(std::string(who) + ": invalid arguments");
  const size_t n = static_cast<size_t>(p.width) * p.height;
  DevBuf<float4> fb;
  if (!fb.alloc(n)) return fail(std::string(who) + ": device allocation failed");
  if (bh_launch_geodesic_kernel(fb.p, &p, -1, nullptr) != 0) return -1;
  if (hipStreamSynchronize(nullptr) != hipSuccess ||
      hipMemcpy(out, fb.p, n * sizeof(float4), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(std::string(who) + ": render or copy failed");
  return 0;
}

int render_orbit(float spin, float r, float inc, float fov, int w, int h, float* out, const char* who) {
  BH_LaunchParams p;
  bridge_defaults(&p, spin, fov, w, h);
  // camera at (0, r sin i, r cos i) looking at the origin; right = forward x (0, 1, 0), up = right x forward
  const float ci = std::cos(inc), si = std::sin(inc);
  p.cam_pos[0] = 0.0f;
  p.cam_pos[1] = r * si;
  p.cam_pos[2] = r * ci;
  const float fx = 0.0f, fy = -si, fz = -ci;
  float rx = -fz, ry = 0.0f, rz = fx;
  float rl = std::sqrt(rx * rx + ry * ry + rz * rz);
  if (rl < 1e-6f) {
    rx = 1.0f;
    ry = rz = 0.0f;
    rl = 1.0f;
  }
  rx /= rl;
  ry /= rl;
  rz /= rl;
  const float ux = ry * fz - rz * fy, uy = rz * fx - rx * fz, uz = rx * fy - ry * fx;
  const float basis[9] = {rx, ry, rz, ux, uy, uz, fx, fy, fz};
  std::memcpy(p.cam_basis, basis, sizeof(basis));
  return render_to_host(p, out, who);
}

}  // namespace
}  // namespace bhk

using namespace bhk;

extern "C" {

int bhbVersionMajor(void) { return 1; }
int bhbVersionMinor(void) { return 0; }
int bhbHasCuda(void) { return 1; }
int bhbHasBoost(void) { return 0; }
int bhbSizeofSourceParams(void) { return static_cast<int>(sizeof(struct BhbSourceParams)); }
int bhbSizeofDiskParams(void) { return static_cast<int>(sizeof(struct BhbDiskParams)); }
int bhbCudaResetDevice(void) {
  return hipDeviceSynchronize() == hipSuccess ? 0 : fail("bhbCudaResetDevice: device error");
}

int bhbTraceGeodesicsImagePlane(double aStar, double observerRRg, double inclinationRad, double alphaMin,
                                double alphaMax, int nAlpha, double betaMin, double betaMax, int nBeta, int maxSteps,
                                double /*stepSize*/, float* outXyz, int* outCounts) {
  return trace_polylines(aStar, observerRRg, inclinationRad, alphaMin, alphaMax, nAlpha, betaMin, betaMax, nBeta,
                         maxSteps, outXyz, outCounts, "bhbTraceGeodesicsImagePlane");
}

int bhbCudaTraceGeodesics(double aStar, double observerRRg, double inclinationRad, double alphaMin, double alphaMax,
                          int nAlpha, double betaMin, double betaMax, int nBeta, int maxSteps, double /*stepSize*/,
                          float* outXyz, int* outCounts) {
  return trace_polylines(aStar, observerRRg, inclinationRad, alphaMin, alphaMax, nAlpha, betaMin, betaMax, nBeta,
                         maxSteps, outXyz, outCounts, "bhbCudaTraceGeodesics");
}

int bhbCudaRenderLensingMap(double aStar, double observerRRg, double inclinationRad, int width, int height,
                            float* outRgba) {
  return lensing_map(aStar, observerRRg, inclinationRad, width, height, outRgba, "bhbCudaRenderLensingMap");
}

int bhbCudaRenderRaytraced(float spin, float observerR, float inclinationRad, int width, int height, float* outRgba) {
  return render_orbit(spin, observerR, inclinationRad, 1.0f, width, height, outRgba, "bhbCudaRenderRaytraced");
}

int bhbCudaRenderRaytracedCamera(float spin, float observerR, float inclinationRad, float fovScale, int width,
                                 int height, float* outRgba) {
  return render_orbit(spin, observerR, inclinationRad, fovScale, width, height, outRgba,
                      "bhbCudaRenderRaytracedCamera");
}

int bhbCudaRenderRaytracedView(float spin, const float* camPos, const float* camBasis, float fovScale, int width,
                               int height, float* outRgba) {
  if (!camPos || !camBasis) return fail("bhbCudaRenderRaytracedView: NULL camera");
  BH_LaunchParams p;
  bridge_defaults(&p, spin, fovScale, width, height);
  std::memcpy(p.cam_pos, camPos, 3 * sizeof(float));
  std::memcpy(p.cam_basis, camBasis, 9 * sizeof(float));
  return render_to_host(p, outRgba, "bhbCudaRenderRaytracedView");
}

int bhb_version_major(void) { return bhbVersionMajor(); }
int bhb_version_minor(void) { return bhbVersionMinor(); }
int bhb_has_cuda(void) { return bhbHasCuda(); }
int bhb_has_boost(void) { return bhbHasBoost(); }
int bhb_sizeof_source_params(void) { return bhbSizeofSourceParams(); }
int bhb_sizeof_disk_params(void) { return bhbSizeofDiskParams(); }
int bhb_cuda_reset_device(void) { return bhbCudaResetDevice(); }
int bhb_trace_geodesics_image_plane(double a, double r, double i, double amin, double amax, int na, double bmin,
                                    double bmax, int nb, int ms, double h, float* xyz, int* cnt) {
  return bhbTraceGeodesicsImagePlane(a, r, i, amin, amax, na, bmin, bmax, nb, ms, h, xyz, cnt);
}
int bhb_cuda_trace_geodesics(double a, double r, double i, double amin, double amax, int na, double bmin, double bmax,
                             int nb, int ms, double h, float* xyz, int* cnt) {
  return bhbCudaTraceGeodesics(a, r, i, amin, amax, na, bmin, bmax, nb, ms, h, xyz, cnt);
}
int bhb_cuda_render_lensing_map(double a, double r, double i, int w, int h, float* out) {
  return bhbCudaRenderLensingMap(a, r, i, w, h, out);
}
int bhb_cuda_render_raytraced(float s, float r, float i, int w, int h, float* out) {
  return bhbCudaRenderRaytraced(s, r, i, w, h, out);
}
int bhb_cuda_render_raytraced_camera(float s, float r, float i, float f, int w, int h, float* out) {
  return bhbCudaRenderRaytracedCamera(s, r, i, f, w, h, out);
}
int bhb_cuda_render_raytraced_view(float s, const float* p, const float* b, float f, int w, int h, float* out) {
  return bhbCudaRenderRaytracedView(s, p, b, f, w, h, out);
}

}  // extern "C"
