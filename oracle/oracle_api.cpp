// oracle_api.cpp - extern "C" API of the FP64 CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg, as the reference of their checks.  The product
// never loads it.
//
// The per-ray loop restates KerrRaytracer::trace (src/physics/raytracer.h:
// 324-378) -- capture / escape / step-budget termination -- with the
// c1 thresholds (r < horizon_factor r_+, r > r_escape on the way out: a
// camera beyond r_escape sees its rays travel in before any can escape; for a
// camera inside it the two forms agree, since an inbound ray reaches
// r_escape only after its radial turning point), the disk test of
// d_trace_geodesic (src/cuda/device_physics.cuh:912-1017) and the
// front-to-back volumetric transfer of d_trace_geodesic_rte
// (src/cuda/device_physics.cuh:2037-2173, d_rte_step :1988-2015), all in FP64.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "kerr_oracle.hpp"
#include "../include/bh_kernel_launch.h"

using namespace orc;

namespace {

struct Scene {
  const BH_TraceOptions* o;
  const double* luts;
  int lut_n;
  bool vol_on;
  VolumeHost V;
  bool pol_on = false;  // polarised transfer
  double pol_c2 = 1.0, pol_s2 = 0.0, pol_frac = 0.0, faraday = 0.0;
};

inline bool in_support(const BH_Volume& v, double r, double u) {
  return r >= v.sup_r_lo && r <= v.sup_r_hi && std::fabs(u) <= v.sup_u_max;
}

// Radiative state of one ray.
struct RayAcc {
  RtAcc<double> rt{0.0, 1.0, 0.0, 0.0, 0.0, 0.0};
  int ncross = 0;
  int nsamples = 0;
  double cross_r[3] = {0, 0, 0}, cross_g[3] = {0, 0, 0};
  PolAcc pol{0.0, 0.0, 1.0, 0.0, 0.0};
};

// Volume transfer over the accepted step of length h that ended at y.
inline void volume_segment(const Scene& S, const RayConst<double>& c, const double* y, double h,
                           RayAcc& acc) {
  double a1, e1, r1 = 0.0;
  if (volume_point(S.V, c, y[0], y[1], y[2], S.o->nu_obs, S.o->length_unit_cm, &a1, &e1, S.faraday,
                   S.pol_on ? &r1 : nullptr))
    acc.nsamples++;
  const double s1 = affine_rate(c, y[0], y[1]);
  if (S.pol_on) pol_segment(acc.pol, acc.rt, h, a1, e1, r1, s1, S.pol_c2, S.pol_s2, S.pol_frac);
  rt_segment(acc.rt, h, a1, e1, s1);
}

// Returns true when the disk terminated the ray.
inline bool crossing(const Scene& S, const CamConst& k, const RayConst<double>& c,
                     const double* y0, const double* f0, const double* y1, const double* f1,
                     double h, RayAcc& acc) {
  if ((y0[1] < 0.0) == (y1[1] < 0.0)) return false;
  const double rc = crossing_radius(y0, f0, y1, f1, h);
  const double g = (rc >= k.r_isco) ? disk_g(rc, c.a, c.xi) : 0.0;
  if (acc.ncross < S.o->max_crossings && acc.ncross < 3) {
    acc.cross_r[acc.ncross] = rc;
    acc.cross_g[acc.ncross] = g;
  }
  if (acc.ncross < 255) acc.ncross++;
  if (S.o->disk_enabled && rc >= k.r_isco && rc <= S.o->disk_r_out) {
    const double F = lut_linear(S.luts, S.lut_n, k.r_isco, S.o->disk_r_out, rc);
    acc.rt.I = std::fma(acc.rt.Tr, disk_intensity(F, g, S.o->nu_obs, S.o->disk_T_max), acc.rt.I);
    return true;
  }
  return false;
}

inline void finish(const RayConst<double>& c, const double* y, const RayAcc& acc, int status,
                   int steps, int attempts, BH_PixelRecord* rec) {
  std::memset(rec, 0, sizeof(*rec));
  rec->intensity = acc.rt.I;
  rec->r_exit = y[0];
  rec->u_exit = y[1];
  rec->phi_exit = y[2];
  for (int i = 0; i < 3; ++i) {
    rec->cross_r[i] = acc.cross_r[i];
    rec->cross_g[i] = acc.cross_g[i];
  }
  rec->tau = acc.rt.tau;
  const double R = radial_potential(c, y[0]);
  const double U = polar_potential(c, y[1]);
  const double rs = y[0] * y[0] + c.a2;
  rec->drift_null = std::fabs(std::fma(y[3], y[3], -R)) / (rs * rs);
  rec->drift_carter = std::fabs(std::fma(y[4], y[4], -U)) / (std::fabs(c.eta) + c.a2 + c.xi * c.xi);
  rec->steps = steps;
  rec->attempts = attempts;
  rec->hit_class = static_cast<uint8_t>(status);
  rec->n_cross = static_cast<uint8_t>(acc.ncross);
  rec->flags = 0;
  rec->confidence = 1.0f;
  rec->fp64_steps = steps;
  rec->n_samples = acc.nsamples;
}

// (I, Q, U, V) at the observer; V = 0 (no Faraday conversion in the
// simplified Mueller matrix).
inline void write_stokes(const RayAcc& acc, double* st) {
  st[0] = acc.rt.I;
  st[1] = acc.pol.q;
  st[2] = acc.pol.u;
  st[3] = 0.0;
}

inline int terminate_check(const Scene& S, const CamConst& k, const double* y, const RayAcc& acc) {
  if (y[0] < k.r_stop) return BH_HIT_HORIZON;
  if (y[0] > S.o->r_escape && y[3] > 0.0) return BH_HIT_ESCAPED;  // outbound
  if (S.o->transmit_min > 0.0 && acc.rt.Tr < S.o->transmit_min) return BH_HIT_OPAQUE;
  return 0;
}

void trace_rk4(const Scene& S, const CamConst& k, int64_t px, int64_t py, BH_PixelRecord* rec,
               double* st = nullptr) {
  double y[kN], k1[kN], yn[kN], fn[kN];
  const RayConst<double> c = setup_ray<double>(k, px, py, y);
  const double h = S.o->h_fixed;
  RayAcc acc;
  if (S.vol_on) acc.rt.s0 = affine_rate(c, y[0], y[1]);
  int steps = 0, status = 0;
  rhs(c, y, k1);
  for (;;) {
    if (steps >= S.o->max_steps) {
      status = BH_HIT_MAX_STEPS;
      break;
    }
    rk4_step(c, y, k1, h, yn);
    rhs(c, yn, fn);
    project_radial(c, yn, fn);
    ++steps;
    if (S.vol_on) volume_segment(S, c, yn, h, acc);
    const bool hit = crossing(S, k, c, y, k1, yn, fn, h, acc);
    std::memcpy(y, yn, sizeof(y));
    std::memcpy(k1, fn, sizeof(k1));
    if (hit) {
      status = BH_HIT_DISK;
      break;
    }
    status = terminate_check(S, k, y, acc);
    if (status) break;
  }
  finish(c, y, acc, status, steps, steps, rec);
  if (st) write_stokes(acc, st);
}

void trace_dp45(const Scene& S, const CamConst& k, int64_t px, int64_t py, BH_PixelRecord* rec,
                double* st = nullptr) {
  double y[kN], k1[kN], y5[kN], k7[kN];
  const RayConst<double> c = setup_ray<double>(k, px, py, y);
  const BH_TraceOptions& o = *S.o;
  RayAcc acc;
  if (S.vol_on) acc.rt.s0 = affine_rate(c, y[0], y[1]);
  int steps = 0, attempts = 0, status = 0;
  bool last_rej = false;
  double h = o.h_init;
  rhs(c, y, k1);
  for (;;) {
    if (steps >= o.max_steps || attempts >= o.max_attempts) {
      status = BH_HIT_MAX_STEPS;
      break;
    }
    double cap = o.h_max;
    if (S.vol_on && in_support(S.V.v, y[0], y[1]) && o.h_vol < cap) cap = o.h_vol;
    if (h > cap) h = cap;
    ++attempts;
    const double err = dp45_attempt(c, y, k1, h, o.rtol, o.atol, y5, k7);
    if (err <= 1.0) {
      const double fac = dp45_factor(err, true, last_rej);
      project_radial(c, y5, k7);
      ++steps;
      if (S.vol_on) volume_segment(S, c, y5, h, acc);
      const bool hit = crossing(S, k, c, y, k1, y5, k7, h, acc);
      std::memcpy(y, y5, sizeof(y));
      std::memcpy(k1, k7, sizeof(k1));
      h = h * fac;
      last_rej = false;
      if (hit) {
        status = BH_HIT_DISK;
        break;
      }
      status = terminate_check(S, k, y, acc);
      if (status) break;
    } else {
      h = h * dp45_factor(err, false, last_rej);
      last_rej = true;
    }
  }
  finish(c, y, acc, status, steps, attempts, rec);
  if (st) write_stokes(acc, st);
}

int64_t camera_of(const std::vector<CamConst>& ks, int64_t gid) {
  // cameras sorted by record offset; find the last with offset <= gid
  int64_t lo = 0, hi = static_cast<int64_t>(ks.size()) - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (ks[mid].offset <= gid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace

extern "C" {

// Trace the rays with global record indices ids[0..n_ids) (or all rays when
// ids is NULL) of the camera set; out[i] receives the record of ids[i].
// volume_cells is a host array; nthreads <= 0 uses all cores.
static int trace_common(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                        const double* luts, int lut_n, const BH_Volume* vol, const float* volume_cells,
                        const int64_t* ids, int64_t n_ids, BH_PixelRecord* out, int nthreads,
                        const double* pol, double* stokes) {
  if (!cams || n_cams <= 0 || !opts || !out) return -1;
  Scene S;
  if (pol) {
    S.pol_on = true;
    S.pol_c2 = std::cos(2.0 * pol[0]);
    S.pol_s2 = std::sin(2.0 * pol[0]);
    S.pol_frac = pol[1];
    S.faraday = pol[2];
  }
  S.o = opts;
  S.luts = luts;
  S.lut_n = lut_n;
  S.vol_on = opts->volume_enabled && vol && volume_cells;
  if (opts->volume_enabled && !S.vol_on) return -1;
  if (S.vol_on) {
    S.V.cells = volume_cells;
    S.V.v = *vol;
  }
  if (opts->disk_enabled && (!luts || lut_n < 2)) return -1;
  std::vector<CamConst> ks;
  int64_t total = 0;
  for (int i = 0; i < n_cams; ++i) {
    ks.push_back(make_cam_const(cams[i], *opts));
    total = std::max<int64_t>(total, cams[i].record_offset + int64_t(cams[i].width) * cams[i].height);
  }
  const int64_t n = ids ? n_ids : total;
  if (nthreads <= 0) nthreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    for (;;) {
      const int64_t i0 = next.fetch_add(64);
      if (i0 >= n) break;
      const int64_t i1 = std::min<int64_t>(n, i0 + 64);
      for (int64_t i = i0; i < i1; ++i) {
        const int64_t gid = ids ? ids[i] : i;
        const int64_t ci = camera_of(ks, gid);
        const CamConst& k = ks[ci];
        const int64_t local = gid - k.offset;
        double* st = stokes ? stokes + 4 * i : nullptr;
        if (local >= k.w * k.h) {  // gap between camera record ranges
          std::memset(&out[i], 0, sizeof(out[i]));
          if (st) st[0] = st[1] = st[2] = st[3] = 0.0;
          continue;
        }
        const int64_t py = local / k.w, px = local % k.w;
        const double* lut = luts ? luts + ci * lut_n : nullptr;
        Scene Sc = S;
        Sc.luts = lut;
        if (opts->integrator == BH_INTEGRATOR_DP45)
          trace_dp45(Sc, k, px, py, &out[i], st);
        else
          trace_rk4(Sc, k, px, py, &out[i], st);
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; ++t) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
  return 0;
}

int orc_trace(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
              const double* luts, int lut_n, const BH_Volume* vol, const float* volume_cells,
              const int64_t* ids, int64_t n_ids, BH_PixelRecord* out, int nthreads) {
  return trace_common(cams, n_cams, opts, luts, lut_n, vol, volume_cells, ids, n_ids, out, nthreads,
                      nullptr, nullptr);
}

// orc_trace with the polarised transfer: pol = {evpa_b [rad], lin_pol_frac,
// faraday_per_ne}; stokes[4 i] = (I, Q, U, V) of ray i at the observer.
// Needs the volume (polarisation is carried only through volume segments).
int orc_trace_pol(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                  const double* luts, int lut_n, const BH_Volume* vol, const float* volume_cells,
                  const double* pol, BH_PixelRecord* out, double* stokes, int nthreads) {
  if (!pol || !stokes || !opts || !opts->volume_enabled) return -1;
  return trace_common(cams, n_cams, opts, luts, lut_n, vol, volume_cells, nullptr, 0, out, nthreads, pol,
                      stokes);
}

// One pol_segment step from a given state, for closed-form checks:
// st_in = {q, u, c, s, r0, Tr, a0, e0, s0}, args = {h, a1, e1, r1, s1, c2, s2, frac};
// st_out = {q, u, c, s, r0}.
void orc_pol_segment(const double* st_in, const double* args, double* st_out) {
  PolAcc p{st_in[0], st_in[1], st_in[2], st_in[3], st_in[4]};
  RtAcc<double> q{};
  q.Tr = st_in[5];
  q.a0 = st_in[6];
  q.e0 = st_in[7];
  q.s0 = st_in[8];
  pol_segment(p, q, args[0], args[1], args[2], args[3], args[4], args[5], args[6], args[7]);
  st_out[0] = p.q;
  st_out[1] = p.u;
  st_out[2] = p.c;
  st_out[3] = p.s;
  st_out[4] = p.r0;
}

// Trace every pixel of a render launch (struct BH_LaunchParams) in FP64 as
// the render ABI maps it: lengths in M = rs/2, the pinhole camera at cam_pos
// (off the spin axis, r > 2), the disk from isco (or the prograde ISCO when
// isco <= 0) to 30 rs, escape beyond max_dist, DP45 rtol 1e-8 / atol 1e-10,
// h_init 0.01, h_max 0.1, max_steps accepted steps and 4 max_steps + 100
// attempts, stop at 1.01 r_+.  out: width x height records, row-major.
int orc_disk_lut(double a, double r_in, double r_out, int n, double* out);

int orc_trace_launch_params(const BH_LaunchParams* lp, BH_PixelRecord* out, int nthreads) {
  if (!lp || !out || lp->width <= 0 || lp->height <= 0 || !(lp->rs > 0.0f)) return -1;
  const BH_LaunchParams& p = *lp;
  const double M = 0.5 * static_cast<double>(p.rs);
  const double x = static_cast<double>(p.cam_pos[0]) / M;
  const double y = static_cast<double>(p.cam_pos[1]) / M;
  const double z = static_cast<double>(p.cam_pos[2]) / M;
  const double r = std::sqrt(x * x + y * y + z * z);
  const double rho = std::sqrt(x * x + y * y);
  if (!(rho > 0.0) || !(r > 2.0)) return -1;
  PinholeConst pin;
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
  BH_TraceOptions o;
  std::memset(&o, 0, sizeof(o));
  o.integrator = BH_INTEGRATOR_DP45;
  o.precision = BH_PRECISION_FP64;
  o.max_steps = p.max_steps;
  o.max_attempts = 4 * p.max_steps + 100;
  o.disk_enabled = p.adisk_enabled ? 1 : 0;
  o.max_crossings = 3;
  o.h_init = 0.01;
  o.h_max = 0.1;
  o.h_vol = 0.01;
  o.rtol = 1e-8;
  o.atol = 1e-10;
  o.horizon_factor = 1.01;
  o.r_escape = static_cast<double>(p.max_dist) / M;
  o.disk_r_out = 30.0 * static_cast<double>(p.rs) / M;
  o.nu_obs = 230e9;
  o.disk_T_max = 1e7;
  const double spin = static_cast<double>(p.spin);
  CamConst k;
  k.a = spin;
  k.sin_o = pin.sin_t;
  k.cos_o = pin.cos_t;
  k.r_obs = r;
  k.r_stop = o.horizon_factor * outer_horizon(spin);
  k.r_isco = p.isco > 0.0f ? static_cast<double>(p.isco) / M : isco_prograde(spin);
  k.alpha_min = 0.0;
  k.dalpha = 1.0;
  k.beta_max = 0.0;
  k.dbeta = -1.0;
  k.w = p.width;
  k.h = p.height;
  k.offset = 0;
  k.pin = &pin;
  std::vector<double> lut(256);
  if (orc_disk_lut(spin, k.r_isco, o.disk_r_out, 256, lut.data()) != 0) return -1;
  Scene S;
  S.o = &o;
  S.luts = lut.data();
  S.lut_n = 256;
  S.vol_on = false;
  const int64_t n = int64_t(p.width) * p.height;
  if (nthreads <= 0) nthreads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    for (;;) {
      const int64_t i0 = next.fetch_add(64);
      if (i0 >= n) break;
      const int64_t i1 = std::min<int64_t>(n, i0 + 64);
      for (int64_t i = i0; i < i1; ++i) trace_dp45(S, k, i % k.w, i / k.w, &out[i]);
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nthreads; ++t) th.emplace_back(worker);
  worker();
  for (auto& t : th) t.join();
  return 0;
}

// Setup of one pinhole pixel of a render launch: y[5] and {xi, eta}.
int orc_pinhole_setup(const BH_LaunchParams* lp, int64_t px, int64_t py, double* y, double* xi_eta) {
  if (!lp) return -1;
  const BH_LaunchParams& p = *lp;
  const double M = 0.5 * static_cast<double>(p.rs);
  const double x = static_cast<double>(p.cam_pos[0]) / M;
  const double yy = static_cast<double>(p.cam_pos[1]) / M;
  const double z = static_cast<double>(p.cam_pos[2]) / M;
  const double r = std::sqrt(x * x + yy * yy + z * z);
  const double rho = std::sqrt(x * x + yy * yy);
  if (!(rho > 0.0) || !(r > 2.0)) return -1;
  PinholeConst pin;
  pin.r = r;
  pin.cos_t = z / r;
  pin.sin_t = rho / r;
  pin.phi = std::atan2(yy, x);
  const double cp = x / rho, sp = yy / rho;
  pin.e_r[0] = x / r;
  pin.e_r[1] = yy / r;
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
  const RayConst<double> c =
      setup_pinhole(pin, static_cast<double>(p.spin), static_cast<double>(px) + 0.5, static_cast<double>(py) + 0.5, y);
  xi_eta[0] = c.xi;
  xi_eta[1] = c.eta;
  return 0;
}

// --- pieces exposed for unit tests -------------------------------------------

void orc_rhs(double a, double xi, double eta, const double* y, double* f) {
  const RayConst<double> c = make_const(a, xi, eta);
  rhs(c, y, f);
}
void orc_rk4_step(double a, double xi, double eta, const double* y, double h, double* yn) {
  const RayConst<double> c = make_const(a, xi, eta);
  double k1[kN];
  rhs(c, y, k1);
  rk4_step(c, y, k1, h, yn);
}
double orc_dp45_attempt(double a, double xi, double eta, const double* y, double h, double rtol,
                        double atol, double* y5) {
  const RayConst<double> c = make_const(a, xi, eta);
  double k1[kN], k7[kN];
  rhs(c, y, k1);
  return dp45_attempt(c, y, k1, h, rtol, atol, y5, k7);
}
double orc_pow_m0p2(double x) { return pow_m0p2(x); }
double orc_dp45_factor(double err, int accepted, int last_rej) {
  return dp45_factor(err, accepted != 0, last_rej != 0);
}
// Setup of pixel (px, py) of cam: writes y[5] = (r, u, phi, v_r, v_u) and {xi, eta}.
void orc_setup(const BH_Camera* cam, const BH_TraceOptions* o, int64_t px, int64_t py, double* y,
               double* xi_eta) {
  const CamConst k = make_cam_const(*cam, *o);
  const RayConst<double> c = setup_ray<double>(k, px, py, y);
  xi_eta[0] = c.xi;
  xi_eta[1] = c.eta;
}
// R(r), R'(r), U(u), U'(u) and the theta-form Theta(theta), Theta'(theta)
// for constants (a, xi, eta).
void orc_potentials(double a, double xi, double eta, double r, double theta, double* out6) {
  const RayConst<double> c = make_const(a, xi, eta);
  const double u = std::cos(theta);
  out6[0] = radial_potential(c, r);
  const double P = r * r + c.c1;
  out6[1] = 4.0 * r * P - (2.0 * r - 2.0) * c.K;
  out6[2] = polar_potential(c, u);
  out6[3] = 2.0 * u * (c.c2 - c.two_a2 * u * u);
  const double s = std::sin(theta);
  out6[4] = eta + a * a * u * u - xi * xi * u * u / (s * s);
  out6[5] = -2.0 * a * a * u * s + 2.0 * xi * xi * u / (s * s * s);
}
double orc_outer_horizon(double a) { return outer_horizon(a); }
double orc_isco(double a) { return isco_prograde(a); }
double orc_photon_orbit(double a, int prograde) {
  return prograde ? photon_orbit_prograde(a) : photon_orbit_retrograde(a);
}
void orc_critical_params(double r, double a, double* xi_eta) {
  critical_params(r, a, &xi_eta[0], &xi_eta[1]);
}
double orc_disk_g(double r, double a, double xi) { return disk_g(r, a, xi); }
double orc_fluid_g(double r, double u, double a, double xi, double ell) {
  return fluid_g(r, u, a, xi, ell);
}
double orc_planck(double nu, double T) { return planck(nu, T); }
double orc_synch_j(double nu, double B, double ne, double th, double k2ratio) {
  return synch_j(nu, B, ne, th, k2ratio);
}
double orc_disk_intensity(double F, double g, double nu, double T_max) {
  return disk_intensity(F, g, nu, T_max);
}
double orc_crossing_radius(const double* y0, const double* f0, const double* y1, const double* f1,
                           double h) {
  return crossing_radius(y0, f0, y1, f1, h);
}
double orc_lut_linear(const double* tab, int n, double x0, double x1, double x) {
  return lut_linear(tab, n, x0, x1, x);
}
void orc_volume_fetch(const BH_Volume* v, const float* cells, double r, double u, double phi,
                      double* out4) {
  VolumeHost V;
  V.cells = cells;
  V.v = *v;
  volume_fetch(V, r, u, phi, out4);
}
// Observer-frame absorption / emission per affine length at a point
// (volume_point); returns 1 if a sample was taken.
int orc_volume_point(const BH_Volume* v, const float* cells, double a, double xi, double eta,
                     double r, double u, double phi, double nu_obs, double L, double* ae) {
  VolumeHost V;
  V.cells = cells;
  V.v = *v;
  const RayConst<double> c = make_const(a, xi, eta);
  return volume_point(V, c, r, u, phi, nu_obs, L, &ae[0], &ae[1]) ? 1 : 0;
}
// One trapezoidal transfer segment (rt_segment): st = {I, T, tau, a0, e0, s0}.
void orc_rt_segment(double* st, double h, double a1, double e1, double s1) {
  RtAcc<double> q{st[0], st[1], st[2], st[3], st[4], st[5]};
  rt_segment(q, h, a1, e1, s1);
  st[0] = q.I;
  st[1] = q.Tr;
  st[2] = q.tau;
  st[3] = q.a0;
  st[4] = q.e0;
  st[5] = q.s0;
}

// Normalised Novikov-Thorne flux table (the reference's NT approximation,
// src/physics/thin_disk.h:243-266 / scripts/generate_luts.py:42-48, M = 1):
// F(r) = (1 - sqrt(r_in/r)) (1 + 0.5 a sqrt(1/r)) / r^3, sampled at
// r_i = r_in + (i/(n-1)) (r_out - r_in), divided by the largest sample
// (src/physics/lut.h:50-86).
int orc_disk_lut(double a, double r_in, double r_out, int n, double* out) {
  if (n < 2 || !(r_out > r_in)) return -1;
  double mx = 0.0;
  for (int i = 0; i < n; ++i) {
    const double u = static_cast<double>(i) / static_cast<double>(n - 1);
    const double r = r_in + u * (r_out - r_in);
    const double basic = 1.0 - std::sqrt(r_in / r);
    const double spin = 1.0 + 0.5 * a * std::sqrt(1.0 / r);
    out[i] = (r < r_in) ? 0.0 : basic * spin / (r * r * r);
    mx = std::max(mx, out[i]);
  }
  if (mx > 0.0)
    for (int i = 0; i < n; ++i) out[i] /= mx;
  return 0;
}

// K_nu(z) = int_0^inf exp(-z cosh t) cosh(nu t) dt (Watson 1944, 6.22(5)),
// trapezoid rule (exponentially convergent for this analytic integrand).
double orc_bessel_k(double nu, double z) {
  const double dt = 0.01;
  double s = 0.5 * std::exp(-z);
  for (int i = 1; i < 20000; ++i) {
    const double t = i * dt;
    const double ch = std::cosh(t);
    if (z * ch > 745.0) break;
    s += std::exp(-z * ch) * std::cosh(nu * t);
  }
  return s * dt;
}

// ln(K_2(1/Theta) / (2 Theta^2)) on n points of ln(Theta) in [lo, hi].
int orc_k2_table(int n, double lo, double hi, double* out) {
  if (n < 2 || !(hi > lo)) return -1;
  for (int i = 0; i < n; ++i) {
    const double lt = lo + (hi - lo) * static_cast<double>(i) / static_cast<double>(n - 1);
    const double th = std::exp(lt);
    out[i] = std::log(orc_bessel_k(2.0, 1.0 / th) / (2.0 * th * th));
  }
  return 0;
}

// Reference formal-solution step rteStep (src/physics/rte_integrator.h:138-175)
// restated, for the RTE known-answer tests.
void orc_rte_step(double* iNu_tau_s, double jNu, double alphaNu, double ds) {
  double iNu = iNu_tau_s[0], tau_acc = iNu_tau_s[1], s = iNu_tau_s[2];
  if (ds <= 0.0) return;
  const double tau = (alphaNu > 0.0) ? alphaNu * ds : 0.0;
  if (alphaNu > 0.0) {
    const double om = -std::expm1(-tau);
    iNu = iNu * (1.0 - om) + (jNu / alphaNu) * om;
  } else {
    iNu += jNu * ds;
  }
  iNu_tau_s[0] = iNu;
  iNu_tau_s[1] = tau_acc + tau;
  iNu_tau_s[2] = s + ds;
}

}  // extern "C"
