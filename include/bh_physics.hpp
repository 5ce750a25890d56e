// bh_physics.hpp - the reference's physics C++ interface for the geodesic
// path (src/physics/kerr.h:91-546, src/physics/raytracer.h:46-79,309-396,
// src/physics/verified/rk4.hpp:49-437), header-only C++17 over the library.
//
// Same names, signatures and CGS units as the reference (mass [g], lengths
// [cm], spin parameter a [cm]); value semantics, noexcept, no allocation in
// the step functions.  The closed forms, the potentials and the single Mino
// step are host utilities; whole-ray tracing (KerrRaytracer) runs on the GPU
// through bh_trace_rays_host (libbhk.so must be linked; it fails loudly
// without a device).
//
// Differences from the reference, by design:
//  * kerrStepMino advances the second-order form r'' = R'/2, theta'' =
//    Theta'/2 (classic RK4 in Mino time) with the velocities' signs carried
//    by the state, instead of the square-root form, which stalls at turning
//    points (kerr.cpp:99-134);
//  * KerrRaytracer traces with the library's adaptive FP64 Dormand-Prince
//    5(4) (rtol 1e-8, atol 1e-10); setStepSize is accepted and not used, and
//    totalDistance stays 0 (the Mino parameter is not accumulated);
//    it traces rays that start inbound (signR < 0); an outbound start is
//    returned untraced (status PROPAGATING, stepsTaken 0); path recording
//    (setRecordPath) is not supported here (bhbCudaTraceGeodesics records
//    polylines).
#ifndef BH_PHYSICS_HPP
#define BH_PHYSICS_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <limits>
#include <utility>
#include <vector>

#include "bhk_trace.h"

namespace physics {

inline constexpr double C = 2.99792458e10;   // cm s^-1
inline constexpr double C2 = C * C;
inline constexpr double G = 6.67430e-8;      // cm^3 g^-1 s^-2
inline constexpr double PI = 3.14159265358979323846;

[[nodiscard]] inline double schwarzschildRadius(double mass) noexcept { return 2.0 * G * mass / C2; }

// kerr.h:98-117
[[nodiscard]] inline double kerrSigma(double r, double a, double theta) noexcept {
  const double c = std::cos(theta);
  return r * r + a * a * c * c;
}
[[nodiscard]] inline double kerrDelta(double r, double a, double rS) noexcept { return r * r - rS * r + a * a; }

// kerr.h:132-186 (NaN beyond extremality)
[[nodiscard]] inline double kerrOuterHorizon(double mass, double a) noexcept {
  const double m = G * mass / C2, d = m * m - a * a;
  return d < 0.0 ? std::numeric_limits<double>::quiet_NaN() : m + std::sqrt(d);
}
[[nodiscard]] inline double kerrInnerHorizon(double mass, double a) noexcept {
  const double m = G * mass / C2, d = m * m - a * a;
  return d < 0.0 ? std::numeric_limits<double>::quiet_NaN() : m - std::sqrt(d);
}
[[nodiscard]] inline double ergosphereRadius(double mass, double a, double theta) noexcept {
  const double m = G * mass / C2, c = std::cos(theta), d = m * m - a * a * c * c;
  return d < 0.0 ? std::numeric_limits<double>::quiet_NaN() : m + std::sqrt(d);
}

// kerr.h:218-296: Bardeen-Press-Teukolsky ISCO and the circular photon orbits
[[nodiscard]] inline double kerrIscoRadius(double mass, double a, bool prograde = true) noexcept {
  const double m = G * mass / C2, s = a / m;
  if (std::fabs(s) > 1.0) return std::numeric_limits<double>::quiet_NaN();
  const double z1 = 1.0 + std::cbrt(1.0 - s * s) * (std::cbrt(1.0 + s) + std::cbrt(1.0 - s));
  const double z2 = std::sqrt(3.0 * s * s + z1 * z1);
  const double k = std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
  return m * (3.0 + z2 + (prograde ? -k : k));
}
[[nodiscard]] inline double kerrPhotonOrbitPrograde(double mass, double a) noexcept {
  const double m = G * mass / C2, s = a / m;
  if (std::fabs(s) > 1.0) return std::numeric_limits<double>::quiet_NaN();
  return 2.0 * m * (1.0 + std::cos((2.0 / 3.0) * std::acos(-s)));
}
[[nodiscard]] inline double kerrPhotonOrbitRetrograde(double mass, double a) noexcept {
  const double m = G * mass / C2, s = a / m;
  if (std::fabs(s) > 1.0) return std::numeric_limits<double>::quiet_NaN();
  return 2.0 * m * (1.0 + std::cos((2.0 / 3.0) * std::acos(s)));
}
[[nodiscard]] inline double kerrPhotonOrbit(double mass, double a, bool prograde = true) noexcept {
  return prograde ? kerrPhotonOrbitPrograde(mass, a) : kerrPhotonOrbitRetrograde(mass, a);
}

// kerr.h:316-330: omega = 2 M a r c / (Sigma (r^2 + a^2) + 2 M a^2 r sin^2)
[[nodiscard]] inline double frameDraggingOmega(double r, double theta, double mass, double a) noexcept {
  const double m = G * mass / C2, s = std::sin(theta);
  const double den = kerrSigma(r, a, theta) * (r * r + a * a) + 2.0 * m * a * a * r * s * s;
  return den < 1e-30 ? 0.0 : 2.0 * m * a * r * C / den;
}

// kerr.h:380-390: static-observer redshift z = 1/sqrt(1 - r_s r / Sigma) - 1
[[nodiscard]] inline double kerrRedshift(double r, double theta, double mass, double a) noexcept {
  const double f = 1.0 - schwarzschildRadius(mass) * r / kerrSigma(r, a, theta);
  return f <= 0.0 ? std::numeric_limits<double>::infinity() : 1.0 / std::sqrt(f) - 1.0;
}

// kerr.h:396-444
struct KerrGeodesicConsts {
  double e;   // energy
  double lz;  // angular momentum about the spin axis
  double q;   // Carter constant Q (R = P^2 - Delta (q + (lz - a e)^2))
};
[[nodiscard]] inline KerrGeodesicConsts kerrEquatorialConsts(double impactParam, double energy = 1.0) noexcept {
  return KerrGeodesicConsts{energy, impactParam * energy, 0.0};
}
struct KerrGeodesicState {
  double r, theta, phi, t;
  double signR, signTheta;  // +1 or -1
};
[[nodiscard]] inline KerrGeodesicState kerrEquatorialState(double r, double phi, double signR) noexcept {
  return KerrGeodesicState{r, 0.5 * PI, phi, 0.0, signR >= 0.0 ? 1.0 : -1.0, 1.0};
}
struct KerrPotentials {
  double rPot, dRdr, thetaPot, dThetadtheta;
};

// kerr.cpp:69-97: R = ((r^2 + a^2) e - a lz)^2 - Delta (q + (lz - a e)^2),
// thetaPot = q + a^2 e^2 cos^2 - lz^2 / sin^2 (sin^2 floored at 1e-12), and
// their derivatives.  thetaPot is the reference's expression: it is the
// polar potential (Sigma dtheta/dtau)^2 = q + cos^2 (a^2 e^2 - lz^2 / sin^2)
// minus lz^2 (same derivative).
[[nodiscard]] inline KerrPotentials kerrPotentials(double r, double theta, double mass, double a,
                                                   const KerrGeodesicConsts& c) noexcept {
  const double m = G * mass / C2, rr = r * r, aa = a * a;
  const double st = std::sin(theta), ct = std::cos(theta);
  const double s2 = std::max(st * st, 1e-12), c2 = ct * ct;
  const double delta = rr - 2.0 * m * r + aa;
  const double A = (rr + aa) * c.e - a * c.lz;
  const double k = c.q + (c.lz - a * c.e) * (c.lz - a * c.e);
  KerrPotentials p{};
  p.rPot = A * A - delta * k;
  p.dRdr = 2.0 * A * (2.0 * r * c.e) - (2.0 * r - 2.0 * m) * k;
  p.thetaPot = c.q + aa * c.e * c.e * c2 - c.lz * c.lz / s2;
  p.dThetadtheta = -2.0 * aa * c.e * c.e * ct * st + 2.0 * c.lz * c.lz * ct / (s2 * st);
  return p;
}

namespace detail {

// Mino-time rates of (r, theta, phi, t, v_r, v_theta) with v = d/dlambda.
inline void mino_rates(const double* s, double m, double a, const KerrGeodesicConsts& c, double mass,
                       double* f) noexcept {
  const double r = s[0], th = s[1];
  const KerrPotentials p = kerrPotentials(r, th, mass, a, c);
  const double st = std::sin(th), s2 = std::max(st * st, 1e-12);
  const double delta = r * r - 2.0 * m * r + a * a;
  const double ds = std::fabs(delta) < 1e-30 ? (delta < 0.0 ? -1e-30 : 1e-30) : delta;
  const double A = (r * r + a * a) * c.e - a * c.lz;
  f[0] = s[4];
  f[1] = s[5];
  f[2] = c.lz / s2 - a * c.e + a * A / ds;
  f[3] = (r * r + a * a) * A / ds + a * (c.lz - a * c.e * s2);
  f[4] = 0.5 * p.dRdr;
  f[5] = 0.5 * p.dThetadtheta;
}

}  // namespace detail

// One Mino-time step of length dlam (second-order form, classic RK4); the
// signs of the returned state are those of the final velocities.
[[nodiscard]] inline KerrGeodesicState kerrStepMino(const KerrGeodesicState& state, double mass, double a,
                                                    const KerrGeodesicConsts& c, double dlam) noexcept {
  const double m = G * mass / C2;
  const KerrPotentials p0 = kerrPotentials(state.r, state.theta, mass, a, c);
  const double theta_pot = p0.thetaPot + c.lz * c.lz;  // the polar potential proper
  double y[6] = {state.r, state.theta, state.phi, state.t, state.signR * std::sqrt(std::max(p0.rPot, 0.0)),
                 state.signTheta * std::sqrt(std::max(theta_pot, 0.0))};
  double k1[6], k2[6], k3[6], k4[6], t[6];
  detail::mino_rates(y, m, a, c, mass, k1);
  for (int i = 0; i < 6; ++i) t[i] = y[i] + 0.5 * dlam * k1[i];
  detail::mino_rates(t, m, a, c, mass, k2);
  for (int i = 0; i < 6; ++i) t[i] = y[i] + 0.5 * dlam * k2[i];
  detail::mino_rates(t, m, a, c, mass, k3);
  for (int i = 0; i < 6; ++i) t[i] = y[i] + dlam * k3[i];
  detail::mino_rates(t, m, a, c, mass, k4);
  for (int i = 0; i < 6; ++i) y[i] += dlam / 6.0 * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
  KerrGeodesicState out{y[0], y[1], y[2], y[3], y[4] >= 0.0 ? 1.0 : -1.0, y[5] >= 0.0 ? 1.0 : -1.0};
  return out;
}

// kerr.h:461-546
class Kerr {
 public:
  explicit Kerr(double mass, double spinParam = 0.0) noexcept
      : mass_(mass), a_(spinParam), mGeom_(G * mass / C2), rS_(schwarzschildRadius(mass)),
        aStar_(spinParam / (G * mass / C2)) {
    if (std::fabs(aStar_) > 1.0) {  // the Thorne limit, as the reference clamps
      aStar_ = aStar_ > 0.0 ? 0.998 : -0.998;
      a_ = aStar_ * mGeom_;
    }
    rPlus_ = kerrOuterHorizon(mass, a_);
    rMinus_ = kerrInnerHorizon(mass, a_);
    rIscoPro_ = kerrIscoRadius(mass, a_, true);
    rIscoRet_ = kerrIscoRadius(mass, a_, false);
    rPhPro_ = kerrPhotonOrbitPrograde(mass, a_);
    rPhRet_ = kerrPhotonOrbitRetrograde(mass, a_);
  }
  [[nodiscard]] static Kerr fromDimensionlessSpin(double mass, double aStar) noexcept {
    return Kerr(mass, aStar * (G * mass / C2));
  }
  [[nodiscard]] double mass() const noexcept { return mass_; }
  [[nodiscard]] double spin() const noexcept { return a_; }
  [[nodiscard]] double dimensionlessSpin() const noexcept { return aStar_; }
  [[nodiscard]] double outerHorizon() const noexcept { return rPlus_; }
  [[nodiscard]] double innerHorizon() const noexcept { return rMinus_; }
  [[nodiscard]] double iscoPrograde() const noexcept { return rIscoPro_; }
  [[nodiscard]] double iscoRetrograde() const noexcept { return rIscoRet_; }
  [[nodiscard]] double photonOrbitPrograde() const noexcept { return rPhPro_; }
  [[nodiscard]] double photonOrbitRetrograde() const noexcept { return rPhRet_; }
  [[nodiscard]] double sigma(double r, double theta) const noexcept { return kerrSigma(r, a_, theta); }
  [[nodiscard]] double delta(double r) const noexcept { return kerrDelta(r, a_, rS_); }
  [[nodiscard]] double ergosphere(double theta) const noexcept { return ergosphereRadius(mass_, a_, theta); }
  [[nodiscard]] double frameDragging(double r, double theta) const noexcept {
    return frameDraggingOmega(r, theta, mass_, a_);
  }
  [[nodiscard]] double redshift(double r, double theta) const noexcept { return kerrRedshift(r, theta, mass_, a_); }
  [[nodiscard]] KerrPotentials potentials(double r, double theta, const KerrGeodesicConsts& c) const noexcept {
    return kerrPotentials(r, theta, mass_, a_, c);
  }
  [[nodiscard]] KerrGeodesicState stepMino(const KerrGeodesicState& s, const KerrGeodesicConsts& c,
                                           double dlam) const noexcept {
    return kerrStepMino(s, mass_, a_, c, dlam);
  }
  [[nodiscard]] bool isExterior(double r) const noexcept { return r > rPlus_; }
  [[nodiscard]] bool inErgosphere(double r, double theta) const noexcept {
    return r < ergosphere(theta) && r > rPlus_;
  }

 private:
  double mass_, a_, mGeom_, rS_, aStar_;
  double rPlus_ = 0.0, rMinus_ = 0.0, rIscoPro_ = 0.0, rIscoRet_ = 0.0, rPhPro_ = 0.0, rPhRet_ = 0.0;
};

// raytracer.h:46-79
enum class RayStatus { PROPAGATING, ESCAPED, CAPTURED, HitDisk, MaxSteps };

struct RayTraceResult {
  std::array<double, 3> finalPosition{};  // (r [cm], theta, phi)
  double totalDistance = 0.0;
  double redshift = 0.0;
  int stepsTaken = 0;
  RayStatus status = RayStatus::PROPAGATING;
  std::vector<std::array<double, 3>> path;
};

// The explicit ray the GPU tracer starts from for (state, c) around a black
// hole of geometric mass m [cm] and spin a [cm]: Bardeen's image-plane
// point of the constants at the start, xi = lz / (e m), eta = q / (e m)^2,
// alpha = -xi / sin(theta),
// beta = -signTheta sqrt(Theta(theta)) with
// Theta = eta + a*^2 cos^2 - xi^2 cot^2; false for an outbound start.
[[nodiscard]] inline bool raySpecFor(const KerrGeodesicState& s, const KerrGeodesicConsts& c, double m, double a,
                                     BH_RaySpec* out) noexcept {
  if (s.signR >= 0.0 || !(s.theta > 0.0 && s.theta < PI) || !(c.e > 0.0)) return false;
  const double xi = c.lz / (c.e * m);
  const double eta = c.q / (c.e * c.e * m * m);
  const double as = a / m, st = std::sin(s.theta), ct = std::cos(s.theta);
  const double th = eta + as * as * ct * ct - xi * xi * ct * ct / (st * st);
  out->alpha = -xi / st;
  out->beta = -(s.signTheta >= 0.0 ? 1.0 : -1.0) * std::sqrt(std::max(th, 0.0));
  out->spin = as;
  out->theta_obs = s.theta;
  out->r_obs = s.r / m;
  return true;
}

// raytracer.h:309-396, on the GPU (one launch per batch of rays).
class KerrRaytracer {
 public:
  KerrRaytracer(double mass, double spinParam) noexcept
      : kerr_(mass, spinParam), rS_(schwarzschildRadius(mass)), rHorizon_(kerr_.outerHorizon()),
        rEscape_(1000.0 * schwarzschildRadius(mass)), stepSize_(0.02 * schwarzschildRadius(mass)) {}

  void setStepSize(double h) noexcept { stepSize_ = h; }
  void setMaxSteps(int n) noexcept { maxSteps_ = n; }
  void setEscapeRadius(double r) noexcept { rEscape_ = r; }
  void setRecordPath(bool record) noexcept { recordPath_ = record; }

  [[nodiscard]] std::vector<RayTraceResult> traceBatch(
      const std::vector<std::pair<KerrGeodesicState, KerrGeodesicConsts>>& rays) const {
    const double m = G * kerr_.mass() / C2;
    std::vector<RayTraceResult> res(rays.size());
    std::vector<BH_RaySpec> specs;
    std::vector<size_t> which;
    for (size_t i = 0; i < rays.size(); ++i) {
      BH_RaySpec sp;
      if (raySpecFor(rays[i].first, rays[i].second, m, kerr_.spin(), &sp)) {
        specs.push_back(sp);
        which.push_back(i);
      } else {
        res[i].finalPosition = {rays[i].first.r, rays[i].first.theta, rays[i].first.phi};
      }
    }
    if (specs.empty()) return res;
    BH_TraceOptions o;
    bh_trace_options_default(&o);
    o.integrator = BH_INTEGRATOR_DP45;
    o.disk_enabled = 0;
    o.max_steps = maxSteps_;
    o.max_attempts = 10 * maxSteps_ + 1000;
    o.horizon_factor = 1.001;
    o.r_escape = rEscape_ / m;
    std::vector<BH_PixelRecord> rec(specs.size());
    if (bh_trace_rays_host(specs.data(), static_cast<int64_t>(specs.size()), &o, rec.data()) != 0) {
      for (size_t k = 0; k < which.size(); ++k) res[which[k]].stepsTaken = -1;  // device error (bh_last_error)
      return res;
    }
    for (size_t k = 0; k < which.size(); ++k) {
      const BH_PixelRecord& r = rec[k];
      RayTraceResult& out = res[which[k]];
      const double theta = std::acos(std::max(-1.0, std::min(1.0, r.u_exit)));
      out.finalPosition = {r.r_exit * m, theta, rays[which[k]].first.phi + r.phi_exit};
      out.stepsTaken = r.steps;
      out.status = r.hit_class == BH_HIT_HORIZON   ? RayStatus::CAPTURED
                   : r.hit_class == BH_HIT_ESCAPED ? RayStatus::ESCAPED
                   : r.hit_class == BH_HIT_DISK    ? RayStatus::HitDisk
                                                   : RayStatus::MaxSteps;
      out.redshift = out.status == RayStatus::CAPTURED ? 1.0 : 1.0 + kerr_.redshift(out.finalPosition[0], theta);
    }
    return res;
  }

  [[nodiscard]] RayTraceResult trace(const KerrGeodesicState& initial, const KerrGeodesicConsts& c) const {
    return traceBatch({{initial, c}})[0];
  }

  [[nodiscard]] RayTraceResult traceEquatorial(double r, double phi, double impactParam, double signR,
                                               double energy = 1.0) const {
    return trace(kerrEquatorialState(r, phi, signR), kerrEquatorialConsts(impactParam, energy));
  }

 private:
  Kerr kerr_;
  double rS_, rHorizon_, rEscape_, stepSize_;
  int maxSteps_{10000};
  bool recordPath_{false};
};

}  // namespace physics

// src/physics/verified/rk4.hpp:49-491: the generic integrators of the
// verified layer (8-component Boyer-Lindquist state, user right-hand side).
namespace verified {

struct StateVector {
  double x0, x1, x2, x3;  // t, r, theta, phi
  double v0, v1, v2, v3;  // their d/dlambda
  constexpr StateVector() noexcept : x0(0), x1(0), x2(0), x3(0), v0(0), v1(0), v2(0), v3(0) {}
  constexpr StateVector(double t, double r, double th, double ph, double dt, double dr, double dth,
                        double dph) noexcept
      : x0(t), x1(r), x2(th), x3(ph), v0(dt), v1(dr), v2(dth), v3(dph) {}
  [[nodiscard]] constexpr double r() const noexcept { return x1; }
  [[nodiscard]] constexpr double theta() const noexcept { return x2; }
};

[[nodiscard]] constexpr StateVector sv_add(const StateVector& a, const StateVector& b) noexcept {
  return StateVector(a.x0 + b.x0, a.x1 + b.x1, a.x2 + b.x2, a.x3 + b.x3, a.v0 + b.v0, a.v1 + b.v1, a.v2 + b.v2,
                     a.v3 + b.v3);
}
[[nodiscard]] constexpr StateVector sv_scale(double k, const StateVector& a) noexcept {
  return StateVector(k * a.x0, k * a.x1, k * a.x2, k * a.x3, k * a.v0, k * a.v1, k * a.v2, k * a.v3);
}

// Classic RK4 step of y' = f(y).
template <typename F>
[[nodiscard]] constexpr StateVector rk4_step(F&& f, double h, const StateVector& y) noexcept {
  const StateVector k1 = f(y);
  const StateVector k2 = f(sv_add(y, sv_scale(h / 2.0, k1)));
  const StateVector k3 = f(sv_add(y, sv_scale(h / 2.0, k2)));
  const StateVector k4 = f(sv_add(y, sv_scale(h, k3)));
  return sv_add(y, sv_scale(h / 6.0, sv_add(k1, sv_add(sv_scale(2.0, k2), sv_add(sv_scale(2.0, k3), k4)))));
}

enum class GeodesicStatus { Propagating, Captured, Escaped, MaxSteps };

[[nodiscard]] constexpr GeodesicStatus check_termination(const StateVector& s, double r_horizon, double r_escape,
                                                         std::size_t step, std::size_t max_steps) noexcept {
  if (s.x1 < r_horizon) return GeodesicStatus::Captured;
  if (s.x1 > r_escape) return GeodesicStatus::Escaped;
  if (step >= max_steps) return GeodesicStatus::MaxSteps;
  return GeodesicStatus::Propagating;
}

template <typename F>
[[nodiscard]] constexpr std::pair<StateVector, GeodesicStatus> integrate_with_termination(
    F&& f, double h, StateVector s, double r_horizon, double r_escape, std::size_t max_steps) noexcept {
  for (std::size_t step = 0; step < max_steps; ++step) {
    const GeodesicStatus st = check_termination(s, r_horizon, r_escape, step, max_steps);
    if (st != GeodesicStatus::Propagating) return {s, st};
    s = rk4_step(f, h, s);
  }
  return {s, GeodesicStatus::MaxSteps};
}

// h (tol / err)^(1/5) with safety 0.9 (rk4.hpp:485-491).
[[nodiscard]] inline double optimal_step(double h, double err, double tol) noexcept {
  return err <= 0.0 ? h : 0.9 * h * std::pow(tol / err, 0.2);
}

// Dormand & Prince (1980) 5(4) step (the tableau of src/physics/verified/rk4.h:
// 193-244) with the error of Hairer I, II.4.11: RMS over the components of
// (y5 - y4) / (atol + rtol max(|y|, |y5|)).  Returns {y5, err}; accept when
// err <= 1 and continue with dp45_next_step.
template <typename F>
[[nodiscard]] std::pair<StateVector, double> dp45_step(F&& f, double h, const StateVector& y, double rtol,
                                                       double atol) noexcept {
  const StateVector k1 = f(y);
  const StateVector k2 = f(sv_add(y, sv_scale(h * (1.0 / 5.0), k1)));
  const StateVector k3 = f(sv_add(y, sv_add(sv_scale(h * (3.0 / 40.0), k1), sv_scale(h * (9.0 / 40.0), k2))));
  const StateVector k4 = f(sv_add(y, sv_add(sv_scale(h * (44.0 / 45.0), k1),
                                            sv_add(sv_scale(h * (-56.0 / 15.0), k2), sv_scale(h * (32.0 / 9.0), k3)))));
  const StateVector k5 = f(sv_add(
      y, sv_add(sv_scale(h * (19372.0 / 6561.0), k1),
                sv_add(sv_scale(h * (-25360.0 / 2187.0), k2),
                       sv_add(sv_scale(h * (64448.0 / 6561.0), k3), sv_scale(h * (-212.0 / 729.0), k4))))));
  const StateVector k6 = f(sv_add(
      y, sv_add(sv_scale(h * (9017.0 / 3168.0), k1),
                sv_add(sv_scale(h * (-355.0 / 33.0), k2),
                       sv_add(sv_scale(h * (46732.0 / 5247.0), k3),
                              sv_add(sv_scale(h * (49.0 / 176.0), k4), sv_scale(h * (-5103.0 / 18656.0), k5)))))));
  const StateVector y5 = sv_add(
      y, sv_add(sv_scale(h * (35.0 / 384.0), k1),
                sv_add(sv_scale(h * (500.0 / 1113.0), k3),
                       sv_add(sv_scale(h * (125.0 / 192.0), k4),
                              sv_add(sv_scale(h * (-2187.0 / 6784.0), k5), sv_scale(h * (11.0 / 84.0), k6))))));
  const StateVector k7 = f(y5);
  const StateVector dy = sv_add(
      sv_scale(h * (71.0 / 57600.0), k1),
      sv_add(sv_scale(h * (-71.0 / 16695.0), k3),
             sv_add(sv_scale(h * (71.0 / 1920.0), k4),
                    sv_add(sv_scale(h * (-17253.0 / 339200.0), k5),
                           sv_add(sv_scale(h * (22.0 / 525.0), k6), sv_scale(h * (-1.0 / 40.0), k7))))));
  const double a[8] = {y.x0, y.x1, y.x2, y.x3, y.v0, y.v1, y.v2, y.v3};
  const double b[8] = {y5.x0, y5.x1, y5.x2, y5.x3, y5.v0, y5.v1, y5.v2, y5.v3};
  const double d[8] = {dy.x0, dy.x1, dy.x2, dy.x3, dy.v0, dy.v1, dy.v2, dy.v3};
  double acc = 0.0;
  for (int i = 0; i < 8; ++i) {
    const double sc = atol + rtol * std::max(std::fabs(a[i]), std::fabs(b[i]));
    acc += (d[i] / sc) * (d[i] / sc);
  }
  return {y5, std::sqrt(acc / 8.0)};
}

// Next step size: h min(5, max(0.2, 0.9 err^(-1/5))) (no growth right after a
// rejection).
[[nodiscard]] inline double dp45_next_step(double h, double err, bool last_rejected) noexcept {
  const double fac = err <= 0.0 ? 5.0 : 0.9 * std::pow(err, -0.2);
  return h * std::min(last_rejected ? 1.0 : 5.0, std::max(0.2, fac));
}

}  // namespace verified

#endif  // BH_PHYSICS_HPP
