// bh_physics.hpp - the reference's physics C++ interface for the geodesic
// path (src/physics/kerr.h:91-546, src/physics/raytracer.h:46-79,309-396,
// src/physics/verified/rk4.hpp:49-437), its conserved quantities
// (src/physics/conservation_monitor.h:30-184), event detection
// (src/physics/event_detection.h:39-241), its Hamiltonian form
// (src/physics/symplectic_integrator.h:40-298, connection.h:127-178) and
// polarised transfer (src/physics/stokes_transport.h:130-747), header-only
// C++17 over the library.
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
//    polylines);
//  * stokesStep uses the exact segment solution at every optical depth (a
//    series for |(alpha_I - i rho_V) ds| < 1e-3) instead of a first-order
//    Taylor step below alpha_I ds = 1e-4; the relativistic Faraday factors
//    are always the Bessel-function forms (the reference falls back to
//    1 / (1 + Theta^2) and 0.5 / Theta^2 when Boost is absent);
//  * kerrForce differentiates H analytically (the reference takes central
//    differences with eps = 1e-5).
#ifndef BH_PHYSICS_HPP
#define BH_PHYSICS_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <limits>
#include <utility>
#include <vector>

#include "bh_symplectic.h"
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

// ---------------------------------------------------------------------------
// Conserved quantities of a geodesic from its contravariant velocity
// (src/physics/conservation_monitor.h:30-184), geometric units, rS = 2 mGeom:
// E = -u_t, L_z = u_phi, mu2 = -g(u, u) (0 for photons), Carter
// Q = u_theta^2 + cos^2 theta (a^2 (mu2 - E^2) + L_z^2 / sin^2 theta).
// ---------------------------------------------------------------------------

struct ConservedQuantities {
  double e = 0.0, lz = 0.0, q = 0.0, mu2 = 0.0;
};

[[nodiscard]] inline ConservedQuantities computeConserved(double r, double theta, double ut, double ur,
                                                          double utheta, double uphi, double mGeom, double a) {
  const double s = std::sin(theta), c = std::cos(theta), s2 = s * s, c2 = c * c;
  const double rS = 2.0 * mGeom;
  const double sig = r * r + a * a * c2, del = r * r - rS * r + a * a;
  // covariant metric components (Boyer-Lindquist)
  const double g_tt = -(1.0 - rS * r / sig);
  const double g_tp = -rS * r * a * s2 / sig;
  const double g_pp = (r * r + a * a + rS * r * a * a * s2 / sig) * s2;
  const double g_rr = sig / del, g_hh = sig;
  ConservedQuantities cq;
  const double u_t = g_tt * ut + g_tp * uphi, u_p = g_tp * ut + g_pp * uphi, u_h = g_hh * utheta;
  cq.e = -u_t;
  cq.lz = u_p;
  cq.mu2 = -(u_t * ut + u_p * uphi + g_rr * ur * ur + u_h * utheta);
  cq.q = u_h * u_h + c2 * (a * a * (cq.mu2 - cq.e * cq.e) + cq.lz * cq.lz / s2);
  return cq;
}

// Largest drift of E, L_z, Q (relative to the first state, floored at 1e-30)
// and of mu2 (absolute) over the states it has been updated with.
struct GeodesicConservationMonitor {
  ConservedQuantities initial;
  bool initialized = false;
  double maxEDrift = 0.0, maxLzDrift = 0.0, maxQDrift = 0.0, maxMu2Drift = 0.0;
  int stepCount = 0;

  void init(const ConservedQuantities& cq) {
    *this = GeodesicConservationMonitor{};
    initial = cq;
    initialized = true;
  }
  // Returns the largest of the current drifts (0 on the first call, which
  // records the initial state).
  [[nodiscard]] double update(const ConservedQuantities& cur) {
    if (!initialized) {
      init(cur);
      return 0.0;
    }
    ++stepCount;
    auto rel = [](double v0, double v1) { return std::fabs(v1 - v0) / std::max(std::fabs(v0), 1e-30); };
    const double dE = rel(initial.e, cur.e), dL = rel(initial.lz, cur.lz), dQ = rel(initial.q, cur.q);
    const double dM = std::fabs(cur.mu2 - initial.mu2);
    maxEDrift = std::max(maxEDrift, dE);
    maxLzDrift = std::max(maxLzDrift, dL);
    maxQDrift = std::max(maxQDrift, dQ);
    maxMu2Drift = std::max(maxMu2Drift, dM);
    return std::max(std::max(dE, dL), std::max(dQ, dM));
  }
  [[nodiscard]] bool violated(double threshold = 1e-8) const { return maxDrift() > threshold; }
  [[nodiscard]] double maxDrift() const {
    return std::max(std::max(maxEDrift, maxLzDrift), std::max(maxQDrift, maxMu2Drift));
  }
};

// ---------------------------------------------------------------------------
// Event detection along a ray (src/physics/event_detection.h:39-241): sign
// changes of r - r_target, v_r or theta - pi/2 between two samples, and
// bisection of the affine parameter to refine a radial crossing or a radial
// turning point.  State is any type with members x1 (r), x2 (theta) and v1
// (dr/dlambda); stepper(l_a, l_b, s_a) returns the state at l_b from s_a.
// ---------------------------------------------------------------------------

enum class EventType { None, HorizonCrossing, PhotonSphereCrossing, TurningPoint, RadialThreshold, EquatorialCrossing };

struct EventResult {
  EventType type = EventType::None;
  double lambda = 0.0;  // affine parameter at the event
  double r = 0.0;
  double theta = 0.0;
  int nBisections = 0;
  bool detected = false;
};

[[nodiscard]] inline bool crossedRadius(double r0, double r1, double rTarget) {
  return (r0 - rTarget) * (r1 - rTarget) < 0.0;
}
[[nodiscard]] inline bool hasTurningPoint(double vr0, double vr1) { return vr0 * vr1 < 0.0; }
[[nodiscard]] inline bool crossedEquator(double theta0, double theta1) {
  return (theta0 - 0.5 * PI) * (theta1 - 0.5 * PI) < 0.0;
}

namespace detail {
// Halves [l0, l1] while keeping the sign change of f(state) in the kept half
// (the left end's state is carried, so the stepper always starts there);
// stops when |f(mid)| < tol or after max_iter halvings.
template <typename State, typename Stepper, typename F>
EventResult bisect_event(EventType type, double l0, double l1, const State& s0, const Stepper& stepper, F f,
                         double tol, int max_iter) {
  EventResult res;
  res.type = type;
  double la = l0, lb = l1;
  State sa = s0;
  for (int i = 0; i < max_iter; ++i) {
    const double lm = 0.5 * (la + lb);
    const State sm = stepper(la, lm, sa);
    if (std::fabs(f(sm)) < tol) {
      res.detected = true;
      res.lambda = lm;
      res.r = sm.x1;
      res.theta = sm.x2;
      res.nBisections = i + 1;
      return res;
    }
    if (f(sa) * f(sm) < 0.0) {
      lb = lm;
    } else {
      la = lm;
      sa = sm;
    }
  }
  const double lm = 0.5 * (la + lb);
  const State sm = stepper(la, lm, sa);
  res.detected = true;  // best estimate after max_iter halvings
  res.lambda = lm;
  res.r = sm.x1;
  res.theta = sm.x2;
  res.nBisections = max_iter;
  return res;
}
}  // namespace detail

// The affine parameter where r passes rTarget within [lambda0, lambda1]
// (event_detection.h:127-191).
template <typename State, typename Stepper>
EventResult bisectRadialCrossing(double lambda0, double lambda1, const State& state0, double rTarget,
                                 const Stepper& stepper, double tolR = 1e-10, int maxIter = 50) {
  return detail::bisect_event(
      EventType::RadialThreshold, lambda0, lambda1, state0, stepper,
      [rTarget](const State& s) { return s.x1 - rTarget; }, tolR, maxIter);
}

// The affine parameter where v_r changes sign within [lambda0, lambda1].
template <typename State, typename Stepper>
EventResult bisectTurningPoint(double lambda0, double lambda1, const State& state0, const Stepper& stepper,
                               double tolVr = 1e-12, int maxIter = 50) {
  return detail::bisect_event(
      EventType::TurningPoint, lambda0, lambda1, state0, stepper, [](const State& s) { return s.v1; }, tolVr,
      maxIter);
}

// The n-th photon subring is labelled by its number of radial turning points.
[[nodiscard]] inline int photonSubringIndex(int turningPointCount) { return turningPointCount; }

// ---------------------------------------------------------------------------
// Hamiltonian form (src/physics/symplectic_integrator.h:40-298 and the
// inverse metric of src/physics/connection.h:127-178), geometric units with
// rS = 2M: phase-space state (q, p) = ((t, r, theta, phi), covariant
// momenta), H = g^{mu nu} p_mu p_nu / 2 (0 on null geodesics), and the
// fourth-order Yoshida composition of drift and kick sub-steps.
// ---------------------------------------------------------------------------

struct KerrParams {
  double a;   // spin J/M
  double rS;  // 2M, > 0
};
struct GeodesicState {
  std::array<double, 4> q;  // t, r, theta, phi
  std::array<double, 4> p;  // p_t, p_r, p_theta, p_phi
};
using Metric4 = std::array<std::array<double, 4>, 4>;

// Contravariant Boyer-Lindquist metric, with the reference's guards: Sigma,
// Delta and A = (r^2 + a^2)^2 - a^2 Delta sin^2 kept at least 1e-30 in
// magnitude, g^{phi phi} = 0 where sin^2 theta <= 1e-30.
[[nodiscard]] inline Metric4 kerrGcon(double r, double theta, double a, double rS) noexcept {
  const double s = std::sin(theta), c = std::cos(theta);
  const double s2 = s * s, a2 = a * a, r2a2 = r * r + a2;
  double sig = r * r + a2 * c * c, del = r * r - rS * r + a2;
  double A = r2a2 * r2a2 - a2 * del * s2;
  if (std::fabs(sig) < 1e-30) sig = 1e-30;
  if (std::fabs(del) < 1e-30) del = del >= 0.0 ? 1e-30 : -1e-30;
  if (std::fabs(A) < 1e-30) A = 1e-30;
  const double ds = del * sig;
  Metric4 g{};
  g[0][0] = -A / ds;
  g[0][3] = g[3][0] = -rS * r * a / ds;
  g[1][1] = del / sig;
  g[2][2] = 1.0 / sig;
  g[3][3] = std::fabs(s2) > 1e-30 ? (del - a2 * s2) / (ds * s2) : 0.0;
  return g;
}

namespace detail {
// Carter's separated form of 2 Sigma H with E = -p_t, L = p_phi:
//   Delta p_r^2 - P^2 / Delta + p_theta^2 + W^2,  P = (r^2 + a^2) E - a L,
//   W = L / sin(theta) - a E sin(theta)
// (algebraically g^{mu nu} p_mu p_nu Sigma, since r^2 + a^2 - Delta = rS r),
// and its derivatives in r and theta; out = {2H, d(2H)/dr, d(2H)/dtheta}.
inline void kerr_2h(const GeodesicState& st, const KerrParams& k, double* out) noexcept {
  const double r = st.q[1], th = st.q[2];
  const double E = -st.p[0], pr = st.p[1], pth = st.p[2], L = st.p[3];
  const double s = std::sin(th), c = std::cos(th), a = k.a, a2 = a * a;
  const double sig = r * r + a2 * c * c, del = r * r - k.rS * r + a2;
  const double P = (r * r + a2) * E - a * L, W = L / s - a * E * s;
  const double Rp = del * pr * pr - P * P / del;
  const double Tp = pth * pth + W * W;
  const double dRp = (2.0 * r - k.rS) * (pr * pr + P * P / (del * del)) - 4.0 * r * E * P / del;
  const double dTp = -2.0 * W * c * (L / (s * s) + a * E);
  out[0] = (Rp + Tp) / sig;
  out[1] = dRp / sig - (Rp + Tp) * 2.0 * r / (sig * sig);
  out[2] = dTp / sig + (Rp + Tp) * 2.0 * a2 * c * s / (sig * sig);
}
}  // namespace detail

// H = g^{mu nu} p_mu p_nu / 2 (symplectic_integrator.h:123-133).
[[nodiscard]] inline double kerrHamiltonian(const GeodesicState& s, const KerrParams& k) noexcept {
  const Metric4 g = kerrGcon(s.q[1], s.q[2], k.a, k.rS);
  const double tt = g[0][0] * s.p[0] * s.p[0] + 2.0 * g[0][3] * s.p[0] * s.p[3] + g[3][3] * s.p[3] * s.p[3];
  return 0.5 * (tt + g[1][1] * s.p[1] * s.p[1] + g[2][2] * s.p[2] * s.p[2]);
}

// dq/dlambda = dH/dp = g^{mu nu} p_nu.
[[nodiscard]] inline std::array<double, 4> kerrVelocity(const GeodesicState& s, const KerrParams& k) noexcept {
  const Metric4 g = kerrGcon(s.q[1], s.q[2], k.a, k.rS);
  return {g[0][0] * s.p[0] + g[0][3] * s.p[3], g[1][1] * s.p[1], g[2][2] * s.p[2],
          g[3][0] * s.p[0] + g[3][3] * s.p[3]};
}

// dp/dlambda = -dH/dq: analytic derivatives of the separated form (t and phi
// are cyclic, so their components are 0).  The reference differentiates
// numerically with step eps (central differences, error O(eps^2)); eps is
// accepted for the same signature and not used.
[[nodiscard]] inline std::array<double, 4> kerrForce(const GeodesicState& s, const KerrParams& k,
                                                     double eps = 1.0e-5) noexcept {
  (void)eps;
  double h[3];
  detail::kerr_2h(s, k, h);
  return {0.0, -0.5 * h[1], -0.5 * h[2], 0.0};
}

inline void applyDrift(GeodesicState& s, double h, const KerrParams& k) noexcept {
  const std::array<double, 4> v = kerrVelocity(s, k);
  for (std::size_t i = 0; i < 4; ++i) s.q[i] += h * v[i];
}

inline void applyKick(GeodesicState& s, double h, const KerrParams& k) noexcept {
  const std::array<double, 4> f = kerrForce(s, k);
  for (std::size_t i = 0; i < 4; ++i) s.p[i] += h * f[i];
}

// Yoshida (1990) fourth-order composition D(c1) K(w1) D(c2) K(w0) D(c2) K(w1)
// D(c1), w1 = 1 / (2 - 2^(1/3)), w0 = 1 - 2 w1, c1 = w1 / 2, c2 = (w0 + w1) / 2
// (symplectic_integrator.h:270-298).  H mixes q and p, so each drift or
// kick sub-step is an explicit Euler step of its flow, not the exact flow,
// and the composition is first order: the H drift and the position error
// halve with h (r = 20 M to lambda = 8, a = 0.9), where a separable H
// would give fourth order.  Kept as the reference composes it.
inline void yoshida4Step(GeodesicState& s, double h, const KerrParams& k) noexcept {
  const double w1 = 1.0 / (2.0 - std::cbrt(2.0));
  const double w0 = 1.0 - 2.0 * w1;
  const double c1 = 0.5 * w1, c2 = 0.5 * (w0 + w1);
  applyDrift(s, c1 * h, k);
  applyKick(s, w1 * h, k);
  applyDrift(s, c2 * h, k);
  applyKick(s, w0 * h, k);
  applyDrift(s, c2 * h, k);
  applyKick(s, w1 * h, k);
  applyDrift(s, c1 * h, k);
}

// nSteps Yoshida steps of dLambda / nSteps (symplectic_integrator.h:304-311);
// a negative dLambda integrates backwards.
inline void integrateGeodesic(GeodesicState& s, double dLambda, int nSteps, const KerrParams& k) noexcept {
  if (nSteps <= 0) return;
  const double h = dLambda / static_cast<double>(nSteps);
  for (int i = 0; i < nSteps; ++i) yoshida4Step(s, h, k);
}

// integrateGeodesic for many states at once on the GPU (bh_symplectic.h):
// the same composition, one state per lane.  False on a device error
// (bh_last_error()); the states are then unchanged.
inline bool integrateGeodesicBatch(std::vector<GeodesicState>& states, double dLambda, int nSteps,
                                   const KerrParams& k) {
  if (states.empty() || nSteps <= 0) return true;
  std::vector<double> buf(states.size() * 8);
  for (size_t i = 0; i < states.size(); ++i)
    for (size_t j = 0; j < 4; ++j) {
      buf[i * 8 + j] = states[i].q[j];
      buf[i * 8 + 4 + j] = states[i].p[j];
    }
  if (bh_symplectic_integrate_host(buf.data(), static_cast<int64_t>(states.size()), dLambda, nSteps, k.a,
                                   k.rS) != 0)
    return false;
  for (size_t i = 0; i < states.size(); ++i)
    for (size_t j = 0; j < 4; ++j) {
      states[i].q[j] = buf[i * 8 + j];
      states[i].p[j] = buf[i * 8 + 4 + j];
    }
  return true;
}

// ---------------------------------------------------------------------------
// Polarised transfer (src/physics/stokes_transport.h:130-747): Stokes
// vectors, the exact segment solution for absorption alpha_I plus Faraday
// rotation rho_V, an RK4 step of the full Mueller propagation matrix,
// synchrotron emission and Faraday coefficients, GR transforms.  The GPU
// tracer carries the same segment solution (bh_trace_image_plane_pol).
// ---------------------------------------------------------------------------

inline constexpr double M_ELECTRON = 9.1093837015e-28;  // g
inline constexpr double E_CHARGE = 4.80320425e-10;      // esu

struct StokesVector {
  double i = 0.0, q = 0.0, u = 0.0, v = 0.0;

  [[nodiscard]] double linPolInt() const noexcept { return std::hypot(q, u); }
  [[nodiscard]] double linPolFrac() const noexcept { return i > 0.0 ? linPolInt() / i : 0.0; }
  [[nodiscard]] double circPolFrac() const noexcept { return i > 0.0 ? v / i : 0.0; }
  [[nodiscard]] double totalPolFrac() const noexcept {
    return i > 0.0 ? std::sqrt(q * q + u * u + v * v) / i : 0.0;
  }
  // Electric vector position angle chi = arg(Q + iU) / 2 in (-pi/2, pi/2].
  [[nodiscard]] double evpa() const noexcept { return 0.5 * std::atan2(u, q); }
  // Q + iU -> e^{2 i dchi} (Q + iU).
  [[nodiscard]] StokesVector rotateEvpa(double dchi) const noexcept {
    const double c = std::cos(2.0 * dchi), s = std::sin(2.0 * dchi);
    return {i, c * q - s * u, s * q + c * u, v};
  }
  [[nodiscard]] StokesVector operator+(const StokesVector& o) const noexcept {
    return {i + o.i, q + o.q, u + o.u, v + o.v};
  }
  [[nodiscard]] StokesVector operator*(double k) const noexcept { return {k * i, k * q, k * u, k * v}; }
};

struct StokesEmission {
  double jI = 0.0, jQ = 0.0, jU = 0.0, jV = 0.0;
  [[nodiscard]] double linPolFrac() const noexcept { return jI > 0.0 ? std::hypot(jQ, jU) / jI : 0.0; }
};

// Entries of the propagation matrix K (aligned frame: alpha_U = rho_U = 0)
// and the segment length.
struct FaradayPropagation {
  double alphaI = 0.0, alphaQ = 0.0, alphaV = 0.0, rhoV = 0.0, rhoQ = 0.0, dsCm = 0.0;
};

namespace detail {
// (1 - e^{-w}) / w for complex w = x + i y, returned as (re, im); a Taylor
// series below |w| = 1e-3 (relative error < 1e-13) and the closed form above.
inline void one_minus_exp_over(double x, double y, double* re, double* im) noexcept {
  const double m2 = x * x + y * y;
  if (m2 < 1e-6) {
    // 1 - w/2 + w^2/6 - w^3/24
    const double w2r = x * x - y * y, w2i = 2.0 * x * y;
    const double w3r = w2r * x - w2i * y, w3i = w2r * y + w2i * x;
    *re = 1.0 - 0.5 * x + w2r / 6.0 - w3r / 24.0;
    *im = -0.5 * y + w2i / 6.0 - w3i / 24.0;
    return;
  }
  const double e = x < 745.0 ? std::exp(-x) : 0.0;
  const double nr = 1.0 - e * std::cos(y), ni = e * std::sin(y);  // 1 - e^{-w}
  *re = (nr * x + ni * y) / m2;
  *im = (ni * x - nr * y) / m2;
}
}  // namespace detail

// Exact transfer over one uniform segment with K = alpha_I (all channels)
// plus Faraday rotation rho_V: I and V obey dX/ds = j_X - A X, and P = Q + iU
// obeys dP/ds = j_P - (A - i R) P, so with z = (A - i R) ds,
// P' = e^{-z} P + j_P ds (1 - e^{-z}) / z (same for I, V with z = A ds).
// A = max(alpha_I, 0); dsCm <= 0 returns s unchanged.
[[nodiscard]] inline StokesVector stokesStep(const StokesVector& s, const StokesEmission& em, double alphaI,
                                             double rhoV, double dsCm) noexcept {
  if (!(dsCm > 0.0)) return s;
  const double A = alphaI > 0.0 ? alphaI : 0.0;
  const double x = A * dsCm, y = -rhoV * dsCm;  // z = x + i y
  const double e = x < 745.0 ? std::exp(-x) : 0.0;
  double fr, fi;
  detail::one_minus_exp_over(x, 0.0, &fr, &fi);
  const double gI = dsCm * fr;  // (1 - e^{-x}) / A
  detail::one_minus_exp_over(x, y, &fr, &fi);
  const double cr = dsCm * fr, ci = dsCm * fi;  // (1 - e^{-z}) / (A - i R)
  const double ec = e * std::cos(y), es = -e * std::sin(y);  // e^{-z} = e^{-x} (cos y - i sin y)
  StokesVector o;
  o.i = e * s.i + gI * em.jI;
  o.v = e * s.v + gI * em.jV;
  o.q = (ec * s.q - es * s.u) + (cr * em.jQ - ci * em.jU);
  o.u = (es * s.q + ec * s.u) + (ci * em.jQ + cr * em.jU);
  return o;
}

// One classic RK4 step of dS/ds = J - K S with the full K (alpha_I, alpha_Q,
// alpha_V, rho_V, rho_Q; aligned frame) over k.dsCm.
[[nodiscard]] inline StokesVector stokesStepFull(const StokesVector& s, const StokesEmission& em,
                                                 const FaradayPropagation& k) noexcept {
  if (!(k.dsCm > 0.0)) return s;
  auto f = [&](const StokesVector& x) -> StokesVector {
    return {em.jI - k.alphaI * x.i - k.alphaQ * x.q - k.alphaV * x.v,
            em.jQ - k.alphaQ * x.i - k.alphaI * x.q - k.rhoV * x.u,
            em.jU + k.rhoV * x.q - k.alphaI * x.u - k.rhoQ * x.v,
            em.jV - k.alphaV * x.i + k.rhoQ * x.u - k.alphaI * x.v};
  };
  const double h = k.dsCm;
  const StokesVector k1 = f(s);
  const StokesVector k2 = f(s + k1 * (0.5 * h));
  const StokesVector k3 = f(s + k2 * (0.5 * h));
  const StokesVector k4 = f(s + k3 * h);
  return s + (k1 + k2 * 2.0 + k3 * 2.0 + k4) * (h / 6.0);
}

// stokesStep over the segments in order (far end first, observer last).
[[nodiscard]] inline StokesVector integrateStokesPath(const std::vector<StokesEmission>& emissions,
                                                      const std::vector<FaradayPropagation>& props,
                                                      StokesVector initial = {}) noexcept {
  const std::size_t n = std::min(emissions.size(), props.size());
  for (std::size_t k = 0; k < n; ++k)
    initial = stokesStep(initial, emissions[k], props[k].alphaI, props[k].rhoV, props[k].dsCm);
  return initial;
}

// I^2 + tol >= Q^2 + U^2 + V^2.
[[nodiscard]] inline bool stokesPolarizationBound(const StokesVector& s, double tol = 1.0e-10) noexcept {
  return s.i * s.i + tol >= s.q * s.q + s.u * s.u + s.v * s.v;
}

// Optically thin synchrotron polarisation of a power law N ~ gamma^-p:
// (p + 1) / (p + 7/3) (Rybicki & Lightman 1979, eq. 6.38); 0 for p <= 0.
[[nodiscard]] inline double synchrotronLinearPolarFrac(double p) noexcept {
  return p > 0.0 ? (p + 1.0) / (p + 7.0 / 3.0) : 0.0;
}

// Thermal synchrotron: the power-law fraction at the effective index
// p = 3 + ln(1 + Theta_e) / 2; 0 for Theta_e <= 0.
[[nodiscard]] inline double thermalSynchLinearPolarFrac(double thetaE) noexcept {
  return thetaE > 0.0 ? synchrotronLinearPolarFrac(3.0 + 0.5 * std::log1p(thetaE)) : 0.0;
}

// (j_I, j_Q, j_U, 0) with the E vector perpendicular to the projected field
// at angle chi_B: j_Q + i j_U = -pi_lin j_I e^{2 i chi_B}.
[[nodiscard]] inline StokesEmission synchrotronPolarizedEmission(double jI, double piLin,
                                                                 double chiBRad) noexcept {
  return {jI, -piLin * jI * std::cos(2.0 * chiBRad), -piLin * jI * std::sin(2.0 * chiBRad), 0.0};
}

namespace detail {
// e^z K_nu(z) = int_0^inf exp(-z (cosh t - 1)) cosh(nu t) dt (Watson 1944,
// 6.22(5)); trapezoid rule with step 0.01, exponentially convergent for this
// analytic, double-exponentially decaying integrand.
inline double bessel_k_scaled(double nu, double z) noexcept {
  const double dt = 0.01;
  double s = 0.5;
  for (int i = 1; i < 40000; ++i) {
    const double t = i * dt;
    const double term = std::exp(-z * (std::cosh(t) - 1.0)) * std::cosh(nu * t);
    s += term;
    if (term < 1e-18 * s) break;
  }
  return s * dt;
}
inline double faraday_prefactor() noexcept {
  return E_CHARGE * E_CHARGE * E_CHARGE / (M_ELECTRON * M_ELECTRON * C2 * C2);
}
}  // namespace detail

// Cold-plasma Faraday rotation rho_V = e^3 n_e B_par / (2 pi m_e^2 c^4 nu^2)
// [rad/cm]; 0 for nu <= 0 or n_e <= 0.
[[nodiscard]] inline double faradayRotationCoeff(double nu, double nE, double bParallel) noexcept {
  if (nu <= 0.0 || nE <= 0.0) return 0.0;
  return detail::faraday_prefactor() / (2.0 * PI) * nE * bParallel / (nu * nu);
}

// Thermal (Maxwell-Juettner) rotation: the cold value times
// f = (K0(1/Theta) + K1(1/Theta)) / (2 Theta^2 K2(1/Theta)), the factor the
// reference states (Shcherbakov 2008, B12); Theta <= 0 gives the cold value.
[[nodiscard]] inline double faradayRotationCoeffRelativistic(double nu, double nE, double bParallel,
                                                             double thetaE) noexcept {
  if (thetaE <= 0.0) return faradayRotationCoeff(nu, nE, bParallel);
  const double z = 1.0 / thetaE;
  const double k2 = detail::bessel_k_scaled(2.0, z);
  const double f = (detail::bessel_k_scaled(0.0, z) + detail::bessel_k_scaled(1.0, z)) / (2.0 * thetaE * thetaE * k2);
  return faradayRotationCoeff(nu, nE, bParallel) * f;
}

// Faraday conversion rho_Q = e^3 n_e B_perp f / (pi m_e^2 c^4 nu^2),
// f = max(0, K1(1/Theta) / (Theta K2(1/Theta)) - 1 / (2 Theta^2)); 0 for
// nu, n_e or Theta <= 0.
[[nodiscard]] inline double faradayConversionCoeff(double nu, double nE, double bPerp, double thetaE) noexcept {
  if (nu <= 0.0 || nE <= 0.0 || thetaE <= 0.0) return 0.0;
  const double z = 1.0 / thetaE;
  const double f = std::max(0.0, detail::bessel_k_scaled(1.0, z) / (thetaE * detail::bessel_k_scaled(2.0, z)) -
                                     0.5 / (thetaE * thetaE));
  return detail::faraday_prefactor() / PI * nE * bPerp * f / (nu * nu);
}

// Rotation measure sum rho_V ds [rad].
[[nodiscard]] inline double accumulatedRM(const std::vector<double>& rhoVPerStep,
                                          const std::vector<double>& dsPerStep) noexcept {
  const std::size_t n = std::min(rhoVPerStep.size(), dsPerStep.size());
  double rm = 0.0;
  for (std::size_t k = 0; k < n; ++k) rm += rhoVPerStep[k] * dsPerStep[k];
  return rm;
}

[[nodiscard]] inline StokesVector faradayRotateStokes(const StokesVector& s, double rmRad) noexcept {
  return s.rotateEvpa(rmRad);
}

// Frequency-integrated Stokes intensities scale as g^3 (invariance of
// I_nu / nu^3 integrated over frequency with I ~ g^4 / g); g <= 0 gives 0.
[[nodiscard]] inline StokesVector grTransformStokes(const StokesVector& s, double g) noexcept {
  if (!(g > 0.0)) return {};
  return s * (g * g * g);
}

// Parallel transport of the polarisation basis along the ray rotates the
// EVPA by psi.
[[nodiscard]] inline StokesVector grParallelTransportRotate(const StokesVector& s, double psiRad) noexcept {
  return s.rotateEvpa(psiRad);
}

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
