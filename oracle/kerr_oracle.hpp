// kerr_oracle.hpp - FP64 CPU oracle of the per-pixel Kerr null-geodesic tracer.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is compiled into
// oracle/liboracle.so, which only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load.  The product (blackhole_amd/) never
// includes, links or calls it.
//
// It restates the reference's hot-path algorithm in plain FP64 C++ with the
// reference's own defects fixed (turning-point stall, Theta-potential term,
// uncompilable DP45), written so that every arithmetic step is an IEEE
// operation (+ - * / sqrt fma) whose result is fixed by the standard; the
// GPU FP64 kernel performs the same operations in the same order, so the two
// agree bit for bit on the trajectory.  Transcendentals (acos, exp, expm1,
// cbrt) appear only in the radiative part, never in the stepping.
//
// Templated on the scalar so that oracle/flopcount.cpp can instantiate the
// same code with a counting scalar (FMA = 2, +-*/ sqrt = 1, transcendentals
// tallied separately) and fix the algorithmic FLOP per step.
//
// Pinning (tests/test_oracle_pinning.py, tests/test_stokes.py): golden
// fixtures of the reference (assets/luts/*.csv), known answers of the
// reference's tests, the reference's own functions compiled into
// oracle/_ref, and closed forms.  Every function here has at least one such
// check.  The RK4 / DP45 trajectories have no reference vector (the
// reference's stepper stalls at turning points and its DP45 does not
// compile); they are pinned by closed forms instead: the Schwarzschild
// capture threshold and deflection integral, the Kerr radial turning point,
// the conserved integrals and the local error orders.  pol_segment is pinned
// by composing the reference's compiled stokesStep forward along random
// paths.
#pragma once

#include <cmath>
#include <cstdint>

#include "../include/bhk_trace.h"

namespace orc {

// ---------------------------------------------------------------------------
// Scalar helpers (overloaded for double here; for the counting scalar in
// flopcount.cpp).
// ---------------------------------------------------------------------------
inline double fma_(double a, double b, double c) { return std::fma(a, b, c); }
inline double sqrt_(double x) { return std::sqrt(x); }
inline double fabs_(double x) { return std::fabs(x); }
inline double floor_(double x) { return std::floor(x); }
inline double val_(double x) { return x; }
inline double acos_(double x) { return std::acos(x); }
inline double exp_(double x) { return std::exp(x); }
inline double expm1_(double x) { return std::expm1(x); }
inline double cbrt_(double x) { return std::cbrt(x); }
inline double log_(double x) { return std::log(x); }
inline double frexp_(double x, int* e) { return std::frexp(x, e); }
inline double ldexp_(double x, int e) { return std::ldexp(x, e); }

// CGS constants, values of src/physics/constants.h:23-48 and
// src/physics/synchrotron.h (E_CHARGE, M_ELECTRON).
constexpr double kC = 2.99792458e10;
constexpr double kHbar = 1.054571817e-27;
constexpr double kKB = 1.380649e-16;
constexpr double kECharge = 4.80320425e-10;
constexpr double kMElectron = 9.1093837015e-28;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kH = 2.0 * kPi * kHbar;  // rte_integrator.h:222-224

// ---------------------------------------------------------------------------
// Closed forms (M = 1)
// ---------------------------------------------------------------------------

// r_+ = M + sqrt(M^2 - a^2)   (src/physics/kerr.h:139-150)
inline double outer_horizon(double a) { return 1.0 + std::sqrt(1.0 - a * a); }

// Prograde ISCO, Bardeen-Press-Teukolsky (src/physics/kerr.h:200-229,
// src/physics/iron_kline.h:65-75).
inline double isco_prograde(double a) {
  const double z1 = 1.0 + std::cbrt(1.0 - a * a) * (std::cbrt(1.0 + a) + std::cbrt(1.0 - a));
  const double z2 = std::sqrt(3.0 * a * a + z1 * z1);
  return 3.0 + z2 - std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

// Photon orbits (src/physics/analytic_kerr_geodesic.h:454-470).
inline double photon_orbit_prograde(double a) {
  return 2.0 * (1.0 + std::cos((2.0 / 3.0) * std::acos(-std::fabs(a))));
}
inline double photon_orbit_retrograde(double a) {
  return 2.0 * (1.0 + std::cos((2.0 / 3.0) * std::acos(std::fabs(a))));
}

// Critical impact parameters of the spherical photon orbit at r
// (src/physics/analytic_kerr_geodesic.h:410-448; Bardeen 1973).
inline void critical_params(double r, double a, double* xi, double* eta) {
  const double r2 = r * r, a2 = a * a;
  const double delta = r2 - 2.0 * r + a2;
  const double dp = 2.0 * r - 2.0;
  *xi = (r2 + a2) / a - 4.0 * r * delta / (a * dp);
  const double xa = *xi - a;
  *eta = 16.0 * r2 * delta / (dp * dp) - xa * xa;
}

// ---------------------------------------------------------------------------
// Geodesic equations
//
// Null geodesic with E = 1, L = xi, Carter constant Q = eta in Mino time
// lambda (d tau_affine = Sigma d lambda), with u = cos(theta):
//   (dr/dlambda)^2 = R(r) = P(r)^2 - Delta (eta + (xi - a)^2),  P = r^2 + a^2 - a xi
//   (du/dlambda)^2 = U(u) = eta + (a^2 - eta - xi^2) u^2 - a^2 u^4
//   dphi/dlambda  = a P / Delta - a + xi / (1 - u^2)
// (dt/dlambda = (r^2 + a^2) P / Delta + a (xi - a (1 - u^2)) is not
// integrated: coordinate time feeds no reported quantity.)
// R, R' are the reference's kerrPotentials (src/physics/kerr.cpp:69-97);
// U(u) = sin^2(theta) Theta(theta) with the Theta of the reference's Carter
// constant (src/physics/verified/geodesic.hpp:178-188), i.e.
// Theta = eta + a^2 cos^2 - xi^2 cot^2.  The phi, t rates are
// kerrMinoDerivatives (src/physics/kerr.cpp:40-66).
// Second-order form (r'' = R'/2, u'' = U'/2) replaces the reference's
// sqrt-with-sign-flip form, which stalls at turning points.
// Independent variable s with d lambda = ds / r, so that dr/ds ~ r far away
// (log-uniform radial stepping) while the whole right-hand side stays
// rational: one division per evaluation, no transcendental.
// State y = (r, u, phi, v_r, v_u), v = d/dlambda.
constexpr int kN = 5;
// ---------------------------------------------------------------------------
template <class T>
struct RayConst {
  T a, a2, xi, eta, K, c1, c2, two_a2;
};

template <class T>
inline RayConst<T> make_const(T a, T xi, T eta) {
  RayConst<T> c;
  c.a = a;
  c.a2 = a * a;
  c.xi = xi;
  c.eta = eta;
  const T xa = xi - a;
  c.K = eta + xa * xa;
  c.c1 = c.a2 - a * xi;
  c.c2 = (c.a2 - eta) - xi * xi;
  c.two_a2 = T(2.0) * c.a2;
  return c;
}

template <class T>
inline void rhs(const RayConst<T>& c, const T* y, T* f) {
  const T r = y[0];
  const T u = y[1];
  const T r2 = r * r;
  const T P = r2 + c.c1;
  const T Delta = fma_(r - T(2.0), r, c.a2);
  T w = fma_(-u, u, T(1.0));
  if (val_(w) < 1e-14) w = T(1e-14);
  const T q = T(1.0) / (r * Delta * w);
  const T inv_r = Delta * w * q;
  f[0] = y[3] * inv_r;
  f[1] = y[4] * inv_r;
  f[2] = fma_(c.xi, Delta, c.a * w * (P - Delta)) * q;
  f[3] = fma_(T(2.0) * r, P, -((r - T(1.0)) * c.K)) * inv_r;
  f[4] = u * fma_(-c.two_a2, u * u, c.c2) * inv_r;
}

// R(r) and U(u) (first integrals) for the drift report and the setup.
template <class T>
inline T radial_potential(const RayConst<T>& c, T r) {
  const T P = r * r + c.c1;
  const T Delta = fma_(r - T(2.0), r, c.a2);
  return fma_(P, P, -(Delta * c.K));
}
template <class T>
inline T polar_potential(const RayConst<T>& c, T u) {
  const T u2 = u * u;
  return fma_(fma_(-c.a2, u2, c.c2), u2, c.eta);
}

// Radial null-constraint projection (the reference's null renormalisation,
// src/physics/verified/null_constraint.hpp:208-271, restricted to the radial
// first integral): after an accepted step, where R(r) > kappa (r^2 + a^2)^2 --
// far from any radial turning point -- reset v_r = sign(v_r) sqrt(R(r)) and
// rescale the cached derivative dr/ds = v_r / r accordingly.  Without it the
// absolute error of v_r^2 - R accumulated on the inbound leg (where
// R ~ r^4 ~ 1e8) acts near the hole like an error of the Carter constant.
// Near turning points (R small) the second-order form runs unprojected.
constexpr double kProjKappa = 0.0625;
template <class T>
inline void project_radial(const RayConst<T>& c, T* y, T* f) {
  const T R = radial_potential(c, y[0]);
  const T s = fma_(y[0], y[0], c.a2);
  if (val_(R) > kProjKappa * val_(s * s)) {
    const T vn = sqrt_(R);
    const T v = (val_(y[3]) < 0.0) ? -vn : vn;
    f[0] = f[0] * (v / y[3]);
    y[3] = v;
  }
}

// ---------------------------------------------------------------------------
// Ray setup from the image plane (Bardeen 1973 distant-observer mapping):
//   xi = -alpha sin(theta_o),  eta = beta^2 + (alpha^2 - a^2) cos^2(theta_o)
//   r = r_obs, u = cos(theta_o), v_r = -sqrt(R(r_obs)) (inbound),
//   v_u = sin(theta_o) beta  (so that Theta(theta_o) = beta^2 exactly).
// sin/cos of the inclination come from the host (one libm call per camera)
// so CPU and GPU start from identical bits.
// ---------------------------------------------------------------------------
// Pinhole camera at a finite position (the render-launch camera): the
// camera's Boyer-Lindquist position, the flat spherical basis there, the
// column-major camera basis and the pixel mapping (host values).
struct PinholeConst {
  double r, cos_t, sin_t, phi;
  double e_r[3], e_t[3], e_p[3];
  double basis[9];
  double fov, aspect, shift_x, shift_y, inv_w, inv_h;
};

struct CamConst {
  double a, sin_o, cos_o, r_obs, r_stop, r_isco;
  double alpha_min, dalpha, beta_max, dbeta;
  int64_t w, h, offset;
  const PinholeConst* pin = nullptr;  // pinhole camera: pixels start from it
};

inline CamConst make_cam_const(const BH_Camera& cam, const BH_TraceOptions& o) {
  CamConst k;
  k.a = cam.spin;
  k.sin_o = std::sin(cam.theta_obs);
  k.cos_o = std::cos(cam.theta_obs);
  k.r_obs = cam.r_obs;
  k.r_stop = o.horizon_factor * outer_horizon(cam.spin);
  k.r_isco = isco_prograde(cam.spin);
  k.alpha_min = cam.alpha_min;
  k.dalpha = (cam.alpha_max - cam.alpha_min) / static_cast<double>(cam.width);
  k.beta_max = cam.beta_max;
  k.dbeta = (cam.beta_max - cam.beta_min) / static_cast<double>(cam.height);
  k.w = cam.width;
  k.h = cam.height;
  k.offset = cam.record_offset;
  return k;
}

// Ray through pixel centre (xc, yc) of a pinhole camera: the direction as
// d_ray_dir maps pixels (src/cuda/device_physics.cuh:1953-1967), its
// projections on the flat spherical basis taken as k^r, k^theta, k^phi, k^t
// from the null condition of the Kerr metric, E, L and p_theta as in
// d_kerr_init_consts (device_physics.cuh:316-392), normalised to E = 1.
// Magnitudes of the initial Mino velocities from the potentials, signs from
// k^r and -k^theta.  Requires r > 2 (k^t = (hb + sqrt(disc)) / -g_tt).
inline RayConst<double> setup_pinhole(const PinholeConst& p, double a, double xc, double yc, double* y) {
  double u = (2.0 * xc * p.inv_w - 1.0) * p.fov;
  u = u * p.aspect + p.shift_x;
  const double v = (2.0 * yc * p.inv_h - 1.0) * p.fov + p.shift_y;
  const double lx = -u, ly = -v, lz = 1.0;
  const double inv_n = 1.0 / std::sqrt(lx * lx + ly * ly + lz * lz);
  double d[3];
  for (int i = 0; i < 3; ++i) d[i] = (p.basis[i] * lx + p.basis[3 + i] * ly + p.basis[6 + i] * lz) * inv_n;
  const double r = p.r;
  const double kr = d[0] * p.e_r[0] + d[1] * p.e_r[1] + d[2] * p.e_r[2];
  const double kth = (d[0] * p.e_t[0] + d[1] * p.e_t[1] + d[2] * p.e_t[2]) / r;
  const double kph = (d[0] * p.e_p[0] + d[1] * p.e_p[1] + d[2] * p.e_p[2]) / (r * p.sin_t);
  const double cos2 = p.cos_t * p.cos_t, sin2 = p.sin_t * p.sin_t, a2 = a * a;
  const double sigma = r * r + a2 * cos2;
  const double delta = r * r - 2.0 * r + a2;
  const double f = 2.0 * r / sigma;
  const double g_tt = -(1.0 - f);
  const double g_tp = -f * a * sin2;
  const double g_rr = sigma / delta;
  const double g_pp = (r * r + a2 + f * a2 * sin2) * sin2;
  const double spatial = g_rr * kr * kr + sigma * kth * kth + g_pp * kph * kph;
  const double hb = g_tp * kph;
  const double disc = hb * hb - g_tt * spatial;
  const double kt = (hb + std::sqrt(disc)) / (-g_tt);
  const double E = -(g_tt * kt + g_tp * kph);
  const double L = g_tp * kt + g_pp * kph;
  const double xi = L / E;
  const double pth = sigma * kth / E;
  const double eta = pth * pth - a2 * cos2 + xi * xi * cos2 / sin2;
  const RayConst<double> c = make_const(a, xi, eta);
  y[0] = r;
  y[1] = p.cos_t;
  y[2] = p.phi;
  double R = radial_potential(c, r);
  double U = polar_potential(c, p.cos_t);
  if (R < 0.0) R = 0.0;
  if (U < 0.0) U = 0.0;
  y[3] = kr >= 0.0 ? std::sqrt(R) : -std::sqrt(R);
  y[4] = kth >= 0.0 ? -std::sqrt(U) : std::sqrt(U);
  return c;
}

template <class T>
inline RayConst<T> setup_ray(const CamConst& k, int64_t px, int64_t py, T* y);

template <>
inline RayConst<double> setup_ray<double>(const CamConst& k, int64_t px, int64_t py, double* y);

template <class T>
inline RayConst<T> setup_ray_image_plane(const CamConst& k, int64_t px, int64_t py, T* y) {
  const T alpha = T(k.alpha_min) + (T(static_cast<double>(px)) + T(0.5)) * T(k.dalpha);
  const T beta = T(k.beta_max) - (T(static_cast<double>(py)) + T(0.5)) * T(k.dbeta);
  const T a = T(k.a);
  const T xi = -(alpha * T(k.sin_o));
  const T eta = fma_(beta, beta, (alpha * alpha - a * a) * (T(k.cos_o) * T(k.cos_o)));
  RayConst<T> c = make_const(a, xi, eta);
  y[0] = T(k.r_obs);
  y[1] = T(k.cos_o);
  y[2] = T(0.0);
  T R = radial_potential(c, y[0]);
  if (val_(R) < 0.0) R = T(0.0);
  y[3] = -sqrt_(R);
  y[4] = T(k.sin_o) * beta;
  return c;
}

template <class T>
inline RayConst<T> setup_ray(const CamConst& k, int64_t px, int64_t py, T* y) {
  return setup_ray_image_plane<T>(k, px, py, y);
}

template <>
inline RayConst<double> setup_ray<double>(const CamConst& k, int64_t px, int64_t py, double* y) {
  if (k.pin) return setup_pinhole(*k.pin, k.a, static_cast<double>(px) + 0.5, static_cast<double>(py) + 0.5, y);
  return setup_ray_image_plane<double>(k, px, py, y);
}

// ---------------------------------------------------------------------------
// Integrators
// ---------------------------------------------------------------------------

// Classic RK4 (tableau of src/physics/verified/rk4.hpp:247-260).
template <class T>
inline void rk4_step(const RayConst<T>& c, const T* y, const T* k1, T h, T* yn) {
  T k2[kN], k3[kN], k4[kN], yt[kN];
  const T hh = T(0.5) * h;
  const T h6 = h / T(6.0);
  for (int i = 0; i < kN; ++i) yt[i] = fma_(hh, k1[i], y[i]);
  rhs(c, yt, k2);
  for (int i = 0; i < kN; ++i) yt[i] = fma_(hh, k2[i], y[i]);
  rhs(c, yt, k3);
  for (int i = 0; i < kN; ++i) yt[i] = fma_(h, k3[i], y[i]);
  rhs(c, yt, k4);
  for (int i = 0; i < kN; ++i) {
    const T s = fma_(T(2.0), k2[i] + k3[i], k1[i] + k4[i]);
    yn[i] = fma_(h6, s, y[i]);
  }
}

// Dormand & Prince (1980) 5(4) tableau, as src/physics/verified/rk4.h:193-244.
namespace dp {
constexpr double a21 = 1.0 / 5.0;
constexpr double a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
constexpr double a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
constexpr double a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0,
                 a54 = -212.0 / 729.0;
constexpr double a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0,
                 a64 = 49.0 / 176.0, a65 = -5103.0 / 18656.0;
constexpr double b1 = 35.0 / 384.0, b3 = 500.0 / 1113.0, b4 = 125.0 / 192.0,
                 b5 = -2187.0 / 6784.0, b6 = 11.0 / 84.0;
// e = b(5th) - b(4th); b(4th) = rk4.h:279-281.
constexpr double e1 = 71.0 / 57600.0, e3 = -71.0 / 16695.0, e4 = 71.0 / 1920.0,
                 e5 = -17253.0 / 339200.0, e6 = 22.0 / 525.0, e7 = -1.0 / 40.0;
}  // namespace dp

// One DP45 attempt from y with k1 = f(y).  Writes the 5th-order solution y5,
// k7 = f(y5) (FSAL) and returns the scaled RMS error norm (Hairer, Norsett &
// Wanner I, II.4.11) over r, u, phi, v_r, v_u with
// sc_i = atol + rtol * max(|y_i|, |y5_i|).  (The reference's dp45Step uses an
// unscaled 2-norm, src/physics/verified/rk4.h:284-289.)
template <class T>
inline T dp45_attempt(const RayConst<T>& c, const T* y, const T* k1, T h, T rtol, T atol,
                      T* y5, T* k7) {
  using namespace dp;
  T k2[kN], k3[kN], k4[kN], k5[kN], k6[kN], yt[kN];
  for (int i = 0; i < kN; ++i) yt[i] = fma_(h, T(a21) * k1[i], y[i]);
  rhs(c, yt, k2);
  for (int i = 0; i < kN; ++i) yt[i] = fma_(h, fma_(T(a32), k2[i], T(a31) * k1[i]), y[i]);
  rhs(c, yt, k3);
  for (int i = 0; i < kN; ++i)
    yt[i] = fma_(h, fma_(T(a43), k3[i], fma_(T(a42), k2[i], T(a41) * k1[i])), y[i]);
  rhs(c, yt, k4);
  for (int i = 0; i < kN; ++i)
    yt[i] = fma_(h,
                 fma_(T(a54), k4[i], fma_(T(a53), k3[i], fma_(T(a52), k2[i], T(a51) * k1[i]))),
                 y[i]);
  rhs(c, yt, k5);
  for (int i = 0; i < kN; ++i)
    yt[i] = fma_(h,
                 fma_(T(a65), k5[i],
                      fma_(T(a64), k4[i],
                           fma_(T(a63), k3[i], fma_(T(a62), k2[i], T(a61) * k1[i])))),
                 y[i]);
  rhs(c, yt, k6);
  for (int i = 0; i < kN; ++i)
    y5[i] = fma_(h,
                 fma_(T(b6), k6[i],
                      fma_(T(b5), k5[i], fma_(T(b4), k4[i], fma_(T(b3), k3[i], T(b1) * k1[i])))),
                 y[i]);
  rhs(c, y5, k7);
  T acc = T(0.0);
  for (int i = 0; i < kN; ++i) {
    const T e = h * fma_(T(e7), k7[i],
                         fma_(T(e6), k6[i],
                              fma_(T(e5), k5[i],
                                   fma_(T(e4), k4[i], fma_(T(e3), k3[i], T(e1) * k1[i])))));
    const T ay = fabs_(y[i]);
    const T ay5 = fabs_(y5[i]);
    const T sc = fma_(rtol, (val_(ay) > val_(ay5)) ? ay : ay5, atol);
    const T q = e / sc;
    acc = fma_(q, q, acc);
  }
  return sqrt_(acc * T(0.2));
}

// x^(-1/5) for x > 0 from IEEE operations only (exact frexp/ldexp split,
// chord start on [0.5, 1), two Newton steps y <- y (6 - m y^5) / 5); relative
// error < 2e-6, ample for a step-size controller, and bit-reproducible.
template <class T>
inline T pow_m0p2(T x) {
  int e = 0;
  const T m = frexp_(x, &e);  // x = m 2^e, m in [0.5, 1)
  int qd = (e >= 0) ? (e / 5) : -((-e + 4) / 5);
  const int rr = e - 5 * qd;  // 0..4
  static const double tab[5] = {1.0, 0.87055056329612412, 0.75785828325519911,
                                0.65975395538644721, 0.57434917749851755};  // 2^(-rr/5)
  T y = fma_(T(-0.29739670999407), m, T(1.29739670999407));
  for (int it = 0; it < 2; ++it) {
    const T y2 = y * y;
    const T y5 = y2 * y2 * y;
    y = y * fma_(T(-0.2) * m, y5, T(1.2));
  }
  return ldexp_(y * T(tab[rr]), -qd);
}

// Step-size controller (Hairer I, II.4.13; facmax 5, facmin 0.2, safety 0.9,
// facmax 1 right after a rejection).  Returns the factor applied to h.
template <class T>
inline T dp45_factor(T err, bool accepted, bool last_rejected) {
  if (accepted) {
    const T e = (val_(err) < 1e-10) ? T(1e-10) : err;
    T fac = T(0.9) * pow_m0p2(e);
    if (val_(fac) > 5.0) fac = T(5.0);
    if (last_rejected && val_(fac) > 1.0) fac = T(1.0);
    return fac;
  }
  T fac = T(0.9) * pow_m0p2(err);
  if (val_(fac) < 0.2) fac = T(0.2);
  return fac;
}

// ---------------------------------------------------------------------------
// Equatorial crossing (disk plane u = 0) between y0 (s = 0) and y1 (s = h)
// with derivatives f0, f1: cubic Hermite in s, start at the linear root,
// three Newton steps, t clamped to [0, 1].  Returns r at the crossing.
// (Reference: crossedEquator + bisectRadialCrossing,
// src/physics/event_detection.h:96-191.)
// ---------------------------------------------------------------------------
template <class T>
inline T hermite(T p0, T m0, T p1, T m1, T h, T t) {
  const T t2 = t * t;
  const T t3 = t2 * t;
  const T h00 = fma_(T(2.0), t3, fma_(T(-3.0), t2, T(1.0)));
  const T h10 = fma_(T(-2.0), t2, t3 + t);
  const T h01 = fma_(T(-2.0), t3, T(3.0) * t2);
  const T h11 = t3 - t2;
  return fma_(h00, p0, fma_(h10 * h, m0, fma_(h01, p1, (h11 * h) * m1)));
}
template <class T>
inline T hermite_d(T p0, T m0, T p1, T m1, T h, T t) {
  // d/dt of hermite()
  const T t2 = t * t;
  const T d00 = T(6.0) * (t2 - t);
  const T d10 = fma_(T(3.0), t2, fma_(T(-4.0), t, T(1.0)));
  const T d11 = fma_(T(3.0), t2, T(-2.0) * t);
  return fma_(d00, p0 - p1, fma_(d10 * h, m0, (d11 * h) * m1));
}
template <class T>
inline T crossing_radius(const T* y0, const T* f0, const T* y1, const T* f1, T h) {
  T t = y0[1] / (y0[1] - y1[1]);
  for (int it = 0; it < 3; ++it) {
    const T p = hermite(y0[1], f0[1], y1[1], f1[1], h, t);
    const T dp = hermite_d(y0[1], f0[1], y1[1], f1[1], h, t);
    if (val_(dp) != 0.0) t = t - p / dp;
    if (val_(t) < 0.0) t = T(0.0);
    if (val_(t) > 1.0) t = T(1.0);
  }
  return hermite(y0[0], f0[0], y1[0], f1[0], h, t);
}

// Redshift of a photon (E = 1, L = xi) emitted by gas on the prograde
// circular equatorial geodesic at r: g = 1 / (u^t (1 - Omega xi)) with
// u^t = (r^1.5 + a) / sqrt(r^1.5 (r^1.5 - 3 r^0.5 + 2a)), Omega = 1/(r^1.5 + a),
// i.e. g = sqrt(r^1.5 (r^1.5 - 3 sqrt r + 2a)) / (r^1.5 + a - xi).
// The reference's kerrDiskGFactor (src/physics/iron_kline.h:90-109) is the
// same expression with xi approximated by r sin(phi) sin(i) and the u^t
// factor r^1.5/(r^1.5 + a) dropped (equal at a = 0).
template <class T>
inline T disk_g(T r, T a, T xi) {
  const T sr = sqrt_(r);
  const T r32 = r * sr;
  const T D = fma_(T(-3.0), sr, r32) + T(2.0) * a;
  if (val_(D) <= 0.0) return T(0.0);
  const T den = r32 + a - xi;
  if (val_(den) <= 0.0) return T(0.0);
  return sqrt_(r32 * D) / den;
}

// Linear interpolation in a table of n samples over [x0, x1], clamped.
template <class T>
inline T lut_linear(const double* tab, int n, T x0, T x1, T x) {
  T tpos = (x - x0) / (x1 - x0) * T(static_cast<double>(n - 1));
  if (!(val_(tpos) >= 0.0)) tpos = T(0.0);  // also catches NaN
  if (val_(tpos) > static_cast<double>(n - 1)) tpos = T(static_cast<double>(n - 1));
  const T fl = floor_(tpos);
  int i = static_cast<int>(val_(fl));
  if (i > n - 2) i = n - 2;
  T f = tpos - T(static_cast<double>(i));
  if (val_(f) > 1.0) f = T(1.0);
  const T v0 = T(tab[i]);
  const T v1 = T(tab[i + 1]);
  return fma_(f, v1 - v0, v0);
}

// Observed specific intensity of the thin disk at radius r: a blackbody at
// T(r) = T_max F(r)^(1/4) (F = normalised NT flux from the LUT), seen at
// nu_obs, I_obs = g^3 B_nu(nu_obs/g, T) = (2 h nu^3 / c^2) / expm1(h nu / (g k T))
// (planckFunction, src/physics/rte_integrator.h:221-230; invariance of
// I_nu/nu^3, src/physics/rte_integrator.h:465-475).
template <class T>
inline T disk_intensity(T F, T g, double nu, double T_max) {
  if (val_(F) <= 0.0 || val_(g) <= 0.0) return T(0.0);
  const T Td = T(T_max) * sqrt_(sqrt_(F));
  const T x = T(kH * nu / kKB) / (g * Td);
  if (val_(x) > 700.0) return T(0.0);
  return T(2.0 * kH * nu * nu * nu / (kC * kC)) / expm1_(x);
}

// ---------------------------------------------------------------------------
// Volume: trilinear sample of the (r, theta, phi) grid, fluid redshift,
// thermal synchrotron and one front-to-back formal-solution segment.
// ---------------------------------------------------------------------------
struct VolumeHost {
  const float* cells;  // host float4 grid
  BH_Volume v;
};
template <class T>
inline void volume_fetch(const VolumeHost& V, T r, T u, T phi, T* out4) {
  const BH_Volume& v = V.v;
  const T theta = acos_(u);
  const T xr = (r - T(v.r_min)) / T(v.r_max - v.r_min) * T(double(v.n_r)) - T(0.5);
  const T xt = theta / T(kPi) * T(double(v.n_th)) - T(0.5);
  const T xp = phi / T(kTwoPi) * T(double(v.n_ph)) - T(0.5);
  const T fr0 = floor_(xr), ft0 = floor_(xt), fp0 = floor_(xp);
  const T fr = xr - fr0, ft = xt - ft0, fp = xp - fp0;
  long long ir0 = static_cast<long long>(val_(fr0));
  long long it0 = static_cast<long long>(val_(ft0));
  long long ip0 = static_cast<long long>(val_(fp0));
  long long ir1 = ir0 + 1, it1 = it0 + 1;
  auto clampi = [](long long i, long long n) { return i < 0 ? 0 : (i > n - 1 ? n - 1 : i); };
  ir0 = clampi(ir0, v.n_r);
  ir1 = clampi(ir1, v.n_r);
  it0 = clampi(it0, v.n_th);
  it1 = clampi(it1, v.n_th);
  ip0 = ((ip0 % v.n_ph) + v.n_ph) % v.n_ph;
  const long long ip1 = (ip0 + 1) % v.n_ph;
  const long long ir[2] = {ir0, ir1}, it[2] = {it0, it1}, ip[2] = {ip0, ip1};
  for (int f = 0; f < 4; ++f) {
    T cc[2][2];
    for (int b = 0; b < 2; ++b) {
      for (int a = 0; a < 2; ++a) {
        const long long base = (ip[b] * v.n_th + it[a]) * v.n_r;
        const T v0 = T(double(V.cells[4 * (base + ir[0]) + f]));
        const T v1 = T(double(V.cells[4 * (base + ir[1]) + f]));
        cc[b][a] = fma_(fr, v1 - v0, v0);
      }
    }
    const T c0 = fma_(ft, cc[0][1] - cc[0][0], cc[0][0]);
    const T c1 = fma_(ft, cc[1][1] - cc[1][0], cc[1][0]);
    out4[f] = fma_(fp, c1 - c0, c0);
  }
}

// g = nu_obs / nu_fluid for gas rotating with constant specific angular
// momentum ell = -u_phi/u_t: Omega = -(g_tp + ell g_tt) / (g_pp + ell g_tp),
// u^t = 1/sqrt(-(g_tt + 2 Omega g_tp + Omega^2 g_pp)), g = 1/(u^t (1 - Omega xi)).
// Returns 0 where the orbit would be superluminal.
template <class T>
inline T fluid_g(T r, T u, T a, T xi, T ell) {
  const T s2 = fma_(-u, u, T(1.0));
  const T a2 = a * a;
  const T sig = fma_(a2, u * u, r * r);
  const T two_r_sig = T(2.0) * r / sig;
  const T gtt = two_r_sig - T(1.0);
  const T gtp = -(a * two_r_sig * s2);
  const T gpp = fma_(a2 * two_r_sig, s2, fma_(r, r, a2)) * s2;
  const T Om = -fma_(ell, gtt, gtp) / fma_(ell, gtp, gpp);
  const T nrm = -fma_(Om * Om, gpp, fma_(T(2.0) * Om, gtp, gtt));
  if (val_(nrm) <= 0.0) return T(0.0);
  const T den = fma_(-Om, xi, T(1.0));
  if (val_(den) <= 0.0) return T(0.0);
  return sqrt_(nrm) / den;
}

// Thermal synchrotron emissivity of src/physics/rte_integrator.h:291-328
// (Mahadevan 1996 fit), with K_2(1/Theta) = 2 Theta^2 exp(tab(ln Theta)).
template <class T>
inline T synch_j(T nu, T B, T ne, T th, T k2r) {
  // k2r = K_2(1/Theta) / (2 Theta^2)
  const T nuB = T(kECharge) * B / T(kTwoPi * kMElectron * kC);
  const T nuS = T(1.5) * nuB * th * th;
  const T x = nu / nuS;
  const T x13 = cbrt_(x);
  const T xm12 = T(1.0) / sqrt_(x);
  const T xm14 = sqrt_(xm12);
  const T xm16 = T(1.0) / sqrt_(x13);
  const T im = T(4.0505) * xm16 * fma_(T(0.5316), xm12, fma_(T(0.40), xm14, T(1.0))) *
               exp_(T(-1.8899) * x13);
  const T pref = T(1.7320508075688772 * kECharge * kECharge * kECharge / (kMElectron * kC * kC)) *
                 ne * nuB;
  // (pref / K2) * th^2 * im with K2 = 2 th^2 k2r
  return pref * im / (T(2.0) * k2r);
}

template <class T>
inline T planck(T nu, T tempK) {
  const T x = T(kH / kKB) * nu / tempK;
  if (val_(x) > 700.0) return T(0.0);
  return T(2.0 * kH / (kC * kC)) * nu * nu * nu / expm1_(x);
}

// Emission and absorption per unit affine parameter at a point of the
// volume, in the observer's frame: with the fluid redshift g, nu_f = nu/g,
// alpha_nu = j_nu / B_nu(T_e) (Kirchhoff, rte_integrator.h:348-357) and the
// invariants I/nu^3, j/nu^2, nu alpha:
//   a = alpha_nu(nu_f) L / g      (d tau      = a d lambda_affine)
//   e = g^2 j_nu(nu_f) L          (d I_obs    = e d lambda_affine)
// L = GM/c^2 in cm.  Returns false when the point lies outside the grid or
// the support (no sample taken); a = e = 0 where the gas is empty.
template <class T>
inline bool volume_point(const VolumeHost& V, const RayConst<T>& c, T r, T u, T phi,
                         double nu_obs, double L, T* a, T* e, double faraday_per_ne = 0.0,
                         T* rho = nullptr) {
  const BH_Volume& v = V.v;
  *a = T(0.0);
  *e = T(0.0);
  if (rho) *rho = T(0.0);
  if (!(val_(r) >= v.r_min && val_(r) <= v.r_max && val_(r) >= v.sup_r_lo &&
        val_(r) <= v.sup_r_hi && std::fabs(val_(u)) <= v.sup_u_max))
    return false;
  T f4[4];
  volume_fetch(V, r, u, phi, f4);
  const T ne = f4[0], th = f4[1], B = f4[2], k2r = f4[3];
  if (!(val_(ne) > 0.0 && val_(th) > 0.0 && val_(B) > 0.0 && val_(k2r) > 0.0)) return true;
  const T g = fluid_g(r, u, c.a, c.xi, T(v.ell));
  if (!(val_(g) > 0.0)) return true;
  if (rho) *rho = T(faraday_per_ne) * ne * T(L) / g;  // Faraday rate per affine parameter
  const T nu_f = T(nu_obs) / g;
  const T tempK = th * T(kMElectron * kC * kC / kKB);
  const T Bnu = planck(nu_f, tempK);
  if (!(val_(Bnu) > 0.0)) return true;
  const T jL = synch_j(nu_f, B, ne, th, k2r) * T(L);
  *a = jL / (g * Bnu);
  *e = g * g * jL;
  return true;
}

// d lambda_affine / ds = Sigma / r.
template <class T>
inline T affine_rate(const RayConst<T>& c, T r, T u) {
  return fma_(c.a2, u * u, r * r) / r;
}

// Front-to-back formal solution over one step of length h between two
// samples (a0, e0, s0) -> (a1, e1, s1), trapezoidal in the affine parameter:
//   d lambda = h (s0 + s1)/2,  d tau = (a0 + a1) d lambda / 2,
//   S = (e0 + e1)/(a0 + a1),   I += T S (1 - e^-dtau),  T *= e^-dtau.
// Second-order accurate (the reference's d_rte_step, device_physics.cuh:
// 1988-2015, is the first-order right-endpoint form of the same update).
template <class T>
struct RtAcc {
  T I, Tr, tau, a0, e0, s0;
};
template <class T>
inline void rt_segment(RtAcc<T>& q, T h, T a1, T e1, T s1) {
  const T asum = q.a0 + a1;
  const T esum = q.e0 + e1;
  if (val_(asum) > 0.0 || val_(esum) > 0.0) {
    const T dl = T(0.5) * h * (q.s0 + s1);
    if (val_(asum) > 0.0) {
      const T dtau = T(0.5) * asum * dl;
      const T om = -expm1_(-dtau);
      q.I = fma_(q.Tr * (esum / asum), om, q.I);
      q.Tr = q.Tr * (T(1.0) - om);
      q.tau = q.tau + dtau;
    } else {
      q.I = fma_(q.Tr, T(0.5) * esum * dl, q.I);
    }
  }
  q.a0 = a1;
  q.e0 = e1;
  q.s0 = s1;
}

// Polarised transfer beside the scalar one (simplified Mueller matrix of
// stokesStep, src/physics/stokes_transport.h:260-369): the observed (Q, U)
// gains, per step, the source (j_Q + i j_U) = -frac e e^{2 i chi} integrated
// exactly under absorption A and Faraday rate R within the step,
// (1 - e^{-w}) / z with z = A - i R, w = z dl, rotated by the Faraday angle
// Phi accumulated between the step and the observer and scaled by the
// transmittance Tr before the step (called before rt_segment).
struct PolAcc {
  double q, u, c, s, r0;
};

inline void pol_segment(PolAcc& p, const RtAcc<double>& q, double h, double a1, double e1, double r1,
                        double s1, double c2, double s2, double frac) {
  const double esum = q.e0 + e1;
  const double rsum = p.r0 + r1;
  if (esum > 0.0 || rsum != 0.0) {
    const double dl = 0.5 * h * (q.s0 + s1);
    const double A = 0.5 * (q.a0 + a1), R = 0.5 * rsum, J = 0.5 * esum;
    const double x = A * dl, z = R * dl;
    const double sz = std::sin(z), cz = std::cos(z);
    double ic, is;
    if (x * x + z * z < 1e-6) {
      ic = dl * (1.0 - 0.5 * x + (x * x - z * z) * (1.0 / 6.0));
      is = dl * (0.5 * z - x * z * (1.0 / 3.0));
    } else {
      const double E = std::exp(-x);
      const double D = A * A + R * R;
      const double omc = 1.0 - E * cz;
      ic = (A * omc + R * E * sz) / D;
      is = (R * omc - A * E * sz) / D;
    }
    const double jq = -frac * J * c2, ju = -frac * J * s2;
    const double qe = jq * ic - ju * is, ue = jq * is + ju * ic;
    p.q = std::fma(q.Tr, p.c * qe - p.s * ue, p.q);
    p.u = std::fma(q.Tr, p.s * qe + p.c * ue, p.u);
    const double nc = p.c * cz - p.s * sz, ns = p.s * cz + p.c * sz;
    p.c = nc;
    p.s = ns;
  }
  p.r0 = r1;
}

}  // namespace orc
