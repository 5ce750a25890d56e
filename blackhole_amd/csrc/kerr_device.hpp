// kerr_device.hpp - FP64 device math of the Kerr null-geodesic tracer (gfx950).
//
// Null geodesics with E = 1, L = xi, Carter constant Q = eta, integrated in
// Mino time with u = cos(theta) and the step variable s, d(lambda) = ds / r:
//   (dr/dl)^2 = R(r) = P^2 - Delta K,  P = r^2 + a^2 - a xi,  K = eta + (xi - a)^2
//   (du/dl)^2 = U(u) = eta + (a^2 - eta - xi^2) u^2 - a^2 u^4
//   d phi/dl  = a P / Delta - a + xi / (1 - u^2)
// in second-order form r'' = R'/2, u'' = U'/2 (state r, u, phi, v_r, v_u).
// Potentials as src/physics/kerr.cpp:69-97 (R, R') with the Carter-constant
// Theta of src/physics/verified/geodesic.hpp:178-188; phi rate as
// kerrMinoDerivatives (src/physics/kerr.cpp:40-66).
//
// Every stepping operation is +, -, *, /, sqrt or fma (IEEE, correctly
// rounded on gfx950 and on the host), in a fixed order, so the FP64 path is
// reproducible bit for bit by any implementation performing the same
// operations; the file must be compiled with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bhk_trace.h"

namespace bhk {

constexpr int kN = 5;  // r, u, phi, v_r, v_u

constexpr double kC = 2.99792458e10;
constexpr double kHbar = 1.054571817e-27;
constexpr double kKB = 1.380649e-16;
constexpr double kECharge = 4.80320425e-10;
constexpr double kMElectron = 9.1093837015e-28;
constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kH = 2.0 * kPi * kHbar;
constexpr double kProjKappa = 0.0625;

// Scalar helpers: the FP64 instantiation uses correctly rounded operations
// (parity with the oracle); the FP32 instantiation uses the hardware
// reciprocal and square root (v_rcp_f32, v_sqrt_f32: about 1 ulp), which the
// precision form's tolerance admits.  (__fdividef compiles to the full
// IEEE division sequence here, ten instructions.)
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double divT(double a, double b) { return a / b; }
__device__ __forceinline__ float divT(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double sqrtT(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float sqrtT(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double fabsT(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float fabsT(float x) { return __builtin_fabsf(x); }

template <class T>
struct RayCT {
  T a, a2, xi, eta, K, c1, c2, two_a2;
};
using RayC = RayCT<double>;

template <class T>
__device__ __forceinline__ RayCT<T> convert_const(const RayC& c) {
  RayCT<T> o;
  o.a = static_cast<T>(c.a);
  o.a2 = static_cast<T>(c.a2);
  o.xi = static_cast<T>(c.xi);
  o.eta = static_cast<T>(c.eta);
  o.K = static_cast<T>(c.K);
  o.c1 = static_cast<T>(c.c1);
  o.c2 = static_cast<T>(c.c2);
  o.two_a2 = static_cast<T>(c.two_a2);
  return o;
}

// Per-camera constants, computed on the host (one libm sin/cos/cbrt per
// camera) and read by every lane that traces a pixel of the camera.
struct CamDev {
  double a, sin_o, cos_o, r_obs, r_stop, r_isco, r_plus;
  double alpha_min, dalpha, beta_max, dbeta;
  int64_t w, h, offset;
  int32_t lut_index;  // disk LUT row
  int32_t vol_index;  // volume of a volume set (multi-volume launches)
};

__device__ __forceinline__ RayC make_const(double a, double xi, double eta) {
  RayC c;
  c.a = a;
  c.a2 = a * a;
  c.xi = xi;
  c.eta = eta;
  const double xa = xi - a;
  c.K = eta + xa * xa;
  c.c1 = c.a2 - a * xi;
  c.c2 = (c.a2 - eta) - xi * xi;
  c.two_a2 = 2.0 * c.a2;
  return c;
}

template <class T>
__device__ __forceinline__ void rhs(const RayCT<T>& c, const T* y, T* f) {
  const T r = y[0];
  const T u = y[1];
  const T r2 = r * r;
  const T P = r2 + c.c1;
  const T Delta = fmaT(r - T(2.0), r, c.a2);
  T w = fmaT(-u, u, T(1.0));
  if (w < T(1e-14)) w = T(1e-14);
  const T q = divT(T(1.0), r * Delta * w);
  const T inv_r = Delta * w * q;
  f[0] = y[3] * inv_r;
  f[1] = y[4] * inv_r;
  f[2] = fmaT(c.xi, Delta, c.a * w * (P - Delta)) * q;
  f[3] = fmaT(T(2.0) * r, P, -((r - T(1.0)) * c.K)) * inv_r;
  f[4] = u * fmaT(-c.two_a2, u * u, c.c2) * inv_r;
}

template <class T>
__device__ __forceinline__ T radial_potential(const RayCT<T>& c, T r) {
  const T P = r * r + c.c1;
  const T Delta = fmaT(r - T(2.0), r, c.a2);
  return fmaT(P, P, -(Delta * c.K));
}
template <class T>
__device__ __forceinline__ T polar_potential(const RayCT<T>& c, T u) {
  const T u2 = u * u;
  return fmaT(fmaT(-c.a2, u2, c.c2), u2, c.eta);
}

// Radial null-constraint projection where R is far from zero (the
// reference's null renormalisation, src/physics/verified/null_constraint.hpp:
// 208-271, restricted to the radial first integral).
template <class T>
__device__ __forceinline__ void project_radial(const RayCT<T>& c, T* y, T* f) {
  const T R = radial_potential(c, y[0]);
  const T s = fmaT(y[0], y[0], c.a2);
  if (R > T(kProjKappa) * (s * s)) {
    const T vn = sqrtT(R);
    const T v = (y[3] < T(0.0)) ? -vn : vn;
    f[0] = f[0] * divT(v, y[3]);
    y[3] = v;
  }
}

// Image plane -> constants and initial state (Bardeen 1973).
__device__ __forceinline__ RayC setup_ray_ab(const CamDev& k, double alpha, double beta, double* y) {
  const double a = k.a;
  const double xi = -(alpha * k.sin_o);
  const double eta = __builtin_fma(beta, beta, (alpha * alpha - a * a) * (k.cos_o * k.cos_o));
  const RayC c = make_const(a, xi, eta);
  y[0] = k.r_obs;
  y[1] = k.cos_o;
  y[2] = 0.0;
  double R = radial_potential(c, y[0]);
  if (R < 0.0) R = 0.0;
  y[3] = -__builtin_sqrt(R);
  y[4] = k.sin_o * beta;
  return c;
}
__device__ __forceinline__ RayC setup_ray(const CamDev& k, int64_t px, int64_t py, double* y) {
  const double alpha = k.alpha_min + (static_cast<double>(px) + 0.5) * k.dalpha;
  const double beta = k.beta_max - (static_cast<double>(py) + 0.5) * k.dbeta;
  return setup_ray_ab(k, alpha, beta, y);
}

// Pinhole camera at a finite position (the render-launch camera): position
// in Boyer-Lindquist coordinates and the flat spherical basis there (host
// values, so the device and the oracle start from identical bits), the
// column-major camera basis [right, up, forward], and the pixel mapping.
struct PinholeDev {
  double r, cos_t, sin_t, phi;
  double e_r[3], e_t[3], e_p[3];
  double basis[9];
  double fov, aspect, shift_x, shift_y, inv_w, inv_h;
};

// Constants and initial state of the ray through pixel centre (xc, yc) =
// (px + 0.5, py + 0.5) of a pinhole camera: the pixel direction as
// d_ray_dir maps it (src/cuda/device_physics.cuh:1953-1967), its projections
// on the flat spherical basis taken as the coordinate components k^r,
// k^theta, k^phi, k^t from the null condition of the Kerr metric (M = 1),
// E = -(g_tt k^t + g_tphi k^phi), L = g_tphi k^t + g_phiphi k^phi and
// p_theta = Sigma k^theta (d_kerr_init_consts, device_physics.cuh:316-392),
// normalised to E = 1: xi = L/E, eta = (p_theta/E)^2 - a^2 cos^2 + xi^2 cot^2.
// The Mino velocities take their magnitudes from the potentials, sqrt(R) and
// sqrt(U), and their signs from k^r and -k^theta.  The camera must be outside
// r = 2 (outside every ergosphere), so k^t is the root (hb + sqrt(disc)) / -g_tt.
__device__ __forceinline__ RayC setup_pinhole(const PinholeDev& p, double a, double xc, double yc, double* y) {
  double u = (2.0 * xc * p.inv_w - 1.0) * p.fov;
  u = u * p.aspect + p.shift_x;
  const double v = (2.0 * yc * p.inv_h - 1.0) * p.fov + p.shift_y;
  const double lx = -u, ly = -v, lz = 1.0;
  const double inv_n = 1.0 / __builtin_sqrt(lx * lx + ly * ly + lz * lz);
  double d[3];
#pragma unroll
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
  const double kt = (hb + __builtin_sqrt(disc)) / (-g_tt);
  const double E = -(g_tt * kt + g_tp * kph);
  const double L = g_tp * kt + g_pp * kph;
  const double xi = L / E;
  const double pth = sigma * kth / E;
  const double eta = pth * pth - a2 * cos2 + xi * xi * cos2 / sin2;
  const RayC c = make_const(a, xi, eta);
  y[0] = r;
  y[1] = p.cos_t;
  y[2] = p.phi;
  double R = radial_potential(c, r);
  double U = polar_potential(c, p.cos_t);
  if (R < 0.0) R = 0.0;
  if (U < 0.0) U = 0.0;
  y[3] = kr >= 0.0 ? __builtin_sqrt(R) : -__builtin_sqrt(R);
  y[4] = kth >= 0.0 ? -__builtin_sqrt(U) : __builtin_sqrt(U);
  return c;
}

// Classic RK4 step from y with k1 = f(y).
template <class T>
__device__ __forceinline__ void rk4_step(const RayCT<T>& c, const T* y, const T* k1, T h, T* yn) {
  T k2[kN], k3[kN], k4[kN], yt[kN];
  const T hh = T(0.5) * h;
  const T h6 = divT(h, T(6.0));
#pragma unroll
  for (int i = 0; i < kN; ++i) yt[i] = fmaT(hh, k1[i], y[i]);
  rhs(c, yt, k2);
#pragma unroll
  for (int i = 0; i < kN; ++i) yt[i] = fmaT(hh, k2[i], y[i]);
  rhs(c, yt, k3);
#pragma unroll
  for (int i = 0; i < kN; ++i) yt[i] = fmaT(h, k3[i], y[i]);
  rhs(c, yt, k4);
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    const T s = fmaT(T(2.0), k2[i] + k3[i], k1[i] + k4[i]);
    yn[i] = fmaT(h6, s, y[i]);
  }
}

// Dormand & Prince (1980) 5(4) pair (tableau of src/physics/verified/rk4.h:
// 193-244), FSAL, RMS error over the 5 components scaled by
// atol + rtol max(|y|, |y5|).
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
constexpr double e1 = 71.0 / 57600.0, e3 = -71.0 / 16695.0, e4 = 71.0 / 1920.0,
                 e5 = -17253.0 / 339200.0, e6 = 22.0 / 525.0, e7 = -1.0 / 40.0;
}  // namespace dp

template <class T>
__device__ __forceinline__ T dp45_attempt(const RayCT<T>& c, const T* y, const T* k1, T h, T rtol,
                                          T atol, T* y5, T* k7) {
  using namespace dp;
  T k2[kN], k3[kN], k4[kN], k5[kN], k6[kN], yt[kN];
#pragma unroll
  for (int i = 0; i < kN; ++i) yt[i] = fmaT(h, T(a21) * k1[i], y[i]);
  rhs(c, yt, k2);
#pragma unroll
  for (int i = 0; i < kN; ++i) yt[i] = fmaT(h, fmaT(T(a32), k2[i], T(a31) * k1[i]), y[i]);
  rhs(c, yt, k3);
#pragma unroll
  for (int i = 0; i < kN; ++i)
    yt[i] = fmaT(h, fmaT(T(a43), k3[i], fmaT(T(a42), k2[i], T(a41) * k1[i])), y[i]);
  rhs(c, yt, k4);
#pragma unroll
  for (int i = 0; i < kN; ++i)
    yt[i] = fmaT(h, fmaT(T(a54), k4[i], fmaT(T(a53), k3[i], fmaT(T(a52), k2[i], T(a51) * k1[i]))),
                 y[i]);
  rhs(c, yt, k5);
#pragma unroll
  for (int i = 0; i < kN; ++i)
    yt[i] = fmaT(h,
                 fmaT(T(a65), k5[i],
                      fmaT(T(a64), k4[i], fmaT(T(a63), k3[i], fmaT(T(a62), k2[i], T(a61) * k1[i])))),
                 y[i]);
  rhs(c, yt, k6);
#pragma unroll
  for (int i = 0; i < kN; ++i)
    y5[i] = fmaT(h,
                 fmaT(T(b6), k6[i],
                      fmaT(T(b5), k5[i], fmaT(T(b4), k4[i], fmaT(T(b3), k3[i], T(b1) * k1[i])))),
                 y[i]);
  rhs(c, y5, k7);
  T acc = T(0.0);
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    const T e = h * fmaT(T(e7), k7[i],
                         fmaT(T(e6), k6[i],
                              fmaT(T(e5), k5[i], fmaT(T(e4), k4[i], fmaT(T(e3), k3[i], T(e1) * k1[i])))));
    const T ay = fabsT(y[i]);
    const T ay5 = fabsT(y5[i]);
    const T sc = fmaT(rtol, (ay > ay5) ? ay : ay5, atol);
    const T qq = divT(e, sc);
    acc = fmaT(qq, qq, acc);
  }
  return sqrtT(acc * T(0.2));
}

// x^(-1/5) from exact exponent split + two Newton steps (bit-reproducible).
__device__ __forceinline__ double pow_m0p2(double x) {
  int e = 0;
  const double m = __builtin_frexp(x, &e);
  const int qd = (e >= 0) ? (e / 5) : -((-e + 4) / 5);
  const int rr = e - 5 * qd;
  const double t2 = (rr == 0) ? 1.0
                  : (rr == 1) ? 0.87055056329612412
                  : (rr == 2) ? 0.75785828325519911
                  : (rr == 3) ? 0.65975395538644721
                              : 0.57434917749851755;
  double y = __builtin_fma(-0.29739670999407, m, 1.29739670999407);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double y2 = y * y;
    const double y5 = y2 * y2 * y;
    y = y * __builtin_fma(-0.2 * m, y5, 1.2);
  }
  return __builtin_ldexp(y * t2, -qd);
}

__device__ __forceinline__ double dp45_factor(double err, bool accepted, bool last_rejected) {
  if (accepted) {
    const double e = (err < 1e-10) ? 1e-10 : err;
    double fac = 0.9 * pow_m0p2(e);
    if (fac > 5.0) fac = 5.0;
    if (last_rejected && fac > 1.0) fac = 1.0;
    return fac;
  }
  double fac = 0.9 * pow_m0p2(err);
  if (fac < 0.2) fac = 0.2;
  return fac;
}

// Cubic Hermite in s between two accepted states, for the equatorial
// crossing (src/physics/event_detection.h:96-191 locates it by bisection).
__device__ __forceinline__ double hermite(double p0, double m0, double p1, double m1, double h,
                                          double t) {
  const double t2 = t * t;
  const double t3 = t2 * t;
  const double h00 = __builtin_fma(2.0, t3, __builtin_fma(-3.0, t2, 1.0));
  const double h10 = __builtin_fma(-2.0, t2, t3 + t);
  const double h01 = __builtin_fma(-2.0, t3, 3.0 * t2);
  const double h11 = t3 - t2;
  return __builtin_fma(h00, p0, __builtin_fma(h10 * h, m0, __builtin_fma(h01, p1, (h11 * h) * m1)));
}
__device__ __forceinline__ double hermite_d(double p0, double m0, double p1, double m1, double h,
                                            double t) {
  const double t2 = t * t;
  const double d00 = 6.0 * (t2 - t);
  const double d10 = __builtin_fma(3.0, t2, __builtin_fma(-4.0, t, 1.0));
  const double d11 = __builtin_fma(3.0, t2, -2.0 * t);
  return __builtin_fma(d00, p0 - p1, __builtin_fma(d10 * h, m0, (d11 * h) * m1));
}
template <class T>
__device__ __forceinline__ double crossing_radius(const T* y0, const T* f0, const T* y1,
                                                  const T* f1, T h) {
  double t = static_cast<double>(y0[1]) / (static_cast<double>(y0[1]) - static_cast<double>(y1[1]));
  const double p0 = y0[1], m0 = f0[1], p1 = y1[1], m1 = f1[1], hd = h;
#pragma unroll
  for (int it = 0; it < 3; ++it) {
    const double p = hermite(p0, m0, p1, m1, hd, t);
    const double dp = hermite_d(p0, m0, p1, m1, hd, t);
    if (dp != 0.0) t = t - p / dp;
    if (t < 0.0) t = 0.0;
    if (t > 1.0) t = 1.0;
  }
  return hermite(static_cast<double>(y0[0]), static_cast<double>(f0[0]), static_cast<double>(y1[0]),
                 static_cast<double>(f1[0]), hd, t);
}

// g of a prograde Keplerian emitter at radius r for a photon with L = xi
// (exact form of src/physics/iron_kline.h:90-109).
__device__ __forceinline__ double disk_g(double r, double a, double xi) {
  const double sr = __builtin_sqrt(r);
  const double r32 = r * sr;
  const double D = __builtin_fma(-3.0, sr, r32) + 2.0 * a;
  if (D <= 0.0) return 0.0;
  const double den = r32 + a - xi;
  if (den <= 0.0) return 0.0;
  return __builtin_sqrt(r32 * D) / den;
}

__device__ __forceinline__ double lut_linear(const double* tab, int n, double x0, double x1,
                                             double x) {
  double tpos = (x - x0) / (x1 - x0) * static_cast<double>(n - 1);
  if (!(tpos >= 0.0)) tpos = 0.0;  // also catches NaN
  if (tpos > static_cast<double>(n - 1)) tpos = static_cast<double>(n - 1);
  const double fl = floor(tpos);
  int i = static_cast<int>(fl);
  if (i > n - 2) i = n - 2;
  double f = tpos - static_cast<double>(i);
  if (f > 1.0) f = 1.0;
  const double v0 = tab[i];
  const double v1 = tab[i + 1];
  return __builtin_fma(f, v1 - v0, v0);
}

// Observed blackbody intensity of the disk (planckFunction,
// src/physics/rte_integrator.h:221-230, with I/nu^3 invariance).
__device__ __forceinline__ double disk_intensity(double F, double g, double nu, double T_max) {
  if (F <= 0.0 || g <= 0.0) return 0.0;
  const double Td = T_max * __builtin_sqrt(__builtin_sqrt(F));
  const double x = (kH * nu / kKB) / (g * Td);
  if (x > 700.0) return 0.0;
  return (2.0 * kH * nu * nu * nu / (kC * kC)) / expm1(x);
}

// ---- volume ---------------------------------------------------------------
//
// The stepping above is reproducible bit for bit; the volume emission below
// evaluates the same formulas as the FP64 oracle (trilinear fetch, fluid
// redshift, thermal synchrotron j, Kirchhoff absorption) to a few units in
// the last place per sample, not bit for bit: the divisions after the fluid
// redshift share one reciprocal, the cell coordinates are scaled by
// reciprocals, acos and the fractional powers of the synchrotron fit have
// their own evaluations below, and expm1 of a small argument is its Taylor
// polynomial.  The fluid redshift, a difference of large terms near the
// light cylinder, keeps the oracle's operations.

// Global address space for device loads known to hit device memory (the
// compiler emits global rather than flat loads); the host pass only parses
// device code, where the qualifier does not apply.
#if defined(__HIP_DEVICE_COMPILE__)
#define BHK_GLOBAL __attribute__((address_space(1)))
#else
#define BHK_GLOBAL
#endif

struct VolDev {
  const float4* cells;
  int n_r, n_th, n_ph;
  double r_min, r_max, r_span, sup_r_lo, sup_r_hi, sup_u_max, ell;
  double s_r, s_th, s_ph, inv_nph;  // n_r / r_span, n_th / pi, n_ph / 2 pi, 1 / n_ph
};

// Radiative constants of a launch at observing frequency nu and length unit L:
//   x_T = CT / (g Theta_e)            (h nu_f / k T_e, nu_f = nu / g)
//   x_M = Cx / (g |B| Theta_e^2)      (nu_f / nu_s, nu_s = 1.5 nu_B Theta_e^2)
//   j L = CjL n_e |B| I_M(x_M) / k2r  (k2r = K_2(1/Theta_e) / (2 Theta_e^2))
//   B_nu(T_e) at nu_f = SB / (g^3 expm1(x_T)),  iSB = 1 / SB
//   Faraday rate = FL n_e / g
struct RadDev {
  double CT, Cx, CjL, SB, iSB, FL;
};

// acos(u), |u| <= 1: acos = pi/2 - asin(u) for |u| <= 1/2, else
// 2 asin(sqrt((1 - |u|) / 2)) reflected for u < 0, with asin(x) = x + x^3 P(x^2)
// on |x| <= 1/2.  P is a degree-11 least-squares Chebyshev fit of the Taylor
// series of (asin(x) - x) / x^3 in x^2 on [0, 1/4]; the result is within
// 4.3e-16 relative of the correctly rounded acos over [-1, 1].
__device__ __forceinline__ double acos_poly(double u) {
  const double au = __builtin_fabs(u);
  const bool big = au > 0.5;
  const double w = (1.0 - au) * 0.5;
  const double z = big ? w : u * u;
  // sqrt(w) for w in [0, 1/4] (never subnormal unless 0): hardware reciprocal
  // square root refined by two Newton-Raphson steps on the root and the half
  // reciprocal (Goldschmidt form), within 1 ulp.
  const double y0 = __builtin_amdgcn_rsq(w);
  double sq = w * y0, hf = 0.5 * y0;
  const double rr = __builtin_fma(-sq, hf, 0.5);
  sq = __builtin_fma(sq, rr, sq);
  hf = __builtin_fma(hf, rr, hf);
  sq = __builtin_fma(__builtin_fma(-sq, sq, w), hf, sq);
  const double x = big ? ((w > 0.0) ? sq : 0.0) : au;
  double p = 0x1.ce1deb71b173dp-6;
  p = __builtin_fma(p, z, -0x1.61a52073fdd05p-7);
  p = __builtin_fma(p, z, 0x1.0712b4f81819bp-6);
  p = __builtin_fma(p, z, 0x1.ff07c6f495617p-8);
  p = __builtin_fma(p, z, 0x1.8527efdaa7650p-7);
  p = __builtin_fma(p, z, 0x1.c8726e298da99p-7);
  p = __builtin_fma(p, z, 0x1.1c59318cbcadfp-6);
  p = __builtin_fma(p, z, 0x1.6e8b2ca05965dp-6);
  p = __builtin_fma(p, z, 0x1.f1c71f7fec096p-6);
  p = __builtin_fma(p, z, 0x1.6db6db689ba28p-5);
  p = __builtin_fma(p, z, 0x1.3333333336569p-4);
  p = __builtin_fma(p, z, 0x1.5555555555555p-3);
  const double as = __builtin_fma(x * z, p, x);
  if (big) return (u > 0.0) ? 2.0 * as : __builtin_fma(-2.0, as, kPi);
  return 0.5 * kPi - ((u < 0.0) ? -as : as);
}

// expm1(x) for x >= 0: x (1 + x/2 + x^2/6 + x^3/24 + x^4/120) below 1e-3
// (truncation x^6 / 720 < 1.4e-18 x), the library function above.
__device__ __forceinline__ double expm1_pos(double x) {
  if (x < 1e-3) {
    const double q = __builtin_fma(__builtin_fma(__builtin_fma(__builtin_fma(x, 1.0 / 120.0, 1.0 / 24.0), x,
                                                               1.0 / 6.0), x, 0.5), x, 1.0);
    return x * q;
  }
  return expm1(x);
}

// Cell-centred trilinear sample of the cell fields {n_e, Theta_e, |B|,
// K_2(1/Theta_e)/(2 Theta_e^2)} (one float4 per corner): r and theta clamp,
// phi wraps (d_sample_grmhd, src/cuda/device_physics.cuh:1775-1829).
__device__ __forceinline__ void volume_fetch(const VolDev& v, double r, double u, double phi, double* out4) {
  // the grid is in device global memory: global (not flat) loads
  const BHK_GLOBAL float4* __restrict__ cells = (const BHK_GLOBAL float4*)(v.cells);
  const int n_r = v.n_r, n_th = v.n_th, n_ph = v.n_ph;
  const double xr = (r - v.r_min) * v.s_r - 0.5;
  const double xt = acos_poly(u) * v.s_th - 0.5;
  const double xp = phi * v.s_ph - 0.5;
  const double fr0 = floor(xr), ft0 = floor(xt), fp0 = floor(xp);
  const double fr = xr - fr0, ft = xt - ft0, fp = xp - fp0;
  // 32-bit index arithmetic (grids are far below 2^31 cells); phi wraps by a
  // floating-point modulo whose quotient may be off by one, then an integer fix-up.
  int ir0 = static_cast<int>(fr0);
  int it0 = static_cast<int>(ft0);
  const double npd = static_cast<double>(n_ph);
  int ip0 = static_cast<int>(fp0 - floor(fp0 * v.inv_nph) * npd);
  ip0 = (ip0 >= n_ph) ? ip0 - n_ph : ip0;
  ip0 = (ip0 < 0) ? ip0 + n_ph : ip0;
  int ir1 = ir0 + 1, it1 = it0 + 1;
  ir0 = ir0 < 0 ? 0 : (ir0 > n_r - 1 ? n_r - 1 : ir0);
  ir1 = ir1 < 0 ? 0 : (ir1 > n_r - 1 ? n_r - 1 : ir1);
  it0 = it0 < 0 ? 0 : (it0 > n_th - 1 ? n_th - 1 : it0);
  it1 = it1 < 0 ? 0 : (it1 > n_th - 1 ? n_th - 1 : it1);
  ip0 = ip0 < 0 ? 0 : (ip0 > n_ph - 1 ? n_ph - 1 : ip0);
  const int ip1 = (ip0 + 1 == n_ph) ? 0 : ip0 + 1;
  const unsigned b00 = static_cast<unsigned>((ip0 * n_th + it0) * n_r);
  const unsigned b01 = static_cast<unsigned>((ip0 * n_th + it1) * n_r);
  const unsigned b10 = static_cast<unsigned>((ip1 * n_th + it0) * n_r);
  const unsigned b11 = static_cast<unsigned>((ip1 * n_th + it1) * n_r);
  const float4 q000 = cells[b00 + ir0], q001 = cells[b00 + ir1];
  const float4 q010 = cells[b01 + ir0], q011 = cells[b01 + ir1];
  const float4 q100 = cells[b10 + ir0], q101 = cells[b10 + ir1];
  const float4 q110 = cells[b11 + ir0], q111 = cells[b11 + ir1];
#define BHK_TRI(F_, OUT_)                                                                     \
  {                                                                                         \
    const double c00 = __builtin_fma(fr, (double)q001.F_ - (double)q000.F_, (double)q000.F_); \
    const double c01 = __builtin_fma(fr, (double)q011.F_ - (double)q010.F_, (double)q010.F_); \
    const double c10 = __builtin_fma(fr, (double)q101.F_ - (double)q100.F_, (double)q100.F_); \
    const double c11 = __builtin_fma(fr, (double)q111.F_ - (double)q110.F_, (double)q110.F_); \
    const double c0 = __builtin_fma(ft, c01 - c00, c00);                                     \
    const double c1 = __builtin_fma(ft, c11 - c10, c10);                                     \
    OUT_ = __builtin_fma(fp, c1 - c0, c0);                                                   \
  }
  BHK_TRI(x, out4[0])
  BHK_TRI(y, out4[1])
  BHK_TRI(z, out4[2])
  BHK_TRI(w, out4[3])
#undef BHK_TRI
}

// Redshift of gas rotating with constant ell = -u_phi/u_t as the numerator
// sqrt(-(g_tt + 2 Omega g_tphi + Omega^2 g_phiphi)) and the denominator
// 1 - Omega xi of g (false when g is not positive), Omega = -(ell g_tt +
// g_tphi) / (ell g_tphi + g_phiphi).  The norm is a difference of large
// terms near the light cylinder, so it keeps the oracle's order of
// operations.
__device__ __forceinline__ bool fluid_g_ratio(double r, double u, double a, double xi, double ell, double* S,
                                              double* D) {
  const double s2 = __builtin_fma(-u, u, 1.0);
  const double a2 = a * a;
  const double sig = __builtin_fma(a2, u * u, r * r);
  const double two_r_sig = 2.0 * r / sig;
  const double gtt = two_r_sig - 1.0;
  const double gtp = -(a * two_r_sig * s2);
  const double gpp = __builtin_fma(a2 * two_r_sig, s2, __builtin_fma(r, r, a2)) * s2;
  const double Om = -__builtin_fma(ell, gtt, gtp) / __builtin_fma(ell, gtp, gpp);
  const double nrm = -__builtin_fma(Om * Om, gpp, __builtin_fma(2.0 * Om, gtp, gtt));
  const double den = __builtin_fma(-Om, xi, 1.0);
  if (!(nrm > 0.0) || !(den > 0.0)) return false;
  *S = __builtin_sqrt(nrm);
  *D = den;
  return true;
}

// Redshift of gas rotating with constant ell (the value of fluid_g_ratio).
__device__ __forceinline__ double fluid_g(double r, double u, double a, double xi, double ell) {
  double S, D;
  return fluid_g_ratio(r, u, a, xi, ell, &S, &D) ? S / D : 0.0;
}

// Thermal synchrotron fit I_M(x) = 4.0505 x^(-1/6) (1 + 0.40 x^(-1/4) +
// 0.5316 x^(-1/2)) exp(-1.8899 x^(1/3)) (src/physics/rte_integrator.h:291-328).
// For 1e-30 < x < 1e30 every power comes from y = x^(-1/12): an FP32
// estimate 2^(-log2(x)/12) refined by two Newton steps of y^-12 = x
// (y <- y (13 - x y^12) / 12, quadratic from ~1e-7), then x^(-1/6) = y^2,
// x^(-1/4) = y^3, x^(-1/2) = y^6, x^(1/3) = x y^8.
__device__ __forceinline__ double synch_im(double x) {
  if (x > 1e-30 && x < 1e30) {
    const float xf = static_cast<float>(x);
    double y = static_cast<double>(__builtin_amdgcn_exp2f(__builtin_amdgcn_logf(xf) * (-1.0f / 12.0f)));
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const double y2 = y * y;
      const double y4 = y2 * y2;
      const double y12 = y4 * y4 * y4;
      y = y * (__builtin_fma(-x, y12, 13.0) * (1.0 / 12.0));
    }
    const double y2 = y * y;
    const double y3 = y2 * y;
    const double y6 = y3 * y3;
    const double x13 = x * (y6 * y2);
    return 4.0505 * y2 * __builtin_fma(0.5316, y6, __builtin_fma(0.40, y3, 1.0)) * exp(-1.8899 * x13);
  }
  const double x13 = cbrt(x);
  const double xm12 = 1.0 / __builtin_sqrt(x);
  const double xm14 = __builtin_sqrt(xm12);
  const double xm16 = 1.0 / __builtin_sqrt(x13);
  return 4.0505 * xm16 * __builtin_fma(0.5316, xm12, __builtin_fma(0.40, xm14, 1.0)) * exp(-1.8899 * x13);
}

// Emission / absorption per unit affine parameter in the observer's frame
// at a volume point: nu_f = nu/g, alpha = j/B_nu(T_e) (Kirchhoff,
// rte_integrator.h:348-357), a = alpha(nu_f) L / g, e = g^2 j(nu_f) L.
// rho (optional): the Faraday rotation rate per unit affine parameter,
// faraday_per_ne n_e L / g (rad; zero outside the support).
// The sample's remaining divisions share one reciprocal: with g = S / D
// (fluid_g_ratio), T = Theta_e^2 |B| and Q = S D T k2r,
//   g = S^2 T k2r / Q,  1/(g T) = D^2 k2r / Q,  1/g = T / (g T),  1/k2r = S D T / Q.
__device__ __forceinline__ bool volume_point(const VolDev& v, const RayC& c, double r, double u, double phi,
                                             const RadDev& K, double* a, double* e, double* rho = nullptr) {
  *a = 0.0;
  *e = 0.0;
  if (rho) *rho = 0.0;
  if (!(r >= v.r_min && r <= v.r_max && r >= v.sup_r_lo && r <= v.sup_r_hi &&
        __builtin_fabs(u) <= v.sup_u_max))
    return false;
  double f4[4];
  volume_fetch(v, r, u, phi, f4);
  const double ne = f4[0], th = f4[1], B = f4[2], k2r = f4[3];
  if (!(ne > 0.0 && th > 0.0 && B > 0.0 && k2r > 0.0)) return true;
  double S, D;
  if (!fluid_g_ratio(r, u, c.a, c.xi, v.ell, &S, &D)) return true;
  const double T = th * th * B;
  const double SD = S * D;
  const double rQ = 1.0 / (SD * (T * k2r));
  const double inv_gT = (D * D) * k2r * rQ;  // 1 / (g Theta^2 |B|)
  const double inv_g = inv_gT * T;
  const double g = S * S * (T * k2r) * rQ;   // S / D
  const double inv_k2r = SD * T * rQ;
  if (!(g > 0.0)) return true;
  if (rho) *rho = K.FL * ne * inv_g;
  const double xT = K.CT * (inv_gT * (th * B));  // CT / (g Theta)
  if (xT > 700.0) return true;                   // B_nu underflows: no emission
  const double jL = K.CjL * ne * B * synch_im(K.Cx * inv_gT) * inv_k2r;
  if (xT > 600.0) {  // B_nu near underflow: the direct quotient
    const double Bnu = K.SB * (inv_g * inv_g * inv_g) / expm1(xT);
    if (!(Bnu > 0.0)) return true;
    *a = jL / (g * Bnu);
  } else {
    // jL / B_nu scaled first: no intermediate smaller than the result (a
    // product jL expm1(x_T) of a faint sample would go subnormal)
    *a = (jL * K.iSB) * ((g * g) * expm1_pos(xT));
  }
  *e = g * g * jL;
  return true;
}

// d lambda_affine / ds = Sigma / r.
__device__ __forceinline__ double affine_rate(const RayC& c, double r, double u) {
  return __builtin_fma(c.a2, u * u, r * r) / r;
}

// Trapezoidal front-to-back formal solution over one step (second-order
// form of d_rte_step, src/cuda/device_physics.cuh:1988-2015).
struct RtAcc {
  double I, Tr, tau, a0, e0, s0;
};
__device__ __forceinline__ void rt_segment(RtAcc& q, double h, double a1, double e1, double s1) {
  const double asum = q.a0 + a1;
  const double esum = q.e0 + e1;
  if (asum > 0.0 || esum > 0.0) {
    const double dl = 0.5 * h * (q.s0 + s1);
    if (asum > 0.0) {
      const double dtau = 0.5 * asum * dl;
      // 1 - e^-dtau: Taylor polynomial below 1e-3 (truncation dtau^6 / 720)
      const double om = (dtau < 1e-3)
                            ? dtau * __builtin_fma(__builtin_fma(__builtin_fma(__builtin_fma(-dtau, 1.0 / 120.0,
                                                                                               1.0 / 24.0),
                                                                                -dtau, 1.0 / 6.0),
                                                                 -dtau, 0.5),
                                                   -dtau, 1.0)
                            : -expm1(-dtau);
      q.I = __builtin_fma(q.Tr * (esum / asum), om, q.I);
      q.Tr = q.Tr * (1.0 - om);
      q.tau = q.tau + dtau;
    } else {
      q.I = __builtin_fma(q.Tr, 0.5 * esum * dl, q.I);
    }
  }
  q.a0 = a1;
  q.e0 = e1;
  q.s0 = s1;
}

// Polarised transfer beside the scalar one: linear (Q, U) under total
// absorption and Faraday rotation (the simplified Mueller matrix of
// stokesStep, src/physics/stokes_transport.h:260-369; V has no source and
// stays zero), accumulated from the observer backwards like the scalar I.
// Rotations of the (Q, U) plane commute, so the state is the observed
// (Q, U), the transmittance (the scalar's Tr) and the cumulative Faraday
// angle Phi as (cos, sin).  Over one step the source (j_Q, j_U) = -frac e
// (cos 2 chi, sin 2 chi) with the trapezoid-averaged emission e,
// absorption A and Faraday rate R per affine parameter contributes, at the
// step's observer-side end, (j_Q + i j_U) (1 - e^{-w}) / z, z = A - i R,
// w = z dl -- the exact integral of the emission attenuated and rotated
// within the step -- and reaches the observer rotated by Phi and scaled by
// Tr.  Must run before rt_segment updates Tr and the step's left values.
struct PolAcc {
  double q, u, c, s, r0;
};

__device__ __forceinline__ void pol_segment(PolAcc& p, const RtAcc& q, double h, double a1, double e1,
                                            double r1, double s1, double c2, double s2, double frac) {
  const double esum = q.e0 + e1;
  const double rsum = p.r0 + r1;
  if (esum > 0.0 || rsum != 0.0) {
    const double dl = 0.5 * h * (q.s0 + s1);
    const double A = 0.5 * (q.a0 + a1), R = 0.5 * rsum, J = 0.5 * esum;
    const double x = A * dl, z = R * dl;
    double sz, cz;
    sincos(z, &sz, &cz);
    double ic, is;
    if (x * x + z * z < 1e-6) {  // (1 - e^{-w}) / z to third order in w
      ic = dl * (1.0 - 0.5 * x + (x * x - z * z) * (1.0 / 6.0));
      is = dl * (0.5 * z - x * z * (1.0 / 3.0));
    } else {
      const double E = exp(-x);
      const double D = A * A + R * R;
      const double omc = 1.0 - E * cz;
      ic = (A * omc + R * E * sz) / D;
      is = (R * omc - A * E * sz) / D;
    }
    const double jq = -frac * J * c2, ju = -frac * J * s2;
    const double qe = jq * ic - ju * is, ue = jq * is + ju * ic;
    p.q = __builtin_fma(q.Tr, p.c * qe - p.s * ue, p.q);
    p.u = __builtin_fma(q.Tr, p.s * qe + p.c * ue, p.u);
    const double nc = p.c * cz - p.s * sz, ns = p.s * cz + p.c * sz;
    p.c = nc;
    p.s = ns;
  }
  p.r0 = r1;
}

}  // namespace bhk
