// mino.hpp - equatorial crossings of Kerr null geodesics from the Mino-time
// integrals of motion (host and device).
//
// In Mino time l the radial and polar motions separate:
//   (dr/dl)^2 = R(r) = r^4 + A r^2 + B r + C,
//       A = a^2 - eta - xi^2,  B = 2 (eta + (xi - a)^2),  C = -a^2 eta,
//   (du/dl)^2 = U(u) = eta - (eta + xi^2 - a^2) u^2 - a^2 u^4,  u = cos(theta),
// so a ray traced back from the observer (r_o, u_o) reaches its m-th
// equatorial crossing (u = 0) after the Mino time
//   G_m = (F_o or 2K - F_o) / k + m 2K / k
// (F_o = F(asin(|u_o| / sqrt(w+)) | mpar), K = K(mpar), k = sqrt(eta / w+),
// mpar = -a^2 w+^2 / eta, w+ the positive root of a^2 w^2 + (eta + xi^2 -
// a^2) w - eta = 0; F_o when the ray starts toward the equator), and the
// crossing radius solves  integral from r to r_o of dr / sqrt(R) = G_m  along
// the radial path (through the turning point r4, the largest root of R, when
// r4 lies outside the stopping radius).  The polar integrals are Carlson's
// R_F (B. C. Carlson, "Numerical computation of real or complex elliptic
// integrals", Numer. Algorithms 10 (1995) 13-26, duplication algorithm); the
// radial roots come from Ferrari's resolvent cubic; the radial integral is
// taken in x = 1/r -- with x = x4 - s^2 when the path turns, which makes the
// integrand 2 / sqrt(S(x4 - s^2)) (S = x^4 R(1/x) / (x4 - x), a cubic) smooth
// through the turning point -- by adaptive Gauss-Kronrod 7-15 quadrature
// (the QUADPACK rule), and the crossing by a bracketed Newton iteration on
// T(w) - G_m, T the Mino time along the path variable w.
//
// The same geodesic equations are what the step integrators of the library
// integrate (kerr_device.hpp, second-order form); here they are solved in
// quadrature instead, to a relative tolerance far below the integrators'
// crossing interpolation.  Rays the quadrature cannot resolve (the integrand
// nearly singular, as close to the critical curve) report failure.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MINO_HD __host__ __device__ inline
#else
#include <cmath>
#define MINO_HD inline
#endif

namespace mino {

#if !defined(__HIPCC__)
using std::acos;
using std::cbrt;
using std::cos;
using std::fabs;
using std::fma;
using std::fmax;
using std::fmin;
using std::sqrt;
#endif

// 1 / sqrt(v), v > 0: on the device the hardware reciprocal square root
// (v_rsq_f64) and one Newton step y (3/2 - v y^2 / 2), relative error about
// 1e-13 -- far below the quadrature tolerance -- for a fraction of the
// correctly rounded sqrt-and-divide sequence; on the host 1 / sqrt(v).
MINO_HD double inv_sqrt(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(v);
  return y * (1.5 - 0.5 * v * (y * y));
#else
  return 1.0 / sqrt(v);
#endif
}

// sqrt(v), v >= 0: on the device v / sqrt(v) from inv_sqrt (0 for v = 0),
// relative error about 1e-13 like inv_sqrt; on the host sqrt(v).
MINO_HD double fast_sqrt(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return v > 0.0 ? v * inv_sqrt(v) : 0.0;
#else
  return sqrt(v);
#endif
}

// 1 / v, v finite and nonzero: on the device the hardware reciprocal
// (v_rcp_f64) and one Newton step y (2 - v y), relative error about 1e-16,
// for a fraction of the correctly rounded division sequence; on the host 1 / v.
MINO_HD double fast_rcp(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rcp(v);
  return fma(y, fma(-v, y, 1.0), y);
#else
  return 1.0 / v;
#endif
}

// Carlson's R_F(x, y, z) for x, y, z >= 0, at most one zero: duplication until
// the arguments agree to 2.5e-3 (|A - x| < 2.5e-3 A for each, no division),
// then the fifth-order series in X = 1 - x / A, Y, Z:
// (1 - E2/10 + E3/14 + E2^2/24 - 3 E2 E3 / 44) / sqrt(A).
MINO_HD double carlson_rf(double x, double y, double z) {
  double A = (x + y + z) * (1.0 / 3.0);
  for (int it = 0; it < 60; ++it) {
    if (fmax(fabs(A - x), fmax(fabs(A - y), fabs(A - z))) < 2.5e-3 * A) {
      const double iA = 1.0 / A;
      const double X = (A - x) * iA, Y = (A - y) * iA, Z = -(X + Y);
      const double E2 = X * Y - Z * Z, E3 = X * Y * Z;
      return (1.0 - E2 * 0.1 + E3 * (1.0 / 14.0) + E2 * E2 * (1.0 / 24.0) - E2 * E3 * (3.0 / 44.0)) * inv_sqrt(A);
    }
    const double sx = fast_sqrt(x), sy = fast_sqrt(y), sz = fast_sqrt(z);
    const double lam = sx * (sy + sz) + sy * sz;
    x = 0.25 * (x + lam);
    y = 0.25 * (y + lam);
    z = 0.25 * (z + lam);
    A = (x + y + z) * (1.0 / 3.0);
  }
  return inv_sqrt(A);
}

// Mino times of the polar motion: the m-th equatorial crossing comes at
// g0 + m dg.  ok = 0 when the ray never reaches the equator (eta <= 0, or an
// observer in the equatorial plane).
struct PolarTimes {
  double g0, dg;
  // the polar motion itself: u(l) = -sgn(u_o) sqrt(wp) sn(k (l - g0) | mpar),
  // K = K(mpar), s = |u_o| / sqrt(wp)
  double wp, k, mpar, K, s;
  int ok, toward;
};

// u_dir: sign of du/dl at the observer (the tracer's initial v_u = sin(theta_o) beta).
MINO_HD PolarTimes polar_times(double a, double xi, double eta, double u_o, double u_dir) {
  PolarTimes t{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 0};
  if (!(eta > 0.0) || u_o == 0.0) return t;
  const double a2 = a * a;
  const double b = eta + xi * xi - a2;
  const double disc = fast_sqrt(b * b + 4.0 * a2 * eta);
  const double wp = (b >= 0.0) ? 2.0 * eta / (b + disc) : (disc - b) / (2.0 * a2);
  const double k = fast_sqrt(eta / wp);
  const double mpar = -a2 * wp * wp / eta;
  double s = fabs(u_o) / fast_sqrt(wp);
  if (s > 1.0) s = 1.0;
  const double s2 = s * s;
  const double Fo = s * carlson_rf(1.0 - s2, 1.0 - mpar * s2, 1.0);
  const double K = carlson_rf(0.0, 1.0 - mpar, 1.0);
  const bool toward = (u_o > 0.0) ? (u_dir < 0.0) : (u_dir > 0.0);
  t.g0 = (toward ? Fo : 2.0 * K - Fo) / k;
  t.dg = 2.0 * K / k;
  t.wp = wp;
  t.k = k;
  t.mpar = mpar;
  t.K = K;
  t.s = s;
  t.toward = toward;
  t.ok = 1;
  return t;
}

// Carlson's R_C(x, y) = R_F(x, y, y), x >= 0, y > 0: with e = 1 - x / y,
// R_C = f(e) / sqrt(y), f(e) = asin(sqrt(e)) / sqrt(e) for e > 0 and
// asinh(sqrt(-e)) / sqrt(-e) for e < 0, and for |e| < 1e-3 their common
// Maclaurin series 1 + e/6 + 3e^2/40 + 5e^3/112 + 35e^4/1152 + 63e^5/2816
// (the next term is below 2e-20).
MINO_HD double carlson_rc(double x, double y) {
  const double e = 1.0 - x * fast_rcp(y);
  double f;
  if (fabs(e) < 1e-3) {
    f = fma(e, fma(e, fma(e, fma(e, fma(e, 63.0 / 2816.0, 35.0 / 1152.0), 5.0 / 112.0), 3.0 / 40.0), 1.0 / 6.0),
            1.0);
  } else if (e > 0.0) {
    const double ir = inv_sqrt(e);
    f = asin(e * ir) * ir;
  } else {
    const double ir = inv_sqrt(-e);
    f = asinh(-e * ir) * ir;
  }
  return f * inv_sqrt(y);
}

// Carlson's R_J(x, y, z, p), x, y, z >= 0 (at most one zero), p > 0, by
// duplication (B. C. Carlson 1995, the algorithm for R_J): each step adds
// 4^-m R_C(1, 1 + e_m) / d_m with d_m = (sp + sx)(sp + sy)(sp + sz) and
// e_m = 4^-3m (p - x)(p - y)(p - z) / d_m^2; once the arguments agree to
// 1e-2 of their mean A, the fifth-order series in X = 1 - x / A, Y, Z and
// P = -(X + Y + Z) / 2 closes it (truncation of order 1e-12 relative; the
// square roots are fast_sqrt's).
MINO_HD double carlson_rj(double x, double y, double z, double p) {
  const double delta = (p - x) * (p - y) * (p - z);
  double A = (x + y + z + 2.0 * p) * 0.2;
  double sum = 0.0, fac = 1.0;
  for (int it = 0; it < 60; ++it) {
    const double mx = fmax(fmax(fabs(A - x), fabs(A - y)), fmax(fabs(A - z), fabs(A - p)));
    if (mx < 1e-2 * A) break;
    const double sx = fast_sqrt(x), sy = fast_sqrt(y), sz = fast_sqrt(z), sp = fast_sqrt(p);
    const double lam = sx * (sy + sz) + sy * sz;
    const double d = (sp + sx) * (sp + sy) * (sp + sz);
    const double id = fast_rcp(d);
    const double e = fac * fac * fac * delta * (id * id);
    sum += fac * carlson_rc(1.0, 1.0 + e) * id;
    x = 0.25 * (x + lam);
    y = 0.25 * (y + lam);
    z = 0.25 * (z + lam);
    p = 0.25 * (p + lam);
    A = 0.25 * (A + lam);
    fac *= 0.25;
  }
  const double iA = fast_rcp(A);
  const double X = (A - x) * iA, Y = (A - y) * iA, Z = (A - z) * iA;
  const double P = -0.5 * (X + Y + Z);
  const double XYZ = X * Y * Z, P2 = P * P;
  const double E2 = X * Y + X * Z + Y * Z - 3.0 * P2;
  const double E3 = XYZ + 2.0 * E2 * P + 4.0 * P2 * P;
  const double E4 = (2.0 * XYZ + E2 * P + 3.0 * P2 * P) * P;
  const double E5 = XYZ * P2;
  const double ser = 1.0 - 3.0 / 14.0 * E2 + E3 * (1.0 / 6.0) + 9.0 / 88.0 * E2 * E2 - 3.0 / 22.0 * E4 -
                     9.0 / 52.0 * E2 * E3 + 3.0 / 26.0 * E5;
  return fac * ser * iA * inv_sqrt(A) + 6.0 * sum;
}

// Jacobi sn, cn, dn(v | m) for m < 1.  For 0 <= m < 1 the arithmetic-
// geometric mean (descending Landen sequence: a_0 = 1, b_0 = sqrt(1 - m),
// c_0 = sqrt(m), at most eight steps -- c_8 is below 1e-19 for m <= 0.999;
// a step whose c_n falls below 1e-16 a_n ends the sequence -- then
// phi_N = 2^N a_N v and phi_{n-1} = (phi_n + asin(c_n sin(phi_n) / a_n)) / 2;
// sn = sin phi_0, cn = cos phi_0, dn = sqrt(1 - m sn^2)).  For m < 0 the
// negative-parameter transformation: with mu = -m / (1 - m) and
// t = sqrt(1 - m), sn(v | m) = sd(t v | mu) / t, cn = cd(t v | mu),
// dn = nd(t v | mu).  Square roots are fast_sqrt's (1e-13).
struct Jacobi {
  double sn, cn, dn;
};

MINO_HD Jacobi jacobi_snd(double v, double m) {
  double t = 1.0, mu = m;
  if (m < 0.0) {
    t = fast_sqrt(1.0 - m);
    mu = -m * fast_rcp(1.0 - m);
    v *= t;
  }
  constexpr int kN = 8;
  double a[kN + 1], c[kN + 1];
  a[0] = 1.0;
  c[0] = fast_sqrt(mu);
  double b = fast_sqrt(1.0 - mu);
#pragma unroll
  for (int n = 0; n < kN; ++n) {
    // once converged, c stays 0 and a fixed: the backward steps only halve
    const bool done = !(c[n] > 1e-16 * a[n]);
    a[n + 1] = done ? a[n] : 0.5 * (a[n] + b);
    c[n + 1] = done ? 0.0 : 0.5 * (a[n] - b);
    b = done ? b : fast_sqrt(a[n] * b);
  }
  double ph = 256.0 * a[kN] * v;
#pragma unroll
  for (int n = kN; n >= 1; --n) {
    if (c[n] != 0.0) ph += asin(c[n] * fast_rcp(a[n]) * sin(ph));
    ph *= 0.5;
  }
  const double sn = sin(ph), cn = cos(ph), dn = fast_sqrt(1.0 - mu * sn * sn);
  if (m < 0.0) {
    const double idn = fast_rcp(dn);
    return Jacobi{sn * idn * fast_rcp(t), cn * idn, idn};
  }
  return Jacobi{sn, cn, dn};
}

// Legendre's incomplete integral of the third kind in Jacobi form,
//   Pi(n; am(v) | m) = integral over [0, v] of dv' / (1 - n sn^2(v' | m)),
// n < 1, m < 1, for any v: with v = 2 K j + v_r (|v_r| <= K, K = K(m)) it is
// 2 j Pi(n | m) + Pi(n; am(v_r) | m), the reduced part from sn and cn of v_r
// (Carlson's form, DLMF 19.25.14 scaled by sin^2):
//   s R_F(c^2, 1 - m s^2, 1) + n / 3 s^3 R_J(c^2, 1 - m s^2, 1, 1 - n s^2),
// and the complete one R_F(0, 1 - m, 1) + n / 3 R_J(0, 1 - m, 1, 1 - n).
MINO_HD double pi3_reduced(double n, double m, double s, double c) {
  const double s2 = s * s;
  const double y = 1.0 - m * s2;
  return s * carlson_rf(c * c, y, 1.0) + n * (1.0 / 3.0) * s * s2 * carlson_rj(c * c, y, 1.0, 1.0 - n * s2);
}

MINO_HD double pi3_complete(double n, double m) {
  return carlson_rf(0.0, 1.0 - m, 1.0) + n * (1.0 / 3.0) * carlson_rj(0.0, 1.0 - m, 1.0, 1.0 - n);
}

// Largest real root of R when R has real roots r3 <= r4 (has34 = 1), from the
// largest root y0 of the resolvent 8y^3 + 8A y^2 + (2A^2 - 8C) y - B^2 = 0:
// with z = fast_sqrt(y0 / 2), r3,4 = z -+ fast_sqrt(-A/2 - z^2 - B / (4z)) and
// r1,2 = -z -+ fast_sqrt(-A/2 - z^2 + B / (4z)).  r4 is polished by Newton on R.
struct RadialRoots {
  double r3, r4;
  int has34;
};

MINO_HD double quartic_R(double A, double B, double C, double r) { return ((r * r + A) * r + B) * r + C; }

MINO_HD RadialRoots radial_roots(double A, double B, double C) {
  RadialRoots o{0.0, 0.0, 0};
  const double c1 = 0.25 * A * A - C;  // y^3 + A y^2 + c1 y - B^2/8
  const double c0 = -0.125 * B * B;
  const double P = c1 - A * A / 3.0;
  const double Q = 2.0 * A * A * A / 27.0 - A * c1 / 3.0 + c0;
  const double D = 0.25 * Q * Q + P * P * P / 27.0;
  double t;
  if (D > 0.0) {
    const double h = -0.5 * Q;
    const double u1 = cbrt(h + (h >= 0.0 ? fast_sqrt(D) : -fast_sqrt(D)));
    t = (u1 != 0.0) ? u1 - P / (3.0 * u1) : 0.0;
  } else {
    const double rp = fast_sqrt(fmax(-P / 3.0, 0.0));
    double c = (rp > 0.0) ? -Q / (2.0 * rp * rp * rp) : 0.0;
    c = fmin(fmax(c, -1.0), 1.0);
    t = 2.0 * rp * cos(acos(c) / 3.0);
  }
  double y = t - A / 3.0;
  for (int it = 0; it < 3; ++it) {
    const double h = ((y + A) * y + c1) * y + c0;
    const double dh = (3.0 * y + 2.0 * A) * y + c1;
    if (dh != 0.0) y -= h / dh;
  }
  if (!(y > 0.0)) return o;
  const double z = fast_sqrt(0.5 * y);
  const double rad = -0.5 * A - z * z - B / (4.0 * z);
  if (rad < 0.0) return o;
  const double sr = fast_sqrt(rad);
  double r4 = z + sr;
  for (int it = 0; it < 2; ++it) {
    const double f = quartic_R(A, B, C, r4);
    const double df = (4.0 * r4 * r4 + 2.0 * A) * r4 + B;
    if (df > 0.0) r4 -= f / df;
  }
  o.r3 = z - sr;
  o.r4 = r4;
  o.has34 = 1;
  return o;
}

// The radial path of one ray as a function of w >= 0 (Mino time T(w) =
// integral of f from 0 to w):
//   turning:  s = s_o - w, x = x4 - s^2, f = 2 / sqrt(S(x)), S the cubic
//             x^4 R(1/x) / (x4 - x) = -(q3 x^3 + q2 x^2 + q1 x + q0);
//   plunging: x = x_o + w, f = 1 / sqrt(x^4 R(1/x)).
struct Path {
  double A, B, C;
  double x_o, x4, s_o;
  double q3, q2, q1, q0;
  double w_end;  // stopping radius (plunging) or escape radius (turning)
  int turning;
};

MINO_HD double path_x(const Path& p, double w) {
  if (p.turning) {
    const double s = p.s_o - w;
    return p.x4 - s * s;
  }
  return p.x_o + w;
}

MINO_HD double path_f(const Path& p, double w) {
  if (p.turning) {
    const double s = p.s_o - w;
    const double x = fma(-s, s, p.x4);
    const double S = -fma(fma(fma(p.q3, x, p.q2), x, p.q1), x, p.q0);
    return S > 0.0 ? 2.0 * inv_sqrt(S) : 1e300;
  }
  const double x = p.x_o + w;
  const double Rx = fma(x * x, fma(x, fma(p.C, x, p.B), p.A), 1.0);
  return Rx > 0.0 ? inv_sqrt(Rx) : 1e300;
}

// One Gauss-Kronrod 7-15 panel (the QUADPACK qk15 rule: Kronrod abscissae
// xgk, Kronrod weights wgk, Gauss weights wg of the odd abscissae and the
// centre): returns the Kronrod estimate, *err = |K15 - G7|.
MINO_HD double gk15(const Path& p, double a, double b, double* err) {
  constexpr double xgk[7] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                             0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                             0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                             0.207784955007898467600689403773245};
  constexpr double wgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                             0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                             0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                             0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
  constexpr double wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                            0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  const double fc = path_f(p, c);
  double rk = wgk[7] * fc, rg = wg[3] * fc;
#pragma unroll
  for (int j = 0; j < 7; ++j) {
    const double dx = h * xgk[j];
    const double f = path_f(p, c - dx) + path_f(p, c + dx);
    rk = fma(wgk[j], f, rk);
    if (j & 1) rg = fma(wg[j >> 1], f, rg);
  }
  *err = fabs((rk - rg) * h);
  return rk * h;
}

// Gauss-Kronrod 3-7 rule on [a, b] for any integrand f: the 7-point Kronrod
// extension of the 3-point Gauss rule (exact for polynomials of degree 11
// and 5); returns the Kronrod estimate, *err = |K7 - G3|.  Nodes 0,
// +-sqrt(3/5) (Gauss) and the roots +-0.4342..., +-0.9604... of the Stieltjes
// polynomial; weights from the exactness conditions (Kronrod 1965).
template <class F>
MINO_HD double gk7_rule(F f, double a, double b, double* err) {
  constexpr double xk0 = 0.96049126870802028342, xg = 0.77459666924148337704, xk1 = 0.43424374934680255800;
  constexpr double wk0 = 0.10465622602646726519, wkg = 0.26848808986833344073, wk1 = 0.40139741477596222291,
                   wkc = 0.45091653865847414235;
  constexpr double wg = 5.0 / 9.0, wgc = 8.0 / 9.0;
  const double c = 0.5 * (a + b), h = 0.5 * (b - a);
  const double fc = f(c);
  const double fg = f(c - h * xg) + f(c + h * xg);
  const double f0 = f(c - h * xk0) + f(c + h * xk0);
  const double f1 = f(c - h * xk1) + f(c + h * xk1);
  const double rk = wkc * fc + wkg * fg + wk0 * f0 + wk1 * f1;
  const double rg = wgc * fc + wg * fg;
  *err = fabs((rk - rg) * h);
  return rk * h;
}

// Adaptive GK15 over [a, b] (a <= b), marching from a with step control
// (no panel list, so it lives in registers): a panel is accepted when
// |K15 - G7| <= tol |K15| (the relative error of the sum is then below tol),
// the next step doubles; a rejected panel is halved.  *ok = 0 when it needed
// more than kMaxPanels panel evaluations.
constexpr int kMaxPanels = 48;

MINO_HD double integrate(const Path& p, double a, double b, double tol, int* ok) {
  double sum = 0.0, x = a, h = b - a;
  const double hmin = 1e-13 * (b - a);
  for (int it = 0; it < kMaxPanels; ++it) {
    if (!(x < b)) {
      *ok = 1;
      return sum;
    }
    if (h > b - x) h = b - x;
    double e;
    const double v = gk15(p, x, x + h, &e);
    if (e <= tol * fabs(v) || h <= hmin) {
      sum += v;
      x += h;
      h *= 2.0;
    } else {
      h *= 0.5;
    }
  }
  *ok = !(x < b);
  return sum;
}

// integrate() for the short intervals of Newton steps: one Gauss-Kronrod
// 3-7 panel when its estimate passes the tolerance, the adaptive 7-15 march
// otherwise.  Near convergence a step spans a small part of the path, where
// 7 points suffice: 60 instead of 72 integrand evaluations per c5 ray, with
// the same radii to 1e-13.
MINO_HD double integrate_short(const Path& p, double a, double b, double tol, int* ok) {
  double e;
  const double v = gk7_rule([&p](double w) { return path_f(p, w); }, a, b, &e);
  if (e <= tol * fabs(v)) {
    *ok = 1;
    return v;
  }
  return integrate(p, a, b, tol, ok);
}

// Builds the radial path of a ray from r_o inward: turning when R has a root
// r4 above r_stop (then it ends at r_escape on the way out), plunging
// otherwise (ends at r_stop).  Returns 0 when the turning point is too close
// to the observer or the roots degenerate.
MINO_HD int make_path(double a, double xi, double eta, double r_o, double r_stop, double r_escape, Path* p) {
  p->A = a * a - eta - xi * xi;
  const double dx = xi - a;
  p->B = 2.0 * (eta + dx * dx);
  p->C = -a * a * eta;
  p->x_o = 1.0 / r_o;
  const RadialRoots rr = radial_roots(p->A, p->B, p->C);
  p->turning = rr.has34 && rr.r4 > r_stop;
  if (p->turning) {
    if (!(rr.r4 < r_o)) return 0;
    const double x4 = 1.0 / rr.r4;
    p->x4 = x4;
    // x^4 R(1/x) = C x^4 + B x^3 + A x^2 + 0 x + 1, divided by (x - x4)
    p->q3 = p->C;
    p->q2 = p->B + x4 * p->q3;
    p->q1 = p->A + x4 * p->q2;
    p->q0 = x4 * p->q1;
    p->s_o = fast_sqrt(x4 - p->x_o);
    const double x_esc = 1.0 / r_escape;
    p->w_end = p->s_o + (x4 > x_esc ? fast_sqrt(x4 - x_esc) : 0.0);
  } else {
    p->x4 = 0.0;
    p->q3 = p->q2 = p->q1 = p->q0 = 0.0;
    p->s_o = 0.0;
    p->w_end = 1.0 / r_stop - p->x_o;
  }
  return 1;
}

// Solves T(w) = G for w in [w_lo, w_hi] with T(w_lo) = T_lo <= G <= T_hi =
// T(w_hi), starting from w0 (clamped into the bracket): Newton steps
// w += (G - T) / f(w), each T update an adaptive integral from the previous
// iterate, bisection whenever a step leaves the bracket.  Returns the number
// of iterations, or -1 on failure; *w_out the solution.
MINO_HD int solve_crossing(const Path& p, double G, double w_lo, double T_lo, double w_hi, double T_hi, double w0,
                           double tol, double* w_out) {
  double lo = w_lo, Tlo = T_lo, hi = w_hi, Thi = T_hi;
  double w = w0;
  if (!(w > lo && w < hi)) {
    // no usable start: invert the cubic Hermite model of T on the bracket
    // (values and slopes f at both ends), from the linear interpolant
    const double L = hi - lo;
    const double idT = 1.0 / (Thi - Tlo);
    const double d0 = path_f(p, lo) * L * idT, d1 = path_f(p, hi) * L * idT;  // end slopes over the T range
    const double target = (G - Tlo) * idT;
    double t = target;
    for (int k = 0; k < 4; ++k) {
      const double t2 = t * t, t3 = t2 * t;
      const double h = (3.0 * t2 - 2.0 * t3) + ((t3 - 2.0 * t2 + t) * d0 + (t3 - t2) * d1);
      const double dh = (6.0 * t - 6.0 * t2) + ((3.0 * t2 - 4.0 * t + 1.0) * d0 + (3.0 * t2 - 2.0 * t) * d1);
      if (!(dh > 0.0)) break;
      t -= (h - target) / dh;
      t = fmin(fmax(t, 0.0), 1.0);
    }
    w = lo + L * t;
    if (!(w > lo && w < hi)) w = lo + L * target;
  }
  int ok = 1;
  // T at the starting point from the nearer end of the bracket
  double T;
  if (w - lo <= hi - w) {
    T = Tlo + integrate(p, lo, w, tol, &ok);
  } else {
    T = Thi - integrate(p, w, hi, tol, &ok);
  }
  if (!ok) return -1;
  for (int it = 1; it <= 40; ++it) {
    const double dT = G - T;
    if (fabs(dT) <= 1e-12 * G + 1e-300) {
      *w_out = w;
      return it;
    }
    if (dT > 0.0) {
      lo = w;
      Tlo = T;
    } else {
      hi = w;
      Thi = T;
    }
    const double f = path_f(p, w);
    double wn = w + dT / f;
    // close enough that the Newton step's own error (quadratic in dT) is
    // negligible: take it without another quadrature
    if (fabs(dT) <= 1e-6 * G && wn > lo && wn < hi) {
      *w_out = wn;
      return it;
    }
    if (!(wn > lo && wn < hi)) wn = 0.5 * (lo + hi);
    const double inc = (wn > w) ? integrate_short(p, w, wn, tol, &ok) : -integrate_short(p, wn, w, tol, &ok);
    if (!ok) return -1;
    T += inc;
    w = wn;
    if (hi - lo <= 1e-15 * (fabs(hi) + 1e-300)) {
      *w_out = w;
      return it;
    }
  }
  return -1;
}

// Crossing radius -> path variable w on the leg where it lies (inbound when
// w < s_o for a turning path); -1 when r is not on the path.
MINO_HD double w_of_r(const Path& p, double r, bool outbound) {
  const double x = 1.0 / r;
  if (p.turning) {
    if (!(x <= p.x4)) return -1.0;
    const double s = fast_sqrt(p.x4 - x);
    return outbound ? p.s_o + s : p.s_o - s;
  }
  return x - p.x_o;
}

// The adaptive Gauss-Kronrod 7-15 march of integrate() over the radial path,
// carrying beside the Mino time T = integral of f dw the radial part of the
// azimuth, Phi = integral of a x (2 - a xi x) / (1 - 2x + a^2 x^2) f dw
// (x = 1 / r: a (2 r - a xi) / Delta per unit Mino time), on the same nodes;
// a panel is accepted when both estimates pass (Phi with an absolute floor
// of tol per unit length, for a = 0 or small Phi).
MINO_HD double integrate_t_phi(const Path& p, double a, double xi, double lo, double hi, double tol, int* ok,
                               double* phi) {
  constexpr double xgk[7] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                             0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                             0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                             0.207784955007898467600689403773245};
  constexpr double wgk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                             0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                             0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                             0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
  constexpr double wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                            0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
  const double axi = a * xi, a2 = a * a;
  double sum = 0.0, sphi = 0.0, x = lo, h = hi - lo;
  const double hmin = 1e-13 * (hi - lo);
  for (int it = 0; it < kMaxPanels; ++it) {
    if (!(x < hi)) {
      *ok = 1;
      *phi = sphi;
      return sum;
    }
    if (h > hi - x) h = hi - x;
    const double c = x + 0.5 * h, hh = 0.5 * h;
    double tk = 0.0, tg = 0.0, pk = 0.0, pg = 0.0;
#pragma unroll
    for (int j = 0; j < 15; ++j) {
      // node j: offsets -xgk[0..6], 0, +xgk[6..0]
      const int jj = j < 7 ? j : (j == 7 ? 7 : 14 - j);
      const double off = j < 7 ? -xgk[j] : (j == 7 ? 0.0 : xgk[14 - j]);
      const double w = c + hh * off;
      const double f = path_f(p, w);
      const double xr = path_x(p, w);
      const double gf = a * xr * fma(-axi, xr, 2.0) * fast_rcp(fma(a2 * xr, xr, fma(-2.0, xr, 1.0))) * f;
      tk = fma(wgk[jj], f, tk);
      pk = fma(wgk[jj], gf, pk);
      if (jj == 7) {
        tg = fma(wg[3], f, tg);
        pg = fma(wg[3], gf, pg);
      } else if (jj & 1) {
        tg = fma(wg[jj >> 1], f, tg);
        pg = fma(wg[jj >> 1], gf, pg);
      }
    }
    const double vt = tk * hh, vp = pk * hh;
    const bool good = fabs((tk - tg) * hh) <= tol * fabs(vt) && fabs((pk - pg) * hh) <= tol * (fabs(vp) + h);
    if (good || h <= hmin) {
      sum += vt;
      sphi += vp;
      x += h;
      h *= 2.0;
    } else {
      h *= 0.5;
    }
  }
  *ok = !(x < hi);
  *phi = sphi;
  return sum;
}

// The direction (u, phi) where an escaping ray's path ends (r_escape), from
// the integrals of the tracer's equations in Mino time l, dphi/dl = xi / (1 -
// u^2) + a (2 r - a xi) / Delta (phi = 0 at the observer):
//   T_end, phi_r: the radial path's Mino time and radial azimuth
//       (integrate_t_phi over the inbound and outbound legs);
//   u(T_end) = -sgn(u_o) sqrt(wp) sn(v_e | mpar), v = k (l - g0), v_e at T_end;
//   phi_u = xi / k [Pi(wp; am(v_e) | mpar) - Pi(wp; am(v_0) | mpar)],
//       v_0 = -k g0 (1 - u^2 = 1 - wp sn^2, dl = dv / k);
// at the observer v_0 = -F_o (toward the equator: sn = -s, cn = sqrt(1 - s^2))
// or -2K + F_o (away from it: Pi(v_0) = -2 Pi(wp | mpar) + Pi(F_o)).
// Returns 0 (direction unknown) for a path that does not turn (no escape), a
// polar motion without equatorial crossings (eta <= 0, an observer in the
// equatorial plane) or a failed quadrature.
MINO_HD int exit_direction(const Path& p, const PolarTimes& pt, double a, double xi, double u_o, double tol,
                           double* u_exit, double* phi_exit) {
  if (!p.turning || !pt.ok) return 0;
  int ok = 1;
  double phi_in = 0.0, phi_out = 0.0;
  double T_end = integrate_t_phi(p, a, xi, 0.0, p.s_o, tol, &ok, &phi_in);
  if (!ok) return 0;
  T_end += integrate_t_phi(p, a, xi, p.s_o, p.w_end, tol, &ok, &phi_out);
  if (!ok) return 0;
  const double twoK = 2.0 * pt.K;
  const double ve = pt.k * (T_end - pt.g0);
  const double je = rint(ve * fast_rcp(twoK));
  const Jacobi J = jacobi_snd(fma(-twoK, je, ve), pt.mpar);
  const long long ji = static_cast<long long>(je);
  const double sn_e = (ji & 1) ? -J.sn : J.sn;  // sn(v + 2K) = -sn(v)
  // xi = 0: the polar term of dphi/dl vanishes (as in the tracer's equation),
  // while Pi(1; .) would diverge over the pole (w+ = 1)
  double phi_u = 0.0;
  if (xi != 0.0) {
    const double c0 = fast_sqrt(fmax(1.0 - pt.s * pt.s, 0.0));
    const double Pc = pi3_complete(pt.wp, pt.mpar);
    const double Po = pi3_reduced(pt.wp, pt.mpar, pt.s, c0);
    const double P0 = pt.toward ? -Po : fma(-2.0, Pc, Po);
    const double Pe = fma(2.0 * je, Pc, pi3_reduced(pt.wp, pt.mpar, J.sn, J.cn));
    phi_u = xi * fast_rcp(pt.k) * (Pe - P0);
  }
  const double sq_wp = fast_sqrt(pt.wp);
  *u_exit = (u_o > 0.0 ? -sq_wp : sq_wp) * sn_e;
  *phi_exit = (phi_in + phi_out) + phi_u;
  return 1;
}

// Disk classes as the tracer's records number them.
constexpr int kHorizon = 1, kDisk = 2, kEscaped = 3;
constexpr int kMaxCrossings = 16;

// The crossing sequence of one ray traced back from (r_o, u_o) with initial
// du/dl sign u_dir, up to the first crossing inside [r_in, r_out] (an opaque
// disk), the stopping radius r_stop or the escape radius r_escape.
// guess[0 .. n_guess) are starting radii for the first crossings (e.g. a
// learned estimate); without them each crossing starts from the linear
// interpolant of its bracket.  EXIT: an escaped ray also gets its end
// direction (exit_direction).
struct RaySolution {
  double r[3];   // first three crossing radii (0 beyond n)
  double r_hit;  // radius of the disk crossing (0 without one)
  int n;         // crossings before the ray ended (the hit included)
  int hit;       // index of the disk crossing, -1 for none
  int cls;       // kHorizon, kDisk or kEscaped
  int iters;     // Newton iterations over all crossings
  int ok;        // 0: the quadrature or the iteration failed
  // EXIT: the exact end direction of an escaped ray at r_escape (exit_ok = 1)
  double u_exit, phi_exit;
  int exit_ok;
};

template <bool EXIT = false>
MINO_HD RaySolution solve_ray(double a, double xi, double eta, double r_o, double u_o, double u_dir,
                              double r_stop, double r_escape, double r_in, double r_out, const double* guess,
                              int n_guess, double tol) {
  RaySolution out{{0.0, 0.0, 0.0}, 0.0, 0, -1, kHorizon, 0, 0, 0.0, 0.0, 0};
  Path p;
  if (!make_path(a, xi, eta, r_o, r_stop, r_escape, &p)) return out;
  const PolarTimes pt = polar_times(a, xi, eta, u_o, u_dir);
  out.cls = p.turning ? kEscaped : kHorizon;
  if (!pt.ok) {  // never reaches the equator: the path's end decides the class
    out.ok = 1;
    return out;
  }
  // Mino time to the turning point (turning paths) or to the stopping radius
  // (plunging); a turning path's outbound leg is integrated only when a
  // crossing lies beyond the turning point, so rays that hit the disk on the
  // way in never need it.  T_total < 0: not integrated yet.
  int ok = 1;
  double T4 = 0.0, T_total = -1.0;
  if (p.turning) {
    T4 = integrate(p, 0.0, p.s_o, tol, &ok);
  } else {
    T_total = integrate(p, 0.0, p.w_end, tol, &ok);
  }
  if (!ok) {
    out.cls = kHorizon;
    return out;
  }
  double w_prev = 0.0, T_prev = 0.0;
  for (int m = 0; m < kMaxCrossings; ++m) {
    const double G = pt.g0 + m * pt.dg;
    double lo = w_prev, Tlo = T_prev, hi, Thi;
    bool outbound = false;
    if (p.turning && G <= T4) {
      hi = p.s_o;
      Thi = T4;
    } else {
      if (p.turning) {
        if (T_total < 0.0) {
          T_total = T4 + integrate(p, p.s_o, p.w_end, tol, &ok);
          if (!ok) return out;
        }
        outbound = true;
        if (lo < p.s_o) {
          lo = p.s_o;
          Tlo = T4;
        }
      }
      if (!(G < T_total)) {  // the path ends (escape or capture) before this crossing
        out.ok = 1;
        if (EXIT && p.turning) out.exit_ok = exit_direction(p, pt, a, xi, u_o, tol, &out.u_exit, &out.phi_exit);
        return out;
      }
      hi = p.w_end;
      Thi = T_total;
    }
    double w0 = -1.0;
    if (m < n_guess && guess[m] > 0.0) w0 = w_of_r(p, guess[m], outbound);
    double w;
    const int it = solve_crossing(p, G, lo, Tlo, hi, Thi, w0, tol, &w);
    if (it < 0) return out;
    out.iters += it;
    const double r = 1.0 / path_x(p, w);
    if (m < 3) out.r[m] = r;
    out.n = m + 1;
    w_prev = w;
    T_prev = G;
    if (r >= r_in && r <= r_out) {
      out.hit = m;
      out.r_hit = r;
      out.cls = kDisk;
      out.ok = 1;
      return out;
    }
  }
  return out;  // too many crossings: not resolved
}

}  // namespace mino
