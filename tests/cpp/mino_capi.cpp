// C entry points over the Mino-time crossing solver (blackhole_amd/csrc/mino.hpp)
// for the CPU tests: the elliptic integral, the radial roots, the polar
// times and the per-ray solution.
#include <cmath>

#include "../../blackhole_amd/csrc/mino.hpp"

extern "C" {

double mino_carlson_rf(double x, double y, double z) { return mino::carlson_rf(x, y, z); }
double mino_carlson_rc(double x, double y) { return mino::carlson_rc(x, y); }
double mino_carlson_rj(double x, double y, double z, double p) { return mino::carlson_rj(x, y, z, p); }

// sn, cn, dn(v | m) into out[3].
void mino_jacobi(double v, double m, double* out) {
  const mino::Jacobi j = mino::jacobi_snd(v, m);
  out[0] = j.sn;
  out[1] = j.cn;
  out[2] = j.dn;
}

// Pi(n; am(v_r) | m) from sn, cn of v_r (|v_r| <= K), and the complete Pi(n | m).
double mino_pi3_reduced(double n, double m, double s, double c) { return mino::pi3_reduced(n, m, s, c); }
double mino_pi3_complete(double n, double m) { return mino::pi3_complete(n, m); }

// The Gauss-Kronrod 3-7 rule on x^d over [a, b]; err = |K7 - G3|.
double mino_gk7_monomial(int d, double a, double b, double* err) {
  return mino::gk7_rule([d](double x) { return std::pow(x, d); }, a, b, err);
}

int mino_radial_roots(double A, double B, double C, double* r34) {
  const mino::RadialRoots rr = mino::radial_roots(A, B, C);
  r34[0] = rr.r3;
  r34[1] = rr.r4;
  return rr.has34;
}

int mino_polar_times(double a, double xi, double eta, double u_o, double u_dir, double* g) {
  const mino::PolarTimes t = mino::polar_times(a, xi, eta, u_o, u_dir);
  g[0] = t.g0;
  g[1] = t.dg;
  return t.ok;
}

// Mino time from r_o down to r (inbound, no turning point before r).
double mino_radial_time(double a, double xi, double eta, double r_o, double r_stop, double r_escape, double r,
                        int* ok) {
  mino::Path p;
  if (!mino::make_path(a, xi, eta, r_o, r_stop, r_escape, &p)) {
    *ok = 0;
    return 0.0;
  }
  return mino::integrate(p, 0.0, mino::w_of_r(p, r, false), 1e-12, ok);
}

// Image-plane ray (alpha, beta) of an observer at (r_o, theta_o): xi, eta and
// the initial polar direction as the tracer sets them up; out = r[3], and
// ints = {n, hit, cls, iters, ok, turning}.
void mino_solve_pixel(double a, double theta_o, double r_o, double alpha, double beta, double r_stop,
                      double r_escape, double r_in, double r_out, const double* guess, int n_guess, double tol,
                      double* r, int* ints) {
  const double so = std::sin(theta_o), co = std::cos(theta_o);
  const double xi = -(alpha * so);
  const double eta = std::fma(beta, beta, (alpha * alpha - a * a) * (co * co));
  const mino::RaySolution s =
      mino::solve_ray(a, xi, eta, r_o, co, so * beta, r_stop, r_escape, r_in, r_out, guess, n_guess, tol);
  for (int i = 0; i < 3; ++i) r[i] = s.r[i];
  ints[0] = s.n;
  ints[1] = s.hit;
  ints[2] = s.cls;
  ints[3] = s.iters;
  ints[4] = s.ok;
  mino::Path p;
  ints[5] = mino::make_path(a, xi, eta, r_o, r_stop, r_escape, &p) ? p.turning : -1;
}

// As mino_solve_pixel, with the escape direction of escaped rays at
// r_escape: dir = {u_exit, phi_exit}; returns exit_ok.
int mino_solve_pixel_exit(double a, double theta_o, double r_o, double alpha, double beta, double r_stop,
                          double r_escape, double r_in, double r_out, double tol, double* r, int* ints,
                          double* dir) {
  const double so = std::sin(theta_o), co = std::cos(theta_o);
  const double xi = -(alpha * so);
  const double eta = std::fma(beta, beta, (alpha * alpha - a * a) * (co * co));
  const mino::RaySolution s =
      mino::solve_ray<true>(a, xi, eta, r_o, co, so * beta, r_stop, r_escape, r_in, r_out, nullptr, 0, tol);
  for (int i = 0; i < 3; ++i) r[i] = s.r[i];
  ints[0] = s.n;
  ints[1] = s.hit;
  ints[2] = s.cls;
  ints[3] = s.iters;
  ints[4] = s.ok;
  dir[0] = s.u_exit;
  dir[1] = s.phi_exit;
  return s.exit_ok;
}
}
