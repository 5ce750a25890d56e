// C entry points over the Mino-time crossing solver (blackhole_amd/csrc/mino.hpp)
// for the CPU tests: the elliptic integral, the radial roots, the polar
// times and the per-ray solution.
#include <cmath>

#include "../../blackhole_amd/csrc/mino.hpp"

extern "C" {

double mino_carlson_rf(double x, double y, double z) { return mino::carlson_rf(x, y, z); }

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
}
