// flopcount.cpp - algorithmic FLOP per unit of work of the tracer, counted by
// instantiating the oracle's templates with a counting scalar.
//
// TEST / MEASUREMENT INFRASTRUCTURE ONLY: prints one JSON object
// with the FLOP per RHS evaluation, per RK4 step, per DP45 attempt, per
// equatorial crossing, per disk hit, per volume sample and per ray setup.
// Counting rules: + - * / sqrt = 1 each, FMA = 2, negation / fabs /
// comparison = 0; transcendentals (acos, exp, expm1, log, cbrt) and
// floor / frexp / ldexp are tallied separately and not counted as FLOP.
#include <cstdio>

struct F {
  double v;
  F() : v(0.0) {}
  F(double x) : v(x) {}  // NOLINT: implicit by design
};

namespace cnt {
long long add = 0, mul = 0, dv = 0, fma = 0, sq = 0, trans = 0, other = 0;
void reset() { add = mul = dv = fma = sq = trans = other = 0; }
long long flop() { return add + mul + dv + sq + 2 * fma; }
}  // namespace cnt

inline F operator+(F a, F b) { ++cnt::add; return F(a.v + b.v); }
inline F operator-(F a, F b) { ++cnt::add; return F(a.v - b.v); }
inline F operator*(F a, F b) { ++cnt::mul; return F(a.v * b.v); }
inline F operator/(F a, F b) { ++cnt::dv; return F(a.v / b.v); }
inline F operator-(F a) { return F(-a.v); }

namespace orc {
inline F fma_(F a, F b, F c) { ++cnt::fma; return F(a.v * b.v + c.v); }
inline F sqrt_(F x);
inline F fabs_(F x);
inline F floor_(F x);
inline double val_(F x) { return x.v; }
inline F acos_(F x);
inline F exp_(F x);
inline F expm1_(F x);
inline F cbrt_(F x);
inline F log_(F x);
inline F frexp_(F x, int* e);
inline F ldexp_(F x, int e);
}  // namespace orc

#include "kerr_oracle.hpp"

namespace orc {
inline F sqrt_(F x) { ++cnt::sq; return F(std::sqrt(x.v)); }
inline F fabs_(F x) { return F(std::fabs(x.v)); }
inline F floor_(F x) { ++cnt::other; return F(std::floor(x.v)); }
inline F acos_(F x) { ++cnt::trans; return F(std::acos(x.v)); }
inline F exp_(F x) { ++cnt::trans; return F(std::exp(x.v)); }
inline F expm1_(F x) { ++cnt::trans; return F(std::expm1(x.v)); }
inline F cbrt_(F x) { ++cnt::trans; return F(std::cbrt(x.v)); }
inline F log_(F x) { ++cnt::trans; return F(std::log(x.v)); }
inline F frexp_(F x, int* e) { ++cnt::other; return F(std::frexp(x.v, e)); }
inline F ldexp_(F x, int e) { ++cnt::other; return F(std::ldexp(x.v, e)); }
}  // namespace orc

using namespace orc;

struct Tally {
  long long flop, trans, other;
};
static Tally take() {
  Tally t{cnt::flop(), cnt::trans, cnt::other};
  cnt::reset();
  return t;
}

int main() {
  // A representative ray state (the counts do not depend on the values
  // except through branches, which take their common side here).
  const F a(0.94), xi(3.0), eta(20.0);
  RayConst<F> c = make_const(a, xi, eta);
  cnt::reset();
  F y[kN] = {F(8.0), F(0.3), F(1.0), F(-40.0), F(2.0)};
  F f[kN], yn[kN], k1[kN], y5[kN], k7[kN];

  rhs(c, y, f);
  const Tally t_rhs = take();

  // RK4 step as the trace loop runs it: rk4_step (3 RHS + combine) + the RHS
  // at the new point (the next step's k1).
  rhs(c, y, k1);
  cnt::reset();
  rk4_step(c, y, k1, F(0.01), yn);
  rhs(c, yn, f);
  project_radial(c, yn, f);
  const Tally t_rk4 = take();

  // DP45 attempt: 6 RHS (k2..k7, FSAL) + combinations + error norm, plus the
  // step-size controller.
  dp45_attempt(c, y, k1, F(0.01), F(1e-8), F(1e-10), y5, k7);
  dp45_factor(F(0.5), true, false);
  const Tally t_dp = take();
  project_radial(c, y5, k7);
  const Tally t_proj = take();

  // Equatorial crossing: Hermite root + disk g.
  F y0[kN] = {F(8.0), F(0.01), F(1.0), F(-40.0), F(-2.0)};
  F y1[kN] = {F(7.9), F(-0.01), F(1.01), F(-39.0), F(-2.0)};
  F f0[kN], f1[kN];
  rhs(c, y0, f0);
  rhs(c, y1, f1);
  cnt::reset();
  const F rc = crossing_radius(y0, f0, y1, f1, F(0.01));
  disk_g(rc, a, xi);
  const Tally t_cross = take();

  // Disk hit: LUT lookup + observed blackbody intensity.
  static double lut[256];
  for (int i = 0; i < 256; ++i) lut[i] = i / 255.0;
  lut_linear(lut, 256, F(2.0), F(20.0), rc);
  disk_intensity(F(0.5), F(0.9), 2.418e17, 1e7);
  const Tally t_disk = take();

  // Volume sample: trilinear fetch (4 fields) + fluid g + Planck + synchrotron
  // j + observer-frame a, e (volume_point), the affine rate and the
  // trapezoidal segment (rt_segment), as the trace loop runs them per sample.
  static float cells[4 * 8 * 8 * 8];
  for (int i = 0; i < 4 * 512; ++i) cells[i] = 1.0f;
  VolumeHost V;
  V.cells = cells;
  V.v = BH_Volume{nullptr, 8, 8, 8, 0, 1.0, 50.0, 2.0, 40.0, 0.8, 4.0};
  cnt::reset();
  F a1, e1;
  volume_point(V, c, F(8.0), F(0.1), F(1.0), 2.3e11, 5.9e11, &a1, &e1);
  RtAcc<F> q{F(0.0), F(1.0), F(0.0), F(1e-3), F(1e-20), F(9.0)};
  rt_segment(q, F(0.01), a1, e1, affine_rate(c, F(8.0), F(0.1)));
  const Tally t_vol = take();

  // Volume bookkeeping outside the support: only the affine rate and the
  // segment shift (no sample).
  RtAcc<F> q0{F(0.0), F(1.0), F(0.0), F(0.0), F(0.0), F(9.0)};
  rt_segment(q0, F(0.01), F(0.0), F(0.0), affine_rate(c, F(60.0), F(0.1)));
  const Tally t_vol_skip = take();

  // Ray setup (Bardeen mapping + R(r_obs) + first RHS).
  CamConst kc{0.94, 0.29, 0.95, 100.0, 1.4, 2.0, -15.0, 0.03, 15.0, 0.03, 1024, 1024, 0};
  F ys[kN];
  RayConst<F> cs = setup_ray<F>(kc, 10, 10, ys);
  rhs(cs, ys, f);
  const Tally t_setup = take();

  std::printf(
      "{\"rhs\": %lld, \"rk4_step\": %lld, \"dp45_attempt\": %lld, \"crossing\": %lld, "
      "\"disk_hit\": %lld, \"volume_sample\": %lld, \"setup\": %lld, \"projection\": %lld, "
      "\"volume_step_nosample\": %lld, "
      "\"trans\": {\"rhs\": %lld, \"rk4_step\": %lld, \"dp45_attempt\": %lld, \"crossing\": %lld, "
      "\"disk_hit\": %lld, \"volume_sample\": %lld, \"setup\": %lld}}\n",
      t_rhs.flop, t_rk4.flop, t_dp.flop, t_cross.flop, t_disk.flop, t_vol.flop, t_setup.flop,
      t_proj.flop, t_vol_skip.flop,
      t_rhs.trans, t_rk4.trans, t_dp.trans, t_cross.trans, t_disk.trans, t_vol.trans,
      t_setup.trans);
  return 0;
}
