// Prints values of the physics C++ interface (include/bh_physics.hpp) for
// tests/test_physics_api.py, in geometric units: mass chosen so that
// G M / c^2 = 1 cm.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bh_physics.hpp"

int main(int argc, char** argv) {
  const double mass = physics::C2 / physics::G;  // M = 1 cm
  const char* what = argc > 1 ? argv[1] : "";
  if (!std::strcmp(what, "closed")) {
    for (double s : {0.0, 0.3, 0.9, 0.998}) {
      std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", s, physics::kerrOuterHorizon(mass, s),
                  physics::kerrIscoRadius(mass, s, true), physics::kerrIscoRadius(mass, s, false),
                  physics::kerrPhotonOrbitPrograde(mass, s), physics::kerrPhotonOrbitRetrograde(mass, s));
    }
  } else if (!std::strcmp(what, "potentials")) {
    // a xi eta r theta -> R R' Theta Theta'
    double a, xi, eta, r, th;
    while (std::scanf("%lf %lf %lf %lf %lf", &a, &xi, &eta, &r, &th) == 5) {
      const physics::KerrGeodesicConsts c{1.0, xi, eta};
      const physics::KerrPotentials p = physics::kerrPotentials(r, th, mass, a, c);
      std::printf("%.17g %.17g %.17g %.17g\n", p.rPot, p.dRdr, p.thetaPot, p.dThetadtheta);
    }
  } else if (!std::strcmp(what, "turning")) {
    // equatorial a = 0.9, b = 6 inbound from r = 12: smallest r reached, then outbound?
    physics::Kerr k(mass, 0.9);
    physics::KerrGeodesicState s = physics::kerrEquatorialState(12.0, 0.0, -1.0);
    const physics::KerrGeodesicConsts c = physics::kerrEquatorialConsts(6.0);
    double rmin = s.r;
    int turned = 0;
    for (int i = 0; i < 40000; ++i) {
      s = k.stepMino(s, c, 2e-4);
      rmin = std::fmin(rmin, s.r);
      if (s.signR > 0 && s.r > 6.0) { turned = 1; break; }
    }
    std::printf("%.17g %d\n", rmin, turned);
  } else if (!std::strcmp(what, "order")) {
    // RK4 order of kerrStepMino: error of N steps of h against 2N steps of h/2 and 4N of h/4
    physics::Kerr k(mass, 0.7);
    const physics::KerrGeodesicConsts c{1.0, 2.5, 9.0};
    auto run = [&](double h, int n) {
      physics::KerrGeodesicState s{20.0, 1.2, 0.0, 0.0, -1.0, 1.0};
      for (int i = 0; i < n; ++i) s = k.stepMino(s, c, h);
      return s.r;
    };
    const double r1 = run(4e-4, 50), r2 = run(2e-4, 100), r3 = run(1e-4, 200);
    std::printf("%.17g %.17g\n", std::fabs(r1 - r2), std::fabs(r2 - r3));
  } else if (!std::strcmp(what, "verified")) {
    auto f = [](const verified::StateVector& y) {
      return verified::StateVector(y.x0, y.x1, y.x2, y.x3, y.v0, y.v1, y.v2, y.v3);  // y' = y
    };
    verified::StateVector y0(1, 1, 1, 1, 1, 1, 1, 1);
    const double e1 = std::fabs(verified::rk4_step(f, 0.1, y0).x1 - std::exp(0.1));
    const auto d1 = verified::dp45_step(f, 0.2, y0, 0.0, 1.0);
    const auto d2 = verified::dp45_step(f, 0.1, y0, 0.0, 1.0);
    std::printf("%.17g %.17g %.17g %.17g %.17g\n", e1, std::fabs(d1.first.x1 - std::exp(0.2)), d1.second, d2.second,
                verified::dp45_next_step(1.0, 1.0, false));
  } else if (!std::strcmp(what, "raytrace")) {
    // KerrRaytracer on the GPU: equatorial rays from r = 80 M with b in {4, 5, 6, 7}, a = 0.5
    physics::KerrRaytracer rt(mass, 0.5);
    rt.setEscapeRadius(200.0);
    rt.setMaxSteps(20000);
    std::vector<std::pair<physics::KerrGeodesicState, physics::KerrGeodesicConsts>> rays;
    for (double b : {4.0, 5.0, 6.0, 7.0})
      rays.push_back({physics::kerrEquatorialState(80.0, 0.0, -1.0), physics::kerrEquatorialConsts(b)});
    for (const physics::RayTraceResult& r : rt.traceBatch(rays))
      std::printf("%d %d %.17g %.17g %.17g\n", static_cast<int>(r.status), r.stepsTaken, r.finalPosition[0],
                  r.finalPosition[1], r.finalPosition[2]);
  } else if (!std::strcmp(what, "stokes")) {
    // I Q U V jI jQ jU jV alphaI rhoV ds -> stokesStep
    double v[11];
    while (std::scanf("%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", v, v + 1, v + 2, v + 3, v + 4, v + 5, v + 6,
                      v + 7, v + 8, v + 9, v + 10) == 11) {
      const physics::StokesVector o =
          physics::stokesStep({v[0], v[1], v[2], v[3]}, {v[4], v[5], v[6], v[7]}, v[8], v[9], v[10]);
      std::printf("%.17g %.17g %.17g %.17g\n", o.i, o.q, o.u, o.v);
    }
  } else if (!std::strcmp(what, "stokes_full")) {
    // I Q U V jI jQ jU jV alphaI alphaQ alphaV rhoV rhoQ ds -> stokesStepFull
    double v[14];
    while (std::scanf("%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", v, v + 1, v + 2, v + 3, v + 4,
                      v + 5, v + 6, v + 7, v + 8, v + 9, v + 10, v + 11, v + 12, v + 13) == 14) {
      const physics::StokesVector o = physics::stokesStepFull(
          {v[0], v[1], v[2], v[3]}, {v[4], v[5], v[6], v[7]}, {v[8], v[9], v[10], v[11], v[12], v[13]});
      std::printf("%.17g %.17g %.17g %.17g\n", o.i, o.q, o.u, o.v);
    }
  } else if (!std::strcmp(what, "conserved")) {
    // r theta ut ur utheta uphi mGeom a -> e lz q mu2; the monitor's drifts over the rows
    double v[8];
    physics::GeodesicConservationMonitor mon;
    double last = 0.0;
    while (std::scanf("%lf %lf %lf %lf %lf %lf %lf %lf", v, v + 1, v + 2, v + 3, v + 4, v + 5, v + 6, v + 7) == 8) {
      const physics::ConservedQuantities c = physics::computeConserved(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
      last = mon.update(c);
      std::printf("%.17g %.17g %.17g %.17g\n", c.e, c.lz, c.q, c.mu2);
    }
    std::printf("%.17g %.17g %.17g %.17g %.17g %d %d\n", last, mon.maxEDrift, mon.maxLzDrift, mon.maxQDrift,
                mon.maxMu2Drift, mon.stepCount, mon.violated(1e-8) ? 1 : 0);
  } else if (!std::strcmp(what, "events")) {
    // analytic rays: r = 12 - 0.7 l - 0.01 l^3 through r = 8; r = 10 + (l - 2.3)^2 turning at 2.3
    struct St {
      double x1, x2, v1;
    };
    auto ray1 = [](double, double lb, const St&) {
      return St{12.0 - 0.7 * lb - 0.01 * lb * lb * lb, 1.0 + 0.1 * lb, -0.7 - 0.03 * lb * lb};
    };
    auto ray2 = [](double, double lb, const St&) {
      return St{10.0 + (lb - 2.3) * (lb - 2.3), 1.2 - 0.05 * lb, 2.0 * (lb - 2.3)};
    };
    const physics::EventResult e1 = physics::bisectRadialCrossing(0.0, 10.0, St{12.0, 1.0, -0.7}, 8.0, ray1);
    const physics::EventResult e2 = physics::bisectTurningPoint(0.0, 5.0, St{10.0 + 2.3 * 2.3, 1.2, -4.6}, ray2);
    const physics::EventResult e3 = physics::bisectRadialCrossing(0.0, 10.0, St{12.0, 1.0, -0.7}, 8.0, ray1, 1e-30, 7);
    for (const physics::EventResult& e : {e1, e2, e3})
      std::printf("%d %.17g %.17g %.17g %d %d\n", static_cast<int>(e.type), e.lambda, e.r, e.theta, e.nBisections,
                  e.detected ? 1 : 0);
    std::printf("%d %d %d %d %d %d %d\n", physics::crossedRadius(5.0, 7.0, 6.0), physics::crossedRadius(5.0, 7.0, 8.0),
                physics::hasTurningPoint(-1.0, 0.5), physics::hasTurningPoint(0.2, 0.5),
                physics::crossedEquator(1.5, 1.6), physics::crossedEquator(1.0, 1.5), physics::photonSubringIndex(3));
  } else if (!std::strcmp(what, "hamilton")) {
    // a rS t r theta phi p_t p_r p_theta p_phi -> g^{mu nu} (16), H, velocity (4), force (4)
    double v[10];
    while (std::scanf("%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", v, v + 1, v + 2, v + 3, v + 4, v + 5, v + 6, v + 7,
                      v + 8, v + 9) == 10) {
      const physics::KerrParams k{v[0], v[1]};
      const physics::GeodesicState st{{v[2], v[3], v[4], v[5]}, {v[6], v[7], v[8], v[9]}};
      const physics::Metric4 g = physics::kerrGcon(v[3], v[4], v[0], v[1]);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) std::printf("%.17g ", g[i][j]);
      std::printf("%.17g ", physics::kerrHamiltonian(st, k));
      const auto vel = physics::kerrVelocity(st, k);
      const auto f = physics::kerrForce(st, k);
      for (int i = 0; i < 4; ++i) std::printf("%.17g ", vel[i]);
      for (int i = 0; i < 4; ++i) std::printf("%.17g%s", f[i], i == 3 ? "\n" : " ");
    }
  } else if (!std::strcmp(what, "integrate")) {
    // a rS dLambda n t r theta phi p_t p_r p_theta p_phi -> integrateGeodesic state and H
    double v[12];
    while (std::scanf("%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", v, v + 1, v + 2, v + 3, v + 4, v + 5,
                      v + 6, v + 7, v + 8, v + 9, v + 10, v + 11) == 12) {
      const physics::KerrParams k{v[0], v[1]};
      physics::GeodesicState st{{v[4], v[5], v[6], v[7]}, {v[8], v[9], v[10], v[11]}};
      physics::integrateGeodesic(st, v[2], static_cast<int>(v[3]), k);
      for (int i = 0; i < 4; ++i) std::printf("%.17g ", st.q[i]);
      for (int i = 0; i < 4; ++i) std::printf("%.17g ", st.p[i]);
      std::printf("%.17g\n", physics::kerrHamiltonian(st, k));
    }
  } else if (!std::strcmp(what, "batch")) {
    // rows a rS dLambda n t r theta phi p_t p_r p_theta p_phi (a, rS, dLambda, n from the first
    // row) -> integrateGeodesicBatch on the GPU, one line per state (status -1 line on error)
    double v[12];
    std::vector<physics::GeodesicState> states;
    double par[4] = {0.0, 0.0, 0.0, 0.0};
    while (std::scanf("%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", v, v + 1, v + 2, v + 3, v + 4, v + 5,
                      v + 6, v + 7, v + 8, v + 9, v + 10, v + 11) == 12) {
      if (states.empty())
        for (int i = 0; i < 4; ++i) par[i] = v[i];
      states.push_back(physics::GeodesicState{{v[4], v[5], v[6], v[7]}, {v[8], v[9], v[10], v[11]}});
    }
    const physics::KerrParams k{par[0], par[1]};
    if (!physics::integrateGeodesicBatch(states, par[2], static_cast<int>(par[3]), k)) {
      std::printf("-1 %s\n", bh_last_error());
      return 1;
    }
    for (const auto& st : states) {
      for (int i = 0; i < 4; ++i) std::printf("%.17g ", st.q[i]);
      for (int i = 0; i < 4; ++i) std::printf("%.17g%s", st.p[i], i == 3 ? "\n" : " ");
    }
  } else if (!std::strcmp(what, "yoshida")) {
    // a rS h n t r theta phi p_t p_r p_theta p_phi -> the state after n steps and H
    double v[12];
    while (std::scanf("%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", v, v + 1, v + 2, v + 3, v + 4, v + 5,
                      v + 6, v + 7, v + 8, v + 9, v + 10, v + 11) == 12) {
      const physics::KerrParams k{v[0], v[1]};
      physics::GeodesicState st{{v[4], v[5], v[6], v[7]}, {v[8], v[9], v[10], v[11]}};
      const int n = static_cast<int>(v[3]);
      for (int i = 0; i < n; ++i) physics::yoshida4Step(st, v[2], k);
      for (int i = 0; i < 4; ++i) std::printf("%.17g ", st.q[i]);
      for (int i = 0; i < 4; ++i) std::printf("%.17g ", st.p[i]);
      std::printf("%.17g\n", physics::kerrHamiltonian(st, k));
    }
  } else if (!std::strcmp(what, "stokes_path")) {
    // first line I Q U V; then segments jI jQ jU jV alphaI rhoV ds -> the state after each
    // stokesStep, then integrateStokesPath over all segments
    physics::StokesVector s0;
    if (std::scanf("%lf %lf %lf %lf", &s0.i, &s0.q, &s0.u, &s0.v) != 4) return 1;
    std::vector<physics::StokesEmission> em;
    std::vector<physics::FaradayPropagation> pr;
    physics::StokesEmission e;
    physics::FaradayPropagation k;
    while (std::scanf("%lf %lf %lf %lf %lf %lf %lf", &e.jI, &e.jQ, &e.jU, &e.jV, &k.alphaI, &k.rhoV, &k.dsCm) == 7) {
      em.push_back(e);
      pr.push_back(k);
    }
    physics::StokesVector s = s0;
    for (size_t i = 0; i < em.size(); ++i) {
      s = physics::stokesStep(s, em[i], pr[i].alphaI, pr[i].rhoV, pr[i].dsCm);
      std::printf("%.17g %.17g %.17g %.17g\n", s.i, s.q, s.u, s.v);
    }
    const physics::StokesVector f = physics::integrateStokesPath(em, pr, s0);
    std::printf("%.17g %.17g %.17g %.17g\n", f.i, f.q, f.u, f.v);
  } else if (!std::strcmp(what, "stokes_misc")) {
    // one line of scalar functions and vector helpers
    const physics::StokesVector s{2.0, 0.6, 0.8, 0.3};
    const physics::StokesVector r = s.rotateEvpa(0.3), f = physics::faradayRotateStokes(s, 0.3);
    const physics::StokesVector g = physics::grTransformStokes(s, 0.8), z = physics::grTransformStokes(s, -1.0);
    const physics::StokesEmission e = physics::synchrotronPolarizedEmission(2.0, 0.7, 0.4);
    std::printf("%.17g %.17g %.17g %.17g %.17g\n", s.linPolInt(), s.linPolFrac(), s.circPolFrac(), s.totalPolFrac(),
                s.evpa());
    std::printf("%.17g %.17g %.17g %.17g\n", r.i, r.q, r.u, r.v);
    std::printf("%.17g %.17g %.17g %.17g\n", f.i, f.q, f.u, f.v);
    std::printf("%.17g %.17g %.17g %.17g\n", g.i, g.q, g.u, g.v);
    std::printf("%.17g %.17g %.17g %.17g\n", z.i, z.q, z.u, z.v);
    std::printf("%.17g %.17g %.17g %.17g\n", e.jI, e.jQ, e.jU, e.jV);
    std::printf("%.17g %.17g %.17g %.17g\n", physics::synchrotronLinearPolarFrac(2.0),
                physics::synchrotronLinearPolarFrac(3.0), physics::thermalSynchLinearPolarFrac(1.0),
                physics::thermalSynchLinearPolarFrac(10.0));
    std::printf("%d %d %d %.17g\n", physics::stokesPolarizationBound(s) ? 1 : 0,
                physics::stokesPolarizationBound({1.0, 0.9, 0.5, 0.0}) ? 1 : 0,
                physics::stokesPolarizationBound({1.0, 1.0, 0.0, 0.0}) ? 1 : 0,
                physics::accumulatedRM({0.5, 1.5, -0.25}, {2.0, 1.0, 4.0}));
  } else if (!std::strcmp(what, "faraday")) {
    // nu ne B theta -> cold, relativistic, conversion
    double nu, ne, b, th;
    while (std::scanf("%lf %lf %lf %lf", &nu, &ne, &b, &th) == 4)
      std::printf("%.17g %.17g %.17g\n", physics::faradayRotationCoeff(nu, ne, b),
                  physics::faradayRotationCoeffRelativistic(nu, ne, b, th),
                  physics::faradayConversionCoeff(nu, ne, b, th));
  } else if (!std::strcmp(what, "rayspec")) {
    double a, alpha, beta, th, r;
    while (std::scanf("%lf %lf %lf %lf %lf", &a, &alpha, &beta, &th, &r) == 5) {
      const double xi = -alpha * std::sin(th);
      const double eta = beta * beta + (alpha * alpha - a * a) * std::cos(th) * std::cos(th);
      const physics::KerrGeodesicConsts c{1.0, xi, eta};
      const physics::KerrGeodesicState s{r, th, 0.0, 0.0, -1.0, beta >= 0 ? -1.0 : 1.0};
      BH_RaySpec sp;
      const bool ok = physics::raySpecFor(s, c, 1.0, a, &sp);
      std::printf("%d %.17g %.17g %.17g %.17g %.17g\n", ok ? 1 : 0, sp.alpha, sp.beta, sp.spin, sp.theta_obs, sp.r_obs);
    }
  }
  return 0;
}
