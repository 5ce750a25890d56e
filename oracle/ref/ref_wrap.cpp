// ref_wrap.cpp - extern "C" wrapper exposing a few of the reference's own
// FP64 physics functions (compiled from /root/reference/src/physics) so the
// oracle can be checked against them.  Test infrastructure only; built by
// oracle/build_ref.sh into oracle/_ref/libref.so.
#include <cmath>

#include "physics/analytic_kerr_geodesic.h"
#include "physics/iron_kline.h"
#include "physics/kerr.h"
#include "physics/rte_integrator.h"
#include "physics/stokes_transport.h"
#include "physics/symplectic_integrator.h"
#include "physics/event_detection.h"
#include "physics/conservation_monitor.h"
#include "physics/verified/geodesic.hpp"

namespace {
// kerr.h works in CGS with mass in grams; M = 1 (geometric) is mass = c^2/G.
constexpr double kMassM1 = physics::C2 / physics::G;
}  // namespace

extern "C" {

void ref_kerr_potentials(double r, double theta, double a, double e, double lz, double q,
                         double* out4) {
  physics::KerrGeodesicConsts c{e, lz, q};
  const physics::KerrPotentials p = physics::kerrPotentials(r, theta, kMassM1, a, c);
  out4[0] = p.rPot;
  out4[1] = p.dRdr;
  out4[2] = p.thetaPot;
  out4[3] = p.dThetadtheta;
}

double ref_carter_constant(double theta, double a, double e, double lz, double ptheta) {
  return verified::carter_constant(theta, a, e, lz, ptheta);
}

void ref_critical_params(double r, double a, double* xi_eta) {
  const physics::ImpactParams ip = physics::criticalImpactParams(r, a);
  xi_eta[0] = ip.xi;
  xi_eta[1] = ip.eta;
}

double ref_photon_orbit(double a, int prograde) {
  return prograde ? physics::progradePhotonOrbit(a) : physics::retrogradePhotonOrbit(a);
}

double ref_isco_geom(double a) { return physics::iscoGeom(a); }

double ref_kerr_isco(double a, int prograde) {
  return physics::kerrIscoRadius(kMassM1, a * physics::G * kMassM1 / physics::C2, prograde != 0) /
         (physics::G * kMassM1 / physics::C2);
}

double ref_outer_horizon(double a) {
  const double m = physics::G * kMassM1 / physics::C2;
  return physics::kerrOuterHorizon(kMassM1, a * m) / m;
}

double ref_disk_g(double r, double phi, double a, double iota) {
  return physics::kerrDiskGFactor(r, phi, a, iota);
}

void ref_rte_step(double* state3, double j, double alpha, double ds) {
  physics::RteState s{state3[0], state3[1], state3[2]};
  s = physics::rteStep(s, j, alpha, ds);
  state3[0] = s.iNu;
  state3[1] = s.tau;
  state3[2] = s.sCm;
}

double ref_planck(double nu, double T) { return physics::planckFunction(nu, T); }

// rte_integrator.h:465-515: frame transforms of I, j and alpha by g.
double ref_gr_transform_intensity(double i_emit, double g) { return physics::grTransformIntensity(i_emit, g); }
double ref_gr_transform_emission(double j_emit, double g) { return physics::grTransformEmission(j_emit, g); }
double ref_gr_transform_absorption(double a_emit, double g) { return physics::grTransformAbsorption(a_emit, g); }

double ref_synch_j(double nu, double B, double ne, double th) {
  return physics::synchrotronThermalEmissivity(nu, B, ne, th);
}

double ref_synch_alpha(double nu, double B, double ne, double th) {
  return physics::synchrotronThermalAbsorption(nu, B, ne, th);
}

void ref_kerr_step_mino(double a, double e, double lz, double q, double* state6, double dlam) {
  physics::KerrGeodesicConsts c{e, lz, q};
  physics::KerrGeodesicState s{state6[0], state6[1], state6[2], state6[3], state6[4], state6[5]};
  s = physics::kerrStepMino(s, kMassM1, a, c, dlam);
  state6[0] = s.r;
  state6[1] = s.theta;
  state6[2] = s.phi;
  state6[3] = s.t;
  state6[4] = s.signR;
  state6[5] = s.signTheta;
}

// stokes_transport.h: s4 / em4 = (I, Q, U, V) / (jI, jQ, jU, jV).
void ref_stokes_step(const double* s4, const double* em4, double alpha_i, double rho_v, double ds,
                     double* out4) {
  const physics::StokesVector s{s4[0], s4[1], s4[2], s4[3]};
  const physics::StokesEmission em{em4[0], em4[1], em4[2], em4[3]};
  const physics::StokesVector o = physics::stokesStep(s, em, alpha_i, rho_v, ds);
  out4[0] = o.i;
  out4[1] = o.q;
  out4[2] = o.u;
  out4[3] = o.v;
}

// k6 = (alphaI, alphaQ, alphaV, rhoV, rhoQ, ds).
void ref_stokes_step_full(const double* s4, const double* em4, const double* k6, double* out4) {
  const physics::StokesVector s{s4[0], s4[1], s4[2], s4[3]};
  const physics::StokesEmission em{em4[0], em4[1], em4[2], em4[3]};
  const physics::FaradayPropagation k{k6[0], k6[1], k6[2], k6[3], k6[4], k6[5]};
  const physics::StokesVector o = physics::stokesStepFull(s, em, k);
  out4[0] = o.i;
  out4[1] = o.q;
  out4[2] = o.u;
  out4[3] = o.v;
}

void ref_synch_pol_emission(double j_i, double pi_lin, double chi_b, double* out4) {
  const physics::StokesEmission e = physics::synchrotronPolarizedEmission(j_i, pi_lin, chi_b);
  out4[0] = e.jI;
  out4[1] = e.jQ;
  out4[2] = e.jU;
  out4[3] = e.jV;
}

double ref_faraday_rotation_coeff(double nu, double ne, double b) {
  return physics::faradayRotationCoeff(nu, ne, b);
}

double ref_synch_lin_pol_frac(double p) { return physics::synchrotronLinearPolarFrac(p); }

double ref_thermal_lin_pol_frac(double theta_e) { return physics::thermalSynchLinearPolarFrac(theta_e); }

// conservation_monitor.h: out4 = (e, lz, q, mu2).
void ref_compute_conserved(double r, double theta, double ut, double ur, double uth, double uph, double m,
                           double a, double* out4) {
  const physics::ConservedQuantities c = physics::computeConserved(r, theta, ut, ur, uth, uph, m, a);
  out4[0] = c.e;
  out4[1] = c.lz;
  out4[2] = c.q;
  out4[3] = c.mu2;
}

// event_detection.h on the analytic rays r = 12 - 0.7 l - 0.01 l^3 (through
// r = 8) and r = 10 + (l - 2.3)^2 (turning at 2.3); out = 3 x (type, lambda,
// r, theta, nBisections, detected) then the predicates and the subring index.
void ref_events(double* out) {
  struct St {
    double x1, x2, v1;
  };
  auto ray1 = [](double, double lb, const St&) {
    return St{12.0 - 0.7 * lb - 0.01 * lb * lb * lb, 1.0 + 0.1 * lb, -0.7 - 0.03 * lb * lb};
  };
  auto ray2 = [](double, double lb, const St&) {
    return St{10.0 + (lb - 2.3) * (lb - 2.3), 1.2 - 0.05 * lb, 2.0 * (lb - 2.3)};
  };
  const physics::EventResult e[3] = {
      physics::bisectRadialCrossing(0.0, 10.0, St{12.0, 1.0, -0.7}, 8.0, ray1),
      physics::bisectTurningPoint(0.0, 5.0, St{10.0 + 2.3 * 2.3, 1.2, -4.6}, ray2),
      physics::bisectRadialCrossing(0.0, 10.0, St{12.0, 1.0, -0.7}, 8.0, ray1, 1e-30, 7)};
  for (int i = 0; i < 3; ++i) {
    out[6 * i] = static_cast<double>(static_cast<int>(e[i].type));
    out[6 * i + 1] = e[i].lambda;
    out[6 * i + 2] = e[i].r;
    out[6 * i + 3] = e[i].theta;
    out[6 * i + 4] = e[i].nBisections;
    out[6 * i + 5] = e[i].detected ? 1.0 : 0.0;
  }
  out[18] = physics::crossedRadius(5.0, 7.0, 6.0);
  out[19] = physics::crossedRadius(5.0, 7.0, 8.0);
  out[20] = physics::hasTurningPoint(-1.0, 0.5);
  out[21] = physics::hasTurningPoint(0.2, 0.5);
  out[22] = physics::crossedEquator(1.5, 1.6);
  out[23] = physics::crossedEquator(1.0, 1.5);
  out[24] = physics::photonSubringIndex(3);
}

// symplectic_integrator.h: st8 = (t, r, theta, phi, p_t, p_r, p_theta, p_phi).
static physics::GeodesicState ref_state(const double* st8) {
  physics::GeodesicState s;
  for (int i = 0; i < 4; ++i) {
    s.q[i] = st8[i];
    s.p[i] = st8[4 + i];
  }
  return s;
}

void ref_kerr_gcon(double r, double theta, double a, double rS, double* out16) {
  const physics::Metric4 g = physics::kerrGcon(r, theta, a, rS);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out16[4 * i + j] = g[i][j];
}

double ref_kerr_hamiltonian(const double* st8, double a, double rS) {
  return physics::kerrHamiltonian(ref_state(st8), physics::KerrParams{a, rS});
}

void ref_kerr_force(const double* st8, double a, double rS, double* out4) {
  const std::array<double, 4> f = physics::kerrForce(ref_state(st8), physics::KerrParams{a, rS});
  for (int i = 0; i < 4; ++i) out4[i] = f[i];
}

void ref_yoshida4(double* st8, double a, double rS, double h, int n) {
  physics::GeodesicState s = ref_state(st8);
  for (int k = 0; k < n; ++k) physics::yoshida4Step(s, h, physics::KerrParams{a, rS});
  for (int i = 0; i < 4; ++i) {
    st8[i] = s.q[i];
    st8[4 + i] = s.p[i];
  }
}

}  // extern "C"
