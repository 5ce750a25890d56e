// ref_wrap_disk.cpp - extern "C" wrapper of the reference's thin-disk flux
// (src/physics/thin_disk.h), in a translation unit of its own: thin_disk.h
// and rte_integrator.h both define planckFunction, with different exception
// specifications, so the two headers cannot share one.  Test infrastructure
// only; built by oracle/build_ref.sh into oracle/_ref/libref.so.
#include <cmath>

#include "physics/thin_disk.h"

namespace {
constexpr double kMassM1 = physics::C2 / physics::G;
}  // namespace

extern "C" {

// thin_disk.h diskFlux at r (units of M) for spin a (units of M) and a disk
// over [r_in, r_out] (units of M), M = 1 in grams; flux in erg / (cm^2 s)
// for an accretion rate of 1 g/s.  Returns 0 outside the disk.
double ref_disk_flux(double r, double a, double r_in, double r_out) {
  const double rg = physics::G * kMassM1 / physics::C2;
  physics::DiskParams d;
  d.mass = kMassM1;
  d.mDot = 1.0;
  d.a = a * rg;
  d.rIn = r_in * rg;
  d.rOut = r_out * rg;
  return physics::diskFlux(r * rg, d);
}


}  // extern "C"
