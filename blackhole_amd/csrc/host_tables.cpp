// host_tables.cpp - host-side builders of the tracer's input tables:
// the thin-disk emissivity LUT, the thermal-synchrotron K_2 table and the
// synthetic GRMHD-like torus volume.  One-time inputs (uploaded once, not
// part of the timed frame).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bhk_trace.h"

namespace bhk {
int fail(const std::string& msg);
}

namespace {

constexpr double kPi = 3.14159265358979323846;

// K_nu(z) = int_0^inf exp(-z cosh t) cosh(nu t) dt (Watson 1944, 6.22(5)),
// trapezoid rule with step 0.01 (the integrand is analytic and decays
// double-exponentially, so the rule converges exponentially).
double bessel_k(double nu, double z) {
  const double dt = 0.01;
  double s = 0.5 * std::exp(-z);
  for (int i = 1; i < 20000; ++i) {
    const double t = i * dt;
    const double ch = std::cosh(t);
    if (z * ch > 745.0) break;
    s += std::exp(-z * ch) * std::cosh(nu * t);
  }
  return s * dt;
}

// SplitMix64 (Steele, Lea & Flood 2014).
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (static_cast<double>(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
};

// In-place iterative radix-2 FFT over n (power of two) elements with stride.
void fft1d(std::complex<double>* a, int n, long long stride, bool inverse) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i * stride], a[j * stride]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    const double ang = 2.0 * kPi / len * (inverse ? 1.0 : -1.0);
    const std::complex<double> wl(std::cos(ang), std::sin(ang));
    for (int i = 0; i < n; i += len) {
      std::complex<double> w(1.0, 0.0);
      for (int k = 0; k < len / 2; ++k) {
        const std::complex<double> u = a[(i + k) * stride];
        const std::complex<double> v = a[(i + k + len / 2) * stride] * w;
        a[(i + k) * stride] = u + v;
        a[(i + k + len / 2) * stride] = u - v;
        w *= wl;
      }
    }
  }
}

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// Fishbone & Moncrief (1976) constant-l torus in Kerr (M = 1):
// l = u_phi u^t of the prograde circular geodesic at r_max,
// ln h(r, theta) = 1/2 ln[(1 + sqrt(1 + 4 l^2 S^2 D / (A^2 sin^2))) / (S D / A)]
//                - 1/2 sqrt(1 + 4 l^2 S^2 D / (A^2 sin^2)) - 2 a r l / A
// minus its value at (r_in, pi/2); S = Sigma, D = Delta, A = (r^2+a^2)^2 - D a^2 sin^2.
double fm_l(double a, double r) {
  const double sr = std::sqrt(r);
  const double r32 = r * sr;
  return (r * r - 2.0 * a * sr + a * a) * (r32 + a) / (r32 * (r32 - 3.0 * sr + 2.0 * a));
}
double fm_lnh_raw(double a, double l, double r, double th) {
  const double s = std::sin(th), c = std::cos(th);
  const double S = r * r + a * a * c * c;
  const double D = r * r - 2.0 * r + a * a;
  const double A = (r * r + a * a) * (r * r + a * a) - D * a * a * s * s;
  const double q = std::sqrt(1.0 + 4.0 * l * l * S * S * D / (A * A * s * s));
  return 0.5 * std::log((1.0 + q) / (S * D / A)) - 0.5 * q - 2.0 * a * r * l / A;
}

}  // namespace

extern "C" {

int bh_disk_lut_build(double a, double r_in, double r_out, int n, double* out) {
  if (!out || n < 2 || !(r_out > r_in) || !(r_in > 0.0))
    return bhk::fail("bh_disk_lut_build: invalid arguments");
  double mx = 0.0;
  for (int i = 0; i < n; ++i) {
    const double u = static_cast<double>(i) / static_cast<double>(n - 1);
    const double r = r_in + u * (r_out - r_in);
    const double basic = 1.0 - std::sqrt(r_in / r);
    const double spin = 1.0 + 0.5 * a * std::sqrt(1.0 / r);
    out[i] = (r < r_in) ? 0.0 : basic * spin / (r * r * r);
    mx = std::max(mx, out[i]);
  }
  if (mx > 0.0)
    for (int i = 0; i < n; ++i) out[i] /= mx;
  return 0;
}

int bh_volume_generate_torus(const BH_TorusParams* p, float* cells, BH_Volume* vol) {
  if (!p || !cells || !vol) return bhk::fail("bh_volume_generate_torus: NULL argument");
  const int nr = p->n_r, nt = p->n_th, np = p->n_ph;
  if (!pow2(nr) || !pow2(nt) || !pow2(np))
    return bhk::fail("bh_volume_generate_torus: grid sizes must be powers of two");
  if (!(p->spin >= 0.0 && p->spin < 1.0) || !(p->r_max > p->r_in) || !(p->r_hi > p->r_lo) ||
      !(p->r_in > 1.0 + std::sqrt(1.0 - p->spin * p->spin)))
    return bhk::fail("bh_volume_generate_torus: invalid torus geometry");
  const double a = p->spin;
  const double l = fm_l(a, p->r_max);
  const double lnh_in = fm_lnh_raw(a, l, p->r_in, 0.5 * kPi);
  const double gam = p->gamma_ad;
  const double dr = (p->r_hi - p->r_lo) / nr;
  const long long N = static_cast<long long>(nr) * nt * np;

  // Gaussian random field with P(k) ~ |k|^spectral_index on the periodic
  // index grid (spectral synthesis), normalised to zero mean, unit variance.
  std::vector<std::complex<double>> F(static_cast<size_t>(N));
  SplitMix64 rng{p->seed};
  for (int ip = 0; ip < np; ++ip) {
    const double kp = (ip <= np / 2) ? ip : ip - np;
    for (int it = 0; it < nt; ++it) {
      const double kt = (it <= nt / 2) ? it : it - nt;
      for (int ir = 0; ir < nr; ++ir) {
        const double kr = (ir <= nr / 2) ? ir : ir - nr;
        const double k2 = kr * kr + kt * kt + kp * kp;
        const double u1 = rng.uniform(), u2 = rng.uniform();
        const double rad = std::sqrt(-2.0 * std::log(u1));
        const std::complex<double> z(rad * std::cos(2.0 * kPi * u2), rad * std::sin(2.0 * kPi * u2));
        const double amp = (k2 > 0.0) ? std::pow(k2, 0.25 * p->spectral_index) : 0.0;
        F[(static_cast<size_t>(ip) * nt + it) * nr + ir] = amp * z;
      }
    }
  }
  for (int ip = 0; ip < np; ++ip)
    for (int it = 0; it < nt; ++it) fft1d(&F[(static_cast<size_t>(ip) * nt + it) * nr], nr, 1, true);
  for (int ip = 0; ip < np; ++ip)
    for (int ir = 0; ir < nr; ++ir) fft1d(&F[static_cast<size_t>(ip) * nt * nr + ir], nt, nr, true);
  for (int it = 0; it < nt; ++it)
    for (int ir = 0; ir < nr; ++ir)
      fft1d(&F[static_cast<size_t>(it) * nr + ir], np, static_cast<long long>(nt) * nr, true);
  double mean = 0.0, var = 0.0;
  for (long long i = 0; i < N; ++i) mean += F[i].real();
  mean /= N;
  for (long long i = 0; i < N; ++i) {
    const double d = F[i].real() - mean;
    var += d * d;
  }
  var /= N;
  const double inv_sd = var > 0.0 ? 1.0 / std::sqrt(var) : 0.0;

  // Torus profile: h = exp(ln h) inside (ln h > 0, r > r_in);
  // rho ~ (h - 1)^(1/(gamma-1)), p ~ rho^gamma, Theta_e ~ p/rho ~ (h - 1),
  // |B| ~ sqrt(p L) (constant plasma beta on the turbulent density).
  std::vector<double> hm1(static_cast<size_t>(nr) * nt, 0.0);
  double hm1_max = 0.0;
  for (int it = 0; it < nt; ++it) {
    const double th = (it + 0.5) * kPi / nt;
    for (int ir = 0; ir < nr; ++ir) {
      const double r = p->r_lo + (ir + 0.5) * dr;
      if (r < p->r_in) continue;
      const double lnh = fm_lnh_raw(a, l, r, th) - lnh_in;
      if (!(lnh > 0.0)) continue;
      const double v = std::exp(lnh) - 1.0;
      hm1[static_cast<size_t>(it) * nr + ir] = v;
      hm1_max = std::max(hm1_max, v);
    }
  }
  if (!(hm1_max > 0.0)) return bhk::fail("bh_volume_generate_torus: empty torus");
  const double b_max = p->b_max;
  // K_2(1/Theta_e) / (2 Theta_e^2) per (r, theta) cell (Theta_e does not
  // depend on phi): the thermal-synchrotron normalisation, stored in the
  // fourth lane of each cell.
  std::vector<double> k2r(static_cast<size_t>(nr) * nt, 0.0);
  for (size_t i = 0; i < k2r.size(); ++i) {
    const double x = hm1[i] / hm1_max;
    if (x > 0.0) {
      const double th_e = p->theta_e_max * x;
      k2r[i] = bessel_k(2.0, 1.0 / th_e) / (2.0 * th_e * th_e);
    }
  }
  const double s = p->sigma;
  double sup_lo = 1e300, sup_hi = -1e300, sup_u = 0.0;
  for (int ip = 0; ip < np; ++ip) {
    for (int it = 0; it < nt; ++it) {
      for (int ir = 0; ir < nr; ++ir) {
        const size_t ci = (static_cast<size_t>(ip) * nt + it) * nr + ir;
        const double x = hm1[static_cast<size_t>(it) * nr + ir] / hm1_max;  // (h-1)/(h-1)_max
        float* c = cells + 4 * ci;
        if (x <= 0.0) {
          c[0] = c[1] = c[2] = c[3] = 0.0f;
          continue;
        }
        const double rho = std::pow(x, 1.0 / (gam - 1.0));
        const double pr = std::pow(x, gam / (gam - 1.0));
        const double G = (F[ci].real() - mean) * inv_sd;
        const double L = std::exp(s * G - 0.5 * s * s);
        c[0] = static_cast<float>(p->n_e_max * rho * L);
        c[1] = static_cast<float>(p->theta_e_max * x);
        c[2] = static_cast<float>(b_max * std::sqrt(pr * L));
        c[3] = static_cast<float>(k2r[static_cast<size_t>(it) * nr + ir]);
        const double r_lo = p->r_lo + ir * dr, r_hi = r_lo + dr;
        sup_lo = std::min(sup_lo, r_lo);
        sup_hi = std::max(sup_hi, r_hi);
        sup_u = std::max(sup_u, std::max(std::fabs(std::cos(it * kPi / nt)),
                                         std::fabs(std::cos((it + 1) * kPi / nt))));
      }
    }
  }
  // Trilinear samples see one cell beyond the support: widen by one cell.
  vol->n_r = nr;
  vol->n_th = nt;
  vol->n_ph = np;
  vol->pad0 = 0;
  vol->r_min = p->r_lo;
  vol->r_max = p->r_hi;
  vol->sup_r_lo = std::max(p->r_lo, sup_lo - dr);
  vol->sup_r_hi = std::min(p->r_hi, sup_hi + dr);
  vol->sup_u_max = std::min(1.0, sup_u + std::sin(kPi / nt));
  // Fluid rotation: constant ell = -u_phi/u_t of the circular geodesic at r_max.
  const double sr = std::sqrt(p->r_max), r32 = p->r_max * sr;
  vol->ell = (p->r_max * p->r_max - 2.0 * a * sr + a * a) / (r32 - 2.0 * sr + a);
  return 0;
}

}  // extern "C"
