// sur_features.hpp - analytic input embedding of the geodesic surrogate,
// shared by the training driver (surrogate.hip) and the fused inference
// kernel (sur_infer.hip) so both see bit-identical features.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/bhk_surrogate.h"

namespace bhk {

// Observer distance the surrogate is trained for (M).
constexpr float kSurRobs = 100.0f;

// First equatorial crossing radius of the ray through image point
// (alpha, beta) for an observer at distance D, from Beloborodov's
// light-bending relation 1 - cos(a_e) = (1 - cos psi)(1 - 2/r) (Schwarzschild,
// A. M. Beloborodov 2002, ApJ 566 L85).  With the conserved impact parameter
// b = r sin(a_e) / sqrt(1 - 2/r) the relation is a quadratic in r,
//     k (2 - k) r^2 + 2 k^2 r = b^2,   k = 1 - cos psi,
// solved as r = b^2 / (k^2 + sqrt(k^4 + k (2 - k) b^2)).  psi is the angle
// between the crossing point and the photon's asymptotic direction: the angle
// psi_o between the crossing point and the observer (the crossing point lies
// in the plane of the line of sight and the image direction, where its z
// component vanishes) plus the angle psi_D the photon still turns between r
// = D and infinity (the same relation at r = D, where it moves outward).
__host__ __device__ inline float bel_radius(float alpha, float beta, float sth, float cth, float D) {
  const float b2 = alpha * alpha + beta * beta;
  const float b = sqrtf(b2);
  const float s = beta / fmaxf(b, 1e-6f);  // sine of the image polar angle
  const float den = sqrtf(cth * cth + sth * sth * s * s);
  const float psi_o = atan2f(cth / den, -sth * s / den);
  const float u = 1.0f - 2.0f / D;
  const float sa = fminf(b * sqrtf(u) / D, 1.0f);
  const float ca = sqrtf(1.0f - sa * sa);
  const float kD = (1.0f - ca) / u;
  const float psi_D = acosf(fmaxf(1.0f - kD, -1.0f));
  const float psi = fminf(psi_o + psi_D, 3.14159265f);
  const float k = 1.0f - cosf(psi);
  return b2 / (k * k + sqrtf(k * k * k * k + k * (2.0f - k) * b2));
}

// Embedding of (alpha, beta, a, cos theta_obs) into BH_SUR_FEATURES features:
// scaled inputs, the Bardeen constants xi, eta, polar coordinates of the
// image point, products with the spin, and the logarithms of two estimates of
// the first crossing radius: feature 14 = ln(r_bel)/3 (the light-bending
// estimate above, the base of the ln r outputs) and feature 15 =
// ln(r_flat)/3 with the flat-space r_flat = sqrt(alpha^2 + (beta/cos)^2).
__host__ __device__ inline void surrogate_features(float alpha, float beta, float a, float cth, float* f) {
  const float sth = sqrtf(fmaxf(1.0f - cth * cth, 0.0f));
  const float rho = sqrtf(alpha * alpha + beta * beta);
  const float xi = -alpha * sth;
  const float eta = beta * beta + (alpha * alpha - a * a) * cth * cth;
  const float inv = 1.0f / fmaxf(rho, 1e-6f);
  f[0] = alpha * 0.05f;
  f[1] = beta * 0.05f;
  f[2] = a;
  f[3] = cth;
  f[4] = sth;
  f[5] = xi * 0.05f;
  f[6] = eta * 0.0025f;
  f[7] = rho * 0.05f;
  f[8] = (rho - 5.196152422706632f) * 0.2f;
  f[9] = beta * inv;
  f[10] = alpha * inv;
  f[11] = a * cth;
  f[12] = a * sth;
  f[13] = 1.0f / (1.0f + rho * rho * 0.04f);
  f[14] = logf(fmaxf(bel_radius(alpha, beta, sth, cth, kSurRobs), 0.05f)) * 0.333333333f;
  const float cc = fmaxf(cth, 0.05f);
  const float rflat = sqrtf(alpha * alpha + (beta / cc) * (beta / cc));
  f[15] = logf(fmaxf(rflat, 0.05f)) * 0.333333333f;
}

// Feature whose 3x is the base of the ln r outputs.
constexpr int kSurBaseFeature = 14;

}  // namespace bhk
