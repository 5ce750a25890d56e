// symplectic.hip - the Hamiltonian form of the null geodesic on the GPU
// (bh_symplectic.h): batches of phase-space states (q, p) advanced by the
// fourth-order Yoshida composition of drifts and kicks, one state per lane.
//
// The device functions perform the operations of the host interface
// (physics::kerrGcon, detail::kerr_2h, kerrVelocity, kerrForce, applyDrift,
// applyKick, yoshida4Step in bh_physics.hpp; after
// src/physics/symplectic_integrator.h:40-311 and the inverse metric of
// src/physics/connection.h:127-178) in the same order (the library is built
// with -ffp-contract=off); the composition weights are computed on the host
// and passed in, so only sin / cos can differ from the host by an ulp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../include/bh_symplectic.h"

namespace bhk {
int fail(const std::string& msg);
}  // namespace bhk

using bhk::fail;

namespace {

struct Gcon {
  double g00, g03, g11, g22, g33;
};

// Contravariant Boyer-Lindquist metric with the reference's guards (Sigma,
// Delta and A at least 1e-30 in magnitude, g^{phi phi} = 0 on the axis).
__device__ __forceinline__ Gcon gcon(double r, double theta, double a, double rS) {
  const double s = sin(theta), c = cos(theta);
  const double s2 = s * s, a2 = a * a, r2a2 = r * r + a2;
  double sig = r * r + a2 * c * c, del = r * r - rS * r + a2;
  double A = r2a2 * r2a2 - a2 * del * s2;
  if (fabs(sig) < 1e-30) sig = 1e-30;
  if (fabs(del) < 1e-30) del = del >= 0.0 ? 1e-30 : -1e-30;
  if (fabs(A) < 1e-30) A = 1e-30;
  const double ds = del * sig;
  Gcon g;
  g.g00 = -A / ds;
  g.g03 = -rS * r * a / ds;
  g.g11 = del / sig;
  g.g22 = 1.0 / sig;
  g.g33 = fabs(s2) > 1e-30 ? (del - a2 * s2) / (ds * s2) : 0.0;
  return g;
}

// d(2H)/dr and d(2H)/dtheta of Carter's separated form
//   2 Sigma H = Delta p_r^2 - P^2 / Delta + p_theta^2 + W^2,
//   P = (r^2 + a^2) E - a L,  W = L / sin(theta) - a E sin(theta).
__device__ __forceinline__ void grad_2h(const double* q, const double* p, double a, double rS, double* d_r,
                                        double* d_th) {
  const double r = q[1], th = q[2];
  const double E = -p[0], pr = p[1], pth = p[2], L = p[3];
  const double s = sin(th), c = cos(th), a2 = a * a;
  const double sig = r * r + a2 * c * c, del = r * r - rS * r + a2;
  const double P = (r * r + a2) * E - a * L, W = L / s - a * E * s;
  const double Rp = del * pr * pr - P * P / del;
  const double Tp = pth * pth + W * W;
  const double dRp = (2.0 * r - rS) * (pr * pr + P * P / (del * del)) - 4.0 * r * E * P / del;
  const double dTp = -2.0 * W * c * (L / (s * s) + a * E);
  *d_r = dRp / sig - (Rp + Tp) * 2.0 * r / (sig * sig);
  *d_th = dTp / sig + (Rp + Tp) * 2.0 * a2 * c * s / (sig * sig);
}

// q += h dH/dp = h g^{mu nu} p_nu
__device__ __forceinline__ void drift(double* q, const double* p, double h, double a, double rS) {
  const Gcon g = gcon(q[1], q[2], a, rS);
  const double v0 = g.g00 * p[0] + g.g03 * p[3];
  const double v1 = g.g11 * p[1];
  const double v2 = g.g22 * p[2];
  const double v3 = g.g03 * p[0] + g.g33 * p[3];
  q[0] += h * v0;
  q[1] += h * v1;
  q[2] += h * v2;
  q[3] += h * v3;
}

// p += h (-dH/dq); t and phi are cyclic (zero force components)
__device__ __forceinline__ void kick(const double* q, double* p, double h, double a, double rS) {
  double dr, dth;
  grad_2h(q, p, a, rS, &dr, &dth);
  const double f1 = -0.5 * dr, f2 = -0.5 * dth;
  p[0] += h * 0.0;
  p[1] += h * f1;
  p[2] += h * f2;
  p[3] += h * 0.0;
}

struct Weights {
  double c1, c2, w0, w1;
};

__global__ void __launch_bounds__(256) k_yoshida4(double* __restrict__ states, long long n, double h, int n_steps,
                                                  double a, double rS, Weights w) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* st = states + i * 8;
  double q[4], p[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    q[k] = st[k];
    p[k] = st[4 + k];
  }
  for (int s = 0; s < n_steps; ++s) {
    drift(q, p, w.c1 * h, a, rS);
    kick(q, p, w.w1 * h, a, rS);
    drift(q, p, w.c2 * h, a, rS);
    kick(q, p, w.w0 * h, a, rS);
    drift(q, p, w.c2 * h, a, rS);
    kick(q, p, w.w1 * h, a, rS);
    drift(q, p, w.c1 * h, a, rS);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    st[k] = q[k];
    st[4 + k] = p[k];
  }
}

// H = g^{mu nu} p_mu p_nu / 2
__global__ void __launch_bounds__(256) k_hamiltonian(const double* __restrict__ states, long long n, double a,
                                                     double rS, double* __restrict__ H) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* st = states + i * 8;
  const double* p = st + 4;
  const Gcon g = gcon(st[1], st[2], a, rS);
  const double tt = g.g00 * p[0] * p[0] + 2.0 * g.g03 * p[0] * p[3] + g.g33 * p[3] * p[3];
  H[i] = 0.5 * (tt + g.g11 * p[1] * p[1] + g.g22 * p[2] * p[2]);
}

// Yoshida (1990) weights, as the host computes them.
Weights yoshida_weights() {
  Weights w;
  w.w1 = 1.0 / (2.0 - std::cbrt(2.0));
  w.w0 = 1.0 - 2.0 * w.w1;
  w.c1 = 0.5 * w.w1;
  w.c2 = 0.5 * (w.w0 + w.w1);
  return w;
}

int check_args(const void* states, int64_t n, double a, double rS, const char* who) {
  if (n < 0) return fail(std::string(who) + ": negative state count");
  if (n > 0 && !states) return fail(std::string(who) + ": NULL state buffer");
  if (!(rS > 0.0) || !std::isfinite(rS) || !std::isfinite(a))
    return fail(std::string(who) + ": rS must be positive and a finite");
  return 0;
}

unsigned grid_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

extern "C" {

int bh_symplectic_integrate(double* states, int64_t n, double dLambda, int32_t nSteps, double a, double rS,
                            void* stream) {
  const char* who = "bh_symplectic_integrate";
  if (check_args(states, n, a, rS, who) != 0) return -1;
  if (!std::isfinite(dLambda)) return fail(std::string(who) + ": dLambda must be finite");
  if (n == 0 || nSteps <= 0) return 0;
  const double h = dLambda / static_cast<double>(nSteps);
  hipLaunchKernelGGL(k_yoshida4, dim3(grid_for(n)), dim3(256), 0, static_cast<hipStream_t>(stream), states,
                     static_cast<long long>(n), h, static_cast<int>(nSteps), a, rS, yoshida_weights());
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string(who) + ": launch failed: " + hipGetErrorString(e));
  return 0;
}

int bh_symplectic_hamiltonian(const double* states, int64_t n, double a, double rS, double* H, void* stream) {
  const char* who = "bh_symplectic_hamiltonian";
  if (check_args(states, n, a, rS, who) != 0) return -1;
  if (n > 0 && !H) return fail(std::string(who) + ": NULL output");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_hamiltonian, dim3(grid_for(n)), dim3(256), 0, static_cast<hipStream_t>(stream), states,
                     static_cast<long long>(n), a, rS, H);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string(who) + ": launch failed: " + hipGetErrorString(e));
  return 0;
}

int bh_symplectic_integrate_host(double* states, int64_t n, double dLambda, int32_t nSteps, double a, double rS) {
  const char* who = "bh_symplectic_integrate_host";
  if (check_args(states, n, a, rS, who) != 0) return -1;
  if (!std::isfinite(dLambda)) return fail(std::string(who) + ": dLambda must be finite");
  if (n == 0 || nSteps <= 0) return 0;
  const size_t bytes = static_cast<size_t>(n) * 8 * sizeof(double);
  double* d = nullptr;
  if (hipMalloc(&d, bytes) != hipSuccess) return fail(std::string(who) + ": hipMalloc failed");
  int rc = 0;
  if (hipMemcpy(d, states, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(std::string(who) + ": upload failed");
  } else if (bh_symplectic_integrate(d, n, dLambda, nSteps, a, rS, nullptr) != 0) {
    rc = -1;
  } else if (hipMemcpy(states, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
    rc = fail(std::string(who) + ": download failed");
  }
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
