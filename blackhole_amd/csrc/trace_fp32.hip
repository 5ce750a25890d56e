// trace_fp32.hip - FP32 precision form of the persistent Kerr tracer (gfx950).
//
// Same geodesic formulation, refill scheme and record as the FP64 kernel
// (trace.hip), with:
//   * the bulk of the stepping in FP32: state (r, u, phi, v_r, v_u),
//     constants and right-hand side in FP32, the state advanced by
//     Kahan-compensated sums of the step increments;
//   * FP64 islands: the ray setup (Carter constant and impact parameters),
//     every step that starts inside r < island_r_factor * r_+, every step
//     that starts within island_turn_dist (M) of a radial turning point
//     (v_r^2 / |R'(r)|) or a polar one (r |delta u| / sin(theta) with
//     delta u = v_u^2 / |U'(u)|), and every step that starts within
//     kPoleIsland of the polar axis (sin^2 theta < 1e-3): there
//     dphi/dlambda = xi / sin^2 theta is singular, FP32 cannot resolve
//     sin^2 theta = 1 - u^2 of a ray that turns close to the axis, and its
//     step-size control would stall.  An island step's FP64 result is kept
//     in the compensated pair (y, cy), so consecutive island steps carry
//     the FP64 state without rounding it to FP32;
//   * FP32 radiative transfer with Kahan-compensated optical depth and
//     intensity, in host-prescaled units (every intermediate in FP32 range);
//   * island steps gathered per wave (a wave runs the FP32 body or the FP64
//     body in one iteration, never both), the FP32 steps' size control in
//     FP32, and three waves per SIMD.
// Replaces the reference's FP32 render kernel (geodesicTraceFp32,
// src/cuda/kernels_fp32.cu:32-111) for the image-plane tracer.
#include <hip/hip_runtime.h>

#include "trace_common.hpp"

namespace bhk {
namespace f32 {

struct LaneF {
  RayC cd;           // FP64 constants (setup and islands)
  RayCT<float> cf;   // FP32 constants
  float y[kN], cy[kN], k1[kN];
  float h;
  float I, Ic, Tr, tau, tauc, a0, e0, s0;
  float r_stop, r_isco, island_r;
  long long gid;
  int cam, steps, attempts, ncross, nsamp, fp64_steps;
  bool last_rej;
};

// Kahan (1965) compensated sum: sum - comp tracks the exact running sum.
__device__ __forceinline__ void kahan_add(float& sum, float& comp, float x) {
  const float yk = x - comp;
  const float t = sum + yk;
  comp = (t - sum) - yk;
  sum = t;
}

__device__ __forceinline__ float affine_rate_f(const RayCT<float>& c, float r, float u) {
  return divT(fmaT(c.a2, u * u, r * r), r);
}

template <bool VOL>
__device__ __forceinline__ bool lane_setup(LaneF& L, const DevParams& P, long long q) {
  CamDev k;
  double alpha, beta;
  if (!resolve_ray(P, q, &L.gid, &L.cam, &k, &alpha, &beta)) return false;
  double yd[kN], fd[kN];
  // FP64 island: constants of motion (a pinhole camera's table entry maps
  // pixels to their centres)
  L.cd = P.pin ? setup_pinhole(*P.pin, k.a, alpha, beta, yd) : setup_ray_ab(k, alpha, beta, yd);
  L.cf = convert_const<float>(L.cd);
  rhs(L.cd, yd, fd);
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    L.y[i] = static_cast<float>(yd[i]);
    L.cy[i] = 0.0f;
    L.k1[i] = static_cast<float>(fd[i]);
  }
  L.h = static_cast<float>(P.h_init);
  L.I = L.Ic = 0.0f;
  L.Tr = 1.0f;
  L.tau = L.tauc = 0.0f;
  L.a0 = L.e0 = 0.0f;
  L.s0 = VOL ? affine_rate_f(L.cf, L.y[0], L.y[1]) : 0.0f;
  clear_crossing_slots(P.out + L.gid);
  L.r_stop = static_cast<float>(k.r_stop);
  L.r_isco = static_cast<float>(k.r_isco);
  L.island_r = static_cast<float>(P.island_r_factor * k.r_plus);
  L.steps = L.attempts = L.ncross = L.nsamp = L.fp64_steps = 0;
  L.last_rej = false;
  return true;
}

// sin^2 theta below which steps run in FP64 (theta within ~1.8 deg of the axis).
constexpr float kPoleIsland = 1e-3f;

// FP64 island test at the start of a step.
__device__ __forceinline__ bool needs_island(const LaneF& L, const DevParams& P) {
  const float r = L.y[0], u = L.y[1];
  if (r < L.island_r) return true;
  if (fmaT(-u, u, 1.0f) < kPoleIsland) return true;
  const float D = static_cast<float>(P.island_turn_dist);
  const float Pp = r * r + L.cf.c1;
  const float Rp = fmaT(4.0f * r, Pp, -((2.0f * r - 2.0f) * L.cf.K));  // R'(r)
  const float vr2 = L.y[3] * L.y[3];
  if (vr2 < D * fabsT(Rp)) return true;  // v_r^2 / |R'| < D
  const float Up = 2.0f * u * fmaT(-L.cf.two_a2, u * u, L.cf.c2);  // U'(u)
  const float w = fmaT(-u, u, 1.0f);
  const float sth = sqrtT(w > 1e-12f ? w : 1e-12f);
  const float vu2 = L.y[4] * L.y[4];
  return r * vu2 < D * sth * fabsT(Up);  // r (v_u^2 / |U'|) / sin(theta) < D
}

// ---- FP32 radiative transfer ---------------------------------------------

__device__ __forceinline__ void volume_fetch_f(const VolDev& v, float r, float u, float phi,
                                               float* out4) {
  const float theta = acosf(u);
  const float xr = (r - static_cast<float>(v.r_min)) * static_cast<float>(v.s_r) - 0.5f;
  const float xt = theta * static_cast<float>(v.s_th) - 0.5f;
  const float xp = phi * static_cast<float>(v.s_ph) - 0.5f;
  const float fr0 = floorf(xr), ft0 = floorf(xt), fp0 = floorf(xp);
  const float fr = xr - fr0, ft = xt - ft0, fp = xp - fp0;
  int ir0 = static_cast<int>(fr0), it0 = static_cast<int>(ft0);
  const float npf = static_cast<float>(v.n_ph);
  int ip0 = static_cast<int>(fp0 - floorf(fp0 * static_cast<float>(v.inv_nph)) * npf);
  ip0 = (ip0 >= v.n_ph) ? ip0 - v.n_ph : ip0;  // the quotient may be off by one
  ip0 = (ip0 < 0) ? ip0 + v.n_ph : ip0;
  int ir1 = ir0 + 1, it1 = it0 + 1;
  const int n_r = v.n_r, n_th = v.n_th, n_ph = v.n_ph;
  ir0 = ir0 < 0 ? 0 : (ir0 > n_r - 1 ? n_r - 1 : ir0);
  ir1 = ir1 < 0 ? 0 : (ir1 > n_r - 1 ? n_r - 1 : ir1);
  it0 = it0 < 0 ? 0 : (it0 > n_th - 1 ? n_th - 1 : it0);
  it1 = it1 < 0 ? 0 : (it1 > n_th - 1 ? n_th - 1 : it1);
  ip0 = ip0 < 0 ? 0 : (ip0 > n_ph - 1 ? n_ph - 1 : ip0);
  const int ip1 = (ip0 + 1 == n_ph) ? 0 : ip0 + 1;
  const BHK_GLOBAL float4* __restrict__ cells = (const BHK_GLOBAL float4*)(v.cells);
  const unsigned b00 = static_cast<unsigned>((ip0 * n_th + it0) * n_r);
  const unsigned b01 = static_cast<unsigned>((ip0 * n_th + it1) * n_r);
  const unsigned b10 = static_cast<unsigned>((ip1 * n_th + it0) * n_r);
  const unsigned b11 = static_cast<unsigned>((ip1 * n_th + it1) * n_r);
  const float4 q000 = cells[b00 + ir0], q001 = cells[b00 + ir1];
  const float4 q010 = cells[b01 + ir0], q011 = cells[b01 + ir1];
  const float4 q100 = cells[b10 + ir0], q101 = cells[b10 + ir1];
  const float4 q110 = cells[b11 + ir0], q111 = cells[b11 + ir1];
#define BHK_TRIF(F_, OUT_)                                 \
  {                                                       \
    const float c00 = fmaT(fr, q001.F_ - q000.F_, q000.F_); \
    const float c01 = fmaT(fr, q011.F_ - q010.F_, q010.F_); \
    const float c10 = fmaT(fr, q101.F_ - q100.F_, q100.F_); \
    const float c11 = fmaT(fr, q111.F_ - q110.F_, q110.F_); \
    const float c0 = fmaT(ft, c01 - c00, c00);             \
    const float c1 = fmaT(ft, c11 - c10, c10);             \
    OUT_ = fmaT(fp, c1 - c0, c0);                          \
  }
  BHK_TRIF(x, out4[0])
  BHK_TRIF(y, out4[1])
  BHK_TRIF(z, out4[2])
  BHK_TRIF(w, out4[3])
#undef BHK_TRIF
}

__device__ __forceinline__ float fluid_g_f(float r, float u, float a, float xi, float ell) {
  const float s2 = fmaT(-u, u, 1.0f);
  const float a2 = a * a;
  const float sig = fmaT(a2, u * u, r * r);
  const float two_r_sig = divT(2.0f * r, sig);
  const float gtt = two_r_sig - 1.0f;
  const float gtp = -(a * two_r_sig * s2);
  const float gpp = fmaT(a2 * two_r_sig, s2, fmaT(r, r, a2)) * s2;
  const float Om = -divT(fmaT(ell, gtt, gtp), fmaT(ell, gtp, gpp));
  const float nrm = -fmaT(Om * Om, gpp, fmaT(2.0f * Om, gtp, gtt));
  if (nrm <= 0.0f) return 0.0f;
  const float den = fmaT(-Om, xi, 1.0f);
  if (den <= 0.0f) return 0.0f;
  return divT(sqrtT(nrm), den);
}

// expm1(x) for x >= 0 in FP32: x (1 + x/2 + x^2/6) below 1e-2 (truncation
// x^3/24 < 4.2e-8 x), the library function above.
__device__ __forceinline__ float expm1_pos_f(float x) {
  return (x < 1e-2f) ? x * fmaT(fmaT(x, 1.0f / 6.0f, 0.5f), x, 1.0f) : expm1f(x);
}

// a (absorption) and e (emission) per affine length in the observer frame,
// as volume_point in kerr_device.hpp, in FP32 with prescaled constants.  The
// powers of the synchrotron fit come from the hardware base-2 logarithm and
// exponential: x^(1/3) = 2^(l/3) and x^(-1/6), x^(-1/4), x^(-1/2) as
// y^2, y^3, y^6 with y = 2^(-l/12), l = log2 x (a few ulp of FP32).
__device__ __forceinline__ bool volume_point_f(const DevParams& P, const RayCT<float>& c, float r,
                                               float u, float phi, float* a, float* e) {
  const VolDev& v = P.vol;
  *a = 0.0f;
  *e = 0.0f;
  if (!(r >= static_cast<float>(v.r_min) && r <= static_cast<float>(v.r_max) &&
        r >= static_cast<float>(v.sup_r_lo) && r <= static_cast<float>(v.sup_r_hi) &&
        fabsT(u) <= static_cast<float>(v.sup_u_max)))
    return false;
  float f4[4];
  volume_fetch_f(v, r, u, phi, f4);
  const float ne = f4[0], th = f4[1], B = f4[2], k2r = f4[3];
  if (!(ne > 0.0f && th > 0.0f && B > 0.0f && k2r > 0.0f)) return true;
  const float g = fluid_g_f(r, u, c.a, c.xi, static_cast<float>(v.ell));
  if (!(g > 0.0f)) return true;
  const float xT = divT(P.CT, g * th);
  const float em = expm1_pos_f(xT);
  if (!(em > 0.0f) || xT > 80.0f) return true;
  const float g3 = g * g * g;
  const float Bnu = divT(P.SB, g3 * em);
  const float x = P.Cx / (g * B * th * th);  // |B| th^2 can underflow: IEEE division
  const float lx = __builtin_amdgcn_logf(x);
  const float x13 = __builtin_amdgcn_exp2f(lx * (1.0f / 3.0f));
  const float y = __builtin_amdgcn_exp2f(lx * (-1.0f / 12.0f));
  const float xm16 = y * y;
  const float xm14 = xm16 * y;
  const float xm12 = xm14 * xm14;
  const float im = 4.0505f * xm16 * fmaT(0.5316f, xm12, fmaT(0.40f, xm14, 1.0f)) * __expf(-1.8899f * x13);
  const float jL = (P.CjL * ne * B * im) / k2r;  // k2r of cold gas can be subnormal: IEEE division
  *a = divT(jL, g * Bnu);
  *e = g * g * jL;
  return true;
}

__device__ __forceinline__ void rt_segment_f(LaneF& L, float h, float a1, float e1, float s1) {
  const float asum = L.a0 + a1;
  const float esum = L.e0 + e1;
  if (asum > 0.0f || esum > 0.0f) {
    const float dl = 0.5f * h * (L.s0 + s1);
    if (asum > 0.0f) {
      const float dtau = 0.5f * asum * dl;
      // 1 - e^-dtau: dtau (1 - dtau/2 + dtau^2/6) below 1e-2
      const float om = (dtau < 1e-2f) ? dtau * fmaT(fmaT(dtau, 1.0f / 6.0f, -0.5f), dtau, 1.0f) : -expm1f(-dtau);
      kahan_add(L.I, L.Ic, L.Tr * (esum / asum) * om);  // asum can be subnormal: IEEE division
      L.Tr = L.Tr * (1.0f - om);
      kahan_add(L.tau, L.tauc, dtau);
    } else {
      kahan_add(L.I, L.Ic, L.Tr * 0.5f * esum * dl);
    }
  }
  L.a0 = a1;
  L.e0 = e1;
  L.s0 = s1;
}

__device__ __forceinline__ void volume_segment_f(LaneF& L, const DevParams& P, float h) {
  float a1, e1;
  if (volume_point_f(P, L.cf, L.y[0], L.y[1], L.y[2], &a1, &e1)) L.nsamp++;
  rt_segment_f(L, h, a1, e1, affine_rate_f(L.cf, L.y[0], L.y[1]));
}

// Crossing of u = 0 between two accepted states (refined in FP64).
template <class T>
__device__ __forceinline__ bool crossing_f(LaneF& L, const DevParams& P, const T* y0, const T* f0,
                                           const T* y1, const T* f1, T h) {
  if ((y0[1] < T(0)) == (y1[1] < T(0))) return false;
  const double rc = crossing_radius(y0, f0, y1, f1, h);
  const double g = (rc >= static_cast<double>(L.r_isco)) ? disk_g(rc, L.cd.a, L.cd.xi) : 0.0;
  if (L.ncross < P.max_crossings) store_crossing(P.out + L.gid, L.ncross, rc, g);
  if (L.ncross < 255) L.ncross++;
  if (P.disk_enabled && rc >= static_cast<double>(L.r_isco) && rc <= P.disk_r_out) {
    const double* lut = P.luts + static_cast<long long>(P.cams[L.cam].lut_index) * P.lut_n;
    const float F = static_cast<float>(lut_linear(lut, P.lut_n, static_cast<double>(L.r_isco), P.disk_r_out, rc));
    if (F > 0.0f && g > 0.0) {
      const float x = divT(P.Cd, static_cast<float>(g) * sqrtT(sqrtT(F)));
      if (x < 80.0f) kahan_add(L.I, L.Ic, L.Tr * divT(P.SB, expm1f(x)));
    }
    return true;
  }
  return false;
}

__device__ __forceinline__ int terminate_f(const LaneF& L, const DevParams& P) {
  if (L.y[0] < L.r_stop) return BH_HIT_HORIZON;
  if (static_cast<double>(L.y[0]) > P.r_escape && L.y[3] > 0.0f) return BH_HIT_ESCAPED;  // outbound
  if (P.transmit_min > 0.0 && static_cast<double>(L.Tr) < P.transmit_min) return BH_HIT_OPAQUE;
  return 0;
}

__device__ __forceinline__ void write_record_f(const LaneF& L, const DevParams& P, int status) {
  BH_PixelRecord* __restrict__ rec = P.out + L.gid;
  double yd[kN];
#pragma unroll
  for (int i = 0; i < kN; ++i) yd[i] = static_cast<double>(L.y[i]) - static_cast<double>(L.cy[i]);
  const double R = radial_potential(L.cd, yd[0]);
  const double U = polar_potential(L.cd, yd[1]);
  const double rs = yd[0] * yd[0] + L.cd.a2;
  double2* d2 = reinterpret_cast<double2*>(rec);
  d2[0] = make_double2(static_cast<double>(L.I) - static_cast<double>(L.Ic), yd[0]);
  d2[1] = make_double2(yd[1], yd[2]);
  d2[5] = make_double2(static_cast<double>(L.tau) - static_cast<double>(L.tauc),
                       __builtin_fabs(__builtin_fma(yd[3], yd[3], -R)) / (rs * rs));
  const double dc = __builtin_fabs(__builtin_fma(yd[4], yd[4], -U)) /
                    (__builtin_fabs(L.cd.eta) + L.cd.a2 + L.cd.xi * L.cd.xi);
  int* i4 = reinterpret_cast<int*>(rec);
  reinterpret_cast<double*>(rec)[12] = dc;
  i4[26] = L.steps;
  i4[27] = L.attempts;
  i4[28] = status | (L.ncross << 8) | ((L.fp64_steps > 0 ? BH_FLAG_FP64_ISLAND : 0) << 16);
  i4[29] = __float_as_int(1.0f);
  i4[30] = L.fp64_steps;
  i4[31] = L.nsamp;
}

// FP64 island step (RK4 or DP45 with the FP64 tolerances) from the
// compensated FP32 state; returns false for a rejected DP45 attempt.
template <int INTEG>
__device__ __forceinline__ bool island_step(LaneF& L, const DevParams& P, double* y0d, double* f0d,
                                            double* y1d, double* f1d, double* hd) {
#pragma unroll
  for (int i = 0; i < kN; ++i) y0d[i] = static_cast<double>(L.y[i]) - static_cast<double>(L.cy[i]);
  rhs(L.cd, y0d, f0d);
  if (INTEG == BH_INTEGRATOR_RK4) {
    *hd = P.h_fixed;
    rk4_step(L.cd, y0d, f0d, *hd, y1d);
    rhs(L.cd, y1d, f1d);
    project_radial(L.cd, y1d, f1d);
    return true;
  }
  double cap = static_cast<double>(L.h);
  *hd = cap;
  const double err = dp45_attempt(L.cd, y0d, f0d, *hd, P.rtol, P.atol, y1d, f1d);
  if (err <= 1.0) {
    L.h = static_cast<float>(*hd * dp45_factor(err, true, L.last_rej));
    L.last_rej = false;
    project_radial(L.cd, y1d, f1d);
    return true;
  }
  L.h = static_cast<float>(*hd * dp45_factor(err, false, L.last_rej));
  L.last_rej = true;
  return false;
}

// FP32 RK4 step with compensated state update; writes the step start
// (y0, f0) for the crossing test.
__device__ __forceinline__ void rk4_step_f(LaneF& L, float h, float* y0, float* f0) {
  float k2[kN], k3[kN], k4[kN], yt[kN];
  const RayCT<float>& c = L.cf;
  const float hh = 0.5f * h;
  const float h6 = h * (1.0f / 6.0f);
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    y0[i] = L.y[i];
    f0[i] = L.k1[i];
    yt[i] = fmaT(hh, L.k1[i], L.y[i]);
  }
  rhs(c, yt, k2);
#pragma unroll
  for (int i = 0; i < kN; ++i) yt[i] = fmaT(hh, k2[i], L.y[i]);
  rhs(c, yt, k3);
#pragma unroll
  for (int i = 0; i < kN; ++i) yt[i] = fmaT(h, k3[i], L.y[i]);
  rhs(c, yt, k4);
#pragma unroll
  for (int i = 0; i < kN; ++i)
    kahan_add(L.y[i], L.cy[i], h6 * fmaT(2.0f, k2[i] + k3[i], L.k1[i] + k4[i]));
  rhs(c, L.y, L.k1);
  const float v_old = L.y[3];
  project_radial(c, L.y, L.k1);
  if (L.y[3] != v_old) L.cy[3] = 0.0f;
}

// FP32 DP45 attempt with compensated update on acceptance.  Returns the
// scaled error; on acceptance the state advances.
__device__ __forceinline__ float dp45_step_f(LaneF& L, const DevParams& P, float h, float* y0,
                                             float* f0) {
  using namespace dp;
  const RayCT<float>& c = L.cf;
  float k2[kN], k3[kN], k4[kN], k5[kN], k6[kN], k7[kN], yt[kN], d5[kN], y5[kN];
  const float* y = L.y;
  const float* k1 = L.k1;
#pragma unroll
  for (int i = 0; i < kN; ++i) yt[i] = fmaT(h, float(a21) * k1[i], y[i]);
  rhs(c, yt, k2);
#pragma unroll
  for (int i = 0; i < kN; ++i) yt[i] = fmaT(h, fmaT(float(a32), k2[i], float(a31) * k1[i]), y[i]);
  rhs(c, yt, k3);
#pragma unroll
  for (int i = 0; i < kN; ++i)
    yt[i] = fmaT(h, fmaT(float(a43), k3[i], fmaT(float(a42), k2[i], float(a41) * k1[i])), y[i]);
  rhs(c, yt, k4);
#pragma unroll
  for (int i = 0; i < kN; ++i)
    yt[i] = fmaT(h, fmaT(float(a54), k4[i], fmaT(float(a53), k3[i], fmaT(float(a52), k2[i], float(a51) * k1[i]))), y[i]);
  rhs(c, yt, k5);
#pragma unroll
  for (int i = 0; i < kN; ++i)
    yt[i] = fmaT(h,
                 fmaT(float(a65), k5[i],
                      fmaT(float(a64), k4[i], fmaT(float(a63), k3[i], fmaT(float(a62), k2[i], float(a61) * k1[i])))),
                 y[i]);
  rhs(c, yt, k6);
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    d5[i] = h * fmaT(float(b6), k6[i],
                     fmaT(float(b5), k5[i], fmaT(float(b4), k4[i], fmaT(float(b3), k3[i], float(b1) * k1[i]))));
    y5[i] = y[i] + d5[i];
  }
  rhs(c, y5, k7);
  float acc = 0.0f;
#pragma unroll
  for (int i = 0; i < kN; ++i) {
    const float e = h * fmaT(float(e7), k7[i],
                             fmaT(float(e6), k6[i],
                                  fmaT(float(e5), k5[i], fmaT(float(e4), k4[i], fmaT(float(e3), k3[i], float(e1) * k1[i])))));
    const float ay = fabsT(y[i]), ay5 = fabsT(y5[i]);
    const float sc = fmaT(P.rtol32, ay > ay5 ? ay : ay5, P.atol32);
    const float qq = divT(e, sc);
    acc = fmaT(qq, qq, acc);
  }
  const float err = sqrtT(acc * 0.2f);
  if (err <= 1.0f) {
#pragma unroll
    for (int i = 0; i < kN; ++i) {
      y0[i] = L.y[i];
      f0[i] = L.k1[i];
      kahan_add(L.y[i], L.cy[i], d5[i]);
      L.k1[i] = k7[i];
    }
    const float v_old = L.y[3];
    project_radial(c, L.y, L.k1);
    if (L.y[3] != v_old) L.cy[3] = 0.0f;
  }
  return err;
}

// DP45 step-size factor of the FP32 steps, in FP32: 0.9 err^(-1/5) with the
// limits of dp45_factor (v_log_f32 and v_exp_f32 work in base 2).
__device__ __forceinline__ float dp45_factor_f(float err, bool accepted, bool last_rejected) {
  const float e = accepted ? fmaxf(err, 1e-10f) : err;
  float fac = 0.9f * __builtin_amdgcn_exp2f(-0.2f * __builtin_amdgcn_logf(e));
  if (accepted) {
    fac = fminf(fac, 5.0f);
    if (last_rejected) fac = fminf(fac, 1.0f);
  } else {
    fac = fmaxf(fac, 0.2f);
  }
  return fac;
}

// Island steps are gathered per wave: a wave runs either the FP32 step body
// (for its lanes that need no island) or the FP64 island body (for the
// lanes that do), never both in one iteration.  It takes the island body once
// kIslandGather lanes wait for it, or after kIslandWait iterations with any
// lane waiting.  A lane's own sequence of steps does not change: waiting
// iterations count neither as steps nor as attempts, so the records are
// those of running every step in turn.
constexpr int kIslandGather = 16;
constexpr int kIslandWait = 12;

template <int INTEG, bool VOL>
__device__ __forceinline__ int step_f(LaneF& L, const DevParams& P, int& wait_it) {
  if (L.steps >= P.max_steps || L.attempts >= P.max_attempts) return BH_HIT_MAX_STEPS;
  if (INTEG == BH_INTEGRATOR_DP45) {
    float cap = static_cast<float>(P.h_max);
    if (VOL && in_support(P.vol, L.y[0], L.y[1]) && static_cast<float>(P.h_vol) < cap)
      cap = static_cast<float>(P.h_vol);
    if (L.h > cap) L.h = cap;
  }
  const bool isl = needs_island(L, P);
  // wave-uniform choice of the step body (wait_it is read from the first
  // active lane, so every lane of the wave sees the same count)
  const int n_isl = __popcll(__ballot(isl));
  const int w_it = __builtin_amdgcn_readfirstlane(wait_it);
  const bool island_iter = n_isl > 0 && (n_isl >= kIslandGather || w_it >= kIslandWait);
  wait_it = island_iter ? 0 : (n_isl > 0 ? w_it + 1 : 0);
  if (isl != island_iter) return 0;  // this lane waits for the other body
  L.attempts++;
  bool hit = false;
  if (isl) {
    double y0d[kN], f0d[kN], y1d[kN], f1d[kN], hd;
    if (!island_step<INTEG>(L, P, y0d, f0d, y1d, f1d, &hd)) return 0;  // rejected
    L.steps++;
    L.fp64_steps++;
#pragma unroll
    for (int i = 0; i < kN; ++i) {
      // y - cy carries the FP64 result to ~2^-48: cy is the rounding residual
      L.y[i] = static_cast<float>(y1d[i]);
      L.cy[i] = static_cast<float>(static_cast<double>(L.y[i]) - y1d[i]);
      L.k1[i] = static_cast<float>(f1d[i]);
    }
    if (VOL) volume_segment_f(L, P, static_cast<float>(hd));
    hit = crossing_f<double>(L, P, y0d, f0d, y1d, f1d, hd);
  } else {
    float y0[kN], f0[kN];
    float h;
    if (INTEG == BH_INTEGRATOR_RK4) {
      h = static_cast<float>(P.h_fixed);
      rk4_step_f(L, h, y0, f0);
    } else {
      h = L.h;
      const float err = dp45_step_f(L, P, h, y0, f0);
      if (err > 1.0f) {
        L.h = h * dp45_factor_f(err, false, L.last_rej);
        L.last_rej = true;
        return 0;
      }
      L.h = h * dp45_factor_f(err, true, L.last_rej);
      L.last_rej = false;
    }
    L.steps++;
    if (VOL) volume_segment_f(L, P, h);
    hit = crossing_f<float>(L, P, y0, f0, L.y, L.k1, h);
  }
  if (hit) return BH_HIT_DISK;
  return terminate_f(L, P);
}

// Three waves per SIMD (168 registers; the FP64 island body spills a few
// values): c3 174 against 207 ms at two waves per SIMD.
template <int INTEG, bool VOL>
__global__ void __launch_bounds__(kBlock, 3) k_trace_fp32(const DevParams* __restrict__ Pp) {
  const DevParams& P = *Pp;
  LaneF L;
  L.gid = -1;
  bool done = false;
  int wait_it = 0;
  const long long n_queue = queue_length(P);
  for (;;) {
    const bool need = (L.gid < 0) && !done;
    const long long id = refill_id(need, P.head);
    if (need) {
      if (id < n_queue) {
        if (!lane_setup<VOL>(L, P, id)) L.gid = -1;
      } else {
        done = true;
      }
    }
    if (__ballot(!done) == 0ull) break;
    if (L.gid >= 0) {
      for (int b = 0; b < P.batch; ++b) {
        const int st = step_f<INTEG, VOL>(L, P, wait_it);
        if (st) {
          write_record_f(L, P, st);
          L.gid = -1;
          break;
        }
      }
    }
  }
}

template <int I, bool V>
int max_blocks_f() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
  if (cached[dev] > 0) return cached[dev];
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace_fp32<I, V>, kBlock, 0) != hipSuccess)
    return -1;
  cached[dev] = nb > 0 ? nb : 1;
  return cached[dev];
}

template <int I, bool V>
hipError_t launch_f(const DevParams* d_params, hipStream_t s, int grid) {
  if (grid <= 0) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((k_trace_fp32<I, V>), dim3(grid), dim3(kBlock), 0, s, d_params);
  return hipGetLastError();
}

}  // namespace f32

// Launch of the FP32 form on a grid of `grid` workgroups.
hipError_t launch_fp32(int integrator, bool vol_on, const DevParams* d_params, hipStream_t s, int grid) {
  if (integrator == BH_INTEGRATOR_DP45)
    return vol_on ? f32::launch_f<BH_INTEGRATOR_DP45, true>(d_params, s, grid)
                  : f32::launch_f<BH_INTEGRATOR_DP45, false>(d_params, s, grid);
  return vol_on ? f32::launch_f<BH_INTEGRATOR_RK4, true>(d_params, s, grid)
                : f32::launch_f<BH_INTEGRATOR_RK4, false>(d_params, s, grid);
}

// Workgroups of the FP32 kernel that fit on one CU.
int occupancy_limit_fp32(int integrator, bool vol_on) {
  if (integrator == BH_INTEGRATOR_DP45)
    return vol_on ? f32::max_blocks_f<BH_INTEGRATOR_DP45, true>() : f32::max_blocks_f<BH_INTEGRATOR_DP45, false>();
  return vol_on ? f32::max_blocks_f<BH_INTEGRATOR_RK4, true>() : f32::max_blocks_f<BH_INTEGRATOR_RK4, false>();
}

}  // namespace bhk
