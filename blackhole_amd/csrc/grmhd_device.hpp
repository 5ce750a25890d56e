// grmhd_device.hpp - the GRMHD volume sample of the render launch (device).
//
// gfx950 has no image-sampling hardware, so a GRMHD "texture" is a device
// descriptor of RGBA32F voxels (x = r fastest, then theta, then phi) sampled
// here in software with the semantics of a linear-filtered, normalised-
// coordinate 3-D texture: texel centres at (i + 0.5) / n, trilinear weights,
// and per axis the address mode of the reference's GRMHD textures -- r
// clamped, theta mirrored (edge texel repeated), phi wrapped.
//
// A point p (Cartesian, z along the spin axis, the launch's length unit) maps
// to u = (|p| - r_min) / (r_max - r_min), v = theta / pi, w = phi / (2 pi),
// phi in [0, 2 pi).  With a right-hand volume and alpha > 0 the two frames
// are blended, left + alpha (right - left) (d_sample_grmhd,
// device_physics.cuh:1789-1829).
#pragma once

#include <hip/hip_runtime.h>

namespace bhk {

struct GrmhdTex {
  const float4* data;
  int w, h, d, pad;
};

// Host: whether h is a live volume made by bh_grmhd_texture_create (grmhd.hip).
bool grmhd_volume_known(unsigned long long h);

__device__ __forceinline__ int grmhd_clamp(int i, int n) { return i < 0 ? 0 : (i >= n ? n - 1 : i); }
__device__ __forceinline__ int grmhd_wrap(int i, int n) {
  const int m = i % n;
  return m < 0 ? m + n : m;
}
__device__ __forceinline__ int grmhd_mirror(int i, int n) {
  // period 2n: 0 .. n-1 forward, n .. 2n-1 backward (n-1 .. 0)
  int m = i % (2 * n);
  if (m < 0) m += 2 * n;
  return m < n ? m : 2 * n - 1 - m;
}

__device__ inline float4 grmhd_fetch(const GrmhdTex& t, float u, float v, float w) {
  const float x = u * static_cast<float>(t.w) - 0.5f;
  const float y = v * static_cast<float>(t.h) - 0.5f;
  const float z = w * static_cast<float>(t.d) - 0.5f;
  const float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
  const float ax = x - fx0, ay = y - fy0, az = z - fz0;
  const int ix = static_cast<int>(fx0), iy = static_cast<int>(fy0), iz = static_cast<int>(fz0);
  const int xi[2] = {grmhd_clamp(ix, t.w), grmhd_clamp(ix + 1, t.w)};
  const int yi[2] = {grmhd_mirror(iy, t.h), grmhd_mirror(iy + 1, t.h)};
  const int zi[2] = {grmhd_wrap(iz, t.d), grmhd_wrap(iz + 1, t.d)};
  float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
    const float wgt = (bx ? ax : 1.0f - ax) * (by ? ay : 1.0f - ay) * (bz ? az : 1.0f - az);
    const float4 q = t.data[(static_cast<long long>(zi[bz]) * t.h + yi[by]) * t.w + xi[bx]];
    acc.x += wgt * q.x;
    acc.y += wgt * q.y;
    acc.z += wgt * q.z;
    acc.w += wgt * q.w;
  }
  return acc;
}

// left / right: device addresses of GrmhdTex descriptors (0 = absent).
__device__ inline float4 grmhd_sample(unsigned long long left, unsigned long long right, float alpha, float r_min,
                                      float r_max, float x, float y, float z) {
  const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (left == 0ull) return zero;
  const float range = r_max - r_min;
  if (!(range > 1e-6f)) return zero;
  const float r = sqrtf(x * x + y * y + z * z);
  const float theta = (r > 1e-6f) ? acosf(fminf(fmaxf(z / r, -1.0f), 1.0f)) : 0.0f;
  float phi = atan2f(y, x);
  if (phi < 0.0f) phi += 6.28318530717958648f;
  const float u = (r - r_min) / range;
  const float v = theta * 0.318309886183790672f;
  const float w = phi * 0.159154943091895336f;
  const float4 a = grmhd_fetch(*reinterpret_cast<const GrmhdTex*>(left), u, v, w);
  if (right == 0ull || !(alpha > 1e-6f)) return a;
  const float4 b = grmhd_fetch(*reinterpret_cast<const GrmhdTex*>(right), u, v, w);
  return make_float4(a.x + alpha * (b.x - a.x), a.y + alpha * (b.y - a.y), a.z + alpha * (b.z - a.z),
                     a.w + alpha * (b.w - a.w));
}

}  // namespace bhk
