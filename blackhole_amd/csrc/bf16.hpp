// bf16.hpp - BF16 conversions and MFMA operand types shared by the
// surrogate's training kernels and its fused rendering kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bhk {
namespace mlp {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

// FP32 -> BF16 with round-to-nearest-even (NaN kept quiet).
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return static_cast<uint16_t>(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(uint16_t h) {
  const uint32_t u = static_cast<uint32_t>(h) << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

}  // namespace mlp
}  // namespace bhk
