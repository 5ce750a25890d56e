// peaks.hip - micro-kernels that measure the attainable FP64 / FP32 FMA,
// BF16 MFMA and HBM copy rates of the device (bh_measure_peaks).
//
// Each kernel keeps every SIMD busy with independent dependency chains:
// 2048 workgroups x 256 lanes = 8 waves per SIMD on 256 CUs.  The results
// are stored so that no chain is dead code; their values are meaningless.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/bh_measure.h"

namespace bhk {
int fail(const std::string& msg);

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int kBlocks = 2048;
constexpr int kThreads = 256;
constexpr int kFmaIters = 65536;
constexpr int kMfmaIters = 16384;

template <class T>
__global__ void __launch_bounds__(kThreads) k_peak_fma(T* out, int iters, T b, T c) {
  const T x = static_cast<T>(threadIdx.x) * static_cast<T>(1e-6);
  T a0 = x, a1 = x + T(1), a2 = x + T(2), a3 = x + T(3), a4 = x + T(4), a5 = x + T(5), a6 = x + T(6),
    a7 = x + T(7);
  for (int i = 0; i < iters; ++i) {
    a0 = fma(a0, b, c);
    a1 = fma(a1, b, c);
    a2 = fma(a2, b, c);
    a3 = fma(a3, b, c);
    a4 = fma(a4, b, c);
    a5 = fma(a5, b, c);
    a6 = fma(a6, b, c);
    a7 = fma(a7, b, c);
  }
  out[blockIdx.x * kThreads + threadIdx.x] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

__global__ void __launch_bounds__(kThreads) k_peak_mfma(float* out, int iters) {
  const int lane = threadIdx.x & 63;
  bf16x8 a, b;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    a[i] = static_cast<__bf16>(1e-3f * static_cast<float>(lane + i));
    b[i] = static_cast<__bf16>(1e-3f * static_cast<float>(64 - lane + i));
  }
  f32x16 c0 = {}, c1 = {}, c2 = {}, c3 = {};
  for (int i = 0; i < iters; ++i) {
    c0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c0, 0, 0, 0);
    c1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c1, 0, 0, 0);
    c2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c2, 0, 0, 0);
    c3 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c3, 0, 0, 0);
  }
  float s = 0.0f;
#pragma unroll
  for (int e = 0; e < 16; ++e) s += (c0[e] + c1[e]) + (c2[e] + c3[e]);
  out[blockIdx.x * kThreads + threadIdx.x] = s;
}

// Grid-stride copy with four independent 16-byte non-temporal loads in flight
// per lane and non-temporal stores (the fastest of the copy forms measured on
// MI355X: 5.3 TB/s on 4 GiB against 4.9 TB/s with ordinary loads and stores).
__global__ void __launch_bounds__(kThreads) k_peak_copy(const f32x4* __restrict__ src, f32x4* __restrict__ dst,
                                                        long long n) {
  const long long stride = static_cast<long long>(gridDim.x) * kThreads;
  long long i = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const f32x4 v0 = __builtin_nontemporal_load(src + i);
    const f32x4 v1 = __builtin_nontemporal_load(src + i + stride);
    const f32x4 v2 = __builtin_nontemporal_load(src + i + 2 * stride);
    const f32x4 v3 = __builtin_nontemporal_load(src + i + 3 * stride);
    __builtin_nontemporal_store(v0, dst + i);
    __builtin_nontemporal_store(v1, dst + i + stride);
    __builtin_nontemporal_store(v2, dst + i + 2 * stride);
    __builtin_nontemporal_store(v3, dst + i + 3 * stride);
  }
  for (; i < n; i += stride) dst[i] = src[i];
}

// Milliseconds per launch of `launch` over `reps` timed launches after one untimed.
template <class F>
hipError_t time_ms(hipStream_t st, int reps, F launch, float* ms) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    launch();
    e = hipEventRecord(e0, st);
  }
  if (e == hipSuccess) {
    for (int r = 0; r < reps; ++r) launch();
    e = hipEventRecord(e1, st);
  }
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
  if (e == hipSuccess) *ms /= static_cast<float>(reps);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return e;
}

}  // namespace
}  // namespace bhk

using namespace bhk;

extern "C" int bh_measure_peaks(double* out, void* stream) {
  if (!out) return fail("bh_measure_peaks: NULL output");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long lanes = static_cast<long long>(kBlocks) * kThreads;
  const long long n4 = (2ll << 30) / 16;  // 16-byte elements in 2 GiB
  void* res = nullptr;
  f32x4* src = nullptr;
  f32x4* dst = nullptr;
  if (hipMalloc(&res, lanes * sizeof(double)) != hipSuccess) return fail("bh_measure_peaks: hipMalloc failed");
  if (hipMalloc(reinterpret_cast<void**>(&src), n4 * 16) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&dst), n4 * 16) != hipSuccess) {
    (void)hipFree(res);
    if (src) (void)hipFree(src);
    return fail("bh_measure_peaks: hipMalloc of the 4 GiB of copy buffers failed");
  }
  float ms[4] = {0.f, 0.f, 0.f, 0.f};
  hipError_t e = time_ms(st, 3, [&] {
    hipLaunchKernelGGL(k_peak_fma<double>, dim3(kBlocks), dim3(kThreads), 0, st, static_cast<double*>(res),
                       kFmaIters, 0.999999, 1e-7);
  }, &ms[0]);
  if (e == hipSuccess)
    e = time_ms(st, 3, [&] {
      hipLaunchKernelGGL(k_peak_fma<float>, dim3(kBlocks), dim3(kThreads), 0, st, static_cast<float*>(res),
                         kFmaIters, 0.999999f, 1e-7f);
    }, &ms[1]);
  if (e == hipSuccess)
    e = time_ms(st, 3, [&] {
      hipLaunchKernelGGL(k_peak_mfma, dim3(kBlocks), dim3(kThreads), 0, st, static_cast<float*>(res), kMfmaIters);
    }, &ms[2]);
  if (e == hipSuccess)
    e = time_ms(st, 10, [&] {
      hipLaunchKernelGGL(k_peak_copy, dim3(16 * kBlocks), dim3(kThreads), 0, st, src, dst, n4);
    }, &ms[3]);
  (void)hipFree(res);
  (void)hipFree(src);
  (void)hipFree(dst);
  if (e != hipSuccess) return fail(std::string("bh_measure_peaks: ") + hipGetErrorString(e));
  const double fma_flop = 2.0 * 8.0 * static_cast<double>(kFmaIters) * static_cast<double>(lanes);
  const double mfma_flop = 4.0 * 32768.0 * static_cast<double>(kMfmaIters) * static_cast<double>(lanes / 64);
  out[0] = fma_flop / (ms[0] * 1e-3) / 1e12;
  out[1] = fma_flop / (ms[1] * 1e-3) / 1e12;
  out[2] = mfma_flop / (ms[2] * 1e-3) / 1e12;
  out[3] = 2.0 * static_cast<double>(n4) * 16.0 / (ms[3] * 1e-3) / 1e9;
  return 0;
}
