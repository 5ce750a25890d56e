// trace_common.hpp - launch-time data shared by the FP64 and FP32 tracer
// kernels: the device-resident parameter block, the camera lookup and the
// wave-level ray refill.
#pragma once

#include <hip/hip_runtime.h>

#include "kerr_device.hpp"

namespace bhk {

constexpr int kBlock = 256;

// Launch parameters, resident in device memory next to the camera table
// (one upload per launch).  The kernels read them through a pointer rather
// than as ~350 bytes of kernel arguments, which kept ~100 scalar registers
// live across the step loop and spilled them.
struct DevParams {
  double h_fixed, h_init, h_max, h_vol, rtol, atol;
  double r_escape, disk_r_out, nu_obs, disk_T_max, length_unit_cm, transmit_min;
  VolDev vol;
  const CamDev* cams;
  const double* luts;
  BH_PixelRecord* out;
  unsigned long long* head;  // ray-queue head
  long long n_total;
  int n_cams, lut_n, batch, max_steps, max_attempts, max_crossings, disk_enabled, pad;
  // FP32 precision form: FP64 islands and FP32 tolerances.
  double island_r_factor, island_turn_dist;
  float rtol32, atol32;
  // FP32 radiative constants, host-scaled so that every intermediate stays
  // in FP32 range:  x_T = CT / (g Theta_e),  B_nu = SB / (g^3 expm1(x_T)),
  // x_M = Cx / (g |B| Theta_e^2),  j L = CjL n_e |B| I_M(x_M) / k2r,
  // disk: x = Cd / (g F^(1/4)),  I = SB / expm1(x).
  float CT, SB, Cx, CjL, Cd, pad32;
};

__device__ __forceinline__ int find_camera(const CamDev* cams, int n, long long gid) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cams[mid].offset <= gid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ bool in_support(const VolDev& v, double r, double u) {
  return r >= v.sup_r_lo && r <= v.sup_r_hi && __builtin_fabs(u) <= v.sup_u_max;
}

// Wave-level ray refill: the lanes with `need` set take consecutive ids from
// the queue head with a single atomic per wave (ballot, then each lane adds
// the number of needing lanes below it).  Must be called by all lanes.
// Returns the id taken (>= n_total: the queue is empty) on needing lanes.
__device__ __forceinline__ long long refill_id(bool need, unsigned long long* head) {
  const int lane = static_cast<int>(__lane_id());
  const unsigned long long below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const unsigned long long m = __ballot(need);
  long long id = -1;
  if (m) {
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(head, static_cast<unsigned long long>(__popcll(m)));
    const unsigned int lo = __shfl(static_cast<unsigned int>(base & 0xffffffffull), leader);
    const unsigned int hi = __shfl(static_cast<unsigned int>(base >> 32), leader);
    base = (static_cast<unsigned long long>(hi) << 32) | lo;
    if (need) id = static_cast<long long>(base) + __popcll(m & below);
  }
  return id;
}

// FP32 precision form launcher (trace_fp32.hip).
hipError_t launch_fp32(int integrator, bool vol_on, const DevParams* d_params, hipStream_t s,
                       int grid_override);

}  // namespace bhk
