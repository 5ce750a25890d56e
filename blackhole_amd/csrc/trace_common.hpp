// trace_common.hpp - launch-time data shared by the FP64 and FP32 tracer
// kernels: the device-resident parameter block, the camera lookup and the
// wave-level ray refill.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "kerr_device.hpp"

namespace bhk {

constexpr int kBlock = 256;

// Per-volume data of a volume set (volumes sharing one grid geometry and
// support): the cells and the fluid's angular momentum.
struct VolBucket {
  const float4* cells;
  double ell;
};

// Launch parameters, resident in device memory next to the camera table
// (one upload per launch).  The kernels read them through a pointer rather
// than as ~350 bytes of kernel arguments, which kept ~100 scalar registers
// live across the step loop and spilled them.
struct DevParams {
  double h_fixed, h_init, h_max, h_vol, rtol, atol;
  double r_escape, disk_r_out, nu_obs, disk_T_max, length_unit_cm, transmit_min;
  VolDev vol;
  RadDev rad;
  const CamDev* cams;
  const double* luts;
  BH_PixelRecord* out;
  unsigned long long* head;  // ray-queue head
  long long n_total;
  int n_cams, lut_n, batch, max_steps, max_attempts, max_crossings, disk_enabled;
  int compact_ids;  // id list: record of entry q -> out[q] instead of out[id]
  const VolBucket* buckets;  // volume set (camera k samples buckets[k.vol_index]); NULL: P.vol
  const PinholeDev* pin;     // pinhole camera (render launch): camera 0's pixels start from it
  // Path recording (polyline output): the start point and the end point of
  // every accepted step, Cartesian (r sin th cos phi, r sin th sin phi, r cos th),
  // at most path_max points per ray, path[(rec * path_max + k) * 3 + c].
  float* path;
  int* path_count;
  int path_max, pad_path;
  // Ray sources other than every pixel of the cameras: an explicit ray list
  // (rays, n_total entries, record i -> out[i]) or a list of pixel ids
  // (ids, count read from count_dev, record -> out[id]).
  const BH_RaySpec* rays;
  const long long* ids;
  const unsigned long long* count_dev;
  double horizon_factor;
  // FP32 precision form: FP64 islands and FP32 tolerances.
  double island_r_factor, island_turn_dist;
  float rtol32, atol32;
  // FP32 radiative constants, host-scaled so that every intermediate stays
  // in FP32 range:  x_T = CT / (g Theta_e),  B_nu = SB / (g^3 expm1(x_T)),
  // x_M = Cx / (g |B| Theta_e^2),  j L = CjL n_e |B| I_M(x_M) / k2r,
  // disk: x = Cd / (g F^(1/4)),  I = SB / expm1(x).
  float CT, SB, Cx, CjL, Cd, pad32;
  // Polarised transfer (POL kernels): (I, Q, U, V) per record into stokes;
  // the volume's emission is linearly polarised at pol_frac along the sky
  // angle chi (pol_c2 = cos 2 chi, pol_s2 = sin 2 chi) and its Faraday rate
  // is faraday_per_ne n_e per fluid-frame cm.
  double4* stokes;
  double pol_c2, pol_s2, pol_frac, faraday_per_ne;
};

__device__ __forceinline__ int find_camera(const CamDev* cams, int n, long long gid) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cams[mid].offset <= gid) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Resolves queue entry q into the record index, the camera constants and the
// image-plane point of its ray.  Returns false for a pixel id in a gap
// between camera record ranges (no record).  cam = -1 for explicit rays.
// Queue order inside a camera: 4 x 4 pixel tiles (bands of four rows from
// the top, tiles left to right within a band, pixels row-major within a
// tile; the right edge and the last band take the partial tiles), so the
// lanes of a wave trace pixels that are neighbours in two dimensions and
// their rays stay alike for longer (2-4 % faster frames than row order on
// the c2 / c4 / c5 workloads).  A bijection of [0, w h); records are written
// to their pixel's slot, so the output does not depend on the order.
__device__ __forceinline__ void tile_order_pixel(unsigned l, unsigned w, unsigned h, long long* px,
                                                 long long* py) {
  const unsigned full_bands = h / 4, band_px = 4 * w;
  unsigned band = l / band_px, bh = 4, m;
  if (band < full_bands) {
    m = l - band * band_px;
  } else {
    band = full_bands;
    bh = h - 4 * full_bands;
    m = l - full_bands * band_px;
  }
  const unsigned nfull = w / 4, tile_px = 4 * bh;
  if (m < nfull * tile_px) {
    const unsigned b = m / tile_px, wi = m - b * tile_px;
    *px = 4 * b + (wi & 3u);
    *py = 4 * band + (wi >> 2);
  } else {
    const unsigned rw = w - 4 * nfull, m2 = m - nfull * tile_px;
    *px = 4 * nfull + m2 % rw;
    *py = 4 * band + m2 / rw;
  }
}

__device__ __forceinline__ bool resolve_ray(const DevParams& P, long long q, long long* rec,
                                            int* cam, CamDev* k, double* alpha, double* beta) {
  if (P.rays) {
    const BH_RaySpec sp = P.rays[q];
    const double a = sp.spin;
    double s_o, c_o;
    sincos(sp.theta_obs, &s_o, &c_o);
    const double rp = 1.0 + __builtin_sqrt(1.0 - a * a);
    const double z1 = 1.0 + cbrt(1.0 - a * a) * (cbrt(1.0 + a) + cbrt(1.0 - a));
    const double z2 = __builtin_sqrt(3.0 * a * a + z1 * z1);
    k->a = a;
    k->sin_o = s_o;
    k->cos_o = c_o;
    k->r_obs = sp.r_obs;
    k->r_plus = rp;
    k->r_stop = P.horizon_factor * rp;
    k->r_isco = 3.0 + z2 - __builtin_sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
    k->lut_index = 0;
    k->vol_index = 0;
    *alpha = sp.alpha;
    *beta = sp.beta;
    *rec = q;
    *cam = -1;
    return true;
  }
  const long long gid = P.ids ? P.ids[q] : q;
  const int c = find_camera(P.cams, P.n_cams, gid);
  *k = P.cams[c];
  const long long local = gid - k->offset;
  const long long npx = static_cast<long long>(k->w) * k->h;
  if (local < 0 || local >= npx) return false;
  long long px, py, pid = gid;
  if (!P.ids && npx < (1ll << 31)) {
    // the queue runs over the whole image: take its pixels in tile order
    tile_order_pixel(static_cast<unsigned>(local), static_cast<unsigned>(k->w), static_cast<unsigned>(k->h), &px,
                     &py);
    pid = k->offset + py * k->w + px;
  } else {
    py = local / k->w;
    px = local - py * k->w;
  }
  *alpha = k->alpha_min + (static_cast<double>(px) + 0.5) * k->dalpha;
  *beta = k->beta_max - (static_cast<double>(py) + 0.5) * k->dbeta;
  *rec = P.compact_ids ? q : pid;
  *cam = c;
  return true;
}

// The crossing slots live in the ray's output record, not in registers (six
// doubles per lane would cost the volume kernels their second wave per SIMD):
// zeroed when the lane takes the ray, each recorded crossing stored as it
// happens (at most max_crossings <= 3 stores per ray).
__device__ __forceinline__ void clear_crossing_slots(BH_PixelRecord* rec) {
  double2* d2 = reinterpret_cast<double2*>(rec);
  d2[2] = make_double2(0.0, 0.0);
  d2[3] = make_double2(0.0, 0.0);
  d2[4] = make_double2(0.0, 0.0);
}

__device__ __forceinline__ void store_crossing(BH_PixelRecord* rec, int slot, double rc, double g) {
  rec->cross_r[slot] = rc;
  rec->cross_g[slot] = g;
}

// Number of queue entries (an id-list count lives in device memory so that
// a producer kernel can append ids without a host round trip).
__device__ __forceinline__ long long queue_length(const DevParams& P) {
  if (!P.count_dev) return P.n_total;
  const long long c = static_cast<long long>(*P.count_dev);
  return c < P.n_total ? c : P.n_total;  // never beyond the caller's max_ids
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

// Appends id to list (flag set) with one atomic per wave; must be called by
// every lane of the wave.
__device__ __forceinline__ void wave_append(bool flag, long long id, long long* list,
                                            unsigned long long* count, long long cap) {
  const int lane = static_cast<int>(__lane_id());
  const unsigned long long m = __ballot(flag);
  if (!m) return;
  const int leader = __ffsll(static_cast<long long>(m)) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(count, static_cast<unsigned long long>(__popcll(m)));
  const unsigned int lo = __shfl(static_cast<unsigned int>(base & 0xffffffffull), leader);
  const unsigned int hi = __shfl(static_cast<unsigned int>(base >> 32), leader);
  base = (static_cast<unsigned long long>(hi) << 32) | lo;
  const unsigned long long below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  if (flag) {
    const long long slot = static_cast<long long>(base + __popcll(m & below));
    if (slot < cap) list[slot] = id;
  }
}

// Thin-disk records from the Mino-time crossing solver (mino_disk.hip):
// every pixel (refine = 0) or, over surrogate records, the rays not flagged
// for fallback with the record's crossing radii as starting points
// (refine = 1).  Unresolved rays are appended to fb_ids.
struct MinoParams {
  const CamDev* cams;
  const double* luts;
  BH_PixelRecord* out;
  long long* fb_ids;
  unsigned long long* fb_count;
  long long fb_cap, n_total;
  double disk_r_out, nu, T_max, r_escape, tol;
  int n_cams, lut_n, refine, exit_dir;  // exit_dir: exact escape directions of escaped rays
};
// Relative tolerance on the |K15 - G7| estimate of each radial quadrature panel.  The
// estimate is far above the Kronrod rule's actual error for these smooth integrands: at
// 1e-7 the crossing radii of 160 000 c5 rays agree with a 1e-12 solve to 4.7e-10 at most
// (1.2e-14 mean), with 19 % fewer integrand evaluations than at 1e-8.
constexpr double kMinoTol = 1e-7;
hipError_t launch_mino_disk(const MinoParams& P, hipStream_t s);

// Host: device constants of n_cams cameras with their validation (trace.hip);
// -1 with the error set, prefixed by who, on an invalid camera.
int make_cam_devs(const BH_Camera* cams, int n_cams, double horizon_factor, const char* who,
                  std::vector<CamDev>* out, long long* total);

// Host: trace the pixels of one pinhole camera (cam supplies the spin, the
// horizon stop and the pixel grid; r_isco > 0 overrides the disk's inner
// edge) on `stream` (trace.hip).
int trace_pinhole(const BH_Camera& cam, const BH_TraceOptions& opts, const double* disk_lut, int lut_n,
                  const PinholeDev& pin, double r_isco, BH_PixelRecord* out, void* stream, const char* who);

// Host: trace one camera recording every ray's polyline (start point and the
// end of each accepted step, at most path_max points, counts per ray) into
// device buffers; DP45 FP64 without disk or volume (trace.hip).
int trace_paths(const BH_Camera& cam, const BH_TraceOptions& opts, float* d_path, int* d_count, int path_max,
                BH_PixelRecord* out, void* stream, const char* who);

// FP32 precision form (trace_fp32.hip): launch on `grid` workgroups, and
// the workgroups per CU its occupancy allows.
hipError_t launch_fp32(int integrator, bool vol_on, const DevParams* d_params, hipStream_t s, int grid);
int occupancy_limit_fp32(int integrator, bool vol_on);

}  // namespace bhk
