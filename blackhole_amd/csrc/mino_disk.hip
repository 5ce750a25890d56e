// mino_disk.hip - thin-disk records from the Mino-time crossing solver
// (mino.hpp) on gfx950: the refinement pass of the surrogate render (the
// network's crossing radii as starting points, every crossing and the hit
// class solved exactly) and a network-free tracer of the same records.
//
// One ray per lane, 256-lane workgroups over the packed pixel index.  A ray
// the solver cannot resolve (quadrature or iteration limit, typically within
// a few 1e-3 M of the critical curve) is flagged BH_FLAG_FALLBACK and
// appended to the fallback list (one atomic per wave) for the integrator.
// The record of a solved ray is the integrator's thin-disk record: the first
// crossing radii, g from the exact Keplerian formula with the ray's xi (0
// inside r_ISCO), the intensity of the crossing inside [r_ISCO, r_out] from
// the camera's LUT by the integrator's formula, the hit class, the number of
// crossings, steps = Newton iterations, flags BH_FLAG_REFINED.  Escaped
// rays end at r_exit = r_escape; when asked for (the EXIT instantiation) with
// their exact direction there (u_exit, phi_exit from the polar motion's
// Jacobi functions and the azimuth integrals: mino::exit_direction).
// Captured and disk rays have no end direction (NaN); an escaping surrogate
// record without the exact one keeps the network's (flagged
// BH_FLAG_EXIT_NETWORK).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bhk_trace.h"
#include "mino.hpp"
#include "trace_common.hpp"

namespace bhk {
int fail(const std::string& msg);

// EXIT: escaped rays also get their exact end direction (an instantiation of
// its own, so that the crossing solve without it keeps its registers; with it
// inlined the EXIT kernel spills 164 B per lane and still runs faster than
// with an out-of-line call or at three waves per SIMD without spills).
template <bool EXIT>
__global__ void __launch_bounds__(256, 4) k_mino_disk(const MinoParams P) {
  const long long q = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  bool fb = false;
  long long gid = q;  // the pixel this thread solves
  if (q < P.n_total) {
    const int cam = find_camera(P.cams, P.n_cams, q);
    const CamDev& k = P.cams[cam];
    long long local = q - k.offset;
    const long long npx = static_cast<long long>(k.w) * k.h;
    if (local >= 0 && local < npx && npx < (1ll << 31)) {
      // threads take the camera's pixels in 4 x 4 tiles (2-D neighbours in a wave)
      long long px, py;
      tile_order_pixel(static_cast<unsigned>(local), static_cast<unsigned>(k.w), static_cast<unsigned>(k.h), &px,
                       &py);
      local = py * k.w + px;
      gid = k.offset + local;
    }
    BH_PixelRecord* rec = P.out + gid;
    int* i4 = reinterpret_cast<int*>(rec);
    double2* d2 = reinterpret_cast<double2*>(rec);
    int4 tail = make_int4(0, __float_as_int(1.0f), 0, 0);  // class/count/flags, confidence, fp64 steps, samples
    if (P.refine) tail = *reinterpret_cast<const int4*>(i4 + 28);
    const unsigned flags_in = (static_cast<unsigned>(tail.x) >> 16) & 0xffu;
    // fallback rays go to the integrator; rays whose network record is kept
    // (classified with margin away from the disk) are not solved
    if (local >= 0 && local < k.w * k.h && !(flags_in & (BH_FLAG_FALLBACK | BH_FLAG_NETWORK_KEPT))) {
      const long long py = local / k.w, px = local - py * k.w;
      const double alpha = k.alpha_min + (static_cast<double>(px) + 0.5) * k.dalpha;
      const double beta = k.beta_max - (static_cast<double>(py) + 0.5) * k.dbeta;
      const double a = k.a;
      const double xi = -(alpha * k.sin_o);
      const double eta = __builtin_fma(beta, beta, (alpha * alpha - a * a) * (k.cos_o * k.cos_o));
      double guess[3] = {0.0, 0.0, 0.0};
      int n_guess = 0;
      if (P.refine) {
        const double4 cr = *reinterpret_cast<const double4*>(reinterpret_cast<const double*>(rec) + 4);
        guess[0] = cr.x;
        guess[1] = cr.y;
        guess[2] = cr.z;
        n_guess = min((tail.x >> 8) & 0xff, 3);
      }
      const mino::RaySolution s = mino::solve_ray<EXIT>(a, xi, eta, k.r_obs, k.cos_o, k.sin_o * beta, k.r_stop,
                                                        P.r_escape, k.r_isco, P.disk_r_out, guess, n_guess, P.tol);
      if (s.ok) {
        double g[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) g[c] = (c < s.n && s.r[c] >= k.r_isco) ? disk_g(s.r[c], a, xi) : 0.0;
        double I = 0.0, r_exit;
        if (s.hit >= 0) {
          const double rh = s.r_hit;
          const double* lut = P.luts + static_cast<long long>(k.lut_index) * P.lut_n;
          const double F = lut_linear(lut, P.lut_n, k.r_isco, P.disk_r_out, rh);
          I = disk_intensity(F, s.hit < 3 ? g[s.hit] : disk_g(rh, a, xi), P.nu, P.T_max);
          r_exit = rh;
        } else {
          r_exit = (s.cls == mino::kHorizon) ? k.r_stop : P.r_escape;
        }
        // the path's end direction: exact for an escaped ray with EXIT; else an
        // escaping surrogate ray keeps the network's (flagged
        // BH_FLAG_EXIT_NETWORK), and the others are marked unknown (NaN)
        const bool esc = s.cls == mino::kEscaped;
        const bool exact = EXIT && esc && s.exit_ok;
        const bool exit_net = !exact && esc && P.refine && (flags_in & BH_FLAG_SURROGATE);
        const double2 dir = exact      ? make_double2(s.u_exit, s.phi_exit)
                            : exit_net ? d2[1]
                                       : make_double2(__builtin_nan(""), __builtin_nan(""));
        d2[0] = make_double2(I, r_exit);
        d2[1] = dir;
        d2[2] = make_double2(s.r[0], s.r[1]);
        d2[3] = make_double2(s.r[2], g[0]);
        d2[4] = make_double2(g[1], g[2]);
        d2[5] = make_double2(0.0, 0.0);
        *reinterpret_cast<int4*>(i4 + 24) = make_int4(0, 0, s.iters, 0);  // drift_carter, steps, attempts
        const unsigned flags = (flags_in & BH_FLAG_SURROGATE) | BH_FLAG_REFINED | (exit_net ? BH_FLAG_EXIT_NETWORK : 0u);
        tail.x = static_cast<int>(static_cast<unsigned>(s.cls) | (static_cast<unsigned>(min(s.n, 255)) << 8) |
                                  (flags << 16));
        *reinterpret_cast<int4*>(i4 + 28) = tail;
      } else {
        fb = true;
        if (!P.refine) {
          d2[0] = make_double2(0.0, 0.0);
          d2[1] = make_double2(0.0, 0.0);
          d2[2] = make_double2(0.0, 0.0);
          d2[3] = make_double2(0.0, 0.0);
          d2[4] = make_double2(0.0, 0.0);
          d2[5] = make_double2(0.0, 0.0);
          *reinterpret_cast<int4*>(i4 + 24) = make_int4(0, 0, 0, 0);
        }
        tail.x = (tail.x & 0xffff) | static_cast<int>((flags_in | BH_FLAG_FALLBACK) << 16);
        *reinterpret_cast<int4*>(i4 + 28) = tail;
      }
    }
  }
  wave_append(fb, gid, P.fb_ids, P.fb_count, P.fb_cap);
}

__global__ void k_mino_zero(unsigned long long* count) {
  if (threadIdx.x == 0) *count = 0ull;
}

hipError_t launch_mino_disk(const MinoParams& P, hipStream_t s) {
  const long long nblk = (P.n_total + 255) / 256;
  if (nblk <= 0) return hipSuccess;
  if (P.exit_dir)
    hipLaunchKernelGGL(k_mino_disk<true>, dim3(static_cast<unsigned>(nblk)), dim3(256), 0, s, P);
  else
    hipLaunchKernelGGL(k_mino_disk<false>, dim3(static_cast<unsigned>(nblk)), dim3(256), 0, s, P);
  return hipGetLastError();
}

namespace {
// Per (device, stream) device copy of the cameras.
struct CamBuf {
  CamDev* d = nullptr;
  size_t cap = 0;
};
std::mutex g_mu;
std::map<std::pair<int, void*>, CamBuf> g_bufs;
}  // namespace

}  // namespace bhk

using namespace bhk;

namespace {

// Both entry points: refine = 0 solves every pixel and starts a new fallback
// list; refine = 1 solves the surrogate's records in place from their radii
// and appends its failures to the list the render started.
int mino_entry(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts, const double* disk_luts, int disk_lut_n,
               BH_PixelRecord* out, int64_t* fallback_ids, uint64_t* fallback_count, int64_t fallback_cap,
               void* stream, int refine, const char* who) {
  const std::string w(who);
  if (!cams || n_cams <= 0 || !opts || !disk_luts || disk_lut_n < 2 || !out || !fallback_ids || !fallback_count)
    return fail(w + ": NULL argument");
  if (fallback_cap <= 0) return fail(w + ": fallback_cap must be > 0");
  if (!opts->disk_enabled || opts->volume_enabled)
    return fail(w + ": thin disk only (disk_enabled = 1, volume_enabled = 0)");
  if (!(opts->disk_r_out > 0.0) || !(opts->r_escape > 0.0)) return fail(w + ": disk_r_out and r_escape must be > 0");
  if (opts->mino_exit < 0 || opts->mino_exit > 2) return fail(w + ": mino_exit must be 0, 1 or 2");
  std::vector<CamDev> hc;
  long long total = 0;
  if (make_cam_devs(cams, n_cams, opts->horizon_factor, who, &hc, &total) != 0) return -1;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(w + ": no HIP device");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t bytes = hc.size() * sizeof(CamDev);
  CamDev* d_cams;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    CamBuf& b = g_bufs[{dev, stream}];
    if (b.cap < bytes) {
      if (b.d) (void)hipFree(b.d);
      b.d = nullptr;
      b.cap = 0;
      const size_t cap = std::max<size_t>(bytes, 64 * sizeof(CamDev));
      if (hipMalloc(&b.d, cap) != hipSuccess) return fail(w + ": hipMalloc failed");
      b.cap = cap;
    }
    d_cams = b.d;
  }
  // Pageable source: the runtime stages it before returning.
  if (hipMemcpyAsync(d_cams, hc.data(), bytes, hipMemcpyHostToDevice, st) != hipSuccess)
    return fail(w + ": camera upload failed");
  MinoParams P;
  std::memset(&P, 0, sizeof(P));
  P.cams = d_cams;
  P.luts = disk_luts;
  P.out = out;
  P.fb_ids = reinterpret_cast<long long*>(fallback_ids);
  P.fb_count = reinterpret_cast<unsigned long long*>(fallback_count);
  P.fb_cap = fallback_cap;
  P.n_total = total;
  P.disk_r_out = opts->disk_r_out;
  P.nu = opts->nu_obs;
  P.T_max = opts->disk_T_max;
  P.r_escape = opts->r_escape;
  P.tol = kMinoTol;
  P.n_cams = n_cams;
  P.lut_n = disk_lut_n;
  P.refine = refine;
  P.exit_dir = opts->mino_exit == 1 || (opts->mino_exit == 0 && !refine);
  if (total / 256 + 1 > 0x7fffffffll) return fail(w + ": too many pixels");
  if (!refine) hipLaunchKernelGGL(k_mino_zero, dim3(1), dim3(64), 0, st, P.fb_count);
  const hipError_t e = launch_mino_disk(P, st);
  return e == hipSuccess ? 0 : fail(w + ": launch failed: " + hipGetErrorString(e));
}

}  // namespace

extern "C" {

int bh_trace_disk_mino(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts, const double* disk_luts,
                       int disk_lut_n, BH_PixelRecord* out, int64_t* fallback_ids, uint64_t* fallback_count,
                       int64_t fallback_cap, void* stream) {
  return mino_entry(cams, n_cams, opts, disk_luts, disk_lut_n, out, fallback_ids, fallback_count, fallback_cap,
                    stream, 0, "bh_trace_disk_mino");
}

int bh_trace_disk_mino_refine(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                              const double* disk_luts, int disk_lut_n, BH_PixelRecord* out, int64_t* fallback_ids,
                              uint64_t* fallback_count, int64_t fallback_cap, void* stream) {
  return mino_entry(cams, n_cams, opts, disk_luts, disk_lut_n, out, fallback_ids, fallback_count, fallback_cap,
                    stream, 1, "bh_trace_disk_mino_refine");
}

}  // extern "C"
