/*
 * bhk_trace.h - C ABI of the MI355X per-pixel Kerr null-geodesic tracer.
 *
 * Plain C (C99), POD structs only, no torch or HIP types in any signature.
 * Geometric units throughout: G = c = M = 1 (lengths in M); angles in rad.
 *
 * Each ray is an image-plane pixel of a distant-observer camera: impact
 * parameters (alpha, beta) map to the constants of motion of a null geodesic
 * (E = 1, xi = L/E, eta = Q/E^2; Bardeen 1973), and the geodesic is integrated
 * backwards from the observer in Boyer-Lindquist coordinates.
 *
 * Reference interfaces this layer replaces or completes:
 *   - src/cuda/kernel_launch.h:145-148  bh_launch_geodesic_kernel (render ABI;
 *     mirrored in bh_kernel_launch.h)
 *   - src/blender_bridge/blender_bridge.h:165-168,348-350
 *     bhbTraceGeodesicsImagePlane / bhbCudaTraceGeodesics (image-plane tracing;
 *     the reference CUDA entry returns -1, blender_bridge.cpp:709-714)
 *   - src/physics/raytracer.h:46-79  RayStatus / RayTraceResult (hit classes,
 *     per-ray result record)
 *   - src/physics/verified/rk4.hpp:247-260 rk4_step and
 *     src/physics/verified/rk4.h:182-299 dp45Step / dp45NextStep (integrators)
 *
 * Error convention (as the reference's C entry points, kernel_launch.cu:220-269):
 * every function returns 0 on success and -1 on failure, never throws across
 * the ABI, and leaves a message retrievable with bh_last_error().
 * Launches are asynchronous on the given stream (NULL = default stream); the
 * caller synchronises.  All launch parameters travel by value in kernel
 * arguments, so concurrent calls on different streams are safe.
 */
#ifndef BHK_TRACE_H
#define BHK_TRACE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BHK_VERSION_MAJOR 0
#define BHK_VERSION_MINOR 1

/* ---------------------------------------------------------------------- */
/* Enumerations                                                            */
/* ---------------------------------------------------------------------- */

/* Hit classes; the first four follow RayStatus (src/physics/raytracer.h:46-52):
 * Captured -> HORIZON, Escaped -> ESCAPED, MaxSteps -> MAX_STEPS, plus the
 * disk hit of HitResult (src/cuda/device_physics.cuh:806-818). */
enum BH_HitClass {
    BH_HIT_NONE      = 0, /* ray not traced (e.g. skipped by a mask)            */
    BH_HIT_HORIZON   = 1, /* r < horizon_factor * r_+                           */
    BH_HIT_DISK      = 2, /* crossed the thin disk inside [r_isco, disk_r_out]  */
    BH_HIT_ESCAPED   = 3, /* r > r_escape, moving outward                       */
    BH_HIT_MAX_STEPS = 4, /* step (RK4) or attempt (DP45) budget exhausted      */
    BH_HIT_OPAQUE    = 5  /* volume transmittance fell below transmit_min       */
};

enum BH_Integrator {
    BH_INTEGRATOR_RK4  = 0, /* classic fixed-step RK4 in s                       */
    BH_INTEGRATOR_DP45 = 1  /* Dormand-Prince 5(4), FSAL, per-component rtol/atol */
};

enum BH_Precision {
    BH_PRECISION_FP64 = 0,  /* FP64 throughout                                   */
    BH_PRECISION_FP32 = 1   /* FP32 stepping, FP64 islands (setup, turning
                               points, r < island_r_factor * r_+)              */
};

/* Per-record flag bits (BH_PixelRecord.flags). */
#define BH_FLAG_FP64_ISLAND   0x01u /* at least one step ran in an FP64 island   */
#define BH_FLAG_FALLBACK      0x02u /* surrogate deferred this ray to the tracer */
#define BH_FLAG_AUDITED       0x04u /* ray traced for the surrogate audit        */
#define BH_FLAG_SURROGATE     0x08u /* record produced by the surrogate          */
#define BH_FLAG_REFINED       0x10u /* crossings solved by the Mino-time solver   */
#define BH_FLAG_EXIT_NETWORK  0x20u /* u_exit, phi_exit are the surrogate
                                       network's escape direction (approximate:
                                       neither solved nor integrated)          */
#define BH_FLAG_NETWORK_KEPT  0x40u /* the surrogate network's record kept
                                       without the exact solve: class escaped or
                                       captured with margin, crossing radii
                                       approximate, all beyond the disk        */

/* ---------------------------------------------------------------------- */
/* Inputs                                                                  */
/* ---------------------------------------------------------------------- */

/* One distant-observer camera.  Pixel (px, py) has centre
 *   alpha = alpha_min + (px + 0.5) * (alpha_max - alpha_min) / width
 *   beta  = beta_max  - (py + 0.5) * (beta_max  - beta_min)  / height
 * (row 0 at the top).  Records of this camera occupy
 * [record_offset, record_offset + width*height) of the output array,
 * row-major. */
typedef struct BH_Camera {
    double  spin;          /* a/M in [0, 1)                                    */
    double  theta_obs;     /* inclination from the +z spin axis [rad], (0, pi) */
    double  r_obs;         /* observer Boyer-Lindquist radius [M]              */
    double  alpha_min, alpha_max;
    double  beta_min, beta_max;
    int32_t width, height;
    int64_t record_offset;
} BH_Camera; /* 72 bytes */

/* One explicitly specified ray (a pixel of a 1x1 camera): image-plane
 * impact parameters, spin, inclination and observer radius.  Used for
 * training-set generation; per-ray sin/cos/ISCO are evaluated on the device
 * (device libm), so these rays are not bit-identical to the camera path. */
typedef struct BH_RaySpec {
    double  alpha, beta;    /* impact parameters [M]                            */
    double  spin;           /* a/M in [0, 1)                                    */
    double  theta_obs;      /* inclination [rad], (0, pi)                       */
    double  r_obs;          /* observer radius [M]                              */
} BH_RaySpec; /* 40 bytes */

/* Integration and radiative-transfer options (bh_trace_options_default). */
typedef struct BH_TraceOptions {
    int32_t integrator;     /* BH_Integrator                                    */
    int32_t precision;      /* BH_Precision                                     */
    int32_t max_steps;      /* accepted-step budget                             */
    int32_t max_attempts;   /* DP45 attempt budget (accepted + rejected)        */
    int32_t disk_enabled;   /* thin Novikov-Thorne disk in the equatorial plane */
    int32_t volume_enabled; /* volumetric synchrotron transfer through BH_Volume*/
    int32_t max_crossings;  /* equatorial crossings recorded (<= 3)            */
    int32_t refill_batch;   /* steps between lane refills in the persistent
                               kernel (0 = library default)                     */
    double  h_fixed;        /* RK4 step in s, where d(lambda_Mino) = ds / r     */
    double  h_init;         /* DP45 first trial step                            */
    double  h_max;          /* DP45 step cap outside the volume support         */
    double  h_vol;          /* DP45 step cap inside the volume support          */
    double  rtol, atol;     /* DP45 tolerances                                  */
    double  horizon_factor; /* stop when r < horizon_factor * r_+               */
    double  r_escape;       /* stop when r > r_escape on the way out (a camera
                               beyond it traces its rays in first)              */
    double  disk_r_out;     /* thin-disk outer edge [M]; inner edge is r_isco   */
    double  nu_obs;         /* observing frequency at infinity [Hz]             */
    double  disk_T_max;     /* disk effective temperature at the flux peak [K]  */
    double  length_unit_cm; /* GM/c^2 in cm (converts path lengths)             */
    double  transmit_min;   /* stop when volume transmittance < this (0 = off)  */
    double  island_r_factor;/* FP32 form: FP64 steps when r < this * r_+        */
    double  island_turn_dist;/* FP32 form: FP64 steps within this [M] of a
                               radial or polar turning point                    */
    int32_t blocks_per_cu;  /* persistent-kernel workgroups per CU (0 = the
                               autotuned value for this device and kernel if
                               one is cached, else the occupancy limit)         */
    int32_t mino_exit;      /* Mino-time solver (bh_trace_disk_mino[_refine], the
                               surrogate render's refinement): escape directions
                               of escaped rays.  0 (default): exact in
                               bh_trace_disk_mino, the surrogate network's
                               (flagged BH_FLAG_EXIT_NETWORK) in a refinement;
                               1: exact in both; 2: none computed (NaN, or the
                               network's in a refinement)                       */
} BH_TraceOptions;

/* Synthetic GRMHD-like volume on a Boyer-Lindquist (r, theta, phi) grid.
 * cells: device pointer to n_r*n_th*n_ph float4 {n_e [cm^-3], Theta_e
 * (k T_e / m_e c^2), |B| [G], K_2(1/Theta_e) / (2 Theta_e^2)};
 * index ((iph * n_th) + ith) * n_r + ir; fewer than 2^31 cells.
 * Cell-centred samples: cell i of an axis of n cells spans
 * [lo + i*d, lo + (i+1)*d]; r clamps, theta clamps (mirror at the poles is
 * the same for the half-cell overhang), phi wraps (2*pi periodic) -- the
 * address modes of src/cuda/device_physics.cuh:1775-1829. */
typedef struct BH_Volume {
    const void* cells;      /* device float4[n_r*n_th*n_ph]                     */
    int32_t n_r, n_th, n_ph, pad0;
    double  r_min, r_max;   /* radial extent of the grid [M]                    */
    double  sup_r_lo, sup_r_hi, sup_u_max; /* support of n_e > 0:
                               r in [lo, hi], |cos theta| <= u_max              */
    double  ell;            /* constant specific angular momentum -u_phi/u_t of
                               the fluid (rotation law of the torus)            */
} BH_Volume;

/* ---------------------------------------------------------------------- */
/* Output                                                                  */
/* ---------------------------------------------------------------------- */

/* One record per pixel (128 bytes).  Crossing slots beyond n_cross are 0. */
typedef struct BH_PixelRecord {
    double   intensity;     /* I_nu at the observer [erg s^-1 cm^-2 Hz^-1 sr^-1] */
    double   r_exit;        /* r at termination [M]                             */
    double   u_exit;        /* cos(theta) at termination                        */
    double   phi_exit;      /* Boyer-Lindquist phi at termination [rad]         */
    double   cross_r[3];    /* radii of the first equatorial crossings [M]      */
    double   cross_g[3];    /* disk redshift g = nu_obs/nu_emit at each crossing
                               (Keplerian prograde orbit; 0 inside r_isco)     */
    double   tau;           /* optical depth accumulated through the volume     */
    double   drift_null;    /* |v_r^2 - R(r)| / max(R_scale) at termination     */
    double   drift_carter;  /* |v_u^2 - U(u)| / max(U_scale) at termination     */
    int32_t  steps;         /* accepted steps                                   */
    int32_t  attempts;      /* attempts (accepted + rejected; RK4: = steps)     */
    uint8_t  hit_class;     /* BH_HitClass                                      */
    uint8_t  n_cross;       /* number of equatorial crossings recorded          */
    uint8_t  flags;         /* BH_FLAG_*                                        */
    uint8_t  pad0;
    float    confidence;    /* surrogate confidence (1 for traced rays)         */
    int32_t  fp64_steps;    /* FP32 form: steps taken in FP64 islands           */
    int32_t  n_samples;     /* volume samples taken along the ray             */
} BH_PixelRecord;

/* ---------------------------------------------------------------------- */
/* Entry points                                                            */
/* ---------------------------------------------------------------------- */

/* Library version; returns BHK_VERSION_MAJOR * 100 + BHK_VERSION_MINOR. */
int bh_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* bh_last_error(void);

/* Fill *opts with the defaults (RK4, FP64, h_fixed 0.01, 2000 steps,
 * horizon_factor 1.01, r_escape 200, disk to 20 M, nu 230 GHz, ...). */
int bh_trace_options_default(BH_TraceOptions* opts);

/* Outer horizon r_+ = 1 + sqrt(1 - a^2) and prograde ISCO (Bardeen, Press &
 * Teukolsky 1972), as src/physics/kerr.h:139-229 with M = 1. */
double bh_outer_horizon(double spin);
double bh_isco_prograde(double spin);

/* Host: the normalised Novikov-Thorne flux table of the thin disk,
 * F(r) = (1 - sqrt(r_in/r)) (1 + a/(2 sqrt r)) / r^3 over n equally spaced
 * radii in [r_in, r_out], divided by its largest sample
 * (src/physics/lut.h:50-86, src/physics/thin_disk.h:243-266 with M = 1). */
int bh_disk_lut_build(double spin, double r_in, double r_out, int n, double* out);

/* Host: synthetic torus volume (Fishbone & Moncrief 1976 constant-l torus,
 * modulated by a seeded lognormal Gaussian random field with a k^(-11/3)
 * spectrum).  cells_host: float[4 * n_r*n_th*n_ph]; fills vol's scalar fields
 * (not vol->cells). */
typedef struct BH_TorusParams {
    double  spin;
    double  r_in, r_max;    /* torus inner edge and pressure maximum [M]        */
    double  r_lo, r_hi;     /* grid radial extent [M]                            */
    int32_t n_r, n_th, n_ph, pad0;
    double  sigma;          /* lognormal width of the turbulence                 */
    double  spectral_index; /* P(k) ~ k^spectral_index (-11/3)                  */
    uint64_t seed;
    double  n_e_max;        /* peak electron density [cm^-3]                     */
    double  theta_e_max;    /* peak dimensionless electron temperature           */
    double  b_max;          /* field strength at the pressure maximum [G]        */
    double  gamma_ad;       /* adiabatic index (4/3)                             */
} BH_TorusParams;
int bh_volume_generate_torus(const BH_TorusParams* p, float* cells_host, BH_Volume* vol);

/* Trace every pixel of n_cameras cameras into out (device array of
 * BH_PixelRecord, caller-owned, sized by the cameras' record ranges; records
 * in gaps between ranges are left untouched).  cams and opts are host
 * pointers.  disk_luts: device double array holding n_cameras tables of
 * disk_lut_n entries (camera c uses disk_luts + c*disk_lut_n, spanning
 * [r_isco(spin_c), opts->disk_r_out]); may be NULL when the disk is
 * disabled.  vol may be NULL when the volume is disabled.  Asynchronous on
 * stream; launches on one stream are serialised, launches on different
 * streams are independent. */
int bh_trace_image_plane(const BH_Camera* cams, int n_cameras,
                         const BH_TraceOptions* opts,
                         const double* disk_luts, int disk_lut_n,
                         const BH_Volume* vol,
                         BH_PixelRecord* out, void* stream);

/* bh_trace_image_plane without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and records as bh_trace_image_plane;
 * the launch block and the ray-queue head live in a ~64 KiB device array that
 * the library holds once per GPU until the process ends (at most 540
 * cameras per call; larger launches return -1 with the limit in
 * bh_last_error()).  All calls on one GPU share that array: calls on one
 * stream are ordered, calls that overlap on one GPU (other streams or host
 * threads) race.  bh_last_kernel_ms() does not time these launches.
 * The launch block is uploaded from a per-host-thread staging buffer by an
 * asynchronous copy; a copy captured into a HIP graph reads that buffer when
 * the graph runs, so a captured call replays its own launch only while no
 * other call of this function on the same host thread comes in between
 * (BH_Frame is the form built for graph replay). */
int bh_trace_image_plane_noalloc(const BH_Camera* cams, int n_cameras,
                                 const BH_TraceOptions* opts,
                                 const double* disk_luts, int disk_lut_n,
                                 const BH_Volume* vol,
                                 BH_PixelRecord* out, void* stream);

/* Trace only the pixels whose global record ids are listed (device int64
 * ids[*n_ids_dev], count in device memory so a producer kernel can append
 * ids without a host round trip; at most max_ids).  Record of id i goes to
 * out[i].  Same cameras / options / tables as bh_trace_image_plane.  Used for
 * the surrogate's fallback and audit rays. */
int bh_trace_pixel_list(const BH_Camera* cams, int n_cameras,
                        const BH_TraceOptions* opts,
                        const double* disk_luts, int disk_lut_n,
                        const BH_Volume* vol,
                        const int64_t* ids, const uint64_t* n_ids_dev, int64_t max_ids,
                        BH_PixelRecord* out, void* stream);

/* As bh_trace_pixel_list, but the record of list entry q goes to out[q]
 * (out holds max_ids records): the audit traces a sample of pixels into a
 * compact buffer beside the full image it checks. */
int bh_trace_pixel_list_compact(const BH_Camera* cams, int n_cameras,
                                const BH_TraceOptions* opts,
                                const double* disk_luts, int disk_lut_n,
                                const BH_Volume* vol,
                                const int64_t* ids, const uint64_t* n_ids_dev, int64_t max_ids,
                                BH_PixelRecord* out, void* stream);

/* Trace n explicitly specified rays (device BH_RaySpec[n]) into out[0..n)
 * (device).  The thin disk must be disabled (crossings are still recorded);
 * the volume may be enabled. */
int bh_trace_rays(const BH_RaySpec* rays, int64_t n, const BH_TraceOptions* opts,
                  const BH_Volume* vol, BH_PixelRecord* out, void* stream);

/* ---------------------------------------------------------------------- */
/* Batched frames: volume sets, prepared launches, graphs, autotuning     */
/* ---------------------------------------------------------------------- */

/* Several volumes sharing one grid geometry and support (n_r, n_th, n_ph,
 * r_min, r_max, sup_*), differing in their cells and fluid angular momentum
 * ell -- e.g. one torus per spin bucket of a camera batch.  Camera i samples
 * vols[cam_vol[i]] (host arrays; cells are device pointers). */
typedef struct BH_VolumeSet {
    const BH_Volume* vols;
    int32_t n_vols;
    int32_t pad0;
    const int32_t* cam_vol;   /* n_cameras entries in [0, n_vols) */
} BH_VolumeSet;

/* bh_trace_image_plane with a volume set (FP64 form; volume_enabled set):
 * every camera of the batch in one persistent launch. */
int bh_trace_image_plane_multi(const BH_Camera* cams, int n_cameras,
                               const BH_TraceOptions* opts,
                               const double* disk_luts, int disk_lut_n,
                               const BH_VolumeSet* vs,
                               BH_PixelRecord* out, void* stream);

/* A prepared frame: cameras, options, tables and the output buffer validated
 * once and kept in device memory, so that launching the frame needs no host
 * work.  bh_frame_launch enqueues the queue reset and the persistent kernel
 * (capturable); bh_frame_graph_launch captures that launch into a hipGraph on
 * its first call (on a stream of the frame's own) and replays the graph on
 * `stream` afterwards.  vol or vs (or neither) as in the one-shot calls.
 * NULL on failure (bh_last_error).  Device pointers must outlive the frame. */
typedef struct BH_Frame BH_Frame;
BH_Frame* bh_frame_create(const BH_Camera* cams, int n_cameras, const BH_TraceOptions* opts,
                          const double* disk_luts, int disk_lut_n,
                          const BH_Volume* vol, const BH_VolumeSet* vs,
                          BH_PixelRecord* out);
int  bh_frame_launch(BH_Frame* f, void* stream);
int  bh_frame_graph_launch(BH_Frame* f, void* stream);
void bh_frame_destroy(BH_Frame* f);

/* Launch shape of the persistent kernel: workgroups per CU (0 = occupancy
 * limit) and steps between lane refills (0 = keep).  Drops a captured graph. */
int bh_frame_set_launch_shape(BH_Frame* f, int blocks_per_cu, int refill_batch);

/* Autotuner: times the frame's kernel on every sample_stride-th pixel (into a
 * buffer of its own) for each workgroups-per-CU value up to the occupancy
 * limit and refill batches {1, 2, 4, 8, 16, 32}; applies the fastest shape to
 * the frame and caches it for this device and kernel (used by later launches
 * whose options leave blocks_per_cu / refill_batch at 0).  Synchronises. */
typedef struct BH_TuneResult {
    int32_t blocks_per_cu, refill_batch;
    int32_t occupancy_limit;    /* workgroups per CU the kernel's registers allow */
    int32_t sample_rays;
    float   sample_ms;          /* best trial time on the sample                  */
    int32_t pad0;
} BH_TuneResult;
int bh_frame_autotune(BH_Frame* f, int sample_stride, BH_TuneResult* best, void* stream);

/* bh_trace_rays with host buffers: copies the n ray specs to the device,
 * traces them and copies the records back (synchronous; device buffers are
 * allocated for the call).  For callers that hold host memory only (FFI
 * bindings, the physics C++ interface). */
int bh_trace_rays_host(const BH_RaySpec* rays, int64_t n, const BH_TraceOptions* opts,
                       BH_PixelRecord* out);

/* Elapsed time of the last bh_trace_image_plane kernel on this thread, from
 * HIP events recorded around the launch on its stream (ms; -1 if none).  The
 * caller must have synchronised the stream. */
float bh_last_kernel_ms(void);

/* Polarised transfer (src/physics/stokes_transport.h; the device path of
 * d_stokes_step / d_trace_geodesic_stokes, device_physics.cuh:2185-2459): the
 * volume's thermal synchrotron emission is linearly polarised at
 * lin_pol_frac along the sky angle evpa_b (j_Q + i j_U = -lin_pol_frac j_I
 * e^{2 i evpa_b}), and the plasma rotates (Q, U) at the Faraday rate
 * rho_V = faraday_per_ne n_e [rad per fluid-frame cm]; total absorption as
 * for I, no dichroism, no conversion (the simplified Mueller matrix of
 * stokesStep), so V stays 0.  The thin disk's emission is unpolarised. */
typedef struct BH_PolParams {
    double evpa_b;          /* polarisation angle chi of the emission [rad]      */
    double lin_pol_frac;    /* linear polarisation fraction of j (0.75: thermal) */
    double faraday_per_ne;  /* rho_V / n_e [rad cm^2]                            */
    double pad0;
} BH_PolParams; /* 32 bytes */

/* bh_trace_image_plane with the polarised transfer: the same records (bit for
 * bit) plus stokes[4 * record] = (I, Q, U, V) at the observer (device double
 * per record).  FP64, volume_enabled, one volume. */
int bh_trace_image_plane_pol(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                             const double* disk_luts, int disk_lut_n, const BH_Volume* vol,
                             const BH_PolParams* pol, BH_PixelRecord* out, double* stokes,
                             void* stream);

/* Thin-disk records without step integration: the equatorial crossings of
 * every pixel's geodesic from the Mino-time integrals of motion (polar
 * elliptic integrals in closed form, the radial integral by adaptive
 * Gauss-Kronrod quadrature through the turning point, each crossing by a
 * bracketed Newton iteration), then the integrator's disk record (hit class,
 * first three crossing radii, Keplerian g, LUT intensity; steps = Newton
 * iterations, flags BH_FLAG_REFINED).  Same cameras, options and LUTs as
 * bh_trace_image_plane with disk_enabled = 1 and volume_enabled = 0.  An
 * escaped ray ends at r_exit = r_escape with its exact direction there
 * (u_exit, phi_exit: Jacobi sn of the polar motion and the azimuth
 * integrals, Legendre's Pi by Carlson's R_F / R_J and the radial part by
 * quadrature) unless opts->mino_exit = 2; captured and disk rays have no
 * end direction (NaN).  In records this solver refines (mino_exit 0), an
 * escaping ray keeps the surrogate network's direction instead, flagged
 * BH_FLAG_EXIT_NETWORK; mino_exit = 1 computes the exact one there too.
 * Rays the solver
 * does not resolve (within about 1e-3 M of the critical curve) get
 * BH_FLAG_FALLBACK and their ids in fallback_ids (count in device memory,
 * zeroed by the call) for bh_trace_pixel_list.  No counterpart in the
 * reference; an exact, integration-free method for the thin disk. */
int bh_trace_disk_mino(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                       const double* disk_luts, int disk_lut_n, BH_PixelRecord* out,
                       int64_t* fallback_ids, uint64_t* fallback_count, int64_t fallback_cap,
                       void* stream);

/* The same solver over records written by bh_surrogate_render with
 * refine = 2: every record not flagged BH_FLAG_FALLBACK is solved in place
 * with its crossing radii and count as Newton starts (flags
 * BH_FLAG_SURROGATE | BH_FLAG_REFINED); failures are flagged and appended
 * to the render's fallback list (the count is not reset).  Separate from
 * the render so that the audit sample can be traced on another stream
 * meanwhile.  No counterpart in the reference. */
int bh_trace_disk_mino_refine(const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                              const double* disk_luts, int disk_lut_n, BH_PixelRecord* out,
                              int64_t* fallback_ids, uint64_t* fallback_count, int64_t fallback_cap,
                              void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BHK_TRACE_H */
