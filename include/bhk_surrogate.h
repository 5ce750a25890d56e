/*
 * bhk_surrogate.h - C ABI of the learned geodesic surrogate (BF16 MFMA MLP).
 *
 * No counterpart in the reference (it has no learned component).  The
 * surrogate maps an image-plane ray
 * (alpha, beta, a, cos theta_obs) to the geometry the thin-disk shading
 * needs -- the first three equatorial crossings (radius, g), the crossing
 * count, capture, the escape direction -- plus a confidence output that
 * predicts its own error.  Rays with low confidence or inside the
 * photon-ring band fall back to the integrator (bh_trace_image_plane with an
 * id list); a random fraction is audited against the integrator.
 *
 * Network: 16 input features (an analytic embedding of the four inputs,
 * padded to 32), 6 hidden layers of width 256 with SiLU, 16 outputs.  The
 * 16 -> 256 input layer runs in FP32 (the inputs are not quantised); every
 * hidden layer has BF16 weights and activations with FP32 accumulation on
 * v_mfma_f32_16x16x32_bf16.  FP32 master weights and Adam state for
 * training.  All buffers device-side.
 */
#ifndef BHK_SURROGATE_H
#define BHK_SURROGATE_H

#include <stdint.h>

#include "bhk_trace.h"

#ifdef __cplusplus
extern "C" {
#endif

#define BH_SUR_FEATURES 16   /* embedded input features                      */
#define BH_SUR_IN_PAD   32   /* feature row stride (MFMA K step)              */
#define BH_SUR_OUT      16   /* output row stride                             */

/* Output columns. */
enum BH_SurrogateOutput {
    BH_SUR_LOGR1 = 0, BH_SUR_G1 = 1,   /* first crossing: ln r, g              */
    BH_SUR_LOGR2 = 2, BH_SUR_G2 = 3,
    BH_SUR_LOGR3 = 4, BH_SUR_G3 = 5,
    BH_SUR_NCROSS = 6,                 /* crossing count (0..3)                */
    BH_SUR_CAPTURE = 7,                /* 1 captured, 0 escaped                */
    BH_SUR_ESC_U = 8,                  /* cos(theta) at escape                 */
    BH_SUR_ESC_SINPHI = 9,             /* sin / cos of phi at escape           */
    BH_SUR_ESC_COSPHI = 10,
    BH_SUR_CONF = 11                   /* log10 of the predicted squared error */
};

typedef struct BH_SurrogateConfig {
    int32_t width;        /* hidden width (256)                                 */
    int32_t depth;        /* hidden layers (6)                                  */
    int32_t max_batch;    /* rows per forward / training call                   */
    int32_t pad0;
    uint64_t seed;        /* weight initialisation seed                         */
} BH_SurrogateConfig;

typedef struct BH_Surrogate BH_Surrogate; /* opaque, device-resident */

/* Create / destroy.  Returns NULL on failure (bh_last_error()). */
BH_Surrogate* bh_surrogate_create(const BH_SurrogateConfig* cfg);
void bh_surrogate_destroy(BH_Surrogate* s);

/* Parameters as one FP32 vector: for each layer l = 0..depth, W_l [out][in]
 * (in = 32 for l = 0, out = 16 for l = depth) followed by b_l [out]. */
int64_t bh_surrogate_param_count(const BH_Surrogate* s);
int bh_surrogate_get_params(const BH_Surrogate* s, float* host_out);
int bh_surrogate_set_params(BH_Surrogate* s, const float* host_in);

/* Feature embedding of n rays given as (alpha, beta, spin, cos theta_obs)
 * rows (device FP32 [n][4]) into device FP32 [n][BH_SUR_IN_PAD]. */
int bh_surrogate_features(const float* rays4, int64_t n, float* features, void* stream);

/* Training targets and loss weights (device FP32 [n][16] each) from n traced
 * records (disk disabled) and their features: ln r of the first three
 * crossings minus 3 x feature 14 (the logarithm of the light-bending
 * estimate of the first crossing radius for an observer at r_obs = 100 M),
 * g, crossing count, capture, escape direction.  w5 (host) = loss weights of
 * {ln r, g, count, capture, escape}. */
int bh_surrogate_targets(const BH_PixelRecord* records, const float* features, int64_t n,
                         float* targets, float* weights, const float* w5, void* stream);

/* Forward pass: features (device FP32 [n][32]) -> outputs (device FP32
 * [n][16]); n <= max_batch. */
int bh_surrogate_forward(BH_Surrogate* s, const float* features, int n, float* out, void* stream);

/* One training step (forward, weighted squared-error loss, backward, Adam)
 * on a batch: targets and weights device FP32 [n][16]; loss_out (host,
 * may be NULL) receives {sum w d^2 + conf term, sum w}. Synchronises. */
int bh_surrogate_train_step(BH_Surrogate* s, const float* features, const float* targets,
                            const float* weights, int n, float lr, float* loss_out, void* stream);

/* Loss gradient with respect to every parameter (same layout as the
 * parameter vector) without updating, for gradient checks.  Synchronises. */
int bh_surrogate_gradients(BH_Surrogate* s, const float* features, const float* targets,
                           const float* weights, int n, float* host_grads, void* stream);

/* ---- image rendering with the surrogate ------------------------------------ */

/* Per-camera polar table of the critical curve (the photon-ring edge) used for
 * the fallback band: BH_SUR_RING_FLOATS floats per camera =
 * {alpha_c, beta_c, lo[256], hi[256], obl[256]}: centre, the curve's smallest
 * and largest radius about the centre in each of 256 polar bins, and
 * 1 / cos of the largest angle between the radial direction and the curve's
 * normal in the bin (converts a distance band into a radial band). */
#define BH_SUR_RING_BINS 256
#define BH_SUR_RING_FLOATS (2 + 3 * BH_SUR_RING_BINS)
int bh_surrogate_ring_tables(const BH_Camera* cams, int n_cams, float* host_out);

typedef struct BH_SurrogateRenderParams {
    float ring_band;      /* M: rays closer than this to the critical curve fall back */
    float edge_tol;       /* relative: crossings this close to r_ISCO / r_out fall back */
    float conf_max;       /* fall back when the confidence output exceeds this      */
    float ambiguity;      /* fall back when count / capture outputs are this far from
                             an integer                                              */
    int32_t audit_every;  /* 1-in-N deterministic audit sample (0: none)            */
    int32_t refine;       /* 1: solve every non-fallback ray's crossings exactly from
                             the Mino-time integrals (bh_trace_disk_mino's solver) with
                             the network's radii as starting points; confidence,
                             count and edge criteria are then not applied, only the
                             ring band and the solver's own failures fall back.  The
                             render then decodes the crossing radii, count and class
                             only: g and intensity are left 0 for the solve (and the
                             integrator, for listed rays) to write.
                             2: the same decode, the solve left to the caller
                             (bh_trace_disk_mino_refine, e.g. on the render's
                             stream while another stream traces the audit list).
                             3: as 2, and the ring band is solved too: its rays are
                             not listed, their records carry no crossings (the
                             solve starts them from its own cubic Hermite model),
                             the audit sample is drawn from every ray, and only the
                             solver's failures reach the fallback list */
    float keep_factor;    /* refine 1-3: a ray outside the ring band whose outputs
                             say escaped or captured, with an unambiguous count and
                             capture and every predicted crossing at r >= keep_factor
                             * disk_r_out, keeps the network's record (crossings and
                             g at them, class, escape direction) without the exact
                             solve, flagged BH_FLAG_NETWORK_KEPT: its intensity is 0
                             whatever the crossing radii.  0: every ray is solved.
                             Ignored when opts->mino_exit = 1 (exact escape
                             directions asked for) */
    int32_t pad0;
} BH_SurrogateRenderParams;

/* Render every pixel of the cameras with the fused surrogate kernel
 * (features, 7 layers and the decode in one launch; width 256 only): the
 * thin-disk record (hit class, first crossings, g from the exact Keplerian
 * formula at the predicted radius, intensity from the disk LUT, flags
 * BH_FLAG_SURROGATE and, for rays that must be traced, BH_FLAG_FALLBACK;
 * u_exit, phi_exit the network's escape direction for its escaped rays and,
 * with refine set, for every ray -- the solve keeps it, flagged
 * BH_FLAG_EXIT_NETWORK, where it finds the ray escaping).
 * Ids of fallback rays are appended to fallback_ids (count in device memory,
 * zeroed by the call), ids of a deterministic 1-in-audit_every sample of the
 * other rays to audit_ids; both lists are in the form bh_trace_pixel_list
 * takes.  Cameras must have r_obs = 100 M (the surrogate's training
 * distance).  All pointers except cams, opts, rp are device memory;
 * ring_tables from bh_surrogate_ring_tables, uploaded.
 * Threading: one surrogate may render on several streams (and host threads)
 * at once; its scratch (device cameras, layer-0 and bias tables, swizzled
 * weights) is kept per (device, stream) under a lock.  Training the same
 * surrogate while it renders is not supported. */
int bh_surrogate_render(BH_Surrogate* s, const BH_Camera* cams, int n_cams, const BH_TraceOptions* opts,
                        const double* disk_luts, int disk_lut_n, const float* ring_tables,
                        const BH_SurrogateRenderParams* rp, BH_PixelRecord* out,
                        int64_t* fallback_ids, uint64_t* fallback_count, int64_t fallback_cap,
                        int64_t* audit_ids, uint64_t* audit_count, int64_t audit_cap, void* stream);

/* Audit statistics of the surrogate records out[ids[q]] against integrator
 * records traced[q] (bh_trace_pixel_list_compact of the same ids), over
 * q < min(*count, cap).  stats (device, 8 doubles, zeroed by the call):
 * {n, n_cmp (reference intensity > floor), sum |dI|/I, max |dI|/I,
 *  class disagreements, sum |dI|, sum I_ref, sum |d ln r_1| (both disk hits)}. */
int bh_surrogate_audit(const BH_PixelRecord* out, const int64_t* ids, const uint64_t* count, int64_t cap,
                       const BH_PixelRecord* traced, double floor_I, double* stats, void* stream);

/* BF16 NT GEMM C[M][N] = A[M][K] B[N][K]^T (+ bias[N]) with FP32 output,
 * K a multiple of 32 (the building block of the surrogate; exported for
 * tests). */
int bh_mlp_gemm_nt(int M, int N, int K, const void* A_bf16, int lda, const void* Bt_bf16, int ldb,
                   const float* bias, float* C, int ldc, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BHK_SURROGATE_H */
