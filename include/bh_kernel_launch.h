/*
 * bh_kernel_launch.h - render-launch C ABI of the MI355X tracer (drop-in for
 * the reference's src/cuda/kernel_launch.h:145-190).
 *
 * Plain C, POD only.  struct BH_LaunchParams keeps the reference's field
 * names, types and order (336 bytes; offsets checked by the tests against a
 * compile of the reference header), because callers fill it field by field.
 *
 * bh_launch_geodesic_kernel traces one null geodesic per framebuffer pixel
 * with the library's tracer: a pinhole camera at cam_pos (Cartesian, z along
 * the spin axis, r = |cam_pos|, cos(theta) = z / r, phi = atan2(y, x)), pixel
 * directions as the reference maps them (kernel: device_physics.cuh:1953-1967:
 * u, v over [-1, 1] scaled by fov_scale, u by the aspect ratio, shifted by
 * frame_shift_{x,y}, local direction (-u, -v, 1) through the column-major
 * cam_basis), and the photon's constants from the Kerr metric at the camera
 * (null condition solved for k^t, E = -(g_tt k^t + g_tphi k^phi),
 * L = g_tphi k^t + g_phiphi k^phi, Q from Sigma k^theta; the reference's
 * d_kerr_init_consts, device_physics.cuh:316-392).  Lengths: rs = 2M sets the
 * unit (every length is divided by M = rs/2); spin is a/M.  The thin disk
 * spans [isco, 30 rs] (the reference's r_disk_in/r_disk_out,
 * device_physics.cuh:936-937); rays stop at the horizon or beyond max_dist.
 *
 * Integration uses the library's adaptive Dormand-Prince 5(4) (rtol 1e-8,
 * atol 1e-10) instead of the reference's fixed Mino-time Euler/RK4 steps;
 * max_steps is the accepted-step budget.  Variants: BH_KERNEL_FP32_BASELINE
 * and BH_KERNEL_FP32_COARSENED run the FP32 precision form (FP32 stepping,
 * FP64 islands), BH_KERNEL_FP16_STORAGE and BH_KERNEL_FP16_H2_ILP (no FP16
 * ray state on this target) run the FP64 form; a negative variant takes
 * bh_select_kernel_variant().
 *
 * framebuffer: device float4[width * height], row-major from the top row:
 *   disk hit  -> rgb = blackbody tint of the observed temperature g F^(1/4)
 *                (F the normalised Novikov-Thorne flux) times
 *                adisk_lit * 2 F g^4 (bolometric, redshift_enabled; F without
 *                the g^4 otherwise), alpha = 1;
 *   horizon   -> (0, 0, 0, 1);
 *   escaped   -> background_intensity * background_enabled in rgb, alpha 0.
 * GRMHD modulation (d_shade_hit): with use_luts set and a volume from
 * bh_grmhd_texture_create (bh_grmhd.h) in the grmhd slot of
 * bh_upload_lut_textures, disk colours are scaled by max(rho, 0) max(u, 0)
 * sampled at the hit point (u = (r - grmhd_r_min) / (grmhd_r_max -
 * grmhd_r_min), v = theta / pi, w = phi / 2 pi; r clamped, theta mirrored,
 * phi wrapped), blended with the grmhd_right volume by grmhd_alpha.
 * Cosmetic layers of the reference renderer (star field, glow, wiregrid,
 * debug exports, Stokes) are accepted and ignored.
 *
 * Returns 0 / -1 (bh_last_error); asynchronous on `stream` (NULL = default).
 */
#ifndef BH_KERNEL_LAUNCH_H
#define BH_KERNEL_LAUNCH_H

#ifdef __cplusplus
extern "C" {
#endif

struct BH_LaunchParams {
    /* Black hole parameters */
    float rs;
    float spin;
    float isco;
    float step_size;
    float fov_scale;
    float max_dist;

    /* Camera */
    float cam_pos[3];
    float cam_basis[9];   /* column-major [right, up, forward] */
    float frame_shift_x;
    float frame_shift_y;

    /* Integration */
    int max_steps;
    int width;
    int height;

    /* Feature flags */
    int adisk_enabled;
    int redshift_enabled;
    int kerr_enabled;
    int use_luts;

    /* LUT domain bounds */
    float lut_radius_min;
    float lut_radius_max;
    float redshift_radius_min;
    float redshift_radius_max;
    float spectral_radius_min;
    float spectral_radius_max;

    /* Time */
    float time_sec;

    /* Doppler / beaming */
    float doppler_strength;

    /* Background */
    float background_intensity;
    int background_enabled;
    float photon_glow_strength;
    int debug_pre_redshift_background;
    int debug_pre_shaping_background;
    int debug_post_shaping_background;
    int debug_shaper_inputs;
    int debug_closest_approach_state;
    int debug_closest_approach_timeline;
    int debug_closest_approach_direction;
    int debug_escaped_direction;
    float background_yaw_rad;
    float background_pitch_rad;
    float background_filter_radius;
    float background_layer_params[12];
    float background_layer_lod_bias[3];

    /* Coordinate-grid overlay */
    int   wiregrid_enabled;
    float wiregrid_show_ergo;
    float wiregrid_grid_scale;
    float wiregrid_motion_scale;
    float wiregrid_infall_scale;
    float wiregrid_strength;
    float wiregrid_scene_preserve;
    float wiregrid_color[4];

    /* GRMHD volume bounds */
    float grmhd_r_min;
    float grmhd_r_max;

    /* Volumetric transfer */
    int   rte_enabled;
    float rte_opacity_scale;

    /* GRMHD temporal interpolation */
    float grmhd_alpha;

    /* Polarised transport */
    int   stokes_enabled;
    float stokes_b_field_angle;
    float stokes_ne_scale;

    /* Disk brightness */
    float adisk_lit;
};

enum BH_KernelVariant {
    BH_KERNEL_FP32_BASELINE  = 0,
    BH_KERNEL_FP32_COARSENED = 1,
    BH_KERNEL_FP16_STORAGE   = 2,
    BH_KERNEL_FP16_H2_ILP    = 3,
    BH_KERNEL_COUNT          = 4
};

/* src/cuda/kernel_launch.h:145-148 */
int bh_launch_geodesic_kernel(void* framebuffer,
                              const struct BH_LaunchParams* params,
                              int variant,
                              void* stream);

/* src/cuda/kernel_launch.h:158: returns BH_KERNEL_FP16_STORAGE, i.e. the FP64
 * form -- on gfx950 FP64 issues at the unpacked FP32 rate and the FP32 form's
 * FP64 islands diverge within waves, so the FP64 form is the faster one. */
int bh_select_kernel_variant(void);

/* src/cuda/kernel_launch.h:176-183: texture handles (u64, 0 = absent),
 * stored process-wide (as the reference's constant-memory slots) and reported
 * back by bh_lut_texture_handle.  The tracer computes the disk flux and
 * redshift in closed form and does not sample the LUT slots; the grmhd and
 * grmhd_right slots take volumes from bh_grmhd_texture_create (gfx950 has no
 * image-sampling hardware; other handles are ignored). */
void bh_upload_lut_textures(unsigned long long emissivity,
                            unsigned long long redshift,
                            unsigned long long spectral,
                            unsigned long long grb,
                            unsigned long long galaxy,
                            unsigned long long grmhd,
                            unsigned long long synch_g,
                            unsigned long long grmhd_right);

/* src/cuda/kernel_launch.h:190 */
void bh_upload_bridge_background_texture(unsigned long long background_equirect);

/* The per-pixel records behind a launch (the tracer's BH_PixelRecord, width x
 * height, row-major from the top row; device memory) instead of the shaded
 * framebuffer -- for callers that want the physics (hit class, crossing
 * radii, g) and for parity checks.  No counterpart in the reference. */
struct BH_PixelRecord;
int bh_trace_launch_params(const struct BH_LaunchParams* params, int variant,
                           struct BH_PixelRecord* out, void* stream);

/* Handle stored by the two calls above: slot 0..7 as bh_upload_lut_textures
 * orders them, 8 = the bridge background; 0 when unset or out of range.  No
 * counterpart in the reference. */
unsigned long long bh_lut_texture_handle(int slot);

#ifdef __cplusplus
}
#endif

#endif /* BH_KERNEL_LAUNCH_H */
