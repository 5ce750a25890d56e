"""ctypes mirror of the C ABI in include/bhk_trace.h and include/bh_kernel_launch.h.

The layouts here must match the C headers byte for byte; tests/test_abi.py
checks every size and offset against the compiled library's own
``bh_abi_layout`` report.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

# Hit classes (include/bhk_trace.h BH_HitClass; src/physics/raytracer.h:46-52)
HIT_NONE, HIT_HORIZON, HIT_DISK, HIT_ESCAPED, HIT_MAX_STEPS, HIT_OPAQUE = range(6)
INTEGRATOR_RK4, INTEGRATOR_DP45 = 0, 1
PRECISION_FP64, PRECISION_FP32 = 0, 1

FLAG_FP64_ISLAND = 0x01
FLAG_FALLBACK = 0x02
FLAG_AUDITED = 0x04
FLAG_SURROGATE = 0x08
FLAG_REFINED = 0x10
FLAG_EXIT_NETWORK = 0x20
FLAG_NETWORK_KEPT = 0x40


class BH_Camera(C.Structure):
    _fields_ = [
        ("spin", C.c_double),
        ("theta_obs", C.c_double),
        ("r_obs", C.c_double),
        ("alpha_min", C.c_double),
        ("alpha_max", C.c_double),
        ("beta_min", C.c_double),
        ("beta_max", C.c_double),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("record_offset", C.c_int64),
    ]


class BH_RaySpec(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double), ("spin", C.c_double),
                ("theta_obs", C.c_double), ("r_obs", C.c_double)]


class BH_TraceOptions(C.Structure):
    _fields_ = [
        ("integrator", C.c_int32),
        ("precision", C.c_int32),
        ("max_steps", C.c_int32),
        ("max_attempts", C.c_int32),
        ("disk_enabled", C.c_int32),
        ("volume_enabled", C.c_int32),
        ("max_crossings", C.c_int32),
        ("refill_batch", C.c_int32),
        ("h_fixed", C.c_double),
        ("h_init", C.c_double),
        ("h_max", C.c_double),
        ("h_vol", C.c_double),
        ("rtol", C.c_double),
        ("atol", C.c_double),
        ("horizon_factor", C.c_double),
        ("r_escape", C.c_double),
        ("disk_r_out", C.c_double),
        ("nu_obs", C.c_double),
        ("disk_T_max", C.c_double),
        ("length_unit_cm", C.c_double),
        ("transmit_min", C.c_double),
        ("island_r_factor", C.c_double),
        ("island_turn_dist", C.c_double),
        ("blocks_per_cu", C.c_int32),
        ("mino_exit", C.c_int32),
    ]


class BH_Volume(C.Structure):
    _fields_ = [
        ("cells", C.c_void_p),
        ("n_r", C.c_int32),
        ("n_th", C.c_int32),
        ("n_ph", C.c_int32),
        ("pad0", C.c_int32),
        ("r_min", C.c_double),
        ("r_max", C.c_double),
        ("sup_r_lo", C.c_double),
        ("sup_r_hi", C.c_double),
        ("sup_u_max", C.c_double),
        ("ell", C.c_double),
    ]


class BH_VolumeSet(C.Structure):
    _fields_ = [
        ("vols", C.POINTER(BH_Volume)),
        ("n_vols", C.c_int32),
        ("pad0", C.c_int32),
        ("cam_vol", C.POINTER(C.c_int32)),
    ]


class BH_PolParams(C.Structure):
    """Polarised transfer parameters (bhk_trace.h BH_PolParams, 32 bytes)."""
    _fields_ = [
        ("evpa_b", C.c_double),
        ("lin_pol_frac", C.c_double),
        ("faraday_per_ne", C.c_double),
        ("pad0", C.c_double),
    ]


class BH_TuneResult(C.Structure):
    _fields_ = [
        ("blocks_per_cu", C.c_int32),
        ("refill_batch", C.c_int32),
        ("occupancy_limit", C.c_int32),
        ("sample_rays", C.c_int32),
        ("sample_ms", C.c_float),
        ("pad0", C.c_int32),
    ]


class BH_TorusParams(C.Structure):
    _fields_ = [
        ("spin", C.c_double),
        ("r_in", C.c_double),
        ("r_max", C.c_double),
        ("r_lo", C.c_double),
        ("r_hi", C.c_double),
        ("n_r", C.c_int32),
        ("n_th", C.c_int32),
        ("n_ph", C.c_int32),
        ("pad0", C.c_int32),
        ("sigma", C.c_double),
        ("spectral_index", C.c_double),
        ("seed", C.c_uint64),
        ("n_e_max", C.c_double),
        ("theta_e_max", C.c_double),
        ("b_max", C.c_double),
        ("gamma_ad", C.c_double),
    ]


class BH_PixelRecord(C.Structure):
    _fields_ = [
        ("intensity", C.c_double),
        ("r_exit", C.c_double),
        ("u_exit", C.c_double),
        ("phi_exit", C.c_double),
        ("cross_r", C.c_double * 3),
        ("cross_g", C.c_double * 3),
        ("tau", C.c_double),
        ("drift_null", C.c_double),
        ("drift_carter", C.c_double),
        ("steps", C.c_int32),
        ("attempts", C.c_int32),
        ("hit_class", C.c_uint8),
        ("n_cross", C.c_uint8),
        ("flags", C.c_uint8),
        ("pad0", C.c_uint8),
        ("confidence", C.c_float),
        ("fp64_steps", C.c_int32),
        ("n_samples", C.c_int32),
    ]


# numpy view of an array of BH_PixelRecord (same layout, 128 bytes)
RECORD_DTYPE = np.dtype(
    {
        "names": [
            "intensity", "r_exit", "u_exit", "phi_exit", "cross_r", "cross_g", "tau",
            "drift_null", "drift_carter", "steps", "attempts", "hit_class", "n_cross",
            "flags", "pad0", "confidence", "fp64_steps", "n_samples",
        ],
        "formats": [
            "<f8", "<f8", "<f8", "<f8", ("<f8", 3), ("<f8", 3), "<f8", "<f8", "<f8",
            "<i4", "<i4", "u1", "u1", "u1", "u1", "<f4", "<i4", "<i4",
        ],
        "offsets": [0, 8, 16, 24, 32, 56, 80, 88, 96, 104, 108, 112, 113, 114, 115, 116, 120, 124],
        "itemsize": 128,
    }
)

assert C.sizeof(BH_PixelRecord) == 128
assert C.sizeof(BH_Camera) == 72
assert C.sizeof(BH_PolParams) == 32


# ---------------------------------------------------------------------------
# Reference render ABI (src/cuda/kernel_launch.h:28-116), mirrored field for
# field in include/bh_kernel_launch.h.
# ---------------------------------------------------------------------------
class BH_LaunchParams(C.Structure):
    _fields_ = [
        ("rs", C.c_float),
        ("spin", C.c_float),
        ("isco", C.c_float),
        ("step_size", C.c_float),
        ("fov_scale", C.c_float),
        ("max_dist", C.c_float),
        ("cam_pos", C.c_float * 3),
        ("cam_basis", C.c_float * 9),
        ("frame_shift_x", C.c_float),
        ("frame_shift_y", C.c_float),
        ("max_steps", C.c_int),
        ("width", C.c_int),
        ("height", C.c_int),
        ("adisk_enabled", C.c_int),
        ("redshift_enabled", C.c_int),
        ("kerr_enabled", C.c_int),
        ("use_luts", C.c_int),
        ("lut_radius_min", C.c_float),
        ("lut_radius_max", C.c_float),
        ("redshift_radius_min", C.c_float),
        ("redshift_radius_max", C.c_float),
        ("spectral_radius_min", C.c_float),
        ("spectral_radius_max", C.c_float),
        ("time_sec", C.c_float),
        ("doppler_strength", C.c_float),
        ("background_intensity", C.c_float),
        ("background_enabled", C.c_int),
        ("photon_glow_strength", C.c_float),
        ("debug_pre_redshift_background", C.c_int),
        ("debug_pre_shaping_background", C.c_int),
        ("debug_post_shaping_background", C.c_int),
        ("debug_shaper_inputs", C.c_int),
        ("debug_closest_approach_state", C.c_int),
        ("debug_closest_approach_timeline", C.c_int),
        ("debug_closest_approach_direction", C.c_int),
        ("debug_escaped_direction", C.c_int),
        ("background_yaw_rad", C.c_float),
        ("background_pitch_rad", C.c_float),
        ("background_filter_radius", C.c_float),
        ("background_layer_params", C.c_float * 12),
        ("background_layer_lod_bias", C.c_float * 3),
        ("wiregrid_enabled", C.c_int),
        ("wiregrid_show_ergo", C.c_float),
        ("wiregrid_grid_scale", C.c_float),
        ("wiregrid_motion_scale", C.c_float),
        ("wiregrid_infall_scale", C.c_float),
        ("wiregrid_strength", C.c_float),
        ("wiregrid_scene_preserve", C.c_float),
        ("wiregrid_color", C.c_float * 4),
        ("grmhd_r_min", C.c_float),
        ("grmhd_r_max", C.c_float),
        ("rte_enabled", C.c_int),
        ("rte_opacity_scale", C.c_float),
        ("grmhd_alpha", C.c_float),
        ("stokes_enabled", C.c_int),
        ("stokes_b_field_angle", C.c_float),
        ("stokes_ne_scale", C.c_float),
        ("adisk_lit", C.c_float),
    ]


assert C.sizeof(BH_LaunchParams) == 336
