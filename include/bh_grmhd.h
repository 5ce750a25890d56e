/*
 * bh_grmhd.h - GRMHD packed-volume ingestion and temporal sampling for the
 * MI355X tracer (drop-in for the loading part of src/grmhd_packed_loader.h
 * and the GRMHD texture slots of src/cuda/kernel_launch.h).
 *
 * Datasets are the nubhlight_pack format (tools/nubhlight_pack.cpp): a JSON
 * metadata file and an RGBA32F binary of width x height x depth voxels, x = r
 * fastest, then theta, then phi, 4 channels per voxel (density, internal
 * energy, v1, v2 by default).  A texture made by bh_grmhd_texture_create is
 * what bh_upload_lut_textures takes in its grmhd / grmhd_right slots; the
 * render launch then scales the disk colour at each hit by rho * u sampled at
 * the hit point, blended between the two frames by BH_LaunchParams.grmhd_alpha
 * (d_sample_grmhd / d_shade_hit, device_physics.cuh:1789-1868).
 *
 * Plain C; host buffers; returns 0 / -1 (message in bh_last_error()).
 */
#ifndef BH_GRMHD_H
#define BH_GRMHD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct BH_GrmhdPacked {
    int32_t  width, height, depth;  /* grid_dims: r (fastest), theta, phi          */
    int32_t  checksum_verified;     /* 1: "checksum_fnv1a64" present and matching   */
    float    r_min, r_max;          /* radial extent ("r_min" / "r_max", default 1 / 50) */
    float    min[4], max[4];        /* per-channel extrema from the metadata        */
    char     channels[4][32];       /* channel names (truncated to 31 characters)   */
    uint64_t checksum;              /* FNV-1a 64 of the binary's bytes              */
    int64_t  floats;                /* width * height * depth * 4                   */
} BH_GrmhdPacked; /* 200 bytes */

/* grmhd_packed_loader.h:71-72 (loadGrmhdPackedTexture with upload = false):
 * parses the JSON at json_path, reads the binary named by "bin" (legacy
 * "binary_file") relative to the JSON's directory, and validates: format
 * RGBA32F, grid_dims [w, h, d] > 0, binary size w h d 4 floats, the FNV-1a64
 * checksum when "checksum_fnv1a64" is present, exactly 4 channels (string
 * array with top-level "min" / "max" arrays, or legacy objects with
 * name / min / max).  rgba: host float[meta->floats], or NULL to read the
 * metadata only (the binary is still read and checked). */
int bh_grmhd_load_packed(const char* json_path, BH_GrmhdPacked* meta, float* rgba, int64_t rgba_floats);

/* FNV-1a 64-bit hash of `bytes` bytes (offset basis 14695981039346656037,
 * prime 1099511628211), as the packer's checksum. */
uint64_t bh_fnv1a64(const void* data, int64_t bytes);

/* Time-series metadata of nubhlight_pack's multi-frame output (version
 * "1.0-multi": frameCount, frameSize, frames[{file, time, offset}],
 * frameMetadata{gridDims, channels, format}).  Writes the frame count, the
 * per-frame byte size, the grid dims (3) and up to max_frames times and
 * offsets (either may be NULL). */
int bh_grmhd_timeseries_info(const char* json_path, int32_t* frame_count, int64_t* frame_bytes,
                             int32_t* grid_dims, double* times, int64_t* offsets, int32_t max_frames);

/* Reads frame `frame` of a time series (its offset in the metadata) from
 * bin_path into rgba (frame_bytes / 4 floats). */
int bh_grmhd_load_frame(const char* json_path, const char* bin_path, int32_t frame, float* rgba,
                        int64_t rgba_floats);

/* The two frames around time t and the blend weight: i0 = the last frame
 * with times[i0] <= t (clamped to [0, n - 2]), alpha = (t - times[i0]) /
 * (times[i0 + 1] - times[i0]) clamped to [0, 1]; n = 1 gives (0, 0). */
int bh_grmhd_frame_pair(const double* times, int32_t n, double t, int32_t* i0, float* alpha);

/* Device volume over host RGBA32F data, sampled as a linear-filtered,
 * normalised-coordinate 3-D texture with x = r clamped, y = theta mirrored,
 * z = phi wrapped.  gfx950 has no image-sampling hardware, so the handle is
 * the device address of the library's volume descriptor (filtering in
 * software, same semantics) rather than a texture object; it goes to
 * bh_upload_lut_textures' grmhd / grmhd_right slots.  Handles not made here
 * are ignored by the render launch.  Destroy after the launches using it. */
int bh_grmhd_texture_create(const float* host_rgba, int32_t width, int32_t height, int32_t depth,
                            unsigned long long* tex);
int bh_grmhd_texture_destroy(unsigned long long tex);

/* The render launch's GRMHD sample (handles from bh_grmhd_texture_create or 0)
 * at n Cartesian points (xyz: host float[n][3],
 * in the launch's length unit; out: host float[n][4]): u = (|p| - r_min) /
 * (r_max - r_min), v = theta / pi, w = phi / 2 pi, the right texture blended
 * in by alpha when it is set and alpha > 0.  For checks and tools; no
 * counterpart in the reference (d_sample_grmhd is device-only there). */
int bh_grmhd_sample(unsigned long long tex, unsigned long long tex_right, float alpha, float r_min,
                    float r_max, const float* xyz, int32_t n, float* out);

/* ---- streaming of time series (src/grmhd_streaming.h:125-301) -----------
 * The reference's GRMHDStreamer keeps recently used frames ("tiles" that are
 * whole frames, grmhd_streaming.cpp:355-411) in a host LRU cache under a byte
 * budget, loads missing ones on a background thread and prefetches the next
 * three frames on every seek.  Here the cache holds frames as device volumes
 * (bh_grmhd_texture_create's kind, ready for the render launch's grmhd
 * slots), in HBM, under a device-memory budget: a frame is read from the
 * binary into pinned host memory and uploaded on the streamer's own stream
 * by the loader thread, so neither disk reads nor uploads block the caller.
 * Eviction is least-recently-used and never takes the current frame or the
 * next one; handles of those two stay valid until a later seek moves the
 * current frame away (then they may be freed; launches using them must be
 * synchronised first).  One host thread at a time per streamer. */
typedef struct BH_GrmhdStreamer BH_GrmhdStreamer;

typedef struct BH_GrmhdStreamStats {
    uint64_t hits, misses;        /* frame look-ups that found / missed a resident frame */
    int64_t  resident_bytes;      /* device bytes of resident frames               */
    int64_t  budget_bytes;
    int32_t  resident_frames, queue_depth;
    int32_t  current_frame, frame_count;
} BH_GrmhdStreamStats; /* 48 bytes */

/* GRMHDStreamer(json, bin) + init(): parses the multi-frame metadata
 * (bh_grmhd_timeseries_info), checks the binary holds every frame, and
 * starts the loader thread.  cache_bytes: device budget (at least one frame);
 * prefetch: frames after the current one loaded ahead on a seek (the
 * reference uses 3).  NULL on failure (bh_last_error()). */
BH_GrmhdStreamer* bh_grmhd_streamer_create(const char* json_path, const char* bin_path, int64_t cache_bytes,
                                           int32_t prefetch);

/* shutdown() and free: stops the loader, frees every resident frame; NULL is ignored. */
void bh_grmhd_streamer_destroy(BH_GrmhdStreamer* s);

/* seekFrame: the current frame (clamped to the series) and the prefetch. */
int bh_grmhd_streamer_seek(BH_GrmhdStreamer* s, int32_t frame);

/* getTile (offset 0: the current frame) / getAdjacentTile (offset 1: the
 * next one): *tex = the resident frame's device volume and 0 returned, or
 * *tex = 0 and 1 returned when it is not resident (for the current frame its
 * load is queued; the adjacent frame is only looked up).  wait != 0 blocks
 * until the frame's load has finished.  -1 on invalid arguments. */
int bh_grmhd_streamer_get_frame(BH_GrmhdStreamer* s, int32_t offset, int32_t wait, unsigned long long* tex);

/* The two frames around simulation time t (the rule of bh_grmhd_frame_pair)
 * as device volumes and their blend weight: seeks to the left frame, then
 * looks both up as bh_grmhd_streamer_get_frame does (right = 0 for a
 * one-frame series).  Returns 0 when both are resident, 1 when one is not. */
int bh_grmhd_streamer_frame_pair(BH_GrmhdStreamer* s, double t, int32_t wait, unsigned long long* left,
                                 unsigned long long* right, float* alpha);

int bh_grmhd_streamer_stats(const BH_GrmhdStreamer* s, BH_GrmhdStreamStats* out);

#ifdef __cplusplus
}
#endif

#endif /* BH_GRMHD_H */
