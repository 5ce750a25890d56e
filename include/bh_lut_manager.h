/*
 * bh_lut_manager.h - the LUT slot registry of the render path (drop-in for
 * the slot semantics of src/cuda/lut_manager.h:15-134).
 *
 * The reference registers OpenGL textures as CUDA texture objects in eight
 * slots (BhLutSlot) and hands their handles to the render launch
 * (bh_upload_lut_textures, src/cuda/cuda_backend.cu:27-34).  This target
 * has neither GL interop nor image-sampling hardware, so a slot here holds a
 * device copy of the table itself, registered from host texels:
 *   - the GRMHD slots (BhLutGrmhd, BhLutGrmhdRight) take an RGBA32F volume
 *     and hold a volume of bh_grmhd_texture_create, which the render launch
 *     samples (r clamped, theta mirrored, phi wrapped);
 *   - the galaxy slot takes a 6-layer cubemap (the reference's owned
 *     layered array, lut_manager.h:37-41);
 *   - the other slots take 2-D tables of 1 or 4 channels.
 * A handle (u64) is the device address of the slot's descriptor
 * (BH_LutTableDev), 0 for an empty slot, valid until the slot is
 * unregistered or re-registered.  Registering an occupied slot replaces its
 * table (lutRegister, lut_manager.h:67-82); invalid slots return -1.  The
 * standalone equirect background (lutRegisterBackground, :97-124) is a 2-D
 * table outside the numbered slots.
 *
 * Plain C; host buffers in, device tables out; returns 0 / -1 (message in
 * bh_last_error()).  A manager is not thread-safe: one host thread at a time.
 */
#ifndef BH_LUT_MANAGER_H
#define BH_LUT_MANAGER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Slot indices, in the GL binding order of the reference (lut_manager.h:22-32). */
enum BhLutSlot {
    BhLutEmissivity = 0,  /* accretion-disk emissivity flux                  */
    BhLutRedshift   = 1,  /* gravitational redshift factor                   */
    BhLutSpectral   = 2,  /* spectral modulation                             */
    BhLutGrb        = 3,  /* gamma-ray burst overlay                         */
    BhLutGalaxy     = 4,  /* galaxy cubemap background (6 layers)            */
    BhLutGrmhd      = 5,  /* GRMHD volume, left frame (RGBA32F 3-D)          */
    BhLutSynchG     = 6,  /* synchrotron G(x) table (R32F, height 1)         */
    BhLutGrmhdRight = 7,  /* GRMHD volume, right frame for time interpolation */
    BhLutCount      = 8
};

/* Host texels of a table: width x height x depth texels of `channels`
 * floats each (x fastest), `layers` such images back to back (6 for the
 * galaxy cubemap, else 1). */
typedef struct BH_LutImage {
    const float* host;
    int32_t width, height, depth;
    int32_t channels;   /* 1 (R32F) or 4 (RGBA32F)                          */
    int32_t layers;     /* 1, or 6 for BhLutGalaxy                          */
    int32_t pad0;
} BH_LutImage; /* 32 bytes */

/* Device descriptor a handle points to. */
typedef struct BH_LutTableDev {
    const float* data;  /* device texels, layout of BH_LutImage            */
    int32_t width, height, depth, channels, layers, slot;
} BH_LutTableDev; /* 32 bytes */

typedef struct BH_LutManager BH_LutManager;

/* LutManager(): every slot empty.  NULL on failure. */
BH_LutManager* bh_lut_manager_create(void);

/* lutRegister (lut_manager.h:67-82): copies the table to the device,
 * replacing the slot's previous table.  -1 for a slot outside [0, 8) or a
 * table the slot does not take. */
int bh_lut_register(BH_LutManager* mgr, int slot, const BH_LutImage* img);

/* lutUnregister (:84-90): frees the slot's table (after the device is idle);
 * no effect on an empty slot or a slot outside [0, 8). */
void bh_lut_unregister(BH_LutManager* mgr, int slot);

/* lutGetTex (:92-99): the slot's handle, 0 if empty or out of range. */
unsigned long long bh_lut_get_tex(const BH_LutManager* mgr, int slot);

/* lutRegisterBackground / lutUnregisterBackground / lutGetBackgroundTex
 * (:101-126): the standalone equirect background, a 2-D table. */
int bh_lut_register_background(BH_LutManager* mgr, const BH_LutImage* img);
void bh_lut_unregister_background(BH_LutManager* mgr);
unsigned long long bh_lut_get_background_tex(const BH_LutManager* mgr);

/* lutShutdown (:128-133): every slot and the background emptied. */
void bh_lut_shutdown(BH_LutManager* mgr);

/* Shuts the manager down and frees it; NULL is ignored. */
void bh_lut_manager_destroy(BH_LutManager* mgr);

/* The eight slot handles to bh_upload_lut_textures and the background to
 * bh_upload_bridge_background_texture (the backend's upload,
 * src/cuda/cuda_backend.cu:25-37).  No counterpart as a function. */
int bh_lut_upload(const BH_LutManager* mgr);

/* Reads a handle's table back: its geometry into meta (meta->host = NULL)
 * and, when host_out is not NULL, its texels (host_floats >= width * height
 * * depth * channels * layers).  For checks and tools; no counterpart in the
 * reference. */
int bh_lut_read(unsigned long long handle, BH_LutImage* meta, float* host_out, int64_t host_floats);

#ifdef __cplusplus
}
#endif

#endif /* BH_LUT_MANAGER_H */
