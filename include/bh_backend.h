/*
 * bh_backend.h - the renderer backend ABI (drop-in for the reference's
 * src/cuda/cuda_backend.h:33-103) on the render launch of bh_kernel_launch.h.
 *
 * An opaque handle owns a linear device framebuffer (float4 [width * height],
 * row-major from the top row) and the active kernel variant.  The library has
 * no OpenGL interop: the GL texture ids passed to bhCudaInit / bhCudaResize /
 * bhCudaRegisterLut / bhCudaRegisterBackgroundTexture are recorded and
 * reported back (bhCudaGlTexture) but not mapped, and a frame stays in the
 * linear framebuffer (bhCudaFramebuffer) instead of being blitted.
 * Returns 0 / -1 (bh_last_error); bhCudaShutdown(NULL) is safe.
 */
#ifndef BH_BACKEND_H
#define BH_BACKEND_H

#include "bh_kernel_launch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct BhCudaBackend BH_CudaBackend;

/* cuda_backend.h:33-43: allocates the framebuffer; NULL on failure. */
BH_CudaBackend* bhCudaInit(unsigned int glTexture, int width, int height);

/* cuda_backend.h:45-53: bh_launch_geodesic_kernel into the framebuffer with
 * the active variant (params->width / height must match the backend's);
 * asynchronous on the backend's stream, which bhCudaSynchronize waits for. */
int bhCudaRenderFrame(BH_CudaBackend* backend, const struct BH_LaunchParams* params);

/* cuda_backend.h:55-65: reallocates the framebuffer. */
int bhCudaResize(BH_CudaBackend* backend, unsigned int glTexture, int width, int height);

/* cuda_backend.h:67-79 (slots 0..7) and :81-93: records the GL texture ids. */
int bhCudaRegisterLut(BH_CudaBackend* backend, int slot, unsigned int glTexture, unsigned int target);
int bhCudaRegisterBackgroundTexture(BH_CudaBackend* backend, unsigned int glTexture, unsigned int target);

/* cuda_backend.h:95-103: variant < 0 selects bh_select_kernel_variant(),
 * variant >= BH_KERNEL_COUNT falls back to BH_KERNEL_FP32_BASELINE (as the
 * reference, cuda_backend.cu:144-154); returns the active variant (-1 for NULL). */
int bhCudaSetVariant(BH_CudaBackend* backend, int variant);
int bhCudaGetVariant(BH_CudaBackend* backend);

/* cuda_backend.h:105-110 */
void bhCudaShutdown(BH_CudaBackend* backend);

/* No counterpart in the reference (the library has no GL blit): the device
 * framebuffer and its size, a wait for the last frame, and the recorded GL
 * ids (slot -1 = the framebuffer texture, 0..7 the LUTs, 8 the background). */
void* bhCudaFramebuffer(BH_CudaBackend* backend, int* width, int* height);
int bhCudaSynchronize(BH_CudaBackend* backend);
unsigned int bhCudaGlTexture(BH_CudaBackend* backend, int slot);

#ifdef __cplusplus
}
#endif

#endif /* BH_BACKEND_H */
