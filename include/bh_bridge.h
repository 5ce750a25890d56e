/*
 * bh_bridge.h - the Blender bridge's GPU entry points (drop-in for the
 * tracing and rendering part of src/blender_bridge/blender_bridge.h) on the
 * MI355X tracer.
 *
 * Plain C; caller-allocated host buffers; every call is synchronous and
 * returns 0 / -1 (-1 on invalid arguments or a device error; message in
 * bh_last_error()).  Geometric units (r_g = GM/c^2 = 1) unless noted.
 * The camelCase names of the reference and its snake_case ctypes aliases
 * (blender_bridge.cpp:967-1100) are both exported.
 */
#ifndef BH_BRIDGE_H
#define BH_BRIDGE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Layout twins of the reference's parameter structs (blender_bridge.h:48-57,
 * :180-186), for the sizeof probes ctypes callers check. */
struct BhbSourceParams {
    char   name[64];
    double massMsun, spin, distanceCm, inclinationDeg, freqHz, rGCm, shadowUas;
};
struct BhbDiskParams {
    double aStar, massMsun, mdotEdd, rOutRg, inclinationRad;
};

/* blender_bridge.h:24-36 */
int bhbVersionMajor(void);          /* 1 */
int bhbVersionMinor(void);          /* 0 */
int bhbHasCuda(void);               /* 1: the GPU paths are built (HIP, gfx950) */
int bhbHasBoost(void);              /* 0 */
int bhbSizeofSourceParams(void);
int bhbSizeofDiskParams(void);
int bhbCudaResetDevice(void);       /* waits for the device to go idle; does not reset
                                       it (other libraries in the process share it) */

/* blender_bridge.h:165-168 and :348-350: one null geodesic per image-plane
 * point, alpha_i = alphaMin + (alphaMax - alphaMin) i / (nAlpha - 1) and
 * beta_j likewise (the midpoint when n = 1), observer at (observerRRg,
 * inclinationRad, phi = 0), traced in 3-D (Carter constant) by the library's
 * FP64 Dormand-Prince 5(4) until r < 1.01 r_+, r > 3 observerRRg, or
 * maxSteps points.  outXyz[((j nAlpha + i) maxSteps + k) 3 + c] receives the
 * k-th point (x, y, z) = (r sin th cos phi, r sin th sin phi, r cos th): the
 * start and the end of every accepted step; outCounts[j nAlpha + i] the
 * number of points.  stepSize is not used (the step is adaptive).  Both
 * names run on the GPU. */
int bhbTraceGeodesicsImagePlane(double aStar, double observerRRg, double inclinationRad,
                                double alphaMin, double alphaMax, int nAlpha,
                                double betaMin, double betaMax, int nBeta,
                                int maxSteps, double stepSize, float* outXyz, int* outCounts);
int bhbCudaTraceGeodesics(double aStar, double observerRRg, double inclinationRad,
                          double alphaMin, double alphaMax, int nAlpha,
                          double betaMin, double betaMax, int nBeta,
                          int maxSteps, double stepSize, float* outXyz, int* outCounts);

/* blender_bridge.h:357: lensing map over alpha = 30 (px / width - 0.5),
 * beta = 30 (0.5 - py / height) [r_g] (the reference's grid), traced with a
 * thin disk from r_ISCO to 100 r_g.  outRgba[(py width + px) 4 + c]:
 *   disk hit -> (sqrt(1 - 2/r), g / sqrt(1 - 2/r), 1, 1) at the crossing,
 *               g the exact Keplerian redshift of the emitter;
 *   escaped  -> (sqrt(1 - 2/r_exit), 1, 0, 1);   captured -> (0, 1, 0, 1). */
int bhbCudaRenderLensingMap(double aStar, double observerRRg, double inclinationRad,
                            int width, int height, float* outRgba);

/* blender_bridge.h:384-405: ray-traced images through bh_launch_geodesic_kernel
 * (bh_kernel_launch.h) with the reference bridge's camera (rs = 1; observerR in
 * r_s; camera at (0, r sin i, r cos i) looking at the origin, fovScale 1 for
 * bhbCudaRenderRaytraced) and defaults (disk, redshift, background 0.8, disk
 * brightness 0.35, 500 r_s escape, 600 steps).  The reference's
 * post-processing (bloom, tone map, grain) is not applied.
 * outRgba: width * height * 4 floats, host. */
int bhbCudaRenderRaytraced(float spin, float observerR, float inclinationRad, int width, int height,
                           float* outRgba);
int bhbCudaRenderRaytracedCamera(float spin, float observerR, float inclinationRad, float fovScale,
                                 int width, int height, float* outRgba);
int bhbCudaRenderRaytracedView(float spin, const float* camPos, const float* camBasis, float fovScale,
                               int width, int height, float* outRgba);

/* snake_case aliases (blender_bridge.cpp:967-1100) */
int bhb_version_major(void);
int bhb_version_minor(void);
int bhb_has_cuda(void);
int bhb_has_boost(void);
int bhb_sizeof_source_params(void);
int bhb_sizeof_disk_params(void);
int bhb_cuda_reset_device(void);
int bhb_trace_geodesics_image_plane(double aStar, double observerRRg, double inclinationRad,
                                    double alphaMin, double alphaMax, int nAlpha,
                                    double betaMin, double betaMax, int nBeta,
                                    int maxSteps, double stepSize, float* outXyz, int* outCounts);
int bhb_cuda_trace_geodesics(double aStar, double observerRRg, double inclinationRad,
                             double alphaMin, double alphaMax, int nAlpha,
                             double betaMin, double betaMax, int nBeta,
                             int maxSteps, double stepSize, float* outXyz, int* outCounts);
int bhb_cuda_render_lensing_map(double aStar, double observerRRg, double inclinationRad,
                                int width, int height, float* outRgba);
int bhb_cuda_render_raytraced(float spin, float observerR, float inclinationRad, int width, int height,
                              float* outRgba);
int bhb_cuda_render_raytraced_camera(float spin, float observerR, float inclinationRad, float fovScale,
                                     int width, int height, float* outRgba);
int bhb_cuda_render_raytraced_view(float spin, const float* camPos, const float* camBasis, float fovScale,
                                   int width, int height, float* outRgba);

#ifdef __cplusplus
}
#endif

#endif /* BH_BRIDGE_H */
