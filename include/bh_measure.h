/*
 * bh_measure.h - attainable device peaks, measured with micro-kernels.
 *
 * No counterpart in the reference: a measurement helper for the roofline
 * figures.  The published MI355X peaks (FP64 vector 78.6 TFLOP/s, BF16 dense
 * MFMA 2.5 PFLOP/s, HBM3E 8 TB/s) are what the kernels are priced against;
 * these kernels measure what one process can reach on the device at hand,
 * so a roofline fraction can be read against both.
 */
#ifndef BH_MEASURE_H
#define BH_MEASURE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Runs four micro-kernels on `stream` (a hipStream_t, NULL for the default
 * stream) on the current device, each once untimed and then timed with HIP
 * events (three launches, ten for the copy):
 *   out[0]  FP64 FMA chains (8 independent chains per lane), TFLOP/s
 *   out[1]  FP32 FMA chains, scalar v_fma_f32 (packed FP32 instructions are
 *           disabled in this library), TFLOP/s
 *   out[2]  v_mfma_f32_32x32x16_bf16 chains (4 accumulators per wave),
 *           TFLOP/s
 *   out[3]  a 2 GiB device-to-device copy kernel (16-byte non-temporal
 *           loads and stores), GB/s of bytes read plus bytes written
 * Allocates and frees 4 GiB of device memory for the copy.  Synchronises
 * the stream.  Returns 0, or -1 with bh_last_error() set. */
int bh_measure_peaks(double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BH_MEASURE_H */
