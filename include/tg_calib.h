/* tg_calib.h -- measured VALU ceilings for the roofline of the interaction kernels.
 *
 * No counterpart in the reference (it has no GPU path); used by the benchmark to state how far
 * each kernel is from what the part sustains for the same instruction mix. */
#ifndef TG_CALIB_H
#define TG_CALIB_H

#ifdef __cplusplus
extern "C" {
#endif

/* Runs two register-only loops `repeats` times (after one warmup each) on `device` and returns
 * the best sustained FP32 FMA rate (TFLOP/s, FMA = 2) and the best rate of the PP far-field pair
 * body (pairs/s).  Returns 0 on success, 1 on bad arguments, 2 on a HIP error. */
int tg_calibrate(int device, int repeats, double* fma_tflops, double* pp_pairs_per_s);
const char* tg_calibrate_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
