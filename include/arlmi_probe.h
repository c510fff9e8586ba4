/*
 * arlmi_probe.h - C ABI of libarlmi_probe.so, measurement tooling beside libarlmi.so (no ARL
 * counterpart): the LDS microbenchmark whose rate the gridder's float64 LDS atomics are compared
 * with in the benchmark's roofline.  Not used by the imaging path.
 */
#ifndef ARLMI_PROBE_H
#define ARLMI_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cycles per LDS-add wavefront instruction per CU for pattern 0 (ds_add_f64 on the gridder's 8x8
 * footprints at random origins of a 64x72 tile), 1 (ds_add_f64 on 64 distinct consecutive
 * addresses), 2 (pattern 0 with 64-bit integer adds), 3 (pattern 0 with ds_add_f32) or 4 (pattern 1
 * with ds_add_f32), two work groups of 8 wavefronts per CU, the device clock taken from
 * hipDeviceAttributeClockRate.  Synchronises the device.  Returns 0, or a negative code with the
 * message in arlmi_probe_last_error(). */
int arlmi_probe_lds_add(int32_t pattern, int32_t iters, double* cycles_per_instr, float* ms_out);
const char* arlmi_probe_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* ARLMI_PROBE_H */
