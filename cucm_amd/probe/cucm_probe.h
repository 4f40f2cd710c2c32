/*
 * cucm_probe.h -- measurement kernels of libcucm_probe.so.  They are not part
 * of the codec's drop-in interface (include/cucm.h) and have no counterpart in
 * the reference; bench.py uses them to measure the memory-system ceilings the
 * codec kernels are compared against.
 */
#ifndef CUCM_PROBE_H
#define CUCM_PROBE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
/* Random one-byte gather microbenchmark over `bytes` (a power of two) of
 * device memory `table`: blocks x threads lanes each do `steps` rounds of
 * `width` (1..8) independent random reads (and, if write_back, writes of the
 * same bytes).  Returns 0, or -4/-6 on allocation/device errors. */
int cucm_probe_gather(void *table, uint64_t bytes, uint32_t blocks, uint32_t threads,
                      uint32_t steps, uint32_t width, uint32_t write_back, double *ms,
                      uint64_t *wave_clocks);
/* Streaming device copy of `bytes` (multiple of 16), used to calibrate the
 * HBM traffic counters. */
int cucm_probe_copy(void *dst, const void *src, uint64_t bytes, double *ms);
#ifdef __cplusplus
}
#endif
#endif
