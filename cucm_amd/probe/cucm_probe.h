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
/* The decoder's state-gather stream replayed without its model arithmetic:
 * n chunks (device addresses h_src[i] of h_len[i] bytes), D-bit speculative
 * rounds (2^D - 1 paths x 7 gathers, each round dependent on the last), the
 * unit's slots written back every U bits, `waves` resident waves (a multiple
 * of 4) each with its own model slot of 7 x 2^tb bytes.  flags bit 0: zero
 * the slot per chunk; bit 1: skip the write-backs; bit 2: pack a round's
 * (path, order) gathers across lanes, one instruction per 9 paths; bit 3:
 * re-gather the unit's slots before its write-backs; bit 4: write each
 * round's first-path slots right after their gathers instead of at the unit
 * end; bit 5: an agent-scope release and acquire fence after every chunk.
 * *ms: kernel time. */
int cucm_probe_replay_decode(const uint64_t *h_src, const uint32_t *h_len, uint32_t n,
                             uint32_t tb, uint32_t D, uint32_t U, uint32_t waves,
                             uint32_t flags, double *ms);
/* The encoder's state traffic replayed the same way: per unit of U bits all
 * 7 gathers of every bit at once, then the unit's 7 states per bit written
 * back and drained before the next unit.  flags bit 0: zero the slot per
 * chunk; bit 1: skip the write-backs; bit 2: read each state's whole 32-byte
 * sector and write the sector back whole (U <= 128); bit 3: do not wait for
 * a unit's write-backs before the next unit's gathers. */
int cucm_probe_replay_encode(const uint64_t *h_src, const uint32_t *h_len, uint32_t n,
                             uint32_t tb, uint32_t U, uint32_t waves, uint32_t flags,
                             double *ms);
#ifdef __cplusplus
}
#endif
#endif
