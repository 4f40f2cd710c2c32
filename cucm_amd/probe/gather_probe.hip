// gather_probe.hip -- measurement kernels, built into libcucm_probe.so (not
// part of the codec library or its C ABI): a random one-byte gather
// microbenchmark (the ceiling the codec's state gathers are compared
// against) and a streaming copy (HBM counter calibration).
//
// Two modes over a table of `bytes` bytes:
//   latency:    each lane follows a dependent chain of random one-byte reads
//               (index = f(previous byte)), so one round trip per step.
//   throughput: each lane issues `width` independent random reads per step;
//               every read also writes the byte back (read-modify-write, as
//               the StateMap update does).
// Results (per launch): total touches and elapsed shader clocks per wave.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cucm_probe.h"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

__global__ void gather_kernel(uint8_t *table, uint64_t bytes, uint32_t steps, uint32_t width,
                              uint32_t write_back, uint64_t *clocks, uint32_t *sink) {
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t h = mix64(gid + 1);
  uint32_t acc = 0;
  const uint64_t t0 = clock64();
  for (uint32_t s = 0; s < steps; ++s) {
    uint32_t v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if ((uint32_t)i < width) {
        const uint64_t idx = mix64(h + i * 0x9e3779b97f4a7c15ull + acc) & (bytes - 1);  // bytes: power of 2
        v[i] = table[idx];
        if (write_back) table[idx] = (uint8_t)(v[i] + 1);
      }
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if ((uint32_t)i < width) acc += v[i];
    h = mix64(h);
  }
  const uint64_t t1 = clock64();
  if ((threadIdx.x & 63) == 0) atomicAdd((unsigned long long *)clocks, (unsigned long long)(t1 - t0));
  // keeps the reads live; sink is one word, so every lane adds to it
  if (acc == 0xdeadbeefu) atomicAdd(sink, 1u);
}

__global__ void copy_kernel(const uint4 *src, uint4 *dst, uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

// Plain streaming copy of `bytes` (a multiple of 16): the reference kernel
// whose FETCH_SIZE / WRITE_SIZE readings calibrate the HBM traffic counters.
extern "C" int cucm_probe_copy(void *dst, const void *src, uint64_t bytes, double *ms) {
  if (!dst || !src || !ms || bytes % 16) return -1;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, 0, (const uint4 *)src, (uint4 *)dst,
                     bytes / 16);
  (void)hipEventRecord(e1, 0);
  hipError_t err = hipEventSynchronize(e1);
  float f = 0;
  (void)hipEventElapsedTime(&f, e0, e1);
  *ms = f;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return err == hipSuccess ? 0 : -6;
}

extern "C" int cucm_probe_gather(void *table, uint64_t bytes, uint32_t blocks, uint32_t threads,
                                 uint32_t steps, uint32_t width, uint32_t write_back,
                                 double *ms, uint64_t *wave_clocks) {
  if (!table || bytes == 0 || (bytes & (bytes - 1)) || width == 0 || width > 8 || !ms ||
      !wave_clocks)
    return -1;
  uint64_t *d = nullptr;
  if (hipMalloc(&d, 16) != hipSuccess) return -4;
  (void)hipMemset(d, 0, 16);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(gather_kernel, dim3(blocks), dim3(threads), 0, 0, (uint8_t *)table, bytes,
                     steps, width, write_back, d, (uint32_t *)(d + 1));
  (void)hipEventRecord(e1, 0);
  hipError_t err = hipEventSynchronize(e1);
  float f = 0;
  (void)hipEventElapsedTime(&f, e0, e1);
  *ms = f;
  (void)hipMemcpy(wave_clocks, d, 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return err == hipSuccess ? 0 : -6;
}
