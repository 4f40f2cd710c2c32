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


// Replay of the decoder's state-gather stream without its model arithmetic:
// persistent waves take chunks from a queue; per chunk the wave zeroes its
// model slot (7 order tables of 2^tb bytes), then for every round of D bits
// lane q < 2^D - 1 hashes speculative path q of the round exactly as the
// decoder does (partial byte plus the path's bits as h0, the FNV chain over
// the 7 previous bytes of the chunk, each state at the codec's table
// position (H_{k-1} ^ b_k) & mask) and gathers its 7 states; the next
// round's addresses depend on this round's loads (so rounds are dependent
// round trips, as in the decoder), and after every unit of U bits the 7
// slots of each of the unit's bits are written back.  The taken path of a
// round is the one the chunk's real bits select.  It measures the ceiling
// the decoder's memory traffic allows at a given number of resident chunks.
struct ReplayChunk {
  uint64_t src;
  uint32_t L, pad;
};

__global__ void __launch_bounds__(256) replay_decode_kernel(const ReplayChunk *chunks, uint32_t n,
                                                            uint32_t *counter, uint8_t *pool,
                                                            uint64_t slot_bytes, uint32_t tb,
                                                            uint32_t D, uint32_t U,
                                                            uint32_t zero_slot, uint32_t *sink) {
  __shared__ uint8_t ubs[4][8 + 64 + 8];  // per wave: 8 history bytes + up to 512 unit bits
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  uint8_t *ub = ubs[wid];
  const uint64_t gw = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wid;
  uint8_t *slot = pool + gw * slot_bytes;
  const uint32_t mask = (1u << tb) - 1u;
  const uint32_t P = (1u << D) - 1u;
  uint32_t acc = 0;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= n) break;
    const uint8_t *src = (const uint8_t *)(uintptr_t)chunks[t].src;
    const uint64_t L = chunks[t].L;
    if (zero_slot & 1u) {
      const uint64_t n16 = (7ull << tb) >> 4;
      uint4 *d = (uint4 *)slot;
      for (uint64_t i = lane; i < n16; i += 64) d[i] = make_uint4(0, 0, 0, 0);
    }
    uint32_t dep = 0;
    const uint64_t vbits = 8 * L;
    for (uint64_t base = 0; base < vbits; base += U) {
      const uint32_t nb = (uint32_t)min((uint64_t)U, vbits - base);
      // 8 history bytes and the unit's bytes (zero outside the chunk)
      const int64_t b0 = (int64_t)(base >> 3) - 8;
      for (uint32_t i = lane; i < 8 + U / 8; i += 64) {
        const int64_t p = b0 + i;
        ub[i] = (p >= 0 && p < (int64_t)L) ? src[p] : (uint8_t)0;
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_s_waitcnt(0xc07f);
      for (uint32_t rb = 0; rb < nb; rb += D) {
        const uint32_t bi = rb >> 3, j0 = rb & 7u;
        const uint32_t cur = ub[8 + bi];
        const uint32_t partial = j0 ? (cur >> (8 - j0)) : 0u;
        uint32_t sum = 0;
        if (zero_slot & 4u) {
          // packed: lane = 7 q + (k - 1) gathers order k of path q, one
          // instruction per round for up to 9 paths
          const uint32_t lq = lane / 7u, lk = 1u + lane % 7u;
          const bool on = lq < P;
          const uint32_t q = on ? lq : 0u, kk = on ? lk : 1u;
          const uint32_t plen = 31u - (uint32_t)__clz((int)(q + 1u));
          const uint32_t pbits = q + 1u - (1u << plen);
          const uint32_t h0 = ((1u << (j0 + plen)) | (partial << plen) | pbits) & 255u;
          uint32_t H = (2166136261u ^ h0) * 16777619u, pos = 0;
#pragma unroll
          for (int k = 1; k < 8; ++k) {
            const uint32_t x = H ^ (uint32_t)ub[8 + bi - k];
            pos = (uint32_t)k == kk ? x : pos;
            H = (uint32_t)k < kk ? x * 16777619u : H;
          }
          sum = slot[((kk - 1) << tb) | ((pos + dep) & mask)];
          if (P > 9) {
            // paths 9.. in a second instruction
            const uint32_t lq2 = 9u + lane / 7u;
            const bool on2 = lq2 < P && lane < 63u;
            const uint32_t q2 = on2 ? lq2 : 0u;
            const uint32_t plen2 = 31u - (uint32_t)__clz((int)(q2 + 1u));
            const uint32_t pbits2 = q2 + 1u - (1u << plen2);
            const uint32_t h02 = ((1u << (j0 + plen2)) | (partial << plen2) | pbits2) & 255u;
            uint32_t H2 = (2166136261u ^ h02) * 16777619u, pos2 = 0;
#pragma unroll
            for (int k = 1; k < 8; ++k) {
              const uint32_t x = H2 ^ (uint32_t)ub[8 + bi - k];
              pos2 = (uint32_t)k == kk ? x : pos2;
              H2 = (uint32_t)k < kk ? x * 16777619u : H2;
            }
            sum += slot[((kk - 1) << tb) | ((pos2 + dep) & mask)];
          }
        } else {
          const uint32_t q = lane < P ? lane : 0u;
          const uint32_t plen = 31u - (uint32_t)__clz((int)(q + 1u));
          const uint32_t pbits = q + 1u - (1u << plen);
          const uint32_t h0 = ((1u << (j0 + plen)) | (partial << plen) | pbits) & 255u;
          uint32_t H = (2166136261u ^ h0) * 16777619u;
#pragma unroll
          for (int k = 1; k < 8; ++k) {
            const uint32_t x = H ^ (uint32_t)ub[8 + bi - k];
            H = x * 16777619u;
            sum += slot[((uint32_t)(k - 1) << tb) | ((x + dep) & mask)];
          }
        }
        // the next round waits for this one's loads (sum < 2^31 always)
        dep = __builtin_amdgcn_readfirstlane(sum) == 0x7fffffffu ? 1u : 0u;
        acc += sum;
        if (zero_slot & 16u) {
          // write the round's path-0 slots while their lines are in cache
          const uint32_t h0 = ((1u << j0) | partial) & 255u;
          uint32_t H = (2166136261u ^ h0) * 16777619u;
#pragma unroll
          for (int k = 1; k < 8; ++k) {
            const uint32_t x = H ^ (uint32_t)ub[8 + bi - k];
            H = x * 16777619u;
            if (lane == 0) slot[((uint32_t)(k - 1) << tb) | ((x + dep) & mask)] = (uint8_t)(rb + dep);
          }
        }
      }
      if (zero_slot & 8u) {
        // bring the unit's written lines back into cache before the writes
        uint32_t s2 = 0;
        for (uint32_t r = lane; r < nb; r += 64) {
          const uint32_t bi = r >> 3, j = r & 7u;
          const uint32_t cur = ub[8 + bi];
          const uint32_t h0 = (1u << j) | (j ? (cur >> (8 - j)) : 0u);
          uint32_t H = (2166136261u ^ h0) * 16777619u;
#pragma unroll
          for (int k = 1; k < 8; ++k) {
            const uint32_t x = H ^ (uint32_t)ub[8 + bi - k];
            H = x * 16777619u;
            s2 += slot[((uint32_t)(k - 1) << tb) | (x & mask)];
          }
        }
        dep = __builtin_amdgcn_readfirstlane(s2) == 0x7fffffffu ? 1u : dep;
      }
      // write back the unit's slots: bit r = lane + 64j, the taken path
      for (uint32_t r = lane; r < ((zero_slot & 18u) ? 0u : nb); r += 64) {
        const uint32_t bi = r >> 3, j = r & 7u;
        const uint32_t cur = ub[8 + bi];
        const uint32_t h0 = (1u << j) | (j ? (cur >> (8 - j)) : 0u);
        uint32_t H = (2166136261u ^ h0) * 16777619u;
#pragma unroll
        for (int k = 1; k < 8; ++k) {
          const uint32_t x = H ^ (uint32_t)ub[8 + bi - k];
          H = x * 16777619u;
          slot[((uint32_t)(k - 1) << tb) | (x & mask)] = (uint8_t)(r + dep);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (zero_slot & 32u) {
      // the fence pair a persistent kernel runs per task to publish its
      // results and take the next task: an agent-scope release (L2 write-back
      // on this chip) and an agent-scope acquire (L2 invalidate)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
  }
  if (acc == 0xdeadbeefu) atomicAdd(sink, 1u);
}

// Replay of the encoder's state traffic without its model arithmetic: per
// unit of U bits every bit's 7 states are gathered at once (lane l takes
// bits l + 64j, as the encoder issues them), then the 7 states of every bit
// are written back and drained before the next unit's gathers, as the
// encoder's deferred update does.  Lanes past the unit's bits repeat bit 0
// (in bounds, same lines) and write nothing.
__global__ void __launch_bounds__(256) replay_encode_kernel(const ReplayChunk *chunks, uint32_t n,
                                                            uint32_t *counter, uint8_t *pool,
                                                            uint64_t slot_bytes, uint32_t tb,
                                                            uint32_t U, uint32_t flags,
                                                            uint32_t *sink) {
  __shared__ uint8_t ubs[4][8 + 64 + 8];  // per wave: 8 history bytes + up to 512 unit bits
  const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
  uint8_t *ub = ubs[wid];
  const uint64_t gw = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wid;
  uint8_t *slot = pool + gw * slot_bytes;
  const uint32_t mask = (1u << tb) - 1u;
  uint32_t acc = 0, dep = 0;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= n) break;
    const uint8_t *src = (const uint8_t *)(uintptr_t)chunks[t].src;
    const uint64_t L = chunks[t].L;
    if (flags & 1u) {
      const uint64_t n16 = (7ull << tb) >> 4;
      uint4 *d = (uint4 *)slot;
      for (uint64_t i = lane; i < n16; i += 64) d[i] = make_uint4(0, 0, 0, 0);
    }
    const uint64_t vbits = 8 * L;
    for (uint64_t base = 0; base < vbits; base += U) {
      const uint32_t nb = (uint32_t)min((uint64_t)U, vbits - base);
      const int64_t b0 = (int64_t)(base >> 3) - 8;
      for (uint32_t i = lane; i < 8 + U / 8; i += 64) {
        const int64_t p = b0 + i;
        ub[i] = (p >= 0 && p < (int64_t)L) ? src[p] : (uint8_t)0;
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_s_waitcnt(0xc07f);
      uint32_t pos[8][7];
      uint32_t v = 0;
      if (flags & 4u) {
        // 32-byte sectors: each state's whole sector is read, and written back
        // whole with the state's byte changed (no partial-sector writes)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (64u * j < U) {
            const uint32_t r = lane + 64u * j < nb ? lane + 64u * j : 0u;
            const uint32_t bi = r >> 3, jb = r & 7u;
            const uint32_t cur = ub[8 + bi];
            const uint32_t h0 = (1u << jb) | (jb ? (cur >> (8 - jb)) : 0u);
            uint32_t H = (2166136261u ^ h0) * 16777619u;
            uint4 sa[7], sb[7];
#pragma unroll
            for (int k = 1; k < 8; ++k) {
              const uint32_t x = H ^ (uint32_t)ub[8 + bi - k];
              H = x * 16777619u;
              pos[j][k - 1] = ((uint32_t)(k - 1) << tb) | ((x + dep) & mask);
              const uint4 *sp = (const uint4 *)(slot + (pos[j][k - 1] & ~31u));
              sa[k - 1] = sp[0];
              sb[k - 1] = sp[1];
            }
#pragma unroll
            for (int k = 0; k < 7; ++k) v += sa[k].x + sb[k].w;
            v &= 0xffffu;
            if (!(flags & 2u) && lane + 64u * j < nb) {
#pragma unroll
              for (int k = 0; k < 7; ++k) {
                uint4 *sp = (uint4 *)(slot + (pos[j][k] & ~31u));
                uint4 a2 = sa[k];
                a2.x ^= v & 255u;  // (a stand-in for the changed state byte)
                sp[0] = a2;
                sp[1] = sb[k];
              }
            }
          }
        }
        acc += v;
        dep = __builtin_amdgcn_readfirstlane(v) == 0x7fffffffu ? 1u : 0u;
        __builtin_amdgcn_s_waitcnt(0x0f70);
        __builtin_amdgcn_wave_barrier();
        continue;
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (64u * j < U) {
          const uint32_t r = lane + 64u * j < nb ? lane + 64u * j : 0u;
          const uint32_t bi = r >> 3, jb = r & 7u;
          const uint32_t cur = ub[8 + bi];
          const uint32_t h0 = (1u << jb) | (jb ? (cur >> (8 - jb)) : 0u);
          uint32_t H = (2166136261u ^ h0) * 16777619u;
#pragma unroll
          for (int k = 1; k < 8; ++k) {
            const uint32_t x = H ^ (uint32_t)ub[8 + bi - k];
            H = x * 16777619u;
            pos[j][k - 1] = ((uint32_t)(k - 1) << tb) | ((x + dep) & mask);
            v += slot[pos[j][k - 1]];
          }
        }
      }
      acc += v;
      // the write-backs (and so the next unit's gathers) wait for this unit's
      // gathers (v < 2^31 always, so dep stays 0)
      dep = __builtin_amdgcn_readfirstlane(v) == 0x7fffffffu ? 1u : 0u;
      if (!(flags & 2u)) {
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (64u * j < U && lane + 64u * j < nb)
#pragma unroll
            for (int k = 0; k < 7; ++k) slot[pos[j][k]] = (uint8_t)(v + k + dep);
      }
      // vmcnt(0): the unit's writes land first (flags bit 3: the next unit's
      // gathers go out while they are in flight)
      if (!(flags & 8u)) __builtin_amdgcn_s_waitcnt(0x0f70);
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (acc == 0xdeadbeefu) atomicAdd(sink, 1u);
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

extern "C" int cucm_probe_replay_decode(const uint64_t *h_src, const uint32_t *h_len, uint32_t n,
                                        uint32_t tb, uint32_t D, uint32_t U, uint32_t waves,
                                        uint32_t zero_slot, double *ms) {
  if (!h_src || !h_len || !ms || n == 0 || tb < 8 || tb > 25 || D < 1 || D > 4 || (8 % D) ||
      U == 0 || U > 512 || U % 8 || waves == 0 || waves % 4)
    return -1;
  const uint64_t slot_bytes = ((7ull << tb) + 255) & ~255ull;
  ReplayChunk *d_chunks = nullptr;
  uint32_t *d_ctl = nullptr;
  uint8_t *d_pool = nullptr;
  if (hipMalloc(&d_chunks, sizeof(ReplayChunk) * n) != hipSuccess) return -4;
  if (hipMalloc(&d_ctl, 16) != hipSuccess || hipMalloc(&d_pool, slot_bytes * waves) != hipSuccess) {
    (void)hipFree(d_chunks);
    if (d_ctl) (void)hipFree(d_ctl);
    return -4;
  }
  ReplayChunk *h = (ReplayChunk *)malloc(sizeof(ReplayChunk) * n);
  if (!h) {
    (void)hipFree(d_chunks);
    (void)hipFree(d_ctl);
    (void)hipFree(d_pool);
    return -4;
  }
  for (uint32_t i = 0; i < n; ++i) h[i] = ReplayChunk{h_src[i], h_len[i], 0};
  (void)hipMemcpy(d_chunks, h, sizeof(ReplayChunk) * n, hipMemcpyHostToDevice);
  free(h);
  (void)hipMemset(d_ctl, 0, 16);
  (void)hipMemset(d_pool, 0, slot_bytes * waves);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(replay_decode_kernel, dim3(waves / 4), dim3(256), 0, 0, d_chunks, n, d_ctl,
                     d_pool, slot_bytes, tb, D, U, zero_slot, d_ctl + 1);
  (void)hipEventRecord(e1, 0);
  hipError_t err = hipEventSynchronize(e1);
  float f = 0;
  (void)hipEventElapsedTime(&f, e0, e1);
  *ms = f;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(d_chunks);
  (void)hipFree(d_ctl);
  (void)hipFree(d_pool);
  return err == hipSuccess ? 0 : -6;
}

extern "C" int cucm_probe_replay_encode(const uint64_t *h_src, const uint32_t *h_len, uint32_t n,
                                        uint32_t tb, uint32_t U, uint32_t waves, uint32_t flags,
                                        double *ms) {
  if (!h_src || !h_len || !ms || n == 0 || tb < 8 || tb > 25 || U == 0 || U > 512 || U % 8 ||
      waves == 0 || waves % 4 || ((flags & 4u) && U > 128))
    return -1;
  const uint64_t slot_bytes = ((7ull << tb) + 255) & ~255ull;
  ReplayChunk *d_chunks = nullptr;
  uint32_t *d_ctl = nullptr;
  uint8_t *d_pool = nullptr;
  if (hipMalloc(&d_chunks, sizeof(ReplayChunk) * n) != hipSuccess) return -4;
  if (hipMalloc(&d_ctl, 16) != hipSuccess || hipMalloc(&d_pool, slot_bytes * waves) != hipSuccess) {
    (void)hipFree(d_chunks);
    if (d_ctl) (void)hipFree(d_ctl);
    return -4;
  }
  ReplayChunk *h = (ReplayChunk *)malloc(sizeof(ReplayChunk) * n);
  if (!h) {
    (void)hipFree(d_chunks);
    (void)hipFree(d_ctl);
    (void)hipFree(d_pool);
    return -4;
  }
  for (uint32_t i = 0; i < n; ++i) h[i] = ReplayChunk{h_src[i], h_len[i], 0};
  (void)hipMemcpy(d_chunks, h, sizeof(ReplayChunk) * n, hipMemcpyHostToDevice);
  free(h);
  (void)hipMemset(d_ctl, 0, 16);
  (void)hipMemset(d_pool, 0, slot_bytes * waves);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, 0);
  hipLaunchKernelGGL(replay_encode_kernel, dim3(waves / 4), dim3(256), 0, 0, d_chunks, n, d_ctl,
                     d_pool, slot_bytes, tb, U, flags, d_ctl + 1);
  (void)hipEventRecord(e1, 0);
  hipError_t err = hipEventSynchronize(e1);
  float f = 0;
  (void)hipEventElapsedTime(&f, e0, e1);
  *ms = f;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(d_chunks);
  (void)hipFree(d_ctl);
  (void)hipFree(d_pool);
  return err == hipSuccess ? 0 : -6;
}
