// cucm_kernels.hip -- gfx950 kernels of the CuCM chunked context-mixing
// codec: the persistent codec kernel (its per-wave code is in cucm_wave.h),
// container assembly, device-side container parsing, and the host launchers.
#define CUCM_WITH_LAUNCHERS
#include "cucm_bigunit.h"
#include "cucm_frozen.h"
#include "cucm_team.h"

namespace cucm {

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__device__ __forceinline__ void stage_tables(const DevTables *g, LDS DevTables *s) {
  const uint32_t n = sizeof(DevTables) / 16;
  const u32x4 *src = (const u32x4 *)g;
  LDS u32x4 *dst = (LDS u32x4 *)s;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

// Largest unit of a launch rounded up to what its NJ covers: the LDS layout
// is laid out for it, so every LDS offset is a compile-time constant.
CUCM_HD uint32_t unit_capacity(uint32_t Umax) { return Umax <= 64 ? 64u : (Umax <= 128 ? 128u : 256u); }

// (kProf: the instantiation with phase clocks, launched when a.prof is set)
template <bool kDecode, int NJ, int D, bool kProf>
__global__ void __launch_bounds__(512) codec_kernel(CodecArgs a) {
  constexpr uint32_t Umax = 64u * NJ;
  extern __shared__ __align__(16) uint8_t lds[];
  LDS DevTables *T = (LDS DevTables *)lds;
  stage_tables(a.tabs, T);
  // (the wave index is uniform; readfirstlane lets the compiler keep every
  // per-wave base address in scalar registers)
  const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  constexpr uint32_t P = kDecode ? (1u << D) - 1u : 0u;
  const uint32_t wbytes = wave_lds_bytes(Umax, P, kDecode);
  LDS uint8_t *wbase = (LDS uint8_t *)lds + align16(sizeof(DevTables)) + wid * wbytes;
  const WaveMem M = carve(wbase, Umax, P, kDecode);
  const uint64_t gw = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wid;
  GLB uint8_t *slot = (GLB uint8_t *)a.pool + gw * a.slot_bytes;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= a.ntasks) break;
    const Task tk = a.tasks[t];
    if (tk.flags & (kTaskFrozen | kTaskBigUnit)) continue;
    if constexpr (kDecode)
      decode_task<NJ, D, kProf>(a, tk, t, M, T, slot, Umax, lane);
    else
      encode_task<NJ, kProf>(a, tk, t, M, T, slot, Umax, lane);
  }
}

// The team decoder: one chunk per workgroup of kTeamWaves waves, each
// workgroup owning one model slot and pulling chunk tasks from the queue.
// (kProf: the instantiation with phase clocks, launched when a.prof is set)
template <int NJ, int D, bool kProf>
__global__ void __launch_bounds__(256) decode_team_kernel(CodecArgs a) {
  constexpr uint32_t Umax = 64u * NJ;
  extern __shared__ __align__(16) uint8_t lds[];
  LDS DevTables *T = (LDS DevTables *)lds;
  stage_tables(a.tabs, T);
  const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  LDS uint8_t *tbase = (LDS uint8_t *)lds + align16(sizeof(DevTables));
  const TeamMem X = carve_team(tbase, Umax, D, wid);
  GLB uint8_t *slot = (GLB uint8_t *)a.pool + (uint64_t)blockIdx.x * a.slot_bytes;
  uint32_t rid = 0;  // running round id (identical in every wave)
  for (;;) {
    if (threadIdx.x == 0) X.ctl[0] = atomicAdd(a.counter, 1u);
    team_sync();
    const uint32_t t = __builtin_amdgcn_readfirstlane(X.ctl[0]);
    team_sync();
    if (t >= a.ntasks) break;
    const Task tk = a.tasks[t];
    if (tk.flags & (kTaskFrozen | kTaskBigUnit)) continue;  // uniform: every wave read the same task
    decode_team<NJ, D, kProf>(a, tk, t, X, T, slot, Umax, wid, lane, rid);
  }
}

// The team encoder: one chunk per workgroup of kTeamWaves waves.
template <int NJ, bool kProf>
__global__ void __launch_bounds__(256) encode_team_kernel(CodecArgs a) {
  constexpr uint32_t Umax = 64u * NJ;
  extern __shared__ __align__(16) uint8_t lds[];
  LDS DevTables *T = (LDS DevTables *)lds;
  stage_tables(a.tabs, T);
  const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  LDS uint8_t *tbase = (LDS uint8_t *)lds + align16(sizeof(DevTables));
  const TeamMem X = carve_team_enc(tbase, Umax, wid);
  GLB uint8_t *slot = (GLB uint8_t *)a.pool + (uint64_t)blockIdx.x * a.slot_bytes;
  for (;;) {
    if (threadIdx.x == 0) X.ctl[0] = atomicAdd(a.counter, 1u);
    team_sync();
    const uint32_t t = __builtin_amdgcn_readfirstlane(X.ctl[0]);
    team_sync();
    if (t >= a.ntasks) break;
    const Task tk = a.tasks[t];
    if (tk.flags & (kTaskFrozen | kTaskBigUnit)) continue;  // uniform: every wave read the same task
    encode_team<NJ, kProf>(a, tk, t, X, T, slot, Umax, wid, lane);
  }
}

// The frozen-model codec: one wave per chunk of the tasks marked
// kTaskFrozen, the shared read-only model read in place (no model slots).
template <bool kDecode, int D>
__global__ void __launch_bounds__(512) frozen_kernel(CodecArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  LDS DevTables *T = (LDS DevTables *)lds;
  stage_tables(a.tabs, T);
  // (the wave index is uniform; readfirstlane lets the compiler keep every
  // per-wave base address in scalar registers)
  const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  const FrozenMem F =
      carve_frozen((LDS uint8_t *)lds + align16(sizeof(DevTables)) + wid * frozen_wave_bytes());
  uint32_t loaded = kNoModel;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= a.ntasks) break;
    const Task tk = a.tasks[t];
    if (!(tk.flags & kTaskFrozen) || tk.L == 0) continue;
    const GLB uint8_t *states = frozen_model(a, tk, F, T, loaded, lane);
    if constexpr (kDecode)
      frozen_decode_task<D>(a, tk, t, F, T, states, lane);
    else
      frozen_encode_task(a, tk, t, F, T, states, lane);
  }
}

// The large-unit codec: one wave per chunk of the tasks marked kTaskBigUnit,
// each wave with a model slot and a region of unit records in global memory.
template <bool kDecode>
__global__ void __launch_bounds__(256) bigunit_kernel(CodecArgs a) {
  extern __shared__ __align__(16) uint8_t lds[];
  LDS DevTables *T = (LDS DevTables *)lds;
  stage_tables(a.tabs, T);
  // (the wave index is uniform; readfirstlane lets the compiler keep every
  // per-wave base address in scalar registers)
  const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  const BigMem M =
      carve_big((LDS uint8_t *)lds + align16(sizeof(DevTables)) + wid * big_wave_bytes());
  const uint64_t gw = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wid;
  GLB uint8_t *slot = (GLB uint8_t *)a.pool + gw * a.slot_bytes;
  const BigRec R = big_records((GLB uint8_t *)a.big + gw * a.big_bytes, a.big_umax);
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.counter, 1u);
    t = __builtin_amdgcn_readfirstlane(t);
    if (t >= a.ntasks) break;
    const Task tk = a.tasks[t];
    if (!(tk.flags & kTaskBigUnit) || tk.L == 0 || tk.U > a.big_umax) continue;
    if constexpr (kDecode)
      big_decode_task(a, tk, t, M, T, slot, R, lane);
    else
      big_encode_task(a, tk, t, M, T, slot, R, lane);
  }
}

// ---------------------------------------------------------------------------
// Serving: one persistent launch over a ring of chunk tasks that the host
// keeps appending to (tail, in host memory), encoder and decoder tasks of
// units up to 128 bits mixed.  Requests' inputs, containers, decoded bytes
// and prelearned models live in host memory shared with the device, so the
// kernel reads nothing the host wrote into device memory while it runs.
// The wave that finishes a job's last task assembles its container (compress)
// and raises the job's done word.
// ---------------------------------------------------------------------------

constexpr int kServeNJ = 2;  // units up to 128 bits
constexpr int kServeD = 2;   // decoder: 2-bit speculative rounds

// n bytes from s to d, the wave's lanes striding: head bytes until d is
// 4-aligned, then words assembled from s, then the tail bytes.
__device__ __forceinline__ void wave_copy(uint8_t *d, const uint8_t *s, uint64_t n,
                                          uint32_t lane) {
  const uint32_t head = (uint32_t)((4u - ((uintptr_t)d & 3u)) & 3u);
  const uint64_t hn = head < n ? head : n;
  if (lane < hn) d[lane] = s[lane];
  const uint64_t nwords = (n - hn) >> 2;
  uint32_t *dw = (uint32_t *)(d + hn);
  const uint8_t *sb = s + hn;
  for (uint64_t i = lane; i < nwords; i += 64) {
    const uint8_t *p = sb + 4 * i;
    dw[i] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
            ((uint32_t)p[3] << 24);
  }
  const uint64_t t0 = hn + 4 * nwords;
  if (lane < n - t0) d[t0 + lane] = s[t0 + lane];
}

// A compress job's batch records after its header: per batch the LE blob
// length, the BE end offsets E_c = 4C + sum_{j<=c} size_j - 1 and the
// payloads (cucm_compress_v2.cu:870-914, 1242-1244).  Returns the container
// length, or 0 if a chunk overflowed its staging region.
__device__ uint64_t serve_assemble(const ServeJob &J, const Task *ring, const uint64_t *res,
                                   uint32_t mask, uint32_t lane) {
  uint8_t *o = (uint8_t *)(uintptr_t)J.out_base;
  uint64_t pos = J.header_len;
  uint32_t bad = 0;
  for (uint32_t c0 = 0; c0 < J.ntasks; c0 += J.batch) {
    const uint32_t C = min(J.batch, J.ntasks - c0);
    uint64_t e = 4ull * C;  // running end of the blob
    for (uint32_t g = 0; g < C; g += 64) {
      const uint32_t c = g + lane;
      const uint64_t r = c < C ? res[(J.first + c0 + c) & mask] : 0ull;
      bad |= __ballot((r & kResultOverflow) != 0) != 0ull ? 1u : 0u;
      const uint32_t sz = (uint32_t)(r & ~kResultOverflow);
      uint64_t run = 0, mine = 0;
      for (uint32_t k = 0; k < 64; ++k) {
        run += __builtin_amdgcn_readlane(sz, k);
        mine = k == lane ? run : mine;
      }
      if (c < C) {
        const uint64_t E = e + mine - 1;
        uint8_t *q = o + pos + 4 + 4ull * c;
        q[0] = (uint8_t)(E >> 24);
        q[1] = (uint8_t)(E >> 16);
        q[2] = (uint8_t)(E >> 8);
        q[3] = (uint8_t)E;
      }
      e += run;
    }
    if (lane == 0) {
      o[pos + 0] = (uint8_t)e;
      o[pos + 1] = (uint8_t)(e >> 8);
      o[pos + 2] = (uint8_t)(e >> 16);
      o[pos + 3] = (uint8_t)(e >> 24);
    }
    uint64_t off = 4ull * C;
    for (uint32_t c = 0; c < C; ++c) {
      const uint32_t ri = (J.first + c0 + c) & mask;
      const uint64_t sz = res[ri] & ~kResultOverflow;
      const uint8_t *src = (const uint8_t *)(uintptr_t)ring[ri].dst;
      if (!bad) wave_copy(o + pos + 4 + off, src, sz, lane);
      off += sz;
    }
    pos += 4 + e;
  }
  return bad ? 0ull : pos;
}

__global__ void __launch_bounds__(512) serve_kernel(ServeArgs s) {
  constexpr int NJ = kServeNJ, D = kServeD;
  constexpr uint32_t Umax = 64u * NJ, P = (1u << D) - 1u;
  extern __shared__ __align__(16) uint8_t lds[];
  LDS DevTables *T = (LDS DevTables *)lds;
  stage_tables(s.a.tabs, T);
  // (the wave index is uniform; readfirstlane lets the compiler keep every
  // per-wave base address in scalar registers)
  const uint32_t wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wbytes = wave_lds_bytes(Umax, P, true);
  const WaveMem M = carve((LDS uint8_t *)lds + align16(sizeof(DevTables)) + wid * wbytes, Umax, P,
                          true);
  const uint64_t gw = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wid;
  GLB uint8_t *slot = (GLB uint8_t *)s.a.pool + gw * s.a.slot_bytes;
  uint32_t *mtail = s.mirror, *mstop = s.mirror + 1;
  uint64_t *mtime = (uint64_t *)(s.mirror + 2);
  uint64_t idle0 = __builtin_amdgcn_s_memrealtime();
  uint32_t spins = 0;
  for (;;) {
    // claim the next published task, if any (compare-and-swap on the head,
    // so a wave that leaves never holds an unpublished index), against the
    // device mirror of the tail
    uint32_t t = 0, got = 0;
    if (lane == 0) {
      const uint32_t tl = __hip_atomic_load(mtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t h = __hip_atomic_load(s.a.counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      while ((int32_t)(tl - h) > 0) {
        const uint32_t old = atomicCAS(s.a.counter, h, h + 1u);
        if (old == h) {
          got = 1u;
          t = h;
          break;
        }
        h = old;
      }
    }
    got = __builtin_amdgcn_readfirstlane(got);
    t = __builtin_amdgcn_readfirstlane(t);
    if (!got) {
      // Refresh the mirror from host memory: one wave at a time, at most once
      // per poll_ticks.  Host words are read across the host link, and idle
      // waves reading them all the time delay the working waves' own host
      // traffic many times over; an acquire at system scope on every poll
      // would also invalidate the L2 under the working waves.
      uint32_t stop = 0;
      const uint64_t now = __builtin_amdgcn_s_memrealtime();
      if (lane == 0) {
        const uint64_t last = __hip_atomic_load(mtime, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (now - last >= s.poll_ticks && atomicCAS((unsigned long long *)mtime, last, now) == last) {
          const uint32_t ht = __hip_atomic_load(s.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          const uint32_t hs = __hip_atomic_load(s.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          uint32_t cur = __hip_atomic_load(mtail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if ((int32_t)(ht - cur) > 0) {
            // synchronises with the host's release of the tail, and passes
            // that on to the wave that reads the mirror and claims a task
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            while ((int32_t)(ht - cur) > 0) {
              const uint32_t o = atomicCAS(mtail, cur, ht);
              if (o == cur) break;
              cur = o;
            }
          }
          if (hs) __hip_atomic_store(mstop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        stop = __hip_atomic_load(mstop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      stop = __builtin_amdgcn_readfirstlane(stop);
      if (stop || now - idle0 > s.idle_ticks) break;
      // look often just after a task, then back off: every ~25 us, doubling
      // every 16 looks up to ~0.8 ms.  Every idle wave's look reads the same
      // few words, and hundreds of idle waves looking often slow the working
      // waves' gathers (seven waves per CU: the last tasks of a saturated run
      // took several times as long).  Any idle wave takes a new task, so the
      // chip as a whole still looks every few microseconds.
      ++spins;
      if (spins < 64) {
        __builtin_amdgcn_s_sleep(32);
      } else {
        const uint32_t reps = 8u << min((spins - 64u) >> 4, 5u);
        for (uint32_t k = 0; k < reps; ++k) __builtin_amdgcn_s_sleep(127);
      }
      continue;
    }
    spins = 0;
    // with the mirror's release above: the task's host data is read after
    // the host published it
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t ri = t & s.ring_mask;
    const Task tk = s.a.tasks[ri];
    if (tk.flags & kTaskDecode) {
      // decoded into device staging, then copied to the host buffer at once:
      // a store to host memory per unit would hold up every later wait of
      // the decoder on its gathers (stores count in the same counter)
      decode_task<NJ, D, false>(s.a, tk, ri, M, T, slot, Umax, lane);
      const ServeJob Jd = s.jobs[tk.job];
      uint8_t *hd = (uint8_t *)(uintptr_t)Jd.out_base + Jd.header_len +
                    (uint64_t)((ri - Jd.first) & s.ring_mask) * tk.S;
      wave_copy(hd, (const uint8_t *)(uintptr_t)tk.dst, tk.L, lane);
    } else {
      encode_task<NJ, false>(s.a, tk, ri, M, T, slot, Umax, lane);
    }
    // this task's payload and result word are visible before it is counted
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    uint32_t prev = 0;
    if (lane == 0) prev = atomicAdd(&s.jobcount[tk.job], 1u);
    prev = __builtin_amdgcn_readfirstlane(prev);
    const ServeJob J = s.jobs[tk.job];
    if (prev + 1u == J.ntasks) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every task of the job is done
      uint64_t len = 0;
      if (!(tk.flags & kTaskDecode)) len = serve_assemble(J, s.a.tasks, s.a.result, s.ring_mask, lane);
      if (lane == 0) __hip_atomic_store(&s.jobcount[tk.job], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // the container (or the decoded bytes of every task) reaches the host
      // before the done word
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
      if (lane == 0) s.lens[tk.job] = len;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
      if (lane == 0) {
        const bool failed = !(tk.flags & kTaskDecode) && len == 0;
        __hip_atomic_store(&s.done[tk.job], J.id | (failed ? 0x80000000u : 0u), __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_SYSTEM);
      }
    }
    idle0 = __builtin_amdgcn_s_memrealtime();
  }
}

// ---------------------------------------------------------------------------
// Container assembly: per job, per batch: LE u32 blob length, BE end
// offsets E_c = 4C + sum_{j<=c} size_j - 1, payloads back to back
// (cucm_compress_v2.cu:870-914, 1242-1244).
// ---------------------------------------------------------------------------

__global__ void layout_kernel(const AsmJob *jobs, uint32_t njobs, const uint64_t *result,
                              uint64_t *task_dst, uint64_t *job_len) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  const AsmJob J = jobs[j];
  uint8_t *o = (uint8_t *)(uintptr_t)J.out_base;
  const uint32_t lenbytes = J.blob_only ? 0u : 4u;
  // (the header bytes are copied by copy_headers_kernel)
  uint64_t pos = J.header_len;
  for (uint32_t c0 = 0; c0 < J.ntasks; c0 += J.batch_chunks) {
    const uint32_t C = min(J.batch_chunks, J.ntasks - c0);
    const uint64_t bstart = pos;
    uint64_t e = 4ull * C;
    for (uint32_t c = 0; c < C; ++c) {
      const uint32_t t = J.first_task + c0 + c;
      const uint64_t sz = result[t] & ~kResultOverflow;
      task_dst[t] = J.out_base + bstart + lenbytes + e;
      e += sz;
      const uint64_t E = e - 1;
      uint8_t *q = o + bstart + lenbytes + 4ull * c;
      q[0] = (uint8_t)(E >> 24);
      q[1] = (uint8_t)(E >> 16);
      q[2] = (uint8_t)(E >> 8);
      q[3] = (uint8_t)E;
    }
    if (lenbytes) {
      o[bstart + 0] = (uint8_t)e;
      o[bstart + 1] = (uint8_t)(e >> 8);
      o[bstart + 2] = (uint8_t)(e >> 16);
      o[bstart + 3] = (uint8_t)(e >> 24);
    }
    pos = bstart + lenbytes + e;
  }
  job_len[j] = pos;
}

// One wave per job: its header and prelearn segment (up to 21 + P bytes)
// into the container, the lanes striding over it.
__global__ void copy_headers_kernel(const AsmJob *jobs, uint32_t njobs, const uint8_t *headers) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t j = w0; j < njobs; j += nw) {
    const AsmJob J = jobs[j];
    uint8_t *o = (uint8_t *)(uintptr_t)J.out_base;
    const uint8_t *h = headers + J.header_off;
    for (uint64_t i = lane; i < J.header_len; i += 64) o[i] = h[i];
  }
}

// One wave per task: copy the staged payload to its container position.
__global__ void copy_payload_kernel(const Task *tasks, uint32_t ntasks, const uint64_t *result,
                                    const uint64_t *task_dst) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t t = w0; t < ntasks; t += nw) {
    if (task_dst[t] == 0) continue;  // job not assembled
    const uint64_t sz = result[t] & ~kResultOverflow;
    const uint8_t *s = (const uint8_t *)(uintptr_t)tasks[t].dst;
    uint8_t *d = (uint8_t *)(uintptr_t)task_dst[t];
    // Head bytes until d is 4-aligned, then words assembled from s.
    const uint32_t head = (uint32_t)((4u - ((uintptr_t)d & 3u)) & 3u);
    const uint64_t hn = head < sz ? head : sz;
    if (lane < hn) d[lane] = s[lane];
    const uint64_t rem = sz - hn;
    const uint64_t nwords = rem >> 2;
    uint32_t *dw = (uint32_t *)(d + hn);
    const uint8_t *sb = s + hn;
    for (uint64_t i = lane; i < nwords; i += 64) {
      const uint8_t *p = sb + 4 * i;
      dw[i] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
              ((uint32_t)p[3] << 24);
    }
    const uint64_t tail0 = hn + 4 * nwords;
    if (lane < sz - tail0) d[tail0 + lane] = s[tail0 + lane];
  }
}

// ---------------------------------------------------------------------------
// Device-side container parsing for decompression: headers are gathered
// into one array (one small D2H), the host sizes the task list, then one
// thread per container walks its batch records and emits chunk tasks
// (cucm_decompress_v2.cu:1189-1247).
// ---------------------------------------------------------------------------

__global__ void read_headers_kernel(const uint64_t *src, const uint64_t *len, uint32_t n,
                                    uint8_t *hdr) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint8_t *s = (const uint8_t *)(uintptr_t)src[j];
  const uint64_t m = len[j] < 21 ? len[j] : 21;
  for (uint32_t i = 0; i < 21; ++i) hdr[21ull * j + i] = i < m ? s[i] : (uint8_t)0;
}

__device__ __forceinline__ uint32_t ld_be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
__device__ __forceinline__ uint32_t ld_le32(const uint8_t *p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

__global__ void build_decode_tasks_kernel(const ParseJob *jobs, uint32_t n, Task *tasks,
                                          int32_t *status) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const ParseJob J = jobs[j];
  const uint8_t *s = (const uint8_t *)(uintptr_t)J.src;
  uint64_t pos = J.seg;
  int32_t st = 0;
  for (uint32_t c0 = 0; c0 < J.ntasks && st == 0; c0 += J.batch) {
    const uint32_t C = min(J.batch, J.ntasks - c0);
    if (pos + 4 > J.len) { st = -2; break; }
    const uint64_t blen = ld_le32(s + pos);
    if (pos + 4 + blen > J.len || blen < 4ull * C + C) { st = -2; break; }
    const uint8_t *blob = s + pos + 4;
    uint64_t prev = 4ull * C - 1;
    for (uint32_t c = 0; c < C; ++c) {
      const uint64_t E = ld_be32(blob + 4ull * c);
      if (E <= prev || E >= blen) { st = -2; break; }
      const uint32_t ci = c0 + c;
      Task t;
      t.src = J.src + pos + 4 + prev + 1;
      t.aux = (uint32_t)(E - prev);
      t.dst = J.dst + (uint64_t)ci * J.S;
      t.L = (ci == J.ntasks - 1) ? (uint32_t)(J.body - (uint64_t)ci * J.S) : J.S;
      t.S = J.S;
      t.U = (uint16_t)J.U;
      t.tb = (uint8_t)J.tb;
      t.flags = (uint8_t)J.flags;
      t.model = J.model;
      t.job = 0;
      tasks[J.first_task + ci] = t;
      prev = E;
    }
    if (st == 0 && prev != blen - 1) st = -2;
    pos += 4 + blen;
  }
  if (st != 0) {
    for (uint32_t c = 0; c < J.ntasks; ++c) {
      Task t = {};
      t.U = (uint16_t)J.U;
      t.tb = (uint8_t)J.tb;
      t.S = J.S;
      // L = 0: the decoder skips it; the kind flags keep it with its codec
      t.flags = (uint8_t)(J.flags & (kTaskFrozen | kTaskBigUnit));
      tasks[J.first_task + c] = t;
    }
  }
  status[j] = st;
}

// ---------------------------------------------------------------------------
// Host-callable launchers
// ---------------------------------------------------------------------------

uint32_t codec_lds_bytes(uint32_t Umax, uint32_t spec_bits, bool decode, uint32_t waves) {
  const uint32_t P = decode ? (1u << spec_bits) - 1u : 0u;
  return align16(sizeof(DevTables)) + waves * wave_lds_bytes(unit_capacity(Umax), P, decode);
}

uint32_t team_codec_lds_bytes(uint32_t Umax, uint32_t spec_bits, bool decode) {
  return align16(sizeof(DevTables)) + (decode ? team_lds_bytes(unit_capacity(Umax), spec_bits)
                                              : team_enc_lds_bytes(unit_capacity(Umax)));
}

bool team_supported(bool decode, uint32_t spec_bits) {
  return !decode || (spec_bits >= 1 && spec_bits <= kTeamMaxSpec && (8 % spec_bits) == 0);
}

template <int NJ, bool kProf>
static hipError_t launch_team_enc_one(const CodecArgs &a, uint32_t blocks, hipStream_t stream) {
  const uint32_t lds = team_codec_lds_bytes(64u * NJ, 0, false);
  hipError_t e = hipFuncSetAttribute((const void *)encode_team_kernel<NJ, kProf>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((encode_team_kernel<NJ, kProf>), dim3(blocks), dim3(64 * kTeamWaves), lds,
                     stream, a);
  return hipGetLastError();
}

template <int NJ>
static hipError_t launch_team_enc(const CodecArgs &a, uint32_t blocks, hipStream_t stream) {
  return a.prof ? launch_team_enc_one<NJ, true>(a, blocks, stream)
                : launch_team_enc_one<NJ, false>(a, blocks, stream);
}

template <int NJ, int D, bool kProf>
static hipError_t launch_team_prof(const CodecArgs &a, uint32_t Umax, uint32_t blocks,
                                   hipStream_t stream) {
  const uint32_t lds = team_codec_lds_bytes(Umax, D, true);
  hipError_t e = hipFuncSetAttribute((const void *)decode_team_kernel<NJ, D, kProf>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((decode_team_kernel<NJ, D, kProf>), dim3(blocks), dim3(64 * kTeamWaves), lds,
                     stream, a);
  return hipGetLastError();
}

template <int NJ, int D>
static hipError_t launch_team_one(const CodecArgs &a, uint32_t Umax, uint32_t blocks,
                                  hipStream_t stream) {
  return a.prof ? launch_team_prof<NJ, D, true>(a, Umax, blocks, stream)
                : launch_team_prof<NJ, D, false>(a, Umax, blocks, stream);
}

template <int NJ>
static hipError_t launch_team(const CodecArgs &a, uint32_t Umax, uint32_t blocks,
                              hipStream_t stream) {
  switch (a.spec_bits) {
    case 1: return launch_team_one<NJ, 1>(a, Umax, blocks, stream);
    case 2: return launch_team_one<NJ, 2>(a, Umax, blocks, stream);
    case 4: return launch_team_one<NJ, 4>(a, Umax, blocks, stream);
    default: return hipErrorInvalidValue;
  }
}

template <bool kDecode, int NJ, int D, bool kProf>
static hipError_t launch_one_prof(const CodecArgs &a, uint32_t blocks, uint32_t wpb, uint32_t lds,
                                  hipStream_t stream) {
  hipError_t e = hipFuncSetAttribute((const void *)codec_kernel<kDecode, NJ, D, kProf>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((codec_kernel<kDecode, NJ, D, kProf>), dim3(blocks), dim3(64 * wpb), lds,
                     stream, a);
  return hipGetLastError();
}

template <bool kDecode, int NJ, int D>
static hipError_t launch_one(const CodecArgs &a, uint32_t Umax, uint32_t blocks, uint32_t wpb,
                             uint32_t lds, hipStream_t stream) {
  (void)Umax;
  return a.prof ? launch_one_prof<kDecode, NJ, D, true>(a, blocks, wpb, lds, stream)
                : launch_one_prof<kDecode, NJ, D, false>(a, blocks, wpb, lds, stream);
}

template <int NJ>
static hipError_t launch_decode(const CodecArgs &a, uint32_t Umax, uint32_t blocks, uint32_t wpb,
                                uint32_t lds, hipStream_t stream) {
  switch (a.spec_bits) {
    case 1: return launch_one<true, NJ, 1>(a, Umax, blocks, wpb, lds, stream);
    case 2: return launch_one<true, NJ, 2>(a, Umax, blocks, wpb, lds, stream);
    case 4: return launch_one<true, NJ, 4>(a, Umax, blocks, wpb, lds, stream);
    case 8: return launch_one<true, NJ, 8>(a, Umax, blocks, wpb, lds, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_codec(bool decode, const CodecArgs &a, uint32_t Umax, uint32_t blocks,
                        uint32_t waves_per_block, bool team, hipStream_t stream) {
  const uint32_t lds = codec_lds_bytes(Umax, a.spec_bits, decode, waves_per_block);
  const int nj = Umax <= 64 ? 1 : (Umax <= 128 ? 2 : 4);
  if (team) {
    if (!team_supported(decode, a.spec_bits)) return hipErrorInvalidValue;
    if (!decode) {
      if (nj == 1) return launch_team_enc<1>(a, blocks, stream);
      if (nj == 2) return launch_team_enc<2>(a, blocks, stream);
      return launch_team_enc<4>(a, blocks, stream);
    }
    if (nj == 1) return launch_team<1>(a, Umax, blocks, stream);
    if (nj == 2) return launch_team<2>(a, Umax, blocks, stream);
    return launch_team<4>(a, Umax, blocks, stream);
  }
  if (decode) {
    if (nj == 1) return launch_decode<1>(a, Umax, blocks, waves_per_block, lds, stream);
    if (nj == 2) return launch_decode<2>(a, Umax, blocks, waves_per_block, lds, stream);
    return launch_decode<4>(a, Umax, blocks, waves_per_block, lds, stream);
  }
  if (nj == 1) return launch_one<false, 1, 0>(a, Umax, blocks, waves_per_block, lds, stream);
  if (nj == 2) return launch_one<false, 2, 0>(a, Umax, blocks, waves_per_block, lds, stream);
  return launch_one<false, 4, 0>(a, Umax, blocks, waves_per_block, lds, stream);
}

uint32_t frozen_lds_bytes(uint32_t waves) {
  return align16(sizeof(DevTables)) + waves * frozen_wave_bytes();
}

template <bool kDecode, int D>
static hipError_t launch_frozen_one(const CodecArgs &a, uint32_t blocks, uint32_t wpb,
                                    hipStream_t stream) {
  const uint32_t lds = frozen_lds_bytes(wpb);
  hipError_t e = hipFuncSetAttribute((const void *)frozen_kernel<kDecode, D>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((frozen_kernel<kDecode, D>), dim3(blocks), dim3(64 * wpb), lds, stream, a);
  return hipGetLastError();
}

hipError_t launch_frozen(bool decode, const CodecArgs &a, uint32_t blocks, uint32_t wpb,
                         hipStream_t stream) {
  if (blocks == 0 || wpb == 0 || wpb > 8) return hipErrorInvalidValue;
  if (!decode) return launch_frozen_one<false, 0>(a, blocks, wpb, stream);
  switch (a.spec_bits) {
    case 1: return launch_frozen_one<true, 1>(a, blocks, wpb, stream);
    case 2: return launch_frozen_one<true, 2>(a, blocks, wpb, stream);
    case 4: return launch_frozen_one<true, 4>(a, blocks, wpb, stream);
    case 8: return launch_frozen_one<true, 8>(a, blocks, wpb, stream);
    default: return hipErrorInvalidValue;
  }
}

uint32_t serve_lds_bytes(uint32_t waves) {
  return align16(sizeof(DevTables)) +
         waves * wave_lds_bytes(64u * kServeNJ, (1u << kServeD) - 1u, true);
}

hipError_t launch_serve(const ServeArgs &s, uint32_t blocks, uint32_t wpb, hipStream_t stream) {
  if (blocks == 0 || wpb == 0 || wpb > 8) return hipErrorInvalidValue;
  const uint32_t lds = serve_lds_bytes(wpb);
  hipError_t e = hipFuncSetAttribute((const void *)serve_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(serve_kernel, dim3(blocks), dim3(64 * wpb), lds, stream, s);
  return hipGetLastError();
}

uint32_t big_lds_bytes(uint32_t waves) {
  return align16(sizeof(DevTables)) + waves * big_wave_bytes();
}

hipError_t launch_big(bool decode, const CodecArgs &a, uint32_t blocks, uint32_t wpb,
                      hipStream_t stream) {
  if (blocks == 0 || wpb == 0 || wpb > 4) return hipErrorInvalidValue;
  const uint32_t lds = big_lds_bytes(wpb);
  const void *fn = decode ? (const void *)bigunit_kernel<true> : (const void *)bigunit_kernel<false>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (decode)
    hipLaunchKernelGGL((bigunit_kernel<true>), dim3(blocks), dim3(64 * wpb), lds, stream, a);
  else
    hipLaunchKernelGGL((bigunit_kernel<false>), dim3(blocks), dim3(64 * wpb), lds, stream, a);
  return hipGetLastError();
}

hipError_t launch_read_headers(const uint64_t *src, const uint64_t *len, uint32_t n, uint8_t *hdr,
                               hipStream_t stream) {
  hipLaunchKernelGGL(read_headers_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, src, len, n,
                     hdr);
  return hipGetLastError();
}

hipError_t launch_build_decode_tasks(const ParseJob *jobs, uint32_t n, Task *tasks,
                                     int32_t *status, hipStream_t stream) {
  hipLaunchKernelGGL(build_decode_tasks_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, jobs, n,
                     tasks, status);
  return hipGetLastError();
}

hipError_t launch_assemble(const AsmJob *jobs, uint32_t njobs, const uint8_t *headers,
                           const Task *tasks, uint32_t ntasks, const uint64_t *result,
                           uint64_t *task_dst, uint64_t *job_len, hipStream_t stream) {
  if (njobs) {
    hipLaunchKernelGGL(layout_kernel, dim3((njobs + 63) / 64), dim3(64), 0, stream, jobs, njobs,
                       result, task_dst, job_len);
    const uint32_t hblocks = min((njobs + 3) / 4, 4096u);
    hipLaunchKernelGGL(copy_headers_kernel, dim3(hblocks), dim3(256), 0, stream, jobs, njobs,
                       headers);
  }
  if (ntasks) {
    const uint32_t blocks = min((ntasks + 3) / 4, 4096u);
    hipLaunchKernelGGL(copy_payload_kernel, dim3(blocks), dim3(256), 0, stream, tasks, ntasks,
                       result, task_dst);
  }
  return hipGetLastError();
}

}  // namespace cucm
