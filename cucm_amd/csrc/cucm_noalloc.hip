// cucm_noalloc.hip -- the allocation-free compress entry point.
//
// cucm_compress_device_noalloc codes one device-resident input into a
// device-resident container like cucm_compress_many_device, but takes every
// buffer beyond the caller's input and output -- chunk tasks, result words,
// payload staging, container offsets, the queue heads and the model slots --
// from one file-scope device array, `scratch`, of kNoAllocScratchBytes.
// Nothing is allocated per call (no hipMalloc/hipFree, no host heap), so
// callers that cannot allocate (serving loops with a fixed footprint, code
// that must not call the allocator) can use it.  The chunk tasks are built on
// the device and the container's length comes back in one 16-byte copy.
//
// Each GPU holds its own copy of `scratch` from module load until the process
// ends; its address is looked up once per device with hipGetSymbolAddress at
// context creation and only that address is used from the host.  Calls on
// one GPU share the array: they are serialised by a per-device lock in this
// process, and every call clears the words that must start at zero.
#define CUCM_WITH_LAUNCHERS
#include <hip/hip_runtime.h>

#include <chrono>
#include <mutex>

#include "cucm.h"
#include "cucm_engine.h"
#include "cucm_model.h"

namespace cucm {
namespace {

constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

std::mutex g_setup_mu;
uint8_t *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device
std::mutex g_dev_mu[kMaxDevices];      // serialises calls that share one device's array

inline uint64_t up256(uint64_t x) { return (x + 255) & ~255ull; }

// One thread per chunk: the encoder task of chunk c of an input of n bytes
// (the same fields encode_run fills on the host).
__global__ void noalloc_tasks_kernel(const uint8_t *in, uint64_t n, uint32_t S, uint32_t U,
                                     uint32_t tb, uint32_t C, uint32_t pad, uint32_t frozen,
                                     uint8_t *staging, uint32_t cap, Task *tasks) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  Task t;
  t.src = (uint64_t)(uintptr_t)(in + (uint64_t)c * S);
  t.dst = (uint64_t)(uintptr_t)(staging + (uint64_t)c * cap);
  t.L = (c == C - 1) ? (uint32_t)(n - (uint64_t)c * S) : S;
  t.S = S;
  t.aux = cap;
  t.U = (uint16_t)U;
  t.tb = (uint8_t)tb;
  t.flags = (uint8_t)(((pad && c == C - 1) ? kTaskTailPad : 0) | (frozen ? kTaskFrozen : 0));
  t.model = 0;
  t.job = 0;
  tasks[c] = t;
}

// The container's length (header, then per batch record a 4-byte length,
// 4 bytes per chunk and the payloads) and whether any chunk overflowed its
// staging region or any batch blob exceeds the 32-bit length field.
__global__ void noalloc_size_kernel(const uint64_t *result, uint32_t C, uint32_t batch,
                                    uint64_t head, uint64_t *info) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint64_t len = head, bad = 0;
  for (uint32_t c0 = 0; c0 < C; c0 += batch) {
    const uint32_t Cb = min(batch, C - c0);
    uint64_t blob = 4ull * Cb;
    for (uint32_t k = 0; k < Cb; ++k) {
      const uint64_t r = result[c0 + k];
      bad |= (r & kResultOverflow) ? 1u : 0u;
      blob += r & ~kResultOverflow;
    }
    bad |= blob > 0xffffffffull ? 2u : 0u;
    len += 4 + blob;
  }
  info[0] = len;
  info[1] = bad;
}

}  // namespace
}  // namespace cucm

using namespace cucm;

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return CUCM_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_setup_mu);
  if (g_scratch[device]) return CUCM_OK;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return CUCM_E_DEVICE;
  }
  g_scratch[device] = (uint8_t *)ptr;
  return CUCM_OK;
}

#define NA_TRY(expr)                 \
  do {                               \
    if ((expr) != hipSuccess) {      \
      (void)hipGetLastError();       \
      return CUCM_E_DEVICE;          \
    }                                \
  } while (0)

extern "C" int cucm_compress_device_noalloc(cucm_ctx *ctx, const uint8_t *d_in, uint64_t in_len,
                                            const cucm_params *p, uint8_t *d_out,
                                            uint64_t out_cap, uint64_t *out_len) {
  if (!ctx || !p || !out_len || !d_out || (in_len && !d_in)) return CUCM_E_PARAM;
  *out_len = 0;
  int rc = validate_params(*p, in_len);
  if (rc) return rc;
  if (p->prelearn_size) return CUCM_E_UNSUPPORTED;  // a trained model needs device memory
  // units above 256 bits keep their records in per-wave global memory
  if (p->unit_bits > kMaxUnitBits && !(p->flags & CUCM_FLAG_FROZEN)) return CUCM_E_UNSUPPORTED;
  if (ctx->device < 0 || ctx->device >= kMaxDevices) return CUCM_E_UNSUPPORTED;
  std::lock_guard<std::mutex> dl(g_dev_mu[ctx->device]);
  std::lock_guard<std::mutex> lk(ctx->mu);
  const auto t0 = std::chrono::steady_clock::now();
  NA_TRY(hipSetDevice(ctx->device));
  uint8_t *base = g_scratch[ctx->device];
  if (!base) return CUCM_E_DEVICE;
  hipStream_t s = ctx->stream;
  const uint64_t N = in_len, S = p->chunk_size;
  const uint32_t C = (uint32_t)((N + S - 1) / S);
  const bool frozen = (p->flags & CUCM_FLAG_FROZEN) != 0;
  const bool pad = !(p->flags & CUCM_FLAG_NO_TAIL_PAD) && !frozen;
  uint8_t header[32] = {};
  write_header(header, N, *p, 0, 0);
  if (C == 0) {
    if (out_cap < 21) return CUCM_E_SMALL;
    NA_TRY(hipMemcpyAsync(d_out, header, 21, hipMemcpyHostToDevice, s));
    NA_TRY(hipStreamSynchronize(s));
    *out_len = 21;
    return CUCM_OK;
  }
  // carve the scratch: queue heads and sizes | tasks | results | container
  // offsets | assembly job | header | payload staging | model slots
  const uint64_t cap = staging_cap(S);
  if (cap > 0xfffffff0ull) return CUCM_E_UNSUPPORTED;
  uint64_t off = 0;
  uint32_t *heads = (uint32_t *)(base + off);
  uint64_t *info = (uint64_t *)(base + off + 64);
  off += 256;
  Task *tasks = (Task *)(base + off);
  off += up256(sizeof(Task) * (uint64_t)C);
  uint64_t *results = (uint64_t *)(base + off);
  off += up256(8ull * C);
  uint64_t *taskdst = (uint64_t *)(base + off);
  off += up256(8ull * C);
  AsmJob *asmjob = (AsmJob *)(base + off);
  off += up256(sizeof(AsmJob));
  uint8_t *hdr = base + off;
  off += 256;
  uint8_t *staging = base + off;
  off += up256(cap * C);
  if (off > kNoAllocScratchBytes) return CUCM_E_BUDGET;
  const uint64_t slot_bytes = frozen ? 0 : up256(7ull << p->table_bits);
  uint8_t *pool = base + off;
  const uint64_t pool_bytes = kNoAllocScratchBytes - off;
  PassPlan pl;
  pl.Umax = frozen ? 32u : p->unit_bits;
  pl.slot_bytes = slot_bytes;
  if (frozen) {
    pl.fz = frozen_shape(ctx, C);
  } else {
    pl.sh = pick_shape(ctx, C, p->unit_bits, 0, false, slot_bytes, pool_bytes);
    if (pl.sh.slots() == 0) return CUCM_E_BUDGET;
  }
  const Shape &sh = pl.sh, &fz = pl.fz;
  if (sh.blocks == 0 && fz.blocks == 0) return CUCM_E_BUDGET;
  hipLaunchKernelGGL(noalloc_tasks_kernel, dim3((C + 255) / 256), dim3(256), 0, s, d_in, N,
                     (uint32_t)S, (uint32_t)p->unit_bits, (uint32_t)p->table_bits, C,
                     pad ? 1u : 0u, frozen ? 1u : 0u, staging, (uint32_t)cap, tasks);
  NA_TRY(hipGetLastError());
  CodecArgs a{};
  a.tasks = tasks;
  a.ntasks = C;
  a.spec_bits = 0;
  a.counter = heads;
  a.pool = pool;
  a.slot_bytes = slot_bytes;
  a.result = results;
  a.prof = nullptr;
  a.tabs = (const DevTables *)ctx->tabs.p;
  a.init = nullptr;  // P = 0: no task carries kTaskInitModel
  cucm_stats st{};
  rc = launch_pass(ctx, a, false, pl, true, st);
  if (rc) return rc;
  hipLaunchKernelGGL(noalloc_size_kernel, dim3(1), dim3(64), 0, s, results, C,
                     (uint32_t)p->batch_chunks, (uint64_t)21, info);
  NA_TRY(hipGetLastError());
  uint64_t got[2] = {0, 0};
  NA_TRY(hipMemcpyAsync(got, info, sizeof got, hipMemcpyDeviceToHost, s));
  NA_TRY(hipStreamSynchronize(s));
  float ms = 0;
  NA_TRY(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  // a chunk above 1.5 S + 256 bytes (adversarial input only) or a batch
  // blob above 4 GiB is refused; cucm_compress_many_device handles both
  if (got[1]) return CUCM_E_UNSUPPORTED;
  if (got[0] > out_cap) {
    *out_len = got[0];
    return CUCM_E_SMALL;
  }
  AsmJob J{};
  J.out_base = (uint64_t)(uintptr_t)d_out;
  J.header_off = 0;
  J.header_len = 21;
  J.first_task = 0;
  J.ntasks = C;
  J.batch_chunks = p->batch_chunks;
  J.blob_only = 0;
  NA_TRY(hipMemcpyAsync(asmjob, &J, sizeof J, hipMemcpyHostToDevice, s));
  NA_TRY(hipMemcpyAsync(hdr, header, 21, hipMemcpyHostToDevice, s));
  NA_TRY(launch_assemble(asmjob, 1, hdr, tasks, C, results, taskdst, info + 2, s));
  NA_TRY(hipStreamSynchronize(s));
  *out_len = got[0];
  st.codec_ms = ms;
  st.tasks = C;
  st.waves = sh.waves() + fz.waves();
  st.teams = sh.team ? sh.blocks : 0;
  st.model_bytes = slot_bytes * sh.slots();
  st.in_bytes = N;
  st.out_bytes = got[0];
  st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                   .count();
  ctx->last = st;
  return CUCM_OK;
}
