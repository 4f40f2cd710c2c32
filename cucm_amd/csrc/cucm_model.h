/*
 * cucm_model.h -- model constants and device task layout shared by the HIP
 * kernels and the host engine of the MI355X CuCM codec.
 */
#ifndef CUCM_MODEL_H
#define CUCM_MODEL_H
#include <stdint.h>

namespace cucm {

constexpr int kOrders = 8;        // context orders 0..7 (cucm_compress_v2.cu:552)
constexpr int kMixCtx = 80;       // mixer weight sets (cucm_compress_v2.cu:554)
constexpr int kMaxUnitBits = 256; // largest unit the main kernels' LDS staging is laid out for
constexpr uint32_t kMaxBigUnitBits = 65504; // largest unit the container's U field holds (x 32)
constexpr uint32_t kTeamWaves = 4; // waves of a team workgroup (one chunk per workgroup)
constexpr uint32_t kFnvBasis = 2166136261u;
constexpr uint32_t kFnvPrime = 16777619u;

// Where order k's state (k >= 1) sits in its table.  The reference indexes
// order k by slot = H_k & mask with H_k = (H_{k-1} ^ b_k) * prime
// (cucm_compress_v2.cu:727-743).  Here the state of slot s is stored at
// position s * prime^-1 mod 2^tb, which is (H_{k-1} ^ b_k) & mask: prime is
// odd, so the map is a bijection of the table (two contexts share a state
// exactly when they share a slot, and every coded bit is unchanged), and the
// 256 contexts that differ only in their farthest byte fall into one aligned
// 256-byte block -- order 1's states of a chunk lie in at most 255 such
// blocks, so a far larger share of the state traffic hits in cache.
// Advances H from H_{k-1} to H_k and returns order k's position.
constexpr uint32_t fnv_pos(uint32_t &H, uint32_t b, uint32_t mask) {
  const uint32_t x = H ^ b;
  H = x * kFnvPrime;
  return x & mask;
}

// Bit-history transition nex[s] = next0 | next1 << 8.  Every entry must
// equal the reference's state_table (cucm_compress_v2.cu:35-285): the coded
// stream depends on all of them.
static const uint16_t kNextState[256] = {
  0x0201, 0x0503, 0x0604, 0x0a07, 0x0c08, 0x0d09, 0x0e0b, 0x130f,
  0x1710, 0x1811, 0x1912, 0x1b14, 0x1c15, 0x1d16, 0x1e1a, 0x211f,
  0x2320, 0x2320, 0x2320, 0x2320, 0x2522, 0x2522, 0x2522, 0x2522,
  0x2522, 0x2522, 0x2724, 0x2724, 0x2724, 0x2724, 0x2826, 0x2b29,
  0x2d2a, 0x2d2a, 0x2f2c, 0x2f2c, 0x312e, 0x312e, 0x3330, 0x3330,
  0x3432, 0x2b35, 0x3936, 0x3936, 0x3b38, 0x3b38, 0x3d3a, 0x3d3a,
  0x3f3c, 0x3f3c, 0x413e, 0x413e, 0x4232, 0x3743, 0x3944, 0x3944,
  0x4946, 0x4946, 0x4b48, 0x4b48, 0x4d4a, 0x4d4a, 0x4f4c, 0x4f4c,
  0x513e, 0x513e, 0x5240, 0x4553, 0x4754, 0x4754, 0x4956, 0x4956,
  0x3b2c, 0x3b2c, 0x3d3a, 0x3d3a, 0x313c, 0x313c, 0x594c, 0x594c,
  0x5b4e, 0x5b4e, 0x5c50, 0x455d, 0x575e, 0x575e, 0x2d60, 0x2d60,
  0x6330, 0x6330, 0x6558, 0x6558, 0x6650, 0x4567, 0x5768, 0x5768,
  0x396a, 0x396a, 0x6d3e, 0x6d3e, 0x6f58, 0x6f58, 0x7050, 0x5571,
  0x5772, 0x5772, 0x3974, 0x3974, 0x773e, 0x773e, 0x7958, 0x7958,
  0x7a5a, 0x557b, 0x617c, 0x617c, 0x397e, 0x397e, 0x813e, 0x813e,
  0x8362, 0x8362, 0x845a, 0x5585, 0x6186, 0x6186, 0x3988, 0x3988,
  0x8b3e, 0x8b3e, 0x8d62, 0x8d62, 0x8e5a, 0x5f8f, 0x6190, 0x6190,
  0x3944, 0x3944, 0x513e, 0x513e, 0x9362, 0x9362, 0x9464, 0x5f95,
  0x6b96, 0x6b96, 0x976c, 0x976c, 0x9864, 0x5f99, 0x6b9a, 0x9b6c,
  0x9c64, 0x5f9d, 0x6b9e, 0x9f6c, 0xa064, 0x69a1, 0x6ba2, 0xa36c,
  0xa46e, 0x69a5, 0x75a6, 0xa776, 0xa86e, 0x69a9, 0x75aa, 0xab76,
  0xac6e, 0x69ad, 0x75ae, 0xaf76, 0xb06e, 0x69b1, 0x75b2, 0xb376,
  0xb46e, 0x73b5, 0x75b6, 0xb776, 0xb878, 0x73b9, 0x7fba, 0xbb80,
  0xbc78, 0x73bd, 0x7fbe, 0xbf80, 0xc078, 0x73c1, 0x7fc2, 0xc380,
  0xc478, 0x73c5, 0x7fc6, 0xc780, 0xc878, 0x73c9, 0x7fca, 0xcb80,
  0xcc78, 0x73cd, 0x7fce, 0xcf80, 0xd078, 0x7dd1, 0x7fd2, 0xd380,
  0xd482, 0x7dd5, 0x89d6, 0xd78a, 0xd882, 0x7dd9, 0x89da, 0xdb8a,
  0xdc82, 0x7ddd, 0x89de, 0xdf8a, 0xe082, 0x7de1, 0x89e2, 0xe38a,
  0xe482, 0x7de5, 0x89e6, 0xe78a, 0xe882, 0x7de9, 0x89ea, 0xeb8a,
  0xec82, 0x7ded, 0x89ee, 0xef8a, 0xf082, 0x7df1, 0x89f2, 0xf38a,
  0xf482, 0x87f5, 0x89f6, 0xf78a, 0xf88c, 0x87f9, 0x45fa, 0xfb50,
  0xfc8c, 0x87f9, 0x45fa, 0xfb50, 0xfc8c, 0x0000, 0x0000, 0x0000,
};

// The 33 knots of squash() (cucm_compress_v2.cu:682-685); interpolated in
// 128-wide segments, the stream depends on every value.
static const int kSquashKnots[33] = {1,    2,    3,    6,    10,   16,   27,
                                     45,   73,   120,  194,  310,  488,  747,
                                     1101, 1546, 2047, 2549, 2994, 3348, 3607,
                                     3785, 3901, 3975, 4022, 4050, 4068, 4079,
                                     4085, 4089, 4092, 4093, 4094};

// Lookup tables the kernels stage into LDS once per workgroup.
struct DevTables {
  int16_t stretch[4096];  // ln(p/(1-p)) domain, inverse of squash
  uint16_t squash[4096];  // 12-bit probability of stretch value d-2047, pr bump included
  uint16_t dt[1024];      // 16384 / (2n + 3): StateMap adaptation rate
  uint8_t nex[512];       // nex[2s + y]
};

// One chunk of work; src and dst are device addresses.  Encoder: src = the
// chunk's bytes, dst = its 4-byte aligned staging region.  Decoder: src =
// the chunk's payload, dst = where its L decoded bytes go.
struct Task {
  uint64_t src;
  uint64_t dst;
  uint32_t L;      // valid bytes of this chunk
  uint32_t S;      // nominal chunk size (bits past 8L up to 8S are padding)
  uint32_t aux;    // encoder: staging capacity; decoder: payload length
  uint16_t U;      // unit bits
  uint8_t tb;      // log2 slots per order table
  uint8_t flags;   // kTaskTailPad | kTaskInitModel | kTaskFrozen
  uint32_t model;  // index of the initial model (kTaskInitModel)
  uint32_t job;    // serving kernel: the task's job slot (else 0)
};
constexpr uint8_t kTaskTailPad = 1;
constexpr uint8_t kTaskInitModel = 2;
constexpr uint8_t kTaskFrozen = 4;  // frozen model: run by frozen_kernel, skipped by the others
constexpr uint8_t kTaskBigUnit = 8; // unit above kMaxUnitBits: run by bigunit_kernel only
constexpr uint8_t kTaskDecode = 16; // serving kernel: a decoder task (else an encoder task)

// Bytes of unit records one wave of the large-unit codec keeps in global
// memory for units of up to Umax bits: 8 slots (u32) and pr << 1 | y (u16)
// per unit bit.
constexpr uint64_t big_record_bytes(uint32_t Umax) {
  return ((uint64_t)Umax * (8 * 4 + 2) + 255) & ~255ull;
}

// A trained initial model (prelearn) in device memory.  Order-0 states are
// given by h0 (1..255) since order 0 has exactly 255 distinct slots.
struct DevInitModel {
  const uint8_t *states;  // [7][1 << tb] for orders 1..7
  uint32_t statep[kOrders][256];
  int32_t w[kMixCtx * kOrders];
  uint8_t st0[256];
  uint32_t tb;
  uint32_t pad;
};

// Result word per task: payload bytes (flush included) in the low 63 bits,
// bit 63 set if the staging region overflowed.
constexpr uint64_t kResultOverflow = 1ull << 63;

// Phases of the per-phase cycle counters (cucm_ctx_phase_cycles).
constexpr int kPhases = 12;

struct CodecArgs {
  const Task *tasks;
  uint32_t ntasks;
  uint32_t spec_bits;      // decoder: bits decided per speculative round
  uint32_t *counter;       // work-queue head
  uint8_t *pool;           // model slots, one per resident wave
  uint64_t slot_bytes;     // bytes per model slot (7 << max tb)
  uint64_t *result;        // per task
  uint64_t *prof;          // optional per-phase cycle totals (kPhases), else null
  const DevTables *tabs;
  const DevInitModel *init;
  uint8_t *big;            // large-unit codec: per-wave unit records
  uint64_t big_bytes;      // bytes of records per wave
  uint32_t big_umax;       // unit bits the records are laid out for
  uint32_t pad;
};

// Container assembly for one job (file) of an encode launch.
struct AsmJob {
  uint64_t out_base;      // device address of the container
  uint64_t header_off;    // header bytes in the header arena
  uint32_t header_len;    // 21 + prelearn segment (0 for a bare batch blob)
  uint32_t first_task;    // tasks [first_task, first_task + ntasks)
  uint32_t ntasks;
  uint32_t batch_chunks;  // chunks per batch record
  uint32_t blob_only;     // 1: a single blob without its LE length word
  uint32_t pad;
};

// One container of a decode launch.
struct ParseJob {
  uint64_t src;   // container device address
  uint64_t len;   // container bytes
  uint64_t dst;   // device address of the decoded body (after the prelearn bytes)
  uint64_t body;  // N - P
  uint32_t first_task, ntasks;
  uint32_t seg;   // 21 + prelearn segment bytes
  uint32_t S, U, tb, batch, flags, model, pad;
};

// A job of the serving kernel, in host memory shared with the device,
// written by the host before the job's first task is published.
struct ServeJob {
  uint64_t out_base;    // host memory: the container (compress; header already there)
                        // or the decoded bytes (decompress; prelearn bytes already there)
  uint32_t first;       // ring index of the job's first task
  uint32_t ntasks;
  uint32_t header_len;  // compress: header + prelearn segment bytes at out_base;
                        // decompress: prelearn bytes (chunk c's bytes at + c * S)
  uint32_t batch;       // compress: chunks per batch record
  uint32_t id;          // written to the job's done word when the job is complete
  uint32_t pad;
};

// Arguments of the serving kernel: persistent waves take tasks from a ring
// in host memory as the host publishes them (tail), until the host sets the
// stop word (the ring is drained first) or, as a safeguard, no task came for
// idle_ticks of the 100 MHz real-time clock.  Idle waves watch a mirror of
// the tail and stop words in device memory; one wave at a time refreshes it
// from host memory.
struct ServeArgs {
  CodecArgs a;                 // tasks = the ring, result = per-ring-slot words,
                               // counter = the ring's claim head (device memory)
  const uint32_t *tail;        // host memory: tasks published so far (wrapping)
  const uint32_t *stop;        // host memory: nonzero ends the kernel
  const ServeJob *jobs;        // host memory: [job_mask + 1]
  uint32_t *jobcount;          // device memory: finished tasks per job slot
  uint32_t *done;              // host memory: per job slot, the id of its last completed job
                               // (bit 31 set: a chunk overflowed its staging region)
  uint64_t *lens;              // host memory: per job slot, the container length (compress)
  uint32_t *mirror;            // device memory: [0] the tail as last read from the host,
                               // [1] the stop word likewise, [2..3] the time of that read
  uint32_t ring_mask, job_mask;
  uint64_t idle_ticks;
  uint64_t poll_ticks;         // the host words are read at most once per poll_ticks
};

}  // namespace cucm

#ifdef CUCM_WITH_LAUNCHERS
#include <hip/hip_runtime_api.h>
namespace cucm {
uint32_t codec_lds_bytes(uint32_t Umax, uint32_t spec_bits, bool decode, uint32_t waves);
// LDS of one team workgroup (one chunk per workgroup of several waves) and
// whether a launch can run as teams.
uint32_t team_codec_lds_bytes(uint32_t Umax, uint32_t spec_bits, bool decode);
bool team_supported(bool decode, uint32_t spec_bits);
hipError_t launch_codec(bool decode, const CodecArgs &a, uint32_t Umax, uint32_t blocks,
                        uint32_t waves_per_block, bool team, hipStream_t stream);
// The frozen-model codec (one wave per chunk; tasks without kTaskFrozen are
// skipped): LDS per workgroup of `waves` waves, and the launcher.
uint32_t frozen_lds_bytes(uint32_t waves);
hipError_t launch_frozen(bool decode, const CodecArgs &a, uint32_t blocks,
                         uint32_t waves_per_block, hipStream_t stream);
// The large-unit codec (one wave per chunk of a task with kTaskBigUnit; the
// others are skipped): LDS per workgroup of `waves` waves, and the launcher.
uint32_t big_lds_bytes(uint32_t waves);
hipError_t launch_big(bool decode, const CodecArgs &a, uint32_t blocks, uint32_t waves_per_block,
                      hipStream_t stream);
// The serving kernel (one wave per chunk task, encoder and decoder tasks of
// units up to 128 bits in one launch, 2-bit speculative decoding): LDS per
// workgroup of `waves` waves, and the launcher.
uint32_t serve_lds_bytes(uint32_t waves);
hipError_t launch_serve(const ServeArgs &s, uint32_t blocks, uint32_t waves_per_block,
                        hipStream_t stream);
hipError_t launch_read_headers(const uint64_t *src, const uint64_t *len, uint32_t n, uint8_t *hdr,
                               hipStream_t stream);
hipError_t launch_build_decode_tasks(const ParseJob *jobs, uint32_t n, Task *tasks,
                                     int32_t *status, hipStream_t stream);
hipError_t launch_assemble(const AsmJob *jobs, uint32_t njobs, const uint8_t *headers,
                           const Task *tasks, uint32_t ntasks, const uint64_t *result,
                           uint64_t *task_dst, uint64_t *job_len, hipStream_t stream);
}  // namespace cucm
#endif
#endif
