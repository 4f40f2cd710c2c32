// cucm_wave.h -- the per-wave device code of the CuCM codec (model
// primitives, unit update, encoder and decoder of one chunk), shared by the
// gfx950 kernels (cucm_kernels.hip) and the CPU wave emulator of the tests.
//
// Execution model: a persistent grid of waves.  Each wave (64 lanes) owns
// one model slot in HBM and pulls chunk tasks from a device work queue
// (atomic counter), so thousands of uneven chunks or files run in a single
// launch.  A chunk is processed one unit (U bits, the reference's frozen-model
// "vector") at a time:
//
//   predict  all U bits x 8 orders in parallel across lanes: FNV context
//            hash, one-byte state gather from the order's table in HBM
//            (order 0 lives in LDS, indexed by the partial byte), statep and
//            stretch lookups in LDS, 8-input integer mixer, squash.
//   code     a uniform scalar loop runs the carry-less 32-bit binary
//            arithmetic coder.
//   update   mixer deltas by LDS integer atomics (order independent, exact);
//            per-order StateMap recurrences walked in bit order by 8 lanes,
//            with in-unit slot aliasing resolved through a per-order LDS
//            open-addressing table of the unit's distinct slots; final states
//            of the distinct slots written back to HBM once per unit.
//
// All arithmetic is the reference's fixed-point integer arithmetic
// (cucm_compress_v2.cu:558-705, 711-868; cucm_decompress_v2.cu:698-896),
// so encoder and decoder mirror each other bit for bit.
//
// Cross-lane primitives go through the small wrappers below; with
// CUCM_WAVE_EMULATION defined, the including file supplies CPU versions of
// them instead (64 lanes run as fibers that meet at every cross-lane
// operation), so the same source can be checked on the host.
#ifndef CUCM_WAVE_H
#define CUCM_WAVE_H
#include "cucm_model.h"

#ifndef CUCM_WAVE_EMULATION
#include <hip/hip_runtime.h>

// Pointers into LDS carry the LDS address space so every access is a ds_* op.
#define LDS __attribute__((address_space(3)))
#define GLB __attribute__((address_space(1)))
#define CUCM_DEV __device__ __forceinline__
#define CUCM_HD __host__ __device__ inline
#define CUCM_HDI __host__ __device__ __forceinline__

namespace cucm {
CUCM_DEV uint32_t lane_read(uint32_t v, uint32_t lane) { return __builtin_amdgcn_readlane(v, lane); }
CUCM_DEV uint32_t lane_first(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
CUCM_DEV uint64_t wave_ballot(bool p) { return __ballot(p); }
CUCM_DEV uint32_t clz32(uint32_t x) { return (uint32_t)__clz((int)x); }  // 32 for 0
CUCM_DEV uint64_t shader_clock() { return clock64(); }
CUCM_DEV void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0f70); }    // vmcnt(0)
// Keeps the instruction scheduler from moving code across this point.
CUCM_DEV void sched_fence() { __builtin_amdgcn_sched_barrier(0); }
CUCM_DEV void wait_lgkm() { __builtin_amdgcn_s_waitcnt(0xc07f); }  // lgkmcnt(0)
// One decoded bit on the scalar unit (all operands wave-uniform): y = x <=
// xmid, then x2 = xmid if y, else x1 = xmid + 1; returns y as an integer.
// Written as a compare, the flag outlives the refill branch and is moved
// back to the scalar unit through a vector register on every bit.
CUCM_DEV uint32_t coder_split(uint32_t x, uint32_t xmid, uint32_t &x1, uint32_t &x2) {
  const uint32_t up = xmid + 1u;
  uint32_t y;
  asm volatile(
      "s_cmp_le_u32 %3, %4\n\t"
      "s_cselect_b32 %0, 1, 0\n\t"
      "s_cselect_b32 %2, %4, %2\n\t"
      "s_cselect_b32 %1, %1, %5"
      : "=&s"(y), "+s"(x1), "+s"(x2)
      : "s"(x), "s"(xmid), "s"(up)
      : "scc");
  return y;
}
// The encoder's interval update on the scalar unit: the bit is bit 0 of prm
// (pr << 20 | y); x2 = xmid if it is set, else x1 = xmid + 1.
CUCM_DEV void coder_encode_split(uint32_t prm, uint32_t xmid, uint32_t &x1, uint32_t &x2) {
  const uint32_t up = xmid + 1u;
  asm volatile(
      "s_bitcmp1_b32 %2, 0\n\t"
      "s_cselect_b32 %1, %3, %1\n\t"
      "s_cselect_b32 %0, %0, %4"
      : "+s"(x1), "+s"(x2)
      : "s"(prm), "s"(xmid), "s"(up)
      : "scc");
}
// coder_split that advances a speculative path index instead of returning
// the bit: q = 2q + 1 + y (the path of the next bit of the round).
CUCM_DEV void coder_split_path(uint32_t x, uint32_t xmid, uint32_t &x1, uint32_t &x2,
                               uint32_t &q) {
  const uint32_t up = xmid + 1u;
  uint32_t t;
  asm volatile(
      "s_cmp_le_u32 %4, %5\n\t"
      "s_cselect_b32 %0, 2, 1\n\t"
      "s_cselect_b32 %2, %5, %2\n\t"
      "s_cselect_b32 %1, %1, %6\n\t"
      "s_lshl1_add_u32 %3, %3, %0"
      : "=&s"(t), "+s"(x1), "+s"(x2), "+s"(q)
      : "s"(x), "s"(xmid), "s"(up)
      : "scc");
}
// Orders the wave's LDS traffic across lanes (and stops the compiler from
// moving LDS accesses across it); lanes of a wave run in lock step.
CUCM_DEV void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Marks a step of a chain whose cross-lane LDS order the hardware already
// gives (a wave's LDS operations execute in program order, and a lane's own
// store-then-load of a possibly equal address keeps its order): no wait is
// emitted; the CPU wave emulation keeps its lanes in step here.
CUCM_DEV void lane_step() { __builtin_amdgcn_wave_barrier(); }
// Waits for this wave's outstanding vector memory operations: a wave's state
// write-backs must land before its own next gathers.  No cache maintenance
// is needed for a wave's own accesses.
CUCM_DEV void vm_drain() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  wait_vm();
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// Workgroup barrier of a multi-wave team: LDS and global writes before it
// (global stores completed) are seen by every wave after it.
CUCM_DEV void team_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  wait_vm();
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// Workgroup barrier that orders LDS only: the waves' outstanding global loads
// (a leader's gathers in flight) stay in flight across it.
CUCM_DEV void team_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// An LDS word published by one wave and polled by another: the writes made
// before the store are seen by a wave whose load returns the stored value.
CUCM_DEV void flag_publish(LDS uint32_t *p, uint32_t v) {
  __hip_atomic_store((uint32_t *)p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
CUCM_DEV uint32_t flag_read(const LDS uint32_t *p) {
  return __hip_atomic_load((uint32_t *)p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// flag_read split in two: the load (its latency can overlap other work) and,
// once the value is used, the acquire fence that orders the later LDS reads
// after it.
CUCM_DEV uint32_t flag_peek(const LDS uint32_t *p) {
  return __hip_atomic_load((uint32_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
CUCM_DEV void flag_acquire() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local"); }
// Low 32 bits of a * b for operands that fit 24-bit signed integers (a
// full-rate multiply; the 32-bit product is quarter rate).
CUCM_DEV uint32_t mul_i24(int32_t a, int32_t b) { return (uint32_t)__mul24(a, b); }
CUCM_DEV uint32_t lds_cas(LDS uint32_t *p, uint32_t cmp, uint32_t v) {
  return atomicCAS((uint32_t *)p, cmp, v);
}
CUCM_DEV void lds_add(LDS int32_t *p, int32_t v) { atomicAdd((int32_t *)p, v); }
CUCM_DEV void glb_add64(uint64_t *p, uint64_t v) {
  atomicAdd((unsigned long long *)p, (unsigned long long)v);
}
}  // namespace cucm
#endif

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

namespace cucm {

constexpr uint32_t kEmpty = 0xffffffffu;
constexpr uint16_t kNoState = 0x100;

// Entries per order of a unit's distinct-slot table: teams use a power of
// two of at least 2U; one wave per chunk uses 1.5U (96, 192 or 384), at
// most two thirds full, so its LDS holds a seventh wave per CU at U = 128.
CUCM_HD uint32_t table_size(uint32_t U) {
  uint32_t t = 64;
  while (t < 2 * U) t <<= 1;
  return t;
}
CUCM_HD uint32_t wave_table_size(uint32_t U) { return U * 3u / 2u; }
// Home entry of a table position among ts entries (ts <= 2^16): the high
// half of a multiplicative hash, scaled to ts; probes move to the next entry.
CUCM_HDI uint32_t table_home(uint32_t pos, uint32_t ts) {
  return (((pos * 0x9E3779B1u) >> 16) * ts) >> 16;
}
CUCM_HDI uint32_t table_next(uint32_t h, uint32_t ts) { return h + 1u == ts ? 0u : h + 1u; }

CUCM_HD uint32_t align16(uint32_t x) { return (x + 15u) & ~15u; }

// Per-wave LDS carve-out.  Laid out for the launch's largest unit (Umax)
// and, for the decoder, the speculation width (P paths).
struct WaveMem {
  LDS uint32_t *stp;    // [8][256] statep
  LDS int32_t *w;       // [80][8] mixer weights
  LDS uint32_t *tkey;   // [7][TS] distinct slots of the unit, orders 1..7
  LDS uint32_t *pslot;  // [2][P][8] decoder: slot (order 0: h0) per path, two rounds
  LDS uint16_t *tval;   // teams: [7][TS] running state of each distinct slot (kNoState: untouched)
  LDS uint8_t *tv8;     // one wave per chunk: [7][TS] running state of each distinct slot,
                        // set to the unit-start state by every bit that records the slot
  LDS uint16_t *ridx;   // [8][Umax] table index (order 0: h0) per unit bit (16-bit)
  LDS uint8_t *rix8;    // the same as bytes (one wave per chunk, Umax <= 128)
  LDS uint16_t *rprm;   // [Umax] pr << 1 | y per unit bit (persists: tail padding)
  LDS uint8_t *st0;     // [256] order-0 states by h0
  LDS uint8_t *rg;      // [8][Umax] unit-start state per unit bit
  LDS uint8_t *rcxt;    // [Umax] mixer context per unit bit
  LDS uint8_t *pg;      // [2][P][8] decoder: unit-start state per path, two rounds
  LDS uint8_t *ub;      // [8 + Umax/8] 8 history bytes + the unit's bytes
  LDS uint32_t *dummy;  // [64] scratch target of predicated-off LDS operations
  uint32_t ts;          // entries per order of tkey / tval / tv8
};

// Byte offsets of the per-wave LDS arrays (16-byte aligned each); the
// decoder keeps nbuf rounds of path slots and states.  One wave per chunk
// keeps the distinct slots' running states as bytes (tv8) in a table of
// wave_table_size entries per order and, up to 128-bit units, its per-bit
// table indices as bytes (rix8); teams (wide) keep 16-bit states with an
// untouched marker (tval), power-of-two tables and 16-bit indices.
struct WaveLayout {
  uint32_t stp, w, tkey, pslot, tval, ridx, rprm, st0, rg, rcxt, pg, ub, dummy, total;
};

CUCM_HDI WaveLayout wave_layout(uint32_t Umax, uint32_t P, bool dec, uint32_t nbuf = 2,
                                bool wide = false) {
  const uint32_t TS = wide ? table_size(Umax) : wave_table_size(Umax);
  const bool idx8 = !wide && Umax <= 128u;
  const uint32_t Pd = dec ? P : 0u;
  WaveLayout l;
  uint32_t off = 0;
#define CUCM_TAKE(field, bytes) \
  l.field = off;                \
  off = align16(off + (bytes));
  CUCM_TAKE(stp, 8 * 256 * 4)
  CUCM_TAKE(w, kMixCtx * kOrders * 4)
  CUCM_TAKE(tkey, 7 * TS * 4)
  CUCM_TAKE(pslot, nbuf * Pd * 8 * 4)
  CUCM_TAKE(tval, 7 * TS * (wide ? 2 : 1))
  CUCM_TAKE(ridx, 8 * Umax * (idx8 ? 1 : 2))
  CUCM_TAKE(rprm, Umax * 2)
  CUCM_TAKE(st0, 256)
  CUCM_TAKE(rg, 8 * Umax)
  CUCM_TAKE(rcxt, Umax)
  CUCM_TAKE(pg, nbuf * Pd * 8)
  CUCM_TAKE(ub, 8 + Umax / 8 + 8)
  CUCM_TAKE(dummy, 64 * 4)
#undef CUCM_TAKE
  l.total = off;
  return l;
}

CUCM_HD uint32_t wave_lds_bytes(uint32_t Umax, uint32_t P, bool dec) {
  return wave_layout(Umax, P, dec).total;
}

CUCM_DEV WaveMem carve(LDS uint8_t *base, uint32_t Umax, uint32_t P, bool dec,
                       uint32_t nbuf = 2, bool wide = false) {
  const WaveLayout l = wave_layout(Umax, P, dec, nbuf, wide);
  WaveMem m;
  m.stp = (LDS uint32_t *)(base + l.stp);
  m.w = (LDS int32_t *)(base + l.w);
  m.tkey = (LDS uint32_t *)(base + l.tkey);
  m.pslot = (LDS uint32_t *)(base + l.pslot);
  m.tval = wide ? (LDS uint16_t *)(base + l.tval) : nullptr;
  m.tv8 = wide ? nullptr : base + l.tval;
  const bool idx8 = !wide && Umax <= 128u;
  m.ridx = idx8 ? nullptr : (LDS uint16_t *)(base + l.ridx);
  m.rix8 = idx8 ? base + l.ridx : nullptr;
  m.ts = wide ? table_size(Umax) : wave_table_size(Umax);
  m.rprm = (LDS uint16_t *)(base + l.rprm);
  m.st0 = base + l.st0;
  m.rg = base + l.rg;
  m.rcxt = base + l.rcxt;
  m.pg = base + l.pg;
  m.ub = base + l.ub;
  m.dummy = (LDS uint32_t *)(base + l.dummy);
  return m;
}

// Per-bit table index of the one-wave kernels: bytes up to 128-bit units
// (NJ <= 2), 16-bit words above.
template <int NJ>
CUCM_DEV uint32_t rix_get(const WaveMem &M, uint32_t i) {
  if constexpr (NJ <= 2)
    return M.rix8[i];
  else
    return M.ridx[i];
}
template <int NJ>
CUCM_DEV void rix_set(const WaveMem &M, uint32_t i, uint32_t v) {
  if constexpr (NJ <= 2)
    M.rix8[i] = (uint8_t)v;
  else
    M.ridx[i] = (uint16_t)v;
}

// Per-phase shader-clock totals of one wave (enabled when args.prof != null):
//   0 model init   1 unit staging   2 hash + gather   3 slot tables + mix
//   4 arithmetic coder   5 mixer deltas + collision scan   6 state resolution
//   7 statep recurrences   8 write-back   9 decoder path predictions
//   10 decoder mix + decode   11 decoder records
struct PhaseClock {
  bool on;
  uint64_t last;
  uint64_t acc[kPhases];
  CUCM_DEV explicit PhaseClock(bool enable) : on(enable), last(0) {
#pragma unroll
    for (int i = 0; i < kPhases; ++i) acc[i] = 0;
    if (on) last = shader_clock();
  }
  CUCM_DEV void mark(int ph) {
    if (on) {
      const uint64_t t = shader_clock();
      acc[ph] += t - last;
      last = t;
    }
  }
  CUCM_DEV void flush(uint64_t *out, uint32_t lane) {
    if (on && lane == 0)
#pragma unroll
      for (int i = 0; i < kPhases; ++i) glb_add64(&out[i], acc[i]);
  }
};

// ---------------------------------------------------------------------------
// Model primitives (device)
// ---------------------------------------------------------------------------

// floor(sum_k w[k] * smp[k] / 2^16), the mixer's dot product
// (cucm_compress_v2.cu:636-640), for stretch outputs |smp[k]| <= 2047.  Each
// weight splits into a signed high and an unsigned low 16-bit half, so every
// product is a 24-bit multiply and both partial sums fit in 32 bits
// (|A| <= 8 * 2^15 * 2047, |B| <= 8 * (2^16 - 1) * 2047 < 2^31); then
// (A * 2^16 + B) >> 16 = A + (B >> 16) with arithmetic shifts, exactly.
CUCM_DEV int32_t mix_dot(const int32_t w[8], const int32_t smp[8]) {
  int32_t A = 0, B = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    A += (w[k] >> 16) * smp[k];
    B += (int32_t)((uint32_t)w[k] & 0xffffu) * smp[k];
  }
  return A + (B >> 16);
}

// squash of the clamped dot product plus the pr bump (cucm_compress_v2.cu:
// 641-644, 787-788; the table holds the bumped values).
CUCM_DEV uint32_t mix_squash(int32_t dot, const LDS uint16_t *squash) {
  const int idx = dot < -2047 ? -2047 : (dot > 2048 ? 2048 : dot);
  return squash[idx + 2047];
}

// Mixer::p (cucm_compress_v2.cu:636-644) plus the pr bump (:787-788).
CUCM_DEV uint32_t mix_predict(const LDS int32_t *wrow, const int32_t smp[8],
                                                const LDS uint16_t *squash) {
  int32_t w[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) w[k] = wrow[k];
  return mix_squash(mix_dot(w, smp), squash);
}

// Mixer::p with the weight row already in registers.
CUCM_DEV uint32_t mix_weights(const int32_t w[8], const int32_t smp[8],
                                                const LDS uint16_t *squash) {
  return mix_squash(mix_dot(w, smp), squash);
}

// Carry-less coder interval split (cucm_compress_v2.cu:855-857):
// x1 + (rng >> 12) * pr + (((rng & 0xfff) * pr) >> 12) with rng = x2 - x1.
// Writing rng = 4096a + b, that sum is a*pr + floor(b*pr / 4096) =
// floor(rng * pr / 4096).  With pr given as pr20 = pr << 20 (pr < 2^12) that
// is the high word of rng * pr20: one scalar multiply.
CUCM_DEV uint32_t coder_mid20(uint32_t x1, uint32_t x2, uint32_t pr20) {
  return x1 + (uint32_t)(((uint64_t)(x2 - x1) * pr20) >> 32);
}
// A unit bit's pr << 1 | y (as the encoder records it) as pr << 20 | y.
CUCM_DEV uint32_t prm20(uint32_t prm) { return ((prm >> 1) << 20) | (prm & 1u); }

// Zeroes (or copies the prelearned model into) the wave's model slot and
// resets the LDS-resident parts of the model.
CUCM_DEV void init_model(const CodecArgs &a, const Task &tk, const WaveMem &M, GLB uint8_t *slot,
                           uint32_t Umax, uint32_t lane) {
  const uint64_t bytes = 7ull << tk.tb;
  const uint64_t n16 = bytes >> 4;
  GLB u32x4 *dst = (GLB u32x4 *)slot;
  if (tk.flags & kTaskInitModel) {
    const DevInitModel *im = a.init + tk.model;
    const GLB u32x4 *src = (const GLB u32x4 *)im->states;
    for (uint64_t i = lane; i < n16; i += 64) dst[i] = src[i];
    for (uint32_t i = lane; i < 8 * 256; i += 64) M.stp[i] = im->statep[i >> 8][i & 255];
    for (uint32_t i = lane; i < kMixCtx * kOrders; i += 64) M.w[i] = im->w[i];
    for (uint32_t i = lane; i < 256; i += 64) M.st0[i] = im->st0[i];
  } else {
    const u32x4 z = {0u, 0u, 0u, 0u};
    for (uint64_t i = lane; i < n16; i += 64) dst[i] = z;
    for (uint32_t i = lane; i < 8 * 256; i += 64) M.stp[i] = 1u << 31;
    for (uint32_t i = lane; i < kMixCtx * kOrders; i += 64) M.w[i] = 0;
    for (uint32_t i = lane; i < 256; i += 64) M.st0[i] = 0;
  }
  for (uint32_t i = lane; i < 7 * M.ts; i += 64) M.tkey[i] = kEmpty;
  for (uint32_t i = lane; i < Umax; i += 64) M.rprm[i] = 0;
  for (uint32_t i = lane; i < 8 + Umax / 8 + 8; i += 64) M.ub[i] = 0;
  vm_drain();
  wave_sync();
}

// First part of the deferred update of one unit (cucm_compress_v2.cu:794-839,
// cucm_decompress_v2.cu:870-896), for unit bits r < nvalid (a multiple of 8):
//   1. mixer deltas from the recorded (pr, y, cxt) and unit-start predictions
//      (LDS integer atomics: the sum is order independent);
//   2. state resolution: the state each StateMap update sees, s = state of
//      the slot after all earlier bits of the unit.  Bits of one byte never
//      share an order-0 slot (distinct partial bytes), so a byte's 8 bits x
//      8 orders go in one step over all 64 lanes; a byte where two bits of an
//      order hash to one slot (a hash collision) is stepped bit by bit;
//   3. the final state of every touched slot back to HBM, and the tables
//      cleared for the next unit.
// ym receives the unit's bits y as uniform masks for the statep walk (the
// second part, the statep walk), which may run while other work proceeds.
template <int NJ>
CUCM_DEV void update_states(const WaveMem &M, const LDS DevTables *T, GLB uint8_t *slot,
                            uint32_t tb, uint32_t Umax, uint32_t nvalid, uint32_t lane,
                            PhaseClock &clk, uint64_t ym[NJ]) {
  const uint32_t TS = M.ts;
  const uint32_t nbytes = nvalid >> 3;
  // y of every unit bit as a uniform bit mask
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const uint32_t r = lane + 64u * j;
    const bool y = (r < nvalid) && (M.rprm[r < Umax ? r : 0] & 1u);
    ym[j] = wave_ballot(y);
  }
  // 1. mixer deltas
  for (uint32_t r = lane; r < nvalid; r += 64) {
    const uint32_t prm = M.rprm[r];
    const int32_t err = ((int32_t)((prm & 1u) << 12) - (int32_t)(prm >> 1)) * 7;
    const uint32_t cxt = M.rcxt[r];
    int32_t d[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t g = M.rg[k * Umax + r];
      const int32_t smp = T->stretch[M.stp[k * 256 + g] >> 20];
      d[k] = (int32_t)(((int64_t)err * smp) >> 16);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) lds_add(&M.w[cxt * 8 + k], d[k]);
  }
  // bytes with an in-byte slot collision in some order >= 1
  uint32_t cmask = 0;
  for (uint32_t base = 0; base < 7 * nbytes; base += 64) {
    const uint32_t it = base + lane;
    bool coll = false;
    if (it < 7 * nbytes) {
      const uint32_t b = it / 7, k = 1 + it % 7;
      uint32_t e[8];
      if constexpr (NJ <= 2) {
        const uint64_t v = *(const LDS uint64_t *)(M.rix8 + k * Umax + 8 * b);
#pragma unroll
        for (int x = 0; x < 8; ++x) e[x] = (uint32_t)(v >> (8 * x)) & 255u;
      } else {
        const u32x4 v = *(const LDS u32x4 *)(M.ridx + k * Umax + 8 * b);
        e[0] = v.x & 0xffffu, e[1] = v.x >> 16, e[2] = v.y & 0xffffu, e[3] = v.y >> 16;
        e[4] = v.z & 0xffffu, e[5] = v.z >> 16, e[6] = v.w & 0xffffu, e[7] = v.w >> 16;
      }
#pragma unroll
      for (int x = 0; x < 8; ++x)
#pragma unroll
        for (int z = x + 1; z < 8; ++z) coll |= (e[x] == e[z]);
    }
    uint64_t bal = wave_ballot(coll);
    while (bal) {
      const uint32_t bit = (uint32_t)__builtin_ctzll(bal);
      bal &= bal - 1;
      cmask |= 1u << ((base + bit) / 7);
    }
  }
  wave_sync();
  clk.mark(5);
  // 2. state resolution, one byte of the unit per step
  {
    const uint32_t k = lane >> 3, j = lane & 7u;
    LDS uint8_t *tv = M.tv8 + (k ? (k - 1) * TS : 0);
#pragma unroll
    for (int seg = 0; seg < NJ; ++seg) {
      const uint64_t yseg = ym[seg];
      const uint32_t bend = min(nbytes, 8u * seg + 8u);
      for (uint32_t b = 8u * seg; b < bend; ++b) {
        const uint32_t r = 8 * b + j;
        const uint32_t y = (uint32_t)(yseg >> (r & 63)) & 1u;
        const uint32_t i = rix_get<NJ>(M, k * Umax + r);
        const uint32_t reps = ((cmask >> b) & 1u) ? 8u : 1u;
        for (uint32_t q = 0; q < reps; ++q) {
          if (reps == 1u || q == j) {
            // (tv[i] starts the unit as the slot's gathered state)
            const uint32_t st = k == 0 ? M.st0[i] : tv[i];
            const uint32_t ns = T->nex[2 * st + y];
            if (k == 0)
              M.st0[i] = (uint8_t)ns;
            else
              tv[i] = (uint8_t)ns;
            M.rg[k * Umax + r] = (uint8_t)st;
          }
          wave_sync();
        }
      }
    }
  }
  clk.mark(6);
  // 3. write back the touched slots (every occurrence writes the same final
  // state), then clear the tables
#pragma unroll 1
  for (uint32_t k = 1; k < 8; ++k) {
    GLB uint8_t *tab = slot + ((uint64_t)(k - 1) << tb);
    for (uint32_t r = lane; r < nvalid; r += 64) {
      const uint32_t i = rix_get<NJ>(M, k * Umax + r);
      tab[M.tkey[(k - 1) * TS + i]] = M.tv8[(k - 1) * TS + i];
    }
  }
  wave_sync();
#pragma unroll 1
  for (uint32_t k = 1; k < 8; ++k)
    for (uint32_t r = lane; r < nvalid; r += 64) {
      const uint32_t i = rix_get<NJ>(M, k * Umax + r);
      M.tkey[(k - 1) * TS + i] = kEmpty;
    }
  vm_drain();
  wave_sync();
  clk.mark(8);
}

// Second part of the deferred update: the per-order statep recurrences
// (cucm_compress_v2.cu:588-601 applied in bit order).  Lane k < 8 walks order
// k's updates; lanes 8..63 run the same instructions on order 7's row and
// write a dummy word, so a walk step has no divergent branch and can share a
// loop with other work.  Software-pipelined so that the dependent chain per
// bit is the update arithmetic alone: the statep entry of bit i+2 is read
// before bit i's write (the two latest writes are forwarded instead), and
// the adaptation rate dt[n] of every candidate is looked up one bit before it
// is needed -- the count n in an entry's low 10 bits advances to
// min(n + 1, 1023) on every update, whatever the bit.  Steps go in blocks of
// 8 bits (unit bit counts are multiples of 8).
// The walk's state lives in plain locals of the function running it (a
// struct or a member function would be optimised as memory before inlining,
// and a select between its fields would then keep it out of registers).
#define CUCM_WALK_STATE                                                            \
  LDS uint32_t *w_sp = nullptr;                                                    \
  LDS uint32_t *w_dummy = nullptr;                                                 \
  const LDS uint16_t *w_dtab = nullptr;                                            \
  const LDS uint8_t *w_sv = nullptr;                                               \
  bool w_writer = false, w_f1 = true, w_f2 = false;                                \
  uint32_t w_nblk = 0, w_blk = 0;                                                  \
  uint64_t w_cur = 0, w_nxt = 0;                                                   \
  uint32_t w_s0 = 0, w_s1 = 0, w_sprev = 0x100u, w_p = 0, w_raw1 = 0, w_pnprev = 0, \
           w_dfprev = 0, w_dA = 0, w_dB = 0, w_dC = 0

#define CUCM_WALK_INIT(M, T, Umax, nvalid, lane)                             \
  do {                                                                       \
    const uint32_t wk_ = min((lane), 7u);                                    \
    w_sp = (M).stp + wk_ * 256;                                              \
    w_dummy = (M).dummy + (lane);                                            \
    w_dtab = (T)->dt;                                                        \
    w_sv = (M).rg + wk_ * (Umax);                                            \
    w_writer = (lane) < 8;                                                   \
    w_nblk = (nvalid) >> 3;                                                  \
    w_blk = 0;                                                               \
    w_cur = *(const LDS uint64_t *)w_sv;                                     \
    w_nxt = w_nblk > 1 ? *(const LDS uint64_t *)(w_sv + 8) : w_cur;          \
    w_s0 = (uint32_t)w_cur & 255u;                                           \
    w_s1 = (uint32_t)(w_cur >> 8) & 255u;                                    \
    w_sprev = 0x100u;                                                        \
    w_p = w_sp[w_s0];                                                        \
    w_raw1 = w_sp[w_s1]; /* read before any write of this unit */           \
    w_dA = w_dtab[w_p & 1023u];                                              \
    w_dB = w_dC = 0;                                                         \
    w_f1 = true;                                                             \
    w_f2 = false;                                                            \
    w_pnprev = w_dfprev = 0;                                                 \
  } while (0)

// Bit t (a constant, 0..7) of the current block, with its y.
#define CUCM_WALK_STEP(t, y)                                                               \
  do {                                                                                     \
    const uint32_t wd_ = w_f1 ? w_dA : (w_f2 ? w_dB : w_dC);                               \
    const int wt2_ = (t) + 2;                                                              \
    const uint64_t ww_ = wt2_ < 8 ? w_cur : w_nxt;                                         \
    const uint32_t ws2_ = (uint32_t)(ww_ >> (8 * (wt2_ & 7))) & 255u;                      \
    const uint32_t wraw2_ = w_sp[ws2_]; /* misses the writes of bits i, i+1 */            \
    const uint32_t wn_ = w_p & 1023u;                                                      \
    const uint32_t wnn_ = min(wn_ + 1u, 1023u);                                            \
    const uint32_t wdfwd_ = w_dtab[wnn_];                                                  \
    const uint32_t wdraw_ = w_dtab[w_raw1 & 1023u];                                        \
    const int32_t we_ = (int32_t)(((uint32_t)(y) << 22) - (w_p >> 10));                   \
    /* |we_ >> 3| <= 2^19 and wd_ <= 5461: a 24-bit product */                          \
    const uint32_t wpn_ = w_p + (mul_i24(we_ >> 3, (int32_t)wd_) & 0xfffffc00u) + (wnn_ - wn_); \
    *(w_writer ? &w_sp[w_s0] : w_dummy) = wpn_;                                            \
    const bool wsame1_ = w_s1 == w_s0, wsame2_ = w_s1 == w_sprev;                          \
    w_p = wsame1_ ? wpn_ : (wsame2_ ? w_pnprev : w_raw1);                                  \
    w_dA = wdfwd_;                                                                         \
    w_dB = w_dfprev;                                                                       \
    w_dC = wdraw_;                                                                         \
    w_f1 = wsame1_;                                                                        \
    w_f2 = wsame2_;                                                                        \
    w_sprev = w_s0;                                                                        \
    w_s0 = w_s1;                                                                           \
    w_s1 = ws2_;                                                                           \
    w_pnprev = wpn_;                                                                       \
    w_dfprev = wdfwd_;                                                                     \
    w_raw1 = wraw2_;                                                                       \
    sched_fence(); /* each step's loads stay ahead of the next step's uses */              \
  } while (0)

#define CUCM_WALK_NEXT_BLOCK()                                                            \
  do {                                                                                    \
    ++w_blk;                                                                              \
    w_cur = w_nxt;                                                                        \
    w_nxt = (w_blk + 1 < w_nblk) ? *(const LDS uint64_t *)(w_sv + 8 * (w_blk + 1)) : w_cur; \
  } while (0)

// y of unit bit r from the masks update_states returns.
template <int NJ>
CUCM_DEV uint32_t unit_y(const uint64_t ym[NJ], uint32_t r) {
  uint64_t w = ym[0];
#pragma unroll
  for (int q = 1; q < NJ; ++q)
    if ((r >> 6) == (uint32_t)q) w = ym[q];
  return (uint32_t)(w >> (r & 63u)) & 1u;
}

// The whole statep walk of a unit on its own.
template <int NJ>
CUCM_DEV void statep_walk(const WaveMem &M, const LDS DevTables *T, uint32_t Umax, uint32_t nvalid,
                          uint32_t lane, const uint64_t ym[NJ]) {
  CUCM_WALK_STATE;
  CUCM_WALK_INIT(M, T, Umax, nvalid, lane);
  for (uint32_t blk = 0; blk < nvalid / 8; ++blk) {
    uint64_t w = ym[0];
#pragma unroll
    for (int q = 1; q < NJ; ++q)
      if ((blk >> 3) == (uint32_t)q) w = ym[q];
    const uint32_t yb = (uint32_t)(w >> (8 * (blk & 7u))) & 255u;
#pragma unroll
    for (int t = 0; t < 8; ++t) CUCM_WALK_STEP(t, (yb >> t) & 1u);
    CUCM_WALK_NEXT_BLOCK();
  }
  wave_sync();
}

// ---------------------------------------------------------------------------
// Encoder
// ---------------------------------------------------------------------------

// Stages unit bytes, hashes every bit's 8 slots, issues the 7 state gathers
// per bit and places every slot in the unit's distinct-slot tables (the
// inserts run while the gathers are in flight).  Lane handles unit bits
// r = r0 + lane + 64j; bits past nvalid compute in-bounds garbage and insert
// nothing.
template <int NJ>
CUCM_DEV void encode_issue(const WaveMem &M, const GLB uint8_t *src, uint64_t L,
                                             uint64_t base, uint32_t U, uint32_t nvalid,
                                             const GLB uint8_t *slot, uint32_t tb, uint32_t mask,
                                             uint32_t Umax, uint32_t lane, uint32_t sl[NJ][8],
                                             uint32_t g[NJ][8], uint32_t idx[NJ][8],
                                             uint32_t r0 = 0) {
  const uint32_t TS = M.ts;
  // 8 history bytes and the unit's bytes (zero outside the chunk)
  const int64_t b0 = (int64_t)(base >> 3) - 8;
  for (uint32_t i = lane; i < 8 + U / 8; i += 64) {
    const int64_t p = b0 + i;
    M.ub[i] = (p >= 0 && p < (int64_t)L) ? src[p] : (uint8_t)0;
  }
  wave_sync();
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const uint32_t r = min(r0 + lane + 64u * j, Umax - 1);
    const uint32_t bi = r >> 3, jb = r & 7u;
    const uint32_t cur = M.ub[8 + bi];
    const uint32_t h0 = (1u << jb) | (jb ? (cur >> (8 - jb)) : 0u);
    uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
    sl[j][0] = h0;
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      sl[j][k] = fnv_pos(H, (uint32_t)M.ub[8 + bi - k], mask);
    }
    g[j][0] = M.st0[h0];
#pragma unroll
    for (int k = 1; k < 8; ++k) g[j][k] = slot[((uint32_t)(k - 1) << tb) | sl[j][k]];
  }
  // distinct-slot tables: the first probes of every key go out together;
  // lanes past the valid bits probe a dummy word; keys that meet another
  // key probe on in rounds in which every unplaced key moves together
  uint32_t old[NJ][8];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const bool ok = r0 + lane + 64u * j < nvalid;
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      idx[j][k] = table_home(sl[j][k], TS);
      LDS uint32_t *at = ok ? &M.tkey[(k - 1) * TS + idx[j][k]] : &M.dummy[lane];
      old[j][k] = lds_cas(at, kEmpty, sl[j][k]);
    }
  }
  uint32_t pend = 0;
#pragma unroll
  for (int j = 0; j < NJ; ++j)
#pragma unroll
    for (int k = 1; k < 8; ++k)
      if ((r0 + lane + 64u * j < nvalid) && old[j][k] != kEmpty && old[j][k] != sl[j][k])
        pend |= 1u << (j * 8 + k);
  while (wave_ballot(pend != 0)) {
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const bool pb = (pend >> (j * 8 + k)) & 1u;
        idx[j][k] = pb ? table_next(idx[j][k], TS) : idx[j][k];
        LDS uint32_t *at = pb ? &M.tkey[(k - 1) * TS + idx[j][k]] : &M.dummy[lane];
        old[j][k] = lds_cas(at, kEmpty, sl[j][k]);
      }
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int k = 1; k < 8; ++k)
        if (old[j][k] == kEmpty || old[j][k] == sl[j][k]) pend &= ~(1u << (j * 8 + k));
  }
}

// Completes the unit's predictions once its gathers are back: statep,
// stretch, mixer context, mix (cucm_compress_v2.cu:711-792), and records
// (pr, y), cxt, slots and unit-start states per bit for the coder and the
// update.
template <int NJ>
CUCM_DEV void encode_predict(const WaveMem &M, const LDS DevTables *T,
                                               uint32_t nvalid, uint32_t Umax, uint32_t lane,
                                               const uint32_t sl[NJ][8], const uint32_t g[NJ][8],
                                               const uint32_t idx[NJ][8]) {
  const uint32_t TS = M.ts;
  uint32_t pv[NJ][8];
#pragma unroll
  for (int j = 0; j < NJ; ++j)
#pragma unroll
    for (int k = 0; k < 8; ++k) pv[j][k] = M.stp[k * 256 + g[j][k]];
  int32_t smp[NJ][8];
#pragma unroll
  for (int j = 0; j < NJ; ++j)
#pragma unroll
    for (int k = 0; k < 8; ++k) smp[j][k] = T->stretch[pv[j][k] >> 20];
  uint32_t cxt[NJ];
  int32_t wv[NJ][8];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const uint32_t r = min(lane + 64u * j, Umax - 1);
    uint32_t active = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) active += (smp[j][k] != 1);
    cxt[j] = active + (uint32_t)(M.ub[8 + (r >> 3) - 1] >> 5) * 10u;
#pragma unroll
    for (int k = 0; k < 8; ++k) wv[j][k] = M.w[cxt[j] * 8 + k];
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const uint32_t r = lane + 64u * j;
    const bool ok = r < nvalid;
    const uint32_t rr = min(r, Umax - 1);
    const uint32_t pr = mix_weights(wv[j], smp[j], T->squash);
    const uint32_t y = ((uint32_t)M.ub[8 + (rr >> 3)] >> (7 - (rr & 7u))) & 1u;
    // bits past the data store nothing (their rprm keeps the stale value
    // the tail padding codes)
    if (ok) {
      M.rprm[rr] = (uint16_t)((pr << 1) | y);
      M.rcxt[rr] = (uint8_t)cxt[j];
      rix_set<NJ>(M, rr, sl[j][0]);
      M.rg[rr] = (uint8_t)g[j][0];
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        rix_set<NJ>(M, k * Umax + rr, idx[j][k]);
        M.rg[k * Umax + rr] = (uint8_t)g[j][k];
        // every bit of a slot gathered the same unit-start state
        M.tv8[(k - 1) * TS + idx[j][k]] = (uint8_t)g[j][k];
      }
    }
  }
  wave_sync();
}

// One step of the carry-less binary arithmetic coder (cucm_compress_v2.cu:
// 841-868) for the packed prm = pr << 20 | y (prm20), in the uniform
// registers x1, x2; the rare byte output goes through acc (4 bytes per store)
// to dst.
#define CUCM_CODE_BIT(prm_expr)                                           \
  do {                                                                    \
    const uint32_t prm_ = (prm_expr);                                     \
    const uint32_t xmid_ = coder_mid20(x1, x2, prm_ & 0xfff00000u);       \
    coder_encode_split(prm_, xmid_, x1, x2);                              \
    if (__builtin_expect(((x1 ^ x2) & 0xff000000u) == 0, 0)) {            \
      /* (rare) shift out the leading bytes x1 and x2 agree on */         \
      do {                                                                \
        acc |= (x2 >> 24) << (8 * (opos & 3u));                           \
        ++opos;                                                           \
        if ((opos & 3u) == 0) {                                           \
          if (opos <= cap) {                                              \
            if (lane == 0) *(GLB uint32_t *)(dst + opos - 4) = acc;       \
          } else {                                                        \
            ovf = 1u;                                                     \
          }                                                               \
          acc = 0;                                                        \
        }                                                                 \
        x1 <<= 8;                                                         \
        x2 = (x2 << 8) | 255u;                                            \
      } while (((x1 ^ x2) & 0xff000000u) == 0);                           \
    }                                                                     \
  } while (0)

// One chunk.  Per unit: stage, hash and gather, predictions, the first part
// of the deferred update, then the arithmetic coder and the statep walk.
template <int NJ, bool kProf>
CUCM_DEV void encode_task(const CodecArgs &a, const Task &tk, uint32_t t,
                                            const WaveMem &M, const LDS DevTables *T,
                                            GLB uint8_t *slot, uint32_t Umax, uint32_t lane) {
  const uint32_t U = tk.U, tb = tk.tb, mask = (1u << tb) - 1u;
  const uint64_t L = tk.L;
  const GLB uint8_t *src = (const GLB uint8_t *)(uintptr_t)tk.src;
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  const uint64_t cap = tk.aux;
  PhaseClock clk(kProf && a.prof != nullptr);  // (clocks only when profiling)
  init_model(a, tk, M, slot, Umax, lane);
  clk.mark(0);
  const uint64_t vbits = 8 * L;
  const uint64_t coded_end = (tk.flags & kTaskTailPad) ? 8ull * tk.S : vbits;
  uint32_t x1 = 0, x2 = 0xffffffffu, acc = 0;
  uint32_t opos = 0, ovf = 0;
  for (uint64_t base = 0; base < coded_end; base += U) {
    const uint32_t nvalid = base < vbits ? (uint32_t)min((uint64_t)U, vbits - base) : 0u;
    if (nvalid) {
      uint32_t sl[NJ][8], g[NJ][8], idx[NJ][8];
      encode_issue<NJ>(M, src, L, base, U, nvalid, slot, tb, mask, Umax, lane, sl, g, idx);
      clk.mark(2);
      encode_predict<NJ>(M, T, nvalid, Umax, lane, sl, g, idx);
      clk.mark(3);
    }
    // The deferred update's first part, then the arithmetic coder and the
    // statep walk in one loop: the coder (a uniform scalar chain; (pr, y) of
    // unit bit r is read from lane r & 63 of the register holding bits
    // 64j..64j+63) and the walk (a vector chain) fill each other's stalls.
    uint64_t ym[NJ] = {};
    CUCM_WALK_STATE;
    if (nvalid) {
      update_states<NJ>(M, T, slot, tb, Umax, nvalid, lane, clk, ym);
      CUCM_WALK_INIT(M, T, Umax, nvalid, lane);
    }
    uint32_t pv[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) pv[j] = prm20((lane + 64u * j < Umax) ? M.rprm[lane + 64u * j] : 0u);
    const uint32_t nc = (uint32_t)min((uint64_t)U, coded_end - base);
    const uint32_t nwalk = nvalid >> 3;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const uint32_t lo = 64u * j;
      const uint32_t hi = min(nc, lo + 64u);
      const uint32_t hw = min(hi, max(nwalk * 8u, lo));  // bits with a walk step
      for (uint32_t r0 = lo; r0 < hw; r0 += 8) {
        const uint32_t yb = (uint32_t)(ym[j] >> (r0 - lo)) & 255u;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          CUCM_WALK_STEP(t, (yb >> t) & 1u);
          CUCM_CODE_BIT(lane_read(pv[j], r0 - lo + t));
        }
        CUCM_WALK_NEXT_BLOCK();
      }
      for (uint32_t r = hw; r < hi; ++r) CUCM_CODE_BIT(lane_read(pv[j], r - lo));
    }
    if (nvalid) wave_sync();  // statep written by the walk is read across lanes next
    clk.mark(4);
  }
  acc |= (x1 >> 24) << (8 * (opos & 3u));
  ++opos;
  const uint64_t wstart = (uint64_t)((opos - 1) & ~3u);
  if (wstart + 4 <= cap) {
    if (lane == 0) *(GLB uint32_t *)(dst + wstart) = acc;
  } else {
    ovf = 1u;
  }
  if (lane == 0) a.result[t] = (uint64_t)opos | (ovf ? kResultOverflow : 0ull);
  clk.flush(a.prof, lane);
  wave_sync();
}

// ---------------------------------------------------------------------------
// Decoder
// ---------------------------------------------------------------------------

// Window of the payload in registers: lane l holds bytes 4l..4l+3 after
// winbase as a big-endian word (0xff past the payload, as the reference
// decoder reads), so the coder's refill is a scalar funnel shift.
CUCM_DEV uint32_t load_window(const GLB uint8_t *pay, uint64_t plen, uint64_t wb,
                                                uint32_t lane) {
  if (plen == 0) return 0xffffffffu;
  // all four loads go out unconditionally (clamped into the payload), then
  // the bytes past its end are replaced
  uint32_t b[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint64_t at = wb + 4ull * lane + i;
    b[i] = pay[at < plen ? at : plen - 1];
  }
  uint32_t w = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint64_t at = wb + 4ull * lane + i;
    w |= (at < plen ? b[i] : 255u) << (24 - 8 * i);
  }
  return w;
}

// The 4 payload bytes at window offset off, as a big-endian word.
CUCM_DEV uint32_t window_be32(uint32_t win, uint32_t off) {
  const uint32_t w0 = lane_read(win, off >> 2);
  const uint32_t w1 = lane_read(win, ((off >> 2) + 1u) & 63u);
  return (uint32_t)(((((uint64_t)w0 << 32) | w1) << (8 * (off & 3u))) >> 32);
}

// Speculative paths of one decoder round of D bits: lane q + 64m evaluates
// path q, which fixes the next plen(q) = floor(log2(q + 1)) bits of the
// current byte to pbits(q) = q + 1 - 2^plen(q) (cucm_decompress_v2.cu:698-770).
// The whole FNV chain of a path is hashed once in its lane; the 7 state
// gathers are issued here and consumed by paths_finish, so the caller can do
// other work while they are in flight.  Slots go to pslot[buf] for recording.
template <int NM, uint32_t P>
CUCM_DEV void paths_issue(const WaveMem &M, const GLB uint8_t *slot, uint32_t tb,
                                            uint32_t mask, uint32_t j0, uint32_t partial,
                                            uint64_t hist, uint32_t buf, uint32_t lane,
                                            uint32_t g[NM][8]) {
  LDS uint32_t *ps = M.pslot + buf * (P * 8u);
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const uint32_t q = lane + 64u * m;
    const bool ok = q < P;
    // lanes without a path repeat path 0, so their gathers fall into path
    // 0's lines (one line per order and path per gather instruction)
    const uint32_t qa = ok ? q : 0u;
    const uint32_t plen = 31u - clz32(qa + 1u);
    const uint32_t pbits = qa + 1u - (1u << plen);
    const uint32_t h0 = ((1u << (j0 + plen)) | (partial << plen) | pbits) & 255u;
    uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
    uint32_t sl[8];
    sl[0] = h0;
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      sl[k] = fnv_pos(H, (uint32_t)((hist >> (8 * (k - 1))) & 255u), mask);
    }
    g[m][0] = M.st0[h0];
    // unconditional loads (a predicated load would be waited for at the
    // branch join).  The offset (k - 1) * 2^tb + position < 7 * 2^25 fits
    // 32 bits.
#pragma unroll
    for (int k = 1; k < 8; ++k) g[m][k] = slot[((uint32_t)(k - 1) << tb) | sl[k]];
    if (ok) {
      LDS u32x4 *d = (LDS u32x4 *)(ps + q * 8u);
      d[0] = u32x4{sl[0], sl[1], sl[2], sl[3]};
      d[1] = u32x4{sl[4], sl[5], sl[6], sl[7]};
    }
  }
}

// Completes the paths of a round: statep, stretch, mixer context and mix
// per path, all lane-local (cucm_decompress_v2.cu:774-809).  pc[m] receives
// pr | cxt << 16; the unit-start states go to pg[buf] for recording.
template <int NM, uint32_t P>
CUCM_DEV void paths_finish(const WaveMem &M, const LDS DevTables *T, uint32_t c1,
                                             uint32_t buf, uint32_t lane, const uint32_t g[NM][8],
                                             uint32_t pc[NM]) {
  LDS uint8_t *pgb = M.pg + buf * (P * 8u);
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const uint32_t q = lane + 64u * m;
    uint32_t pv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) pv[k] = M.stp[k * 256 + g[m][k]];
    int32_t smp[8];
    uint32_t active = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      smp[k] = T->stretch[pv[k] >> 20];
      active += (smp[k] != 1);
    }
    const uint32_t cxt = active + c1 * 10u;
    const LDS u32x4 *wr = (const LDS u32x4 *)(M.w + cxt * 8u);
    const u32x4 wa = wr[0], wb = wr[1];
    const int32_t w[8] = {(int32_t)wa.x, (int32_t)wa.y, (int32_t)wa.z, (int32_t)wa.w,
                          (int32_t)wb.x, (int32_t)wb.y, (int32_t)wb.z, (int32_t)wb.w};
    pc[m] = mix_weights(w, smp, T->squash) | (cxt << 16);
    if (q < P) {
      const uint32_t lo = g[m][0] | (g[m][1] << 8) | (g[m][2] << 16) | (g[m][3] << 24);
      const uint32_t hi = g[m][4] | (g[m][5] << 8) | (g[m][6] << 16) | (g[m][7] << 24);
      *(LDS uint64_t *)(pgb + q * 8u) = ((uint64_t)hi << 32) | lo;
    }
  }
}

// Path q's packed pr | cxt << 16 (uniform q).
template <int NM>
CUCM_DEV uint32_t path_value(const uint32_t pc[NM], uint32_t q) {
  if (NM == 1) return lane_read(pc[0], q);
  uint32_t v = lane_read(pc[0], q & 63u);
#pragma unroll
  for (int m = 1; m < NM; ++m) {
    const uint32_t vm = lane_read(pc[m], q & 63u);
    v = (q >> 6) == (uint32_t)m ? vm : v;
  }
  return v;
}

// One chunk.  Each round decides D bits: the paths of the round are mixed,
// the bits decoded by a scalar coder loop, then the next round's gathers are
// issued before the current round's bits are recorded for the unit update.
template <int NJ, int D, bool kProf>
CUCM_DEV void decode_task(const CodecArgs &a, const Task &tk, uint32_t t,
                                            const WaveMem &M, const LDS DevTables *T,
                                            GLB uint8_t *slot, uint32_t Umax, uint32_t lane) {
  constexpr uint32_t P = (1u << D) - 1u;
  constexpr int NM = (int)((P + 63u) / 64u);
  const uint32_t U = tk.U, tb = tk.tb, mask = (1u << tb) - 1u;
  const uint32_t TS = M.ts;
  const uint64_t L = tk.L;
  const GLB uint8_t *pay = (const GLB uint8_t *)(uintptr_t)tk.src;
  const uint64_t plen = tk.aux;
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  PhaseClock clk(kProf && a.prof != nullptr);  // (clocks only when profiling)
  init_model(a, tk, M, slot, Umax, lane);
  clk.mark(0);
  uint64_t winbase = 0;
  uint32_t win = load_window(pay, plen, 0, lane);
  uint32_t x = lane_read(win, 0);
  uint32_t x1 = 0, x2 = 0xffffffffu;
  uint64_t ip = 4;
  uint64_t hist = 0;     // the 8 bytes before the current one, latest lowest
  uint32_t partial = 0;  // decided bits of the current byte
  const uint64_t vbits = 8 * L;
  // the first round's gathers of every unit are issued before the previous
  // unit's statep walk (they read only the states the write-back stored)
  uint32_t g[NM][8];
  if (vbits) paths_issue<NM, P>(M, slot, tb, mask, 0, 0, hist, 0, lane, g);
  for (uint64_t base = 0; base < vbits; base += U) {
    // unit bits: a multiple of 8, so every round decides exactly D bits
    const uint32_t nbits = (uint32_t)min((uint64_t)U, vbits - base);
    uint32_t buf = 0;
    for (uint32_t rb = 0; rb < nbits; rb += D) {
      // the coder consumes at most 4 bytes per bit
      if (ip + 4 * D + 8 > winbase + 256) {
        winbase = ip & ~3ull;
        win = load_window(pay, plen, winbase, lane);
      }
      uint32_t pc[NM];
      paths_finish<NM, P>(M, T, (uint32_t)(hist & 255u) >> 5, buf, lane, g, pc);
      wave_sync();  // pg[buf] is read across lanes by the record step
      clk.mark(9);
      // decode the round's D bits (uniform scalar loop)
      uint32_t pre = 0;
      uint64_t chosen = 0, cxp = 0;
      uint64_t prmp[2] = {0, 0};
#pragma unroll
      for (int i = 0; i < D; ++i) {
        const uint32_t q = (1u << i) - 1u + pre;
        const uint32_t v = path_value<NM>(pc, q);
        const uint32_t pr = v & 0xffffu, cx = v >> 16;
        // (v << 20 is pr << 20: pr < 2^12 and cxt sits above bit 15)
        const uint32_t xmid = coder_mid20(x1, x2, v << 20);
        const uint32_t y = coder_split(x, xmid, x1, x2);
        pre = (pre << 1) | y;
        if (__builtin_expect(((x1 ^ x2) & 0xff000000u) == 0, 0)) {
          // shift out every leading byte x1 and x2 agree on (at most 4)
          const uint32_t nb = min((uint32_t)clz32(x1 ^ x2) >> 3, 4u);
          const uint32_t sh = 8 * nb;
          x1 = (uint32_t)((uint64_t)x1 << sh);
          x2 = (uint32_t)(((uint64_t)x2 << sh) | ((1ull << sh) - 1ull));
          const uint32_t nxt = window_be32(win, (uint32_t)(ip - winbase));
          x = (uint32_t)((((uint64_t)x << 32) | nxt) >> (32 - sh));
          ip += nb;
        }
        chosen |= (uint64_t)q << (8 * i);
        prmp[i >> 2] |= (uint64_t)((pr << 1) | y) << (16 * (i & 3));
        cxp |= (uint64_t)cx << (8 * i);
      }
      clk.mark(10);
      const uint32_t j0 = rb & 7u, bi = rb >> 3;
      partial = (partial << D) | pre;
      if (j0 + D == 8) {
        if (lane == 0) M.ub[8 + bi] = (uint8_t)partial;
        hist = (hist << 8) | partial;
        partial = 0;
      }
      // next round's gathers go out before this round is recorded
      const uint32_t cur = buf;
      buf ^= 1u;
      if (rb + D < nbits) paths_issue<NM, P>(M, slot, tb, mask, (rb + D) & 7u, partial, hist, buf, lane, g);
      clk.mark(1);
      // record the taken paths: slot and unit-start state of every order of
      // every decided bit, and (pr, y), cxt per bit
      {
        const LDS uint8_t *pgb = M.pg + cur * (P * 8u);
        const LDS uint32_t *psb = M.pslot + cur * (P * 8u);
        const bool rec = lane < 8u * D;
        const uint32_t i = lane >> 3, k = lane & 7u;
        const uint32_t r = min(rb + i, Umax - 1);
        const uint32_t q = (uint32_t)(chosen >> (8 * i)) & 255u;
        const uint32_t gq = pgb[q * 8 + k];
        const uint32_t sl = psb[q * 8 + k];
        const bool ins = rec && k != 0;
        uint32_t h = table_home(sl, TS);
        LDS uint32_t *tkk = M.tkey + (k ? (k - 1) * TS : 0);
        uint32_t old = lds_cas((ins ? &tkk[h] : &M.dummy[lane]), kEmpty, sl);
        if (wave_ballot(ins && old != kEmpty && old != sl)) {
          if (ins)
            while (old != kEmpty && old != sl) {
              h = table_next(h, TS);
              old = lds_cas(&tkk[h], kEmpty, sl);
            }
        }
        if (rec) {
          M.rg[k * Umax + r] = (uint8_t)gq;
          rix_set<NJ>(M, k * Umax + r, k == 0 ? sl : h);
        }
        // the slot's unit-start state (every round of the unit gathered the same)
        if (ins) M.tv8[(k - 1) * TS + h] = (uint8_t)gq;
        if (lane < (uint32_t)D) {
          const uint64_t pw = lane < 4 ? prmp[0] : prmp[1];
          M.rprm[rb + lane] = (uint16_t)(pw >> (16 * (lane & 3u)));
          M.rcxt[rb + lane] = (uint8_t)(cxp >> (8 * lane));
        }
      }
      wave_sync();
      clk.mark(11);
    }
    uint64_t ym[NJ];
    update_states<NJ>(M, T, slot, tb, Umax, nbits, lane, clk, ym);
    const uint32_t nb = nbits >> 3;
    for (uint32_t i = lane; i < nb; i += 64) dst[(base >> 3) + i] = M.ub[8 + i];
    if (base + U < vbits) paths_issue<NM, P>(M, slot, tb, mask, 0, 0, hist, 0, lane, g);
    statep_walk<NJ>(M, T, Umax, nbits, lane, ym);
    clk.mark(7);
  }
  if (lane == 0) a.result[t] = ip;
  clk.flush(a.prof, lane);
  wave_sync();
}

}  // namespace cucm

#endif
