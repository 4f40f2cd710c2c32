// cucm_kernels.hip -- gfx950 kernels of the CuCM chunked context-mixing codec.
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
//   code     lane 0 runs the carry-less 32-bit binary arithmetic coder.
//   update   mixer deltas by LDS integer atomics (order independent, exact);
//            per-order StateMap recurrences walked in bit order by 8 lanes,
//            with in-unit slot aliasing resolved through a per-order LDS
//            open-addressing table of the unit's distinct slots; final states
//            of the distinct slots written back to HBM once per unit.
//
// All arithmetic is the reference's fixed-point integer arithmetic
// (cucm_compress_v2.cu:558-705, 711-868; cucm_decompress_v2.cu:698-896),
// so encoder and decoder mirror each other bit for bit.
#include <hip/hip_runtime.h>

#define CUCM_WITH_LAUNCHERS
#include "cucm_model.h"

namespace cucm {

constexpr uint32_t kEmpty = 0xffffffffu;
constexpr uint16_t kNoState = 0x100;
constexpr uint32_t kWindow = 256;  // decoder payload window in LDS

// Pointers into LDS carry the LDS address space so every access is a ds_* op.
#define LDS __attribute__((address_space(3)))
#define GLB __attribute__((address_space(1)))
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__host__ __device__ inline uint32_t table_size(uint32_t U) {
  uint32_t t = 64;
  while (t < 2 * U) t <<= 1;
  return t;
}

__host__ __device__ inline uint32_t align16(uint32_t x) { return (x + 15u) & ~15u; }

// Per-wave LDS carve-out.  Laid out for the launch's largest unit (Umax)
// and, for the decoder, the speculation width (P paths).
struct WaveMem {
  LDS uint32_t *stp;    // [8][256] statep
  LDS int32_t *w;       // [80][8] mixer weights
  LDS uint32_t *tkey;   // [7][TS] distinct slots of the unit, orders 1..7
  LDS uint32_t *pslot;  // [P][8] decoder: slot (order 0: h0) per path
  LDS uint16_t *tval;   // [7][TS] running state of each distinct slot
  LDS uint16_t *ridx;   // [8][Umax] table index (order 0: h0) per unit bit
  LDS uint16_t *rprm;   // [Umax] pr << 1 | y per unit bit (persists: tail padding)
  LDS int16_t *psmp;    // [P][8] decoder: stretched prediction per path
  LDS uint16_t *ppr;    // [P] decoder: mixed probability per path
  LDS uint8_t *st0;     // [256] order-0 states by h0
  LDS uint8_t *rg;      // [8][Umax] unit-start state per unit bit
  LDS uint8_t *rcxt;    // [Umax] mixer context per unit bit
  LDS uint8_t *pg;      // [P][8] decoder: unit-start state per path
  LDS uint8_t *pcxt;    // [P] decoder: mixer context per path
  LDS uint8_t *ub;      // [8 + Umax/8] 8 history bytes + the unit's bytes
  LDS uint8_t *win;     // [kWindow] decoder payload window
  LDS uint8_t *chosen;  // [8] decoder: path taken by each bit of a round
  LDS uint32_t *dummy;  // [64] scratch target of predicated-off LDS operations
};

// Byte offsets of the per-wave LDS arrays (16-byte aligned each).
struct WaveLayout {
  uint32_t stp, w, tkey, pslot, tval, ridx, rprm, psmp, ppr, st0, rg, rcxt, pg, pcxt, ub, win,
      chosen, dummy, total;
};

__host__ __device__ __forceinline__ WaveLayout wave_layout(uint32_t Umax, uint32_t P, bool dec) {
  const uint32_t TS = table_size(Umax);
  const uint32_t Pd = dec ? P : 0u;
  WaveLayout l;
  uint32_t off = 0;
#define CUCM_TAKE(field, bytes) \
  l.field = off;                \
  off = align16(off + (bytes));
  CUCM_TAKE(stp, 8 * 256 * 4)
  CUCM_TAKE(w, kMixCtx * kOrders * 4)
  CUCM_TAKE(tkey, 7 * TS * 4)
  CUCM_TAKE(pslot, Pd * 8 * 4)
  CUCM_TAKE(tval, 7 * TS * 2)
  CUCM_TAKE(ridx, 8 * Umax * 2)
  CUCM_TAKE(rprm, Umax * 2)
  CUCM_TAKE(psmp, Pd * 8 * 2)
  CUCM_TAKE(ppr, Pd * 2)
  CUCM_TAKE(st0, 256)
  CUCM_TAKE(rg, 8 * Umax)
  CUCM_TAKE(rcxt, Umax)
  CUCM_TAKE(pg, Pd * 8)
  CUCM_TAKE(pcxt, Pd)
  CUCM_TAKE(ub, 8 + Umax / 8 + 8)
  CUCM_TAKE(win, dec ? kWindow : 0u)
  CUCM_TAKE(chosen, dec ? 8u : 0u)
  CUCM_TAKE(dummy, 64 * 4)
#undef CUCM_TAKE
  l.total = off;
  return l;
}

__host__ __device__ inline uint32_t wave_lds_bytes(uint32_t Umax, uint32_t P, bool dec) {
  return wave_layout(Umax, P, dec).total;
}

__device__ __forceinline__ WaveMem carve(LDS uint8_t *base, uint32_t Umax, uint32_t P, bool dec) {
  const WaveLayout l = wave_layout(Umax, P, dec);
  WaveMem m;
  m.stp = (LDS uint32_t *)(base + l.stp);
  m.w = (LDS int32_t *)(base + l.w);
  m.tkey = (LDS uint32_t *)(base + l.tkey);
  m.pslot = (LDS uint32_t *)(base + l.pslot);
  m.tval = (LDS uint16_t *)(base + l.tval);
  m.ridx = (LDS uint16_t *)(base + l.ridx);
  m.rprm = (LDS uint16_t *)(base + l.rprm);
  m.psmp = (LDS int16_t *)(base + l.psmp);
  m.ppr = (LDS uint16_t *)(base + l.ppr);
  m.st0 = base + l.st0;
  m.rg = base + l.rg;
  m.rcxt = base + l.rcxt;
  m.pg = base + l.pg;
  m.pcxt = base + l.pcxt;
  m.ub = base + l.ub;
  m.win = base + l.win;
  m.chosen = base + l.chosen;
  m.dummy = (LDS uint32_t *)(base + l.dummy);
  return m;
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
  __device__ __forceinline__ explicit PhaseClock(bool enable) : on(enable), last(0) {
#pragma unroll
    for (int i = 0; i < kPhases; ++i) acc[i] = 0;
    if (on) last = clock64();
  }
  __device__ __forceinline__ void mark(int ph) {
    if (on) {
      const uint64_t t = clock64();
      acc[ph] += t - last;
      last = t;
    }
  }
  __device__ __forceinline__ void flush(uint64_t *out, uint32_t lane) {
    if (on && lane == 0)
#pragma unroll
      for (int i = 0; i < kPhases; ++i) atomicAdd((unsigned long long *)&out[i], (unsigned long long)acc[i]);
  }
};

// ---------------------------------------------------------------------------
// Model primitives (device)
// ---------------------------------------------------------------------------

// Waits for this wave's outstanding vector memory operations (s_waitcnt
// vmcnt(0)): a wave's state write-backs must land before its own next
// gathers.  No cache maintenance is needed for a wave's own accesses.
__device__ __forceinline__ void vm_drain() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0), expcnt/lgkmcnt untouched
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Mixer::p (cucm_compress_v2.cu:636-644) plus the pr bump (:787-788).
__device__ __forceinline__ uint32_t mix_predict(const LDS int32_t *wrow, const int32_t smp[8],
                                                const LDS uint16_t *squash) {
  int64_t dot = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) dot += (int64_t)wrow[k] * (int64_t)smp[k];
  dot >>= 16;
  const int idx = dot < -2047 ? -2047 : (dot > 2048 ? 2048 : (int)dot);
  uint32_t pr = squash[idx + 2047];
  if (pr < 2048) ++pr;
  return pr;
}

// Mixer::p with the weight row already in registers.
__device__ __forceinline__ uint32_t mix_weights(const int32_t w[8], const int32_t smp[8],
                                                const LDS uint16_t *squash) {
  int64_t dot = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) dot += (int64_t)w[k] * (int64_t)smp[k];
  dot >>= 16;
  const int idx = dot < -2047 ? -2047 : (dot > 2048 ? 2048 : (int)dot);
  uint32_t pr = squash[idx + 2047];
  if (pr < 2048) ++pr;
  return pr;
}

// Carry-less coder interval split (cucm_compress_v2.cu:855-857).
__device__ __forceinline__ uint32_t coder_mid(uint32_t x1, uint32_t x2, uint32_t pr) {
  const uint32_t rng = x2 - x1;
  return x1 + (rng >> 12) * pr + (((rng & 0xfffu) * pr) >> 12);
}

// StateMap adaptation rate dt[n] = 16384 / (2n + 3) (cucm_compress_v2.cu:702-704):
// a reciprocal estimate corrected by one remainder step, exact for n < 1024.
__device__ __forceinline__ uint32_t adapt_rate(uint32_t n) {
  const uint32_t m = 2 * n + 3;
  uint32_t q = (uint32_t)(16384.0f * __builtin_amdgcn_rcpf((float)m));
  const int32_t rem = 16384 - (int32_t)(q * m);
  q += (rem >= (int32_t)m) ? 1u : 0u;
  q -= (rem < 0) ? 1u : 0u;
  return q;
}

// Distinct-slot tables (one open-addressing table of TS keys per order
// 1..7): places every pending (j, k) key and leaves its bucket in idx.
// Returning LDS atomics are slow on this hardware, so insertion runs in
// rounds of plain reads and writes: read the bucket; my key -> done; empty
// -> write my key and read it back (the last writer wins, losers probe on);
// another key -> probe the next bucket.  Lanes holding the same key see the
// same values in every round, so they move together and a key is placed in
// exactly one bucket.  Each round fills every contested empty bucket, and a
// table has at least twice as many buckets as keys, so the loop ends.
template <int NJ>
__device__ __forceinline__ void place_slots(LDS uint32_t *tkey, uint32_t TS,
                                            const uint32_t sl[NJ][8], uint32_t idx[NJ][8],
                                            uint32_t pend) {
  while (__ballot(pend != 0)) {
    uint32_t v[NJ][8];
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int k = 1; k < 8; ++k)
        if (pend & (1u << (j * 8 + k))) v[j][k] = tkey[(k - 1) * TS + idx[j][k]];
    uint32_t cand = 0;
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const uint32_t bit = 1u << (j * 8 + k);
        if (pend & bit) {
          if (v[j][k] == sl[j][k]) {
            pend &= ~bit;
          } else if (v[j][k] == kEmpty) {
            tkey[(k - 1) * TS + idx[j][k]] = sl[j][k];
            cand |= bit;
          } else {
            idx[j][k] = (idx[j][k] + 1) & (TS - 1);
          }
        }
      }
    wave_sync();
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int k = 1; k < 8; ++k)
        if (cand & (1u << (j * 8 + k))) v[j][k] = tkey[(k - 1) * TS + idx[j][k]];
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const uint32_t bit = 1u << (j * 8 + k);
        if (cand & bit) {
          if (v[j][k] == sl[j][k])
            pend &= ~bit;
          else
            idx[j][k] = (idx[j][k] + 1) & (TS - 1);
        }
      }
    wave_sync();
  }
}

// Single-key form of place_slots for lanes holding one key of order k (k >= 1);
// lanes with active == false take part in the rounds without a key.
__device__ __forceinline__ uint32_t place_slot(LDS uint32_t *tk, uint32_t TS, uint32_t key,
                                               bool active) {
  uint32_t h = key & (TS - 1);
  bool pend = active;
  while (__ballot(pend)) {
    uint32_t v = 0;
    if (pend) v = tk[h];
    bool cand = false;
    if (pend) {
      if (v == key) {
        pend = false;
      } else if (v == kEmpty) {
        tk[h] = key;
        cand = true;
      } else {
        h = (h + 1) & (TS - 1);
      }
    }
    wave_sync();
    if (cand) {
      if (tk[h] == key)
        pend = false;
      else
        h = (h + 1) & (TS - 1);
    }
    wave_sync();
  }
  return h;
}

// Zeroes (or copies the prelearned model into) the wave's model slot and
// resets the LDS-resident parts of the model.
__device__ __forceinline__ void init_model(const CodecArgs &a, const Task &tk, const WaveMem &M, GLB uint8_t *slot,
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
  const uint32_t TS = table_size(Umax);
  for (uint32_t i = lane; i < 7 * TS; i += 64) {
    M.tkey[i] = kEmpty;
    M.tval[i] = kNoState;
  }
  for (uint32_t i = lane; i < Umax; i += 64) M.rprm[i] = 0;
  for (uint32_t i = lane; i < 8 + Umax / 8 + 8; i += 64) M.ub[i] = 0;
  vm_drain();
  wave_sync();
}

// Applies the deferred updates of one unit (cucm_compress_v2.cu:794-839,
// cucm_decompress_v2.cu:870-896), for unit bits r < nvalid (a multiple of 8):
//   1. mixer deltas from the recorded (pr, y, cxt) and unit-start predictions
//      (LDS integer atomics: the sum is order independent);
//   2. state resolution: the state each StateMap update sees, s = state of
//      the slot after all earlier bits of the unit.  Bits of one byte never
//      share an order-0 slot (distinct partial bytes), so a byte's 8 bits x
//      8 orders go in one step over all 64 lanes; a byte where two bits of an
//      order hash to one slot (a hash collision) is stepped bit by bit;
//   3. per order, the statep recurrence in bit order (8 lanes), with the next
//      bit's statep read issued before the current write and forwarded when
//      both hit the same state;
//   4. the final state of every touched slot back to HBM.
template <int NJ>
__device__ __forceinline__ void update_unit(const WaveMem &M, const LDS DevTables *T, GLB uint8_t *slot, uint32_t tb,
                            uint32_t Umax, uint32_t nvalid, uint32_t lane, PhaseClock &clk) {
  const uint32_t TS = table_size(Umax);
  const uint32_t nbytes = nvalid >> 3;
  // y of every unit bit as a uniform bit mask
  uint64_t ym[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const uint32_t r = lane + 64u * j;
    const bool y = (r < nvalid) && (M.rprm[r < Umax ? r : 0] & 1u);
    ym[j] = __ballot(y);
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
    for (int k = 0; k < 8; ++k) atomicAdd((int32_t *)&M.w[cxt * 8 + k], d[k]);
  }
  // bytes with an in-byte slot collision in some order >= 1
  uint32_t cmask = 0;
  for (uint32_t base = 0; base < 7 * nbytes; base += 64) {
    const uint32_t it = base + lane;
    bool coll = false;
    if (it < 7 * nbytes) {
      const uint32_t b = it / 7, k = 1 + it % 7;
      const u32x4 v = *(const LDS u32x4 *)(M.ridx + k * Umax + 8 * b);
      const uint32_t e[8] = {v.x & 0xffffu, v.x >> 16, v.y & 0xffffu, v.y >> 16,
                             v.z & 0xffffu, v.z >> 16, v.w & 0xffffu, v.w >> 16};
#pragma unroll
      for (int x = 0; x < 8; ++x)
#pragma unroll
        for (int z = x + 1; z < 8; ++z) coll |= (e[x] == e[z]);
    }
    uint64_t bal = __ballot(coll);
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
    LDS uint16_t *tv = M.tval + (k ? (k - 1) * TS : 0);
#pragma unroll
    for (int seg = 0; seg < NJ; ++seg) {
      const uint64_t yseg = ym[seg];
      const uint32_t bend = min(nbytes, 8u * seg + 8u);
      for (uint32_t b = 8u * seg; b < bend; ++b) {
        const uint32_t r = 8 * b + j;
        const uint32_t y = (uint32_t)(yseg >> (r & 63)) & 1u;
        const uint32_t i = M.ridx[k * Umax + r];
        const uint32_t g = M.rg[k * Umax + r];
        const uint32_t reps = ((cmask >> b) & 1u) ? 8u : 1u;
        for (uint32_t q = 0; q < reps; ++q) {
          if (reps == 1u || q == j) {
            uint32_t st;
            if (k == 0) {
              st = M.st0[i];
            } else {
              const uint32_t v = tv[i];
              st = v > 255u ? g : v;
            }
            const uint32_t ns = T->nex[2 * st + y];
            if (k == 0)
              M.st0[i] = (uint8_t)ns;
            else
              tv[i] = (uint16_t)ns;
            M.rg[k * Umax + r] = (uint8_t)st;
          }
          wave_sync();
        }
      }
    }
  }
  clk.mark(6);
  // 3. statep recurrences
  if (lane < 8) {
    LDS uint32_t *sp = M.stp + lane * 256;
    const LDS u32x4 *S4 = (const LDS u32x4 *)(M.rg + lane * Umax);
    const uint32_t nblk = (nvalid + 15) >> 4;
    u32x4 cur = S4[0];
    uint32_t s = cur.x & 255u;
    uint32_t p = sp[s];
    for (uint32_t blk = 0; blk < nblk; ++blk) {
      const u32x4 nxt = (blk + 1 < nblk) ? S4[blk + 1] : cur;
      const uint32_t seg = blk >> 2;
      uint64_t yw = ym[0];
#pragma unroll
      for (int q = 1; q < NJ; ++q)
        if (seg == (uint32_t)q) yw = ym[q];
      yw >>= 16 * (blk & 3u);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const uint32_t r = blk * 16 + i;
        if (r < nvalid) {
          const uint32_t word = (i + 1) < 16 ? cur[(i + 1) >> 2] : nxt.x;
          uint32_t sn = (word >> (8 * ((i + 1) & 3))) & 255u;
          if (r + 1 >= nvalid) sn = s;
          const uint32_t pn_raw = sp[sn];
          const uint32_t y = (uint32_t)(yw >> i) & 1u;
          const uint32_t n = p & 1023u;
          const int32_t e = (int32_t)((y << 22) - (p >> 10));
          const uint32_t adj = ((uint32_t)(e >> 3) * adapt_rate(n)) & 0xfffffc00u;
          const uint32_t pn = p + adj + (n < 1023u ? 1u : 0u);
          sp[s] = pn;
          p = (sn == s) ? pn : pn_raw;
          s = sn;
        }
      }
      cur = nxt;
    }
  }
  wave_sync();
  clk.mark(7);
  // 4. write back the touched slots (every occurrence writes the same final
  // state), then clear the tables
#pragma unroll 1
  for (uint32_t k = 1; k < 8; ++k) {
    GLB uint8_t *tab = slot + ((uint64_t)(k - 1) << tb);
    for (uint32_t r = lane; r < nvalid; r += 64) {
      const uint32_t i = M.ridx[k * Umax + r];
      tab[M.tkey[(k - 1) * TS + i]] = (uint8_t)M.tval[(k - 1) * TS + i];
    }
  }
  wave_sync();
#pragma unroll 1
  for (uint32_t k = 1; k < 8; ++k)
    for (uint32_t r = lane; r < nvalid; r += 64) {
      const uint32_t i = M.ridx[k * Umax + r];
      M.tkey[(k - 1) * TS + i] = kEmpty;
      M.tval[(k - 1) * TS + i] = kNoState;
    }
  vm_drain();
  wave_sync();
  clk.mark(8);
}

// ---------------------------------------------------------------------------
// Encoder
// ---------------------------------------------------------------------------

template <int NJ>
__device__ __forceinline__ void encode_task(const CodecArgs &a, const Task &tk, uint32_t t, const WaveMem &M,
                            const LDS DevTables *T, GLB uint8_t *slot, uint32_t Umax, uint32_t lane) {
  const uint32_t U = tk.U, tb = tk.tb, mask = (1u << tb) - 1u;
  const uint32_t TS = table_size(Umax);
  const uint64_t L = tk.L;
  const GLB uint8_t *src = (const GLB uint8_t *)(uintptr_t)tk.src;
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  const uint64_t cap = tk.aux;
  PhaseClock clk(a.prof != nullptr);
  init_model(a, tk, M, slot, Umax, lane);
  clk.mark(0);
  const uint64_t vbits = 8 * L;
  const uint64_t coded_end = (tk.flags & kTaskTailPad) ? 8ull * tk.S : vbits;
  uint32_t x1 = 0, x2 = 0xffffffffu, acc = 0;
  uint64_t opos = 0;
  bool ovf = false;
  for (uint64_t base = 0; base < coded_end; base += U) {
    const uint32_t nvalid = base < vbits ? (uint32_t)min((uint64_t)U, vbits - base) : 0u;
    if (nvalid) {
      // 8 history bytes and the unit's bytes (zero outside the chunk)
      const int64_t b0 = (int64_t)(base >> 3) - 8;
      for (uint32_t i = lane; i < 8 + U / 8; i += 64) {
        const int64_t p = b0 + i;
        M.ub[i] = (p >= 0 && p < (int64_t)L) ? src[p] : (uint8_t)0;
      }
      wave_sync();
      clk.mark(1);
      // predict: lane handles unit bits r = lane + 64j; each layer of
      // independent memory operations is issued before any result is used
      uint32_t sl[NJ][8], g[NJ][8];
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        // every lane computes (bits past nvalid read in-bounds unit bytes);
        // only the stores below are predicated
        const uint32_t r = min(lane + 64u * j, Umax - 1);
        const uint32_t bi = r >> 3, jb = r & 7u;
        const uint32_t cur = M.ub[8 + bi];
        const uint32_t h0 = (1u << jb) | (jb ? (cur >> (8 - jb)) : 0u);
        uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
        sl[j][0] = h0;
#pragma unroll
        for (int k = 1; k < 8; ++k) {
          H = (H ^ (uint32_t)M.ub[8 + bi - k]) * kFnvPrime;
          sl[j][k] = H & mask;
        }
        g[j][0] = M.st0[h0];
#pragma unroll
        for (int k = 1; k < 8; ++k) g[j][k] = slot[((uint64_t)(k - 1) << tb) + sl[j][k]];
      }
      if (clk.on) __builtin_amdgcn_s_waitcnt(0x0f70);  // attribute the gather wait to phase 2
      clk.mark(2);
      uint32_t idx[NJ][8], pv[NJ][8];
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int k = 0; k < 8; ++k) pv[j][k] = M.stp[k * 256 + g[j][k]];
      {
        // distinct-slot tables: first probes of every key go out together;
        // lanes past the valid bits probe a dummy word; one uniform test
        // guards the (rare) re-probe path
        uint32_t old[NJ][8];
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          const bool ok = lane + 64u * j < nvalid;
#pragma unroll
          for (int k = 1; k < 8; ++k) {
            idx[j][k] = sl[j][k] & (TS - 1);
            LDS uint32_t *at = ok ? &M.tkey[(k - 1) * TS + idx[j][k]] : &M.dummy[lane];
            old[j][k] = atomicCAS((uint32_t *)at, kEmpty, sl[j][k]);
          }
        }
        bool retry = false;
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
          for (int k = 1; k < 8; ++k)
            retry |= (lane + 64u * j < nvalid) && old[j][k] != kEmpty && old[j][k] != sl[j][k];
        if (__ballot(retry)) {
#pragma unroll
          for (int j = 0; j < NJ; ++j)
            if (lane + 64u * j < nvalid)
#pragma unroll
              for (int k = 1; k < 8; ++k)
                while (old[j][k] != kEmpty && old[j][k] != sl[j][k]) {
                  idx[j][k] = (idx[j][k] + 1) & (TS - 1);
                  old[j][k] = atomicCAS((uint32_t *)&M.tkey[(k - 1) * TS + idx[j][k]], kEmpty,
                                        sl[j][k]);
                }
        }
      }
      if (clk.on) __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
      clk.mark(9);
      int32_t smp[NJ][8];
#pragma unroll
      for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int k = 0; k < 8; ++k) smp[j][k] = T->stretch[pv[j][k] >> 20];
      if (clk.on) __builtin_amdgcn_s_waitcnt(0xc07f);
      clk.mark(10);
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
      if (clk.on) __builtin_amdgcn_s_waitcnt(0xc07f);
      clk.mark(11);
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
          M.ridx[rr] = (uint16_t)sl[j][0];
          M.rg[rr] = (uint8_t)g[j][0];
#pragma unroll
          for (int k = 1; k < 8; ++k) {
            M.ridx[k * Umax + rr] = (uint16_t)idx[j][k];
            M.rg[k * Umax + rr] = (uint8_t)g[j][k];
          }
        }
      }
      wave_sync();
      clk.mark(3);
    }
    // Arithmetic coder: a uniform (scalar) loop; (pr, y) of unit bit r is
    // read from lane r & 63 of the register that holds bits 64j..64j+63.
    uint32_t pv[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) pv[j] = (lane + 64u * j < Umax) ? M.rprm[lane + 64u * j] : 0u;
    const uint32_t nc = (uint32_t)min((uint64_t)U, coded_end - base);
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const uint32_t lo = 64u * j;
      const uint32_t hi = min(nc, lo + 64u);
      for (uint32_t r = lo; r < hi; ++r) {
        const uint32_t prm = __builtin_amdgcn_readlane(pv[j], r - lo);
        const uint32_t xmid = coder_mid(x1, x2, prm >> 1);
        if (prm & 1u)
          x2 = xmid;
        else
          x1 = xmid + 1u;
        while (((x1 ^ x2) & 0xff000000u) == 0) {
          acc |= (x2 >> 24) << (8 * (uint32_t)(opos & 3));
          ++opos;
          if ((opos & 3) == 0) {
            if (opos <= cap) {
              if (lane == 0) *(GLB uint32_t *)(dst + opos - 4) = acc;
            } else {
              ovf = true;
            }
            acc = 0;
          }
          x1 <<= 8;
          x2 = (x2 << 8) | 255u;
        }
      }
    }
    clk.mark(4);
    if (nvalid) update_unit<NJ>(M, T, slot, tb, Umax, nvalid, lane, clk);
    wave_sync();
  }
  acc |= (x1 >> 24) << (8 * (uint32_t)(opos & 3));
  ++opos;
  const uint64_t wstart = (opos - 1) & ~3ull;
  if (wstart + 4 <= cap) {
    if (lane == 0) *(GLB uint32_t *)(dst + wstart) = acc;
  } else {
    ovf = true;
  }
  if (lane == 0) a.result[t] = opos | (ovf ? kResultOverflow : 0ull);
  clk.flush(a.prof, lane);
  wave_sync();
}

// ---------------------------------------------------------------------------
// Decoder
// ---------------------------------------------------------------------------

// Window of the payload in registers: lane l holds bytes 4l..4l+3 after
// winbase (0xff past the payload, as the reference decoder reads).
__device__ __forceinline__ uint32_t load_window(const GLB uint8_t *pay, uint64_t plen, uint64_t wb,
                                                uint32_t lane) {
  uint32_t w = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint64_t at = wb + 4ull * lane + i;
    w |= (uint32_t)(at < plen ? pay[at] : (uint8_t)255) << (8 * i);
  }
  return w;
}

__device__ __forceinline__ uint32_t window_byte(uint32_t win, uint64_t off) {
  const uint32_t w = __builtin_amdgcn_readlane(win, (uint32_t)(off >> 2));
  return (w >> (8 * (uint32_t)(off & 3))) & 255u;
}

__device__ __forceinline__ uint32_t lane_value(const uint32_t v[4], uint32_t q) {
  if (q < 64) return __builtin_amdgcn_readlane(v[0], q);
  if (q < 128) return __builtin_amdgcn_readlane(v[1], q - 64);
  if (q < 192) return __builtin_amdgcn_readlane(v[2], q - 128);
  return __builtin_amdgcn_readlane(v[3], q - 192);
}

template <int NJ>
__device__ __forceinline__ void decode_task(const CodecArgs &a, const Task &tk, uint32_t t, const WaveMem &M,
                            const LDS DevTables *T, GLB uint8_t *slot, uint32_t Umax, uint32_t lane) {
  const uint32_t U = tk.U, tb = tk.tb, mask = (1u << tb) - 1u;
  const uint32_t TS = table_size(Umax);
  const uint32_t D = a.spec_bits, P = (1u << D) - 1u;
  const uint64_t L = tk.L;
  const GLB uint8_t *pay = (const GLB uint8_t *)(uintptr_t)tk.src;
  const uint64_t plen = tk.aux;
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  PhaseClock clk(a.prof != nullptr);
  init_model(a, tk, M, slot, Umax, lane);
  clk.mark(0);
  uint64_t winbase = 0;
  uint32_t win = load_window(pay, plen, 0, lane);
  uint32_t x = __builtin_bswap32(__builtin_amdgcn_readlane(win, 0));
  uint32_t x1 = 0, x2 = 0xffffffffu;
  uint64_t ip = 4;
  const uint64_t vbits = 8 * L;
  for (uint64_t base = 0; base < vbits; base += U) {
    const uint32_t nbits = (uint32_t)min((uint64_t)U, vbits - base);
    for (uint32_t rb = 0; rb < nbits; rb += D) {
      // the coder consumes at most 4 bytes per bit
      if (ip + 4 * D + 8 > winbase + 256) {
        winbase = ip & ~3ull;
        win = load_window(pay, plen, winbase, lane);
      }
      const uint32_t j0 = rb & 7u, bi = rb >> 3;
      const uint32_t known = j0 ? ((uint32_t)M.ub[8 + bi] >> (8 - j0)) : 0u;
      // speculative predictions: path q fixes the next plen(q) bits to
      // pbits(q); item it = 8q + k, two items per lane per pass
      uint32_t hb[8];
#pragma unroll
      for (int i = 1; i < 8; ++i) hb[i] = M.ub[8 + bi - i];
      for (uint32_t it0 = lane; it0 < 8 * P; it0 += 128) {
        uint32_t gq[2], sq[2];
        bool ok[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const uint32_t it = it0 + 64u * h;
          const uint32_t q = it >> 3, k = it & 7u;
          const uint32_t plen_q = 31u - __clz(q + 1);
          ok[h] = it < 8 * P && rb + plen_q < nbits;
          if (ok[h]) {
            const uint32_t pbits = q + 1u - (1u << plen_q);
            const uint32_t h0 = (1u << (j0 + plen_q)) | (known << plen_q) | pbits;
            uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
#pragma unroll
            for (int i = 1; i < 8; ++i) {
              const uint32_t Hn = (H ^ hb[i]) * kFnvPrime;
              H = ((uint32_t)i <= k) ? Hn : H;
            }
            sq[h] = k ? (H & mask) : h0;
            gq[h] = k ? (uint32_t)slot[((uint64_t)(k - 1) << tb) + sq[h]] : (uint32_t)M.st0[h0];
          }
        }
        uint32_t pq[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (ok[h]) pq[h] = M.stp[((it0 + 64u * h) & 7u) * 256 + gq[h]];
#pragma unroll
        for (int h = 0; h < 2; ++h)
          if (ok[h]) {
            const uint32_t it = it0 + 64u * h;
            M.psmp[it] = T->stretch[pq[h] >> 20];
            M.pg[it] = (uint8_t)gq[h];
            M.pslot[it] = sq[h];
          }
      }
      wave_sync();
      clk.mark(9);
      // mix every path (lane q + 64m holds path q + 64m)
      uint32_t prq[4] = {0, 0, 0, 0}, cxq[4] = {0, 0, 0, 0};
      const uint32_t c1 = M.ub[8 + bi - 1] >> 5;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const uint32_t q = lane + 64u * m;
        if (q < P) {
          const uint32_t plen_q = 31u - __clz(q + 1);
          if (rb + plen_q < nbits) {
            int32_t smp[8];
            uint32_t active = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
              smp[k] = M.psmp[q * 8 + k];
              active += (smp[k] != 1);
            }
            const uint32_t cxt = active + c1 * 10u;
            prq[m] = mix_predict(M.w + cxt * 8, smp, T->squash);
            cxq[m] = cxt;
          }
        }
      }
      // decode up to D bits (uniform scalar loop)
      const uint32_t nd = min(D, nbits - rb);
      uint32_t pre = 0;
      uint64_t chosen = 0;
      for (uint32_t i = 0; i < nd; ++i) {
        const uint32_t q = (1u << i) - 1u + pre;
        const uint32_t pr = lane_value(prq, q);
        const uint32_t cx = lane_value(cxq, q);
        const uint32_t xmid = coder_mid(x1, x2, pr);
        const uint32_t y = x <= xmid;
        if (y)
          x2 = xmid;
        else
          x1 = xmid + 1u;
        while (((x1 ^ x2) & 0xff000000u) == 0) {
          x1 <<= 8;
          x2 = (x2 << 8) | 255u;
          x = (x << 8) | window_byte(win, ip - winbase);
          ++ip;
        }
        pre = (pre << 1) | y;
        chosen |= (uint64_t)q << (8 * i);
        if (lane == 0) {
          M.rprm[rb + i] = (uint16_t)((pr << 1) | y);
          M.rcxt[rb + i] = (uint8_t)cx;
        }
      }
      if (lane == 0) M.ub[8 + bi] = (uint8_t)(M.ub[8 + bi] | (pre << (8 - j0 - nd)));
      clk.mark(10);
      // record the taken paths' slots and unit-start states
      {
        const bool rec = lane < 8 * nd;
        const uint32_t i = lane >> 3, k = lane & 7u;
        const uint32_t r = min(rb + i, Umax - 1);
        const uint32_t q = (uint32_t)(chosen >> (8 * i)) & 255u;
        const uint32_t gq = M.pg[q * 8 + k];
        const uint32_t sl = M.pslot[q * 8 + k];
        const bool ins = rec && k != 0;
        uint32_t h = sl & (TS - 1);
        LDS uint32_t *tk = M.tkey + (k ? (k - 1) * TS : 0);
        uint32_t old = atomicCAS((uint32_t *)(ins ? &tk[h] : &M.dummy[lane]), kEmpty, sl);
        if (__ballot(ins && old != kEmpty && old != sl)) {
          if (ins)
            while (old != kEmpty && old != sl) {
              h = (h + 1) & (TS - 1);
              old = atomicCAS((uint32_t *)&tk[h], kEmpty, sl);
            }
        }
        if (rec) {
          M.rg[k * Umax + r] = (uint8_t)gq;
          M.ridx[k * Umax + r] = (uint16_t)(k == 0 ? sl : h);
        }
      }
      wave_sync();
      clk.mark(11);
    }
    update_unit<NJ>(M, T, slot, tb, Umax, nbits, lane, clk);
    const uint32_t nb = nbits >> 3;
    for (uint32_t i = lane; i < nb; i += 64) dst[(base >> 3) + i] = M.ub[8 + i];
    uint8_t hist = 0;
    if (lane < 8) hist = M.ub[nb + lane];
    wave_sync();
    if (lane < 8) M.ub[lane] = hist;
    for (uint32_t i = lane; i < Umax / 8 + 8; i += 64) M.ub[8 + i] = 0;
    wave_sync();
  }
  if (lane == 0) a.result[t] = ip;
  clk.flush(a.prof, lane);
  wave_sync();
}

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

template <bool kDecode, int NJ>
__global__ void __launch_bounds__(512) codec_kernel(CodecArgs a, uint32_t Umax) {
  extern __shared__ __align__(16) uint8_t lds[];
  LDS DevTables *T = (LDS DevTables *)lds;
  stage_tables(a.tabs, T);
  const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t P = kDecode ? (1u << a.spec_bits) - 1u : 0u;
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
    if (kDecode)
      decode_task<NJ>(a, tk, t, M, T, slot, Umax, lane);
    else
      encode_task<NJ>(a, tk, t, M, T, slot, Umax, lane);
  }
}

// ---------------------------------------------------------------------------
// Container assembly: per job, per batch: LE u32 blob length, BE end
// offsets E_c = 4C + sum_{j<=c} size_j - 1, payloads back to back
// (cucm_compress_v2.cu:870-914, 1242-1244).
// ---------------------------------------------------------------------------

__global__ void layout_kernel(const AsmJob *jobs, uint32_t njobs, const uint8_t *headers,
                              const uint64_t *result, uint64_t *task_dst, uint64_t *job_len) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs) return;
  const AsmJob J = jobs[j];
  uint8_t *o = (uint8_t *)(uintptr_t)J.out_base;
  const uint32_t lenbytes = J.blob_only ? 0u : 4u;
  for (uint32_t i = 0; i < J.header_len; ++i) o[i] = headers[J.header_off + i];
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
      t.pad = 0;
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
      tasks[J.first_task + c] = t;  // L = 0: the decoder skips it
    }
  }
  status[j] = st;
}

// ---------------------------------------------------------------------------
// Host-callable launchers
// ---------------------------------------------------------------------------

uint32_t codec_lds_bytes(uint32_t Umax, uint32_t spec_bits, bool decode, uint32_t waves) {
  const uint32_t P = decode ? (1u << spec_bits) - 1u : 0u;
  return align16(sizeof(DevTables)) + waves * wave_lds_bytes(Umax, P, decode);
}

template <bool kDecode, int NJ>
static hipError_t launch_one(const CodecArgs &a, uint32_t Umax, uint32_t blocks, uint32_t wpb,
                             uint32_t lds, hipStream_t stream) {
  hipError_t e = hipFuncSetAttribute((const void *)codec_kernel<kDecode, NJ>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((codec_kernel<kDecode, NJ>), dim3(blocks), dim3(64 * wpb), lds, stream, a,
                     Umax);
  return hipGetLastError();
}

hipError_t launch_codec(bool decode, const CodecArgs &a, uint32_t Umax, uint32_t blocks,
                        uint32_t waves_per_block, hipStream_t stream) {
  const uint32_t lds = codec_lds_bytes(Umax, a.spec_bits, decode, waves_per_block);
  const int nj = Umax <= 64 ? 1 : (Umax <= 128 ? 2 : 4);
  if (decode) {
    if (nj == 1) return launch_one<true, 1>(a, Umax, blocks, waves_per_block, lds, stream);
    if (nj == 2) return launch_one<true, 2>(a, Umax, blocks, waves_per_block, lds, stream);
    return launch_one<true, 4>(a, Umax, blocks, waves_per_block, lds, stream);
  }
  if (nj == 1) return launch_one<false, 1>(a, Umax, blocks, waves_per_block, lds, stream);
  if (nj == 2) return launch_one<false, 2>(a, Umax, blocks, waves_per_block, lds, stream);
  return launch_one<false, 4>(a, Umax, blocks, waves_per_block, lds, stream);
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
                       headers, result, task_dst, job_len);
  }
  if (ntasks) {
    const uint32_t blocks = min((ntasks + 3) / 4, 4096u);
    hipLaunchKernelGGL(copy_payload_kernel, dim3(blocks), dim3(256), 0, stream, tasks, ntasks,
                       result, task_dst);
  }
  return hipGetLastError();
}

}  // namespace cucm
