// cucm_bigunit.h -- the codec for units of more than 256 bits (one wave per
// chunk), shared by the gfx950 kernel (cucm_kernels.hip) and the CPU wave
// emulator of the tests.
//
// The reference accepts any unit length U that is a multiple of 32, up to
// what the container's 16-bit field holds (cucm_compress_v2.cu:1203-1204,
// 1343-1349; cucm_decompress_v2.cu:1190-1191).  The main kernels keep a
// unit's records in LDS, laid out for U <= 256; this codec keeps them in a
// per-wave region of global memory instead, so U can be as large as 65,504:
//
//   predict  the unit's bits in blocks of 64 (lane = bit; the decoder one
//            bit at a time, lane = order), every prediction from the model
//            as it stood at the unit start (cucm_compress_v2.cu:711-792);
//            per bit the 8 slots and (pr, y) are recorded in global memory
//            and the mixer deltas summed in LDS (order independent);
//   code     the carry-less coder over the unit's recorded (pr, y) -- the
//            records persist, so a tail-padded last chunk codes the stale
//            (pr, y) of its slots as the reference does (:850-851);
//   update   the mixer deltas added; then the StateMap updates in bit order
//            (cucm_compress_v2.cu:821-839): blocks of 64 bits whose slots'
//            current states are gathered, one lane per order walks the
//            block's bits with a small LDS table of the block's distinct
//            slots (a slot hit again inside the block sees its updated
//            state), and the block's final states are written back before
//            the next block gathers.
//
// All arithmetic is the reference's fixed-point integer arithmetic, as in
// cucm_wave.h, whose primitives (mixer, coder split, payload window) it
// uses.
#ifndef CUCM_BIGUNIT_H
#define CUCM_BIGUNIT_H
#include "cucm_wave.h"

namespace cucm {

constexpr uint32_t kBigBlock = 64;  // bits per StateMap update block
constexpr uint32_t kBigTS = 128;    // distinct-slot table entries per order and block

// Per-wave LDS of the large-unit codec.
struct BigMem {
  LDS uint32_t *stp;    // [8][256] statep
  LDS int32_t *w;       // [80][8] mixer weights
  LDS int32_t *wd;      // [80][8] the unit's summed mixer deltas
  LDS uint32_t *hkey;   // [7][kBigTS] distinct slots of an update block, orders 1..7
  LDS uint32_t *bslot;  // [8][kBigBlock] slots of the block's bits (order 0: h0)
  LDS uint8_t *st0;     // [256] order-0 states by h0
  LDS uint8_t *hval;    // [7][kBigTS] running state of each distinct slot
  LDS uint8_t *bg;      // [8][kBigBlock] current state of each (order, bit) of the block
  LDS uint8_t *by;      // [kBigBlock] y of the block's bits
};

struct BigLayout {
  uint32_t stp, w, wd, hkey, bslot, st0, hval, bg, by, total;
};

CUCM_HDI BigLayout big_layout() {
  BigLayout l;
  uint32_t off = 0;
#define CUCM_BTAKE(field, bytes) \
  l.field = off;                 \
  off = align16(off + (bytes));
  CUCM_BTAKE(stp, 8 * 256 * 4)
  CUCM_BTAKE(w, kMixCtx * kOrders * 4)
  CUCM_BTAKE(wd, kMixCtx * kOrders * 4)
  CUCM_BTAKE(hkey, 7 * kBigTS * 4)
  CUCM_BTAKE(bslot, 8 * kBigBlock * 4)
  CUCM_BTAKE(st0, 256)
  CUCM_BTAKE(hval, 7 * kBigTS)
  CUCM_BTAKE(bg, 8 * kBigBlock)
  CUCM_BTAKE(by, kBigBlock)
#undef CUCM_BTAKE
  l.total = off;
  return l;
}

CUCM_HD uint32_t big_wave_bytes() { return big_layout().total; }

CUCM_DEV BigMem carve_big(LDS uint8_t *base) {
  const BigLayout l = big_layout();
  BigMem m;
  m.stp = (LDS uint32_t *)(base + l.stp);
  m.w = (LDS int32_t *)(base + l.w);
  m.wd = (LDS int32_t *)(base + l.wd);
  m.hkey = (LDS uint32_t *)(base + l.hkey);
  m.bslot = (LDS uint32_t *)(base + l.bslot);
  m.st0 = base + l.st0;
  m.hval = base + l.hval;
  m.bg = base + l.bg;
  m.by = base + l.by;
  return m;
}

// Global records of one wave (big_record_bytes(Umax) bytes, cucm_model.h):
// the 8 slots of every unit bit ([8][Umax], order 0: h0) and its pr << 1 | y
// ([Umax]).
struct BigRec {
  GLB uint32_t *slots;
  GLB uint16_t *prm;
};
CUCM_DEV BigRec big_records(GLB uint8_t *base, uint32_t Umax) {
  BigRec r;
  r.slots = (GLB uint32_t *)base;
  r.prm = (GLB uint16_t *)(base + 32ull * Umax);
  return r;
}

// The StateMap update of one bit (cucm_compress_v2.cu:592-601) on statep
// entry p: count-adaptive step towards y, the count saturating at 1023.
CUCM_DEV uint32_t big_statep_step(uint32_t p, uint32_t y, const LDS uint16_t *dt) {
  const uint32_t n = p & 1023u;
  const int32_t e = (int32_t)((y << 22) - (p >> 10));
  // |e >> 3| <= 2^19 and dt <= 5461: the product fits 31 bits
  const uint32_t adj = (uint32_t)((e >> 3) * (int32_t)dt[n]) & 0xfffffc00u;
  return p + adj + (n < 1023u ? 1u : 0u);
}

// Model slot zeroed (or the prelearned model copied in), LDS model reset,
// the unit records' (pr, y) zeroed (the reference zeroes them once per
// batch, cucm_compress_v2.cu:1039-1040).
CUCM_DEV void big_init(const CodecArgs &a, const Task &tk, const BigMem &M, GLB uint8_t *slot,
                       const BigRec &R, uint32_t lane) {
  const uint64_t n16 = (7ull << tk.tb) >> 4;
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
  for (uint32_t i = lane; i < kMixCtx * kOrders; i += 64) M.wd[i] = 0;
  for (uint32_t i = lane; i < 7 * kBigTS; i += 64) M.hkey[i] = kEmpty;
  for (uint32_t i = lane; i < tk.U; i += 64) R.prm[i] = 0;
  vm_drain();
  wave_sync();
}

// The deferred update of a unit's first nvalid bits: mixer deltas, then the
// StateMap updates in bit order, block by block.
CUCM_DEV void big_update(const BigMem &M, const LDS DevTables *T, GLB uint8_t *slot, uint32_t tb,
                         uint32_t U, uint32_t nvalid, const BigRec &R, uint32_t lane) {
  for (uint32_t i = lane; i < kMixCtx * kOrders; i += 64) {
    M.w[i] = (int32_t)((uint32_t)M.w[i] + (uint32_t)M.wd[i]);
    M.wd[i] = 0;
  }
  const uint32_t kk = min(lane, 7u);  // lane k < 8 walks order k
  for (uint32_t r0 = 0; r0 < nvalid; r0 += kBigBlock) {
    const uint32_t nb = min(kBigBlock, nvalid - r0);
    // the block's slots, y and current states (earlier blocks written back)
    if (lane < nb) {
      const uint32_t r = r0 + lane;
      M.by[lane] = (uint8_t)(R.prm[r] & 1u);
      M.bslot[lane] = R.slots[r];
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const uint32_t sl = R.slots[(uint64_t)k * U + r];
        M.bslot[k * kBigBlock + lane] = sl;
        M.bg[k * kBigBlock + lane] = slot[((uint32_t)(k - 1) << tb) | sl];
      }
    }
    wave_sync();
    // one lane per order walks the block's bits in order
    if (lane < 8) {
      LDS uint32_t *keys = M.hkey + (kk ? (kk - 1) * kBigTS : 0u);
      LDS uint8_t *vals = M.hval + (kk ? (kk - 1) * kBigTS : 0u);
      LDS uint32_t *sp = M.stp + kk * 256;
      for (uint32_t i = 0; i < nb; ++i) {
        const uint32_t y = M.by[i];
        const uint32_t sl = M.bslot[kk * kBigBlock + i];
        uint32_t s, h = 0;
        if (kk == 0) {
          s = M.st0[sl];
        } else {
          h = sl & (kBigTS - 1);
          while (keys[h] != kEmpty && keys[h] != sl) h = (h + 1) & (kBigTS - 1);
          if (keys[h] == kEmpty) {
            keys[h] = sl;
            s = M.bg[kk * kBigBlock + i];
          } else {
            s = vals[h];
          }
        }
        sp[s] = big_statep_step(sp[s], y, T->dt);
        const uint32_t ns = T->nex[2 * s + y];
        if (kk == 0)
          M.st0[sl] = (uint8_t)ns;
        else
          vals[h] = (uint8_t)ns;
      }
    }
    wave_sync();
    // final states of the block's distinct slots back to HBM
    for (uint32_t e = lane; e < 7 * kBigTS; e += 64) {
      const uint32_t key = M.hkey[e];
      if (key != kEmpty) {
        slot[((e / kBigTS) << tb) | key] = M.hval[e];
        M.hkey[e] = kEmpty;
      }
    }
    vm_drain();  // written back before the next block (or unit) gathers
    wave_sync();
  }
}

// One chunk through the large-unit encoder.
CUCM_DEV void big_encode_task(const CodecArgs &a, const Task &tk, uint32_t t, const BigMem &M,
                              const LDS DevTables *T, GLB uint8_t *slot, const BigRec &R,
                              uint32_t lane) {
  const uint32_t U = tk.U, tb = tk.tb, mask = (1u << tb) - 1u;
  const uint64_t L = tk.L;
  const GLB uint8_t *src = (const GLB uint8_t *)(uintptr_t)tk.src;
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  const uint64_t cap = tk.aux;
  big_init(a, tk, M, slot, R, lane);
  const uint64_t vbits = 8 * L;
  const uint64_t coded_end = (tk.flags & kTaskTailPad) ? 8ull * tk.S : vbits;
  uint32_t x1 = 0, x2 = 0xffffffffu, acc = 0;
  uint32_t opos = 0, ovf = 0;
  for (uint64_t base = 0; base < coded_end; base += U) {
    const uint32_t nvalid = base < vbits ? (uint32_t)min((uint64_t)U, vbits - base) : 0u;
    // predictions from the unit-start model, 64 bits at a time
    for (uint32_t r0 = 0; r0 < nvalid; r0 += 64) {
      const uint32_t r = r0 + lane;
      const bool ok = r < nvalid;
      const uint64_t b = base + (ok ? r : 0u);
      const uint64_t bi = b >> 3;
      const uint32_t j = (uint32_t)(b & 7u);
      const uint32_t cur = src[bi];
      const uint32_t h0 = (1u << j) | (j ? (cur >> (8 - j)) : 0u);
      uint32_t sl[8], g[8];
      uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
      sl[0] = h0;
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const uint32_t hk = bi >= (uint64_t)k ? (uint32_t)src[bi - k] : 0u;
        sl[k] = fnv_pos(H, hk, mask);
      }
      g[0] = M.st0[h0];
#pragma unroll
      for (int k = 1; k < 8; ++k) g[k] = slot[((uint32_t)(k - 1) << tb) | sl[k]];
      int32_t smp[8];
      uint32_t active = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        smp[k] = T->stretch[M.stp[k * 256 + g[k]] >> 20];
        active += (smp[k] != 1);
      }
      const uint32_t c1 = bi ? (uint32_t)src[bi - 1] : 0u;
      const uint32_t cxt = active + (c1 >> 5) * 10u;
      int32_t wv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) wv[k] = M.w[cxt * 8 + k];
      const uint32_t pr = mix_weights(wv, smp, T->squash);
      const uint32_t y = (cur >> (7 - j)) & 1u;
      if (ok) {
#pragma unroll
        for (int k = 0; k < 8; ++k) R.slots[(uint64_t)k * U + r] = sl[k];
        R.prm[r] = (uint16_t)((pr << 1) | y);
        const int32_t err = ((int32_t)(y << 12) - (int32_t)pr) * 7;
#pragma unroll
        for (int k = 0; k < 8; ++k)
          lds_add(&M.wd[cxt * 8 + k], (int32_t)(((int64_t)err * smp[k]) >> 16));
      }
    }
    vm_drain();  // the records land before they are read back
    wave_sync();
    // the coder over the unit's recorded (pr, y), stale ones past the data
    const uint32_t nc = (uint32_t)min((uint64_t)U, coded_end - base);
    for (uint32_t r0 = 0; r0 < nc; r0 += 64) {
      const uint32_t pv = prm20(R.prm[min(r0 + lane, U - 1)]);
      const uint32_t n = min(64u, nc - r0);
      for (uint32_t i = 0; i < n; ++i) CUCM_CODE_BIT(lane_read(pv, i));
    }
    if (nvalid) big_update(M, T, slot, tb, U, nvalid, R, lane);
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
  wave_sync();
}

// One chunk through the large-unit decoder: one bit per step (lane k < 8
// evaluates order k), the payload read through the register window.
CUCM_DEV void big_decode_task(const CodecArgs &a, const Task &tk, uint32_t t, const BigMem &M,
                              const LDS DevTables *T, GLB uint8_t *slot, const BigRec &R,
                              uint32_t lane) {
  const uint32_t U = tk.U, tb = tk.tb, mask = (1u << tb) - 1u;
  const uint64_t L = tk.L;
  const GLB uint8_t *pay = (const GLB uint8_t *)(uintptr_t)tk.src;
  const uint64_t plen = tk.aux;
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  big_init(a, tk, M, slot, R, lane);
  uint64_t winbase = 0;
  uint32_t win = load_window(pay, plen, 0, lane);
  uint32_t x = lane_read(win, 0);
  uint32_t x1 = 0, x2 = 0xffffffffu;
  uint64_t ip = 4;
  uint64_t hist = 0;     // the 8 bytes before the current one, latest lowest
  uint32_t partial = 0;  // decided bits of the current byte
  const uint64_t vbits = 8 * L;
  const uint32_t kk = min(lane, 7u);
  for (uint64_t base = 0; base < vbits; base += U) {
    const uint32_t nbits = (uint32_t)min((uint64_t)U, vbits - base);
    for (uint32_t r = 0; r < nbits; ++r) {
      if (ip + 16 > winbase + 256) {
        winbase = ip & ~3ull;
        win = load_window(pay, plen, winbase, lane);
      }
      const uint32_t j = (uint32_t)((base + r) & 7u);
      const uint32_t h0 = (1u << j) | partial;
      uint32_t H = (kFnvBasis ^ h0) * kFnvPrime, pos = 0;
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        const uint32_t x = H ^ (uint32_t)((hist >> (8 * (k - 1))) & 255u);
        pos = (uint32_t)k == kk ? x : pos;
        H = (uint32_t)k < kk ? x * kFnvPrime : H;
      }
      const uint32_t sl = kk ? (pos & mask) : h0;
      // every lane loads (lane 0 a byte of table 0, in bounds) and selects
      const uint32_t gh = slot[((kk ? kk - 1 : 0u) << tb) | sl];
      const uint32_t g = kk ? gh : (uint32_t)M.st0[h0];
      const int32_t smp = T->stretch[M.stp[kk * 256 + g] >> 20];
      int32_t s[8];
      uint32_t active = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        s[k] = (int32_t)lane_read((uint32_t)smp, (uint32_t)k);
        active += (s[k] != 1);
      }
      const uint32_t cxt = active + (uint32_t)((hist & 255u) >> 5) * 10u;
      int32_t wv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) wv[k] = M.w[cxt * 8 + k];
      const uint32_t pr = lane_first(mix_weights(wv, s, T->squash));
      const uint32_t xmid = coder_mid20(x1, x2, pr << 20);
      const uint32_t y = coder_split(x, xmid, x1, x2);
      if (((x1 ^ x2) & 0xff000000u) == 0) {
        // shift out every leading byte x1 and x2 agree on (at most 4)
        const uint32_t nb = min((uint32_t)clz32(x1 ^ x2) >> 3, 4u);
        const uint32_t sh = 8 * nb;
        x1 = (uint32_t)((uint64_t)x1 << sh);
        x2 = (uint32_t)(((uint64_t)x2 << sh) | ((1ull << sh) - 1ull));
        const uint32_t nxt = window_be32(win, (uint32_t)(ip - winbase));
        x = (uint32_t)((((uint64_t)x << 32) | nxt) >> (32 - sh));
        ip += nb;
      }
      if (lane < 8) {
        R.slots[(uint64_t)kk * U + r] = sl;
        const int32_t err = ((int32_t)(y << 12) - (int32_t)pr) * 7;
        lds_add(&M.wd[cxt * 8 + kk], (int32_t)(((int64_t)err * smp) >> 16));
      }
      if (lane == 0) R.prm[r] = (uint16_t)((pr << 1) | y);
      partial = (partial << 1) | y;
      if (j == 7) {
        if (lane == 0) dst[(base + r) >> 3] = (uint8_t)partial;
        hist = (hist << 8) | partial;
        partial = 0;
      }
    }
    vm_drain();  // the records land before the update reads them back
    wave_sync();
    big_update(M, T, slot, tb, U, nbits, R, lane);
  }
  if (lane == 0) a.result[t] = ip;
  wave_sync();
}

}  // namespace cucm

#endif
