// cucm_frozen.h -- the frozen-model codec: every bit of a chunk is predicted
// from the initial model (zero, or prelearned from the container's first P
// bytes) and the model is never updated.
//
// The reference ships no code for this mode: it is the paper's "frozen
// model" (prelearn, then no runtime updates).  The per-bit model is the
// normal path's -- slots (cucm_compress_v2.cu:727-743), StateMap prediction
// (:579-581), stretch, mixer context and mix (:764-792) -- and so is the
// coder (:841-868); what goes is every update and the tail padding, so a
// chunk codes exactly its 8L bits and then the flush byte.  The container
// marks the mode with bit 1 of its flag byte.
//
// With no updates the model is read-only and shared: one copy of the order
// tables serves every chunk of a launch (no per-chunk model slot), so the
// gathers of all chunks hit one table that the L2 and MALL can hold.
//
//   encoder  one wave per chunk, in blocks of 64 bytes: lane l predicts the
//            8 bits of byte l (all 56 HBM gathers of a lane in flight at
//            once), then a uniform scalar loop codes the block's 512 bits;
//            the next block's gathers are issued before the current block
//            is coded, so they arrive while the coder runs.
//   decoder  one wave per chunk, rounds of D speculative bits as in the
//            normal decoder (lane q evaluates path q), without the records
//            and unit updates.
#ifndef CUCM_FROZEN_H
#define CUCM_FROZEN_H
#include "cucm_wave.h"

namespace cucm {

// Per-wave LDS of the frozen codec.
struct FrozenMem {
  LDS int16_t *sm;  // [8][256] stretch(statep >> 20) of the model
  LDS int32_t *w;   // [80][8] mixer weights
  LDS uint8_t *st0; // [256] order-0 states by h0
  LDS uint8_t *ub;  // [8 + 64] encoder: 8 history bytes + the block's bytes
};

CUCM_HD uint32_t frozen_wave_bytes() {
  return align16(8 * 256 * 2) + align16(kMixCtx * kOrders * 4) + align16(256) + align16(8 + 64);
}

CUCM_DEV FrozenMem carve_frozen(LDS uint8_t *base) {
  FrozenMem f;
  uint32_t off = 0;
  f.sm = (LDS int16_t *)(base + off);
  off += align16(8 * 256 * 2);
  f.w = (LDS int32_t *)(base + off);
  off += align16(kMixCtx * kOrders * 4);
  f.st0 = base + off;
  off += align16(256);
  f.ub = base + off;
  return f;
}

constexpr uint32_t kNoModel = 0xffffffffu;    // nothing loaded yet
constexpr uint32_t kZeroModel = 0xfffffffeu;  // the all-zero initial model

// Loads the task's model into the wave's LDS unless it is already there
// (`loaded` keeps the key between tasks) and returns its order-1..7 state
// tables in HBM, or null for the zero model (every state 0: no gathers).
CUCM_DEV const GLB uint8_t *frozen_model(const CodecArgs &a, const Task &tk, const FrozenMem &F,
                                         const LDS DevTables *T, uint32_t &loaded,
                                         uint32_t lane) {
  const bool init = (tk.flags & kTaskInitModel) != 0;
  const uint32_t key = init ? tk.model : kZeroModel;
  const DevInitModel *im = init ? a.init + tk.model : nullptr;
  if (key != loaded) {
    wave_sync();  // every lane is done with the previous model
    if (init) {
      for (uint32_t i = lane; i < 8 * 256; i += 64) F.sm[i] = T->stretch[im->statep[i >> 8][i & 255] >> 20];
      for (uint32_t i = lane; i < kMixCtx * kOrders; i += 64) F.w[i] = im->w[i];
      for (uint32_t i = lane; i < 256; i += 64) F.st0[i] = im->st0[i];
    } else {
      const int16_t s = T->stretch[(1u << 31) >> 20];
      for (uint32_t i = lane; i < 8 * 256; i += 64) F.sm[i] = s;
      for (uint32_t i = lane; i < kMixCtx * kOrders; i += 64) F.w[i] = 0;
      for (uint32_t i = lane; i < 256; i += 64) F.st0[i] = 0;
    }
    wave_sync();
    loaded = key;
  }
  return init ? (const GLB uint8_t *)im->states : nullptr;
}

// ---------------------------------------------------------------------------
// Encoder
// ---------------------------------------------------------------------------

// The block's bytes [base, base + 64) and the 8 before it into ub (zero
// outside the chunk: the chunk-relative zero context, cucm_compress_v2.cu:
// 731-741).
CUCM_DEV void frozen_stage(const FrozenMem &F, const GLB uint8_t *src, uint64_t L, uint64_t base,
                           uint32_t lane) {
  wave_sync();  // the previous block's reads of ub are done
  for (uint32_t i = lane; i < 8 + 64; i += 64) {
    const int64_t p = (int64_t)base - 8 + i;
    F.ub[i] = (p >= 0 && p < (int64_t)L) ? src[p] : (uint8_t)0;
  }
  wave_sync();
}

// Lane l: byte c = ub[8 + l], its previous byte c1, and the state of every
// order of its 8 bits (bit j's partial byte h0 = 1 << j | c >> (8 - j); the
// FNV-1a chain over h0 and the previous bytes gives orders 1..7).  The
// gathers are left in flight.
CUCM_DEV void frozen_block_issue(const FrozenMem &F, const GLB uint8_t *states, uint32_t tb,
                                 uint32_t mask, uint32_t lane, uint32_t &c, uint32_t &c1,
                                 uint32_t g[8][8]) {
  uint32_t h[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) h[k] = F.ub[8 + lane - k];
  c = h[0];
  c1 = h[1];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t h0 = (1u << j) | (j ? (c >> (8 - j)) : 0u);
    g[j][0] = F.st0[h0];
  }
  if (states) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t h0 = (1u << j) | (j ? (c >> (8 - j)) : 0u);
      uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        g[j][k] = states[((uint32_t)(k - 1) << tb) | fnv_pos(H, h[k], mask)];
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int k = 1; k < 8; ++k) g[j][k] = 0;
  }
}

// prm[j] = pr << 20 | y of bit j of the lane's byte: stretch of each order's
// prediction, mixer context (c1 >> 5) * 10 + active orders, mix, pr bump.
CUCM_DEV void frozen_block_mix(const FrozenMem &F, const LDS DevTables *T, uint32_t c,
                               uint32_t c1, const uint32_t g[8][8], uint32_t prm[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    int32_t smp[8];
    uint32_t active = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      smp[k] = F.sm[k * 256 + g[j][k]];
      active += (smp[k] != 1);
    }
    const uint32_t cxt = active + (c1 >> 5) * 10u;
    const LDS u32x4 *wr = (const LDS u32x4 *)(F.w + cxt * 8u);
    const u32x4 wa = wr[0], wb = wr[1];
    const int32_t w[8] = {(int32_t)wa.x, (int32_t)wa.y, (int32_t)wa.z, (int32_t)wa.w,
                          (int32_t)wb.x, (int32_t)wb.y, (int32_t)wb.z, (int32_t)wb.w};
    const uint32_t pr = mix_weights(w, smp, T->squash);
    prm[j] = (pr << 20) | ((c >> (7 - j)) & 1u);
  }
}

// One chunk: blocks of 64 bytes, the next block's gathers in flight while the
// scalar coder codes the current one.  The payload goes to the task's
// 4-byte aligned staging region; result = payload bytes (flush included),
// with kResultOverflow if the region was too small.
CUCM_DEV void frozen_encode_task(const CodecArgs &a, const Task &tk, uint32_t t,
                                 const FrozenMem &F, const LDS DevTables *T,
                                 const GLB uint8_t *states, uint32_t lane) {
  const uint32_t tb = tk.tb, mask = (1u << tb) - 1u;
  const uint64_t L = tk.L;
  const GLB uint8_t *src = (const GLB uint8_t *)(uintptr_t)tk.src;
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  const uint64_t cap = tk.aux;
  uint32_t x1 = 0, x2 = 0xffffffffu, acc = 0;
  uint32_t opos = 0, ovf = 0;
  uint32_t c = 0, c1 = 0, g[8][8];
  frozen_stage(F, src, L, 0, lane);
  frozen_block_issue(F, states, tb, mask, lane, c, c1, g);
  for (uint64_t base = 0; base < L; base += 64) {
    const uint32_t n = (uint32_t)min((uint64_t)64, L - base);
    uint32_t prm[8];
    frozen_block_mix(F, T, c, c1, g, prm);
    if (base + 64 < L) {
      frozen_stage(F, src, L, base + 64, lane);
      frozen_block_issue(F, states, tb, mask, lane, c, c1, g);
    }
    for (uint32_t i = 0; i < n; ++i) {
#pragma unroll
      for (int j = 0; j < 8; ++j) CUCM_CODE_BIT(lane_read(prm[j], i));
    }
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

// ---------------------------------------------------------------------------
// Decoder
// ---------------------------------------------------------------------------

// Path q of a round (the next floor(log2(q + 1)) bits of the current byte
// fixed to q + 1 - 2^plen): its slots and the gathers of orders 1..7, left
// in flight.
template <int NM, uint32_t P>
CUCM_DEV void frozen_paths_issue(const FrozenMem &F, const GLB uint8_t *states, uint32_t tb,
                                 uint32_t mask, uint32_t j0, uint32_t partial, uint64_t hist,
                                 uint32_t lane, uint32_t g[NM][8]) {
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    // lanes without a path repeat path 0 (its lines, in bounds)
    const uint32_t q = lane + 64u * m < P ? lane + 64u * m : 0u;
    const uint32_t plen = 31u - clz32(q + 1u);
    const uint32_t pbits = q + 1u - (1u << plen);
    const uint32_t h0 = ((1u << (j0 + plen)) | (partial << plen) | pbits) & 255u;
    g[m][0] = F.st0[h0];
    if (states) {
      uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
#pragma unroll
      for (int k = 1; k < 8; ++k) {
        // the offset (k - 1) * 2^tb + position < 7 * 2^25 fits 32 bits
        g[m][k] = states[((uint32_t)(k - 1) << tb) |
                         fnv_pos(H, (uint32_t)((hist >> (8 * (k - 1))) & 255u), mask)];
      }
    } else {
#pragma unroll
      for (int k = 1; k < 8; ++k) g[m][k] = 0;
    }
  }
}

// pr of every path of the round (c1 = the previous byte >> 5).
template <int NM, uint32_t P>
CUCM_DEV void frozen_paths_mix(const FrozenMem &F, const LDS DevTables *T, uint32_t c1,
                               const uint32_t g[NM][8], uint32_t pc[NM]) {
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    int32_t smp[8];
    uint32_t active = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      smp[k] = F.sm[k * 256 + g[m][k]];
      active += (smp[k] != 1);
    }
    const uint32_t cxt = active + c1 * 10u;
    const LDS u32x4 *wr = (const LDS u32x4 *)(F.w + cxt * 8u);
    const u32x4 wa = wr[0], wb = wr[1];
    const int32_t w[8] = {(int32_t)wa.x, (int32_t)wa.y, (int32_t)wa.z, (int32_t)wa.w,
                          (int32_t)wb.x, (int32_t)wb.y, (int32_t)wb.z, (int32_t)wb.w};
    pc[m] = mix_weights(w, smp, T->squash);
  }
}

// One chunk, D bits per round: mix the round's paths, decode its D bits
// (uniform scalar loop; y = x <= xmid, cucm_decompress_v2.cu:816-844), store
// each completed byte, issue the next round.
template <int D>
CUCM_DEV void frozen_decode_task(const CodecArgs &a, const Task &tk, uint32_t t,
                                 const FrozenMem &F, const LDS DevTables *T,
                                 const GLB uint8_t *states, uint32_t lane) {
  constexpr uint32_t P = (1u << D) - 1u;
  constexpr int NM = (int)((P + 63u) / 64u);
  const uint32_t tb = tk.tb, mask = (1u << tb) - 1u;
  const uint64_t L = tk.L;
  const GLB uint8_t *pay = (const GLB uint8_t *)(uintptr_t)tk.src;
  const uint64_t plen = tk.aux;
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  uint64_t winbase = 0;
  uint32_t win = load_window(pay, plen, 0, lane);
  uint32_t x = lane_read(win, 0);
  uint32_t x1 = 0, x2 = 0xffffffffu;
  uint64_t ip = 4;
  uint64_t hist = 0;     // the 8 bytes before the current one, latest lowest
  uint32_t partial = 0;  // decided bits of the current byte
  const uint64_t vbits = 8 * L;
  uint32_t g[NM][8];
  if (vbits) frozen_paths_issue<NM, P>(F, states, tb, mask, 0, 0, hist, lane, g);
  for (uint64_t rb = 0; rb < vbits; rb += D) {
    // the coder consumes at most 4 bytes per bit
    if (ip + 4 * D + 8 > winbase + 256) {
      winbase = ip & ~3ull;
      win = load_window(pay, plen, winbase, lane);
    }
    uint32_t pc[NM];
    frozen_paths_mix<NM, P>(F, T, (uint32_t)(hist & 255u) >> 5, g, pc);
#pragma unroll
    for (int m = 0; m < NM; ++m) pc[m] <<= 20;  // pr << 20 (coder_mid20)
    uint32_t pre = 0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const uint32_t q = (1u << i) - 1u + pre;
      const uint32_t pr = path_value<NM>(pc, q);
      const uint32_t xmid = coder_mid20(x1, x2, pr);
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
    }
    partial = (partial << D) | pre;
    if (((rb & 7u) + D) == 8) {
      if (lane == 0) dst[rb >> 3] = (uint8_t)partial;
      hist = (hist << 8) | partial;
      partial = 0;
    }
    if (rb + D < vbits)
      frozen_paths_issue<NM, P>(F, states, tb, mask, (uint32_t)((rb + D) & 7u), partial, hist,
                                lane, g);
  }
  if (lane == 0) a.result[t] = ip;
  wave_sync();
}

}  // namespace cucm

#endif
