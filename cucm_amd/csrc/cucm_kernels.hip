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

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Makes this wave's global stores visible to its own later global loads.
__device__ __forceinline__ void global_sync() {
  __threadfence();
  __builtin_amdgcn_wave_barrier();
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
  uint32_t *stp;    // [8][256] statep
  int32_t *w;       // [80][8] mixer weights
  uint32_t *tkey;   // [7][TS] distinct slots of the unit, orders 1..7
  uint32_t *pslot;  // [P][8] decoder: slot (order 0: h0) per path
  uint16_t *tval;   // [7][TS] running state of each distinct slot
  uint16_t *ridx;   // [8][Umax] table index (order 0: h0) per unit bit
  uint16_t *rprm;   // [Umax] pr << 1 | y per unit bit (persists: tail padding)
  int16_t *psmp;    // [P][8] decoder: stretched prediction per path
  uint16_t *ppr;    // [P] decoder: mixed probability per path
  uint8_t *st0;     // [256] order-0 states by h0
  uint8_t *rg;      // [8][Umax] unit-start state per unit bit
  uint8_t *rcxt;    // [Umax] mixer context per unit bit
  uint8_t *pg;      // [P][8] decoder: unit-start state per path
  uint8_t *pcxt;    // [P] decoder: mixer context per path
  uint8_t *ub;      // [8 + Umax/8] 8 history bytes + the unit's bytes
  uint8_t *win;     // [kWindow] decoder payload window
  uint8_t *chosen;  // [8] decoder: path taken by each bit of a round
};

__host__ __device__ inline uint32_t wave_lds_bytes(uint32_t Umax, uint32_t P, bool dec,
                                                   WaveMem *m = nullptr, uint8_t *base = nullptr) {
  const uint32_t TS = table_size(Umax);
  uint32_t off = 0;
  auto take = [&](uint32_t bytes) {
    uint32_t at = off;
    off = align16(off + bytes);
    return at;
  };
  uint32_t o_stp = take(8 * 256 * 4);
  uint32_t o_w = take(kMixCtx * kOrders * 4);
  uint32_t o_tkey = take(7 * TS * 4);
  uint32_t o_pslot = take(dec ? P * 8 * 4 : 0);
  uint32_t o_tval = take(7 * TS * 2);
  uint32_t o_ridx = take(8 * Umax * 2);
  uint32_t o_rprm = take(Umax * 2);
  uint32_t o_psmp = take(dec ? P * 8 * 2 : 0);
  uint32_t o_ppr = take(dec ? P * 2 : 0);
  uint32_t o_st0 = take(256);
  uint32_t o_rg = take(8 * Umax);
  uint32_t o_rcxt = take(Umax);
  uint32_t o_pg = take(dec ? P * 8 : 0);
  uint32_t o_pcxt = take(dec ? P : 0);
  uint32_t o_ub = take(8 + Umax / 8 + 8);
  uint32_t o_win = take(dec ? kWindow : 0);
  uint32_t o_chosen = take(dec ? 8 : 0);
  if (m) {
    m->stp = (uint32_t *)(base + o_stp);
    m->w = (int32_t *)(base + o_w);
    m->tkey = (uint32_t *)(base + o_tkey);
    m->pslot = (uint32_t *)(base + o_pslot);
    m->tval = (uint16_t *)(base + o_tval);
    m->ridx = (uint16_t *)(base + o_ridx);
    m->rprm = (uint16_t *)(base + o_rprm);
    m->psmp = (int16_t *)(base + o_psmp);
    m->ppr = (uint16_t *)(base + o_ppr);
    m->st0 = base + o_st0;
    m->rg = base + o_rg;
    m->rcxt = base + o_rcxt;
    m->pg = base + o_pg;
    m->pcxt = base + o_pcxt;
    m->ub = base + o_ub;
    m->win = base + o_win;
    m->chosen = base + o_chosen;
  }
  return off;
}

// ---------------------------------------------------------------------------
// Model primitives (device)
// ---------------------------------------------------------------------------

// Mixer::p (cucm_compress_v2.cu:636-644) plus the pr bump (:787-788).
__device__ __forceinline__ uint32_t mix_predict(const int32_t *wrow, const int32_t smp[8],
                                                const uint16_t *squash) {
  int64_t dot = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) dot += (int64_t)wrow[k] * (int64_t)smp[k];
  dot >>= 16;
  const int idx = dot < -2047 ? -2047 : (dot > 2048 ? 2048 : (int)dot);
  uint32_t pr = squash[idx + 2047];
  if (pr < 2048) ++pr;
  return pr;
}

// Carry-less coder interval split (cucm_compress_v2.cu:855-857).
__device__ __forceinline__ uint32_t coder_mid(uint32_t x1, uint32_t x2, uint32_t pr) {
  const uint32_t rng = x2 - x1;
  return x1 + __umul24(rng >> 12, pr) + (__umul24(rng & 0xfffu, pr) >> 12);
}

// Inserts `slot` into the order's distinct-slot table; returns its index.
__device__ __forceinline__ uint32_t table_insert(uint32_t *tkey, uint32_t TS, uint32_t slot) {
  uint32_t h = slot & (TS - 1);
  for (;;) {
    const uint32_t old = atomicCAS(&tkey[h], kEmpty, slot);
    if (old == kEmpty || old == slot) return h;
    h = (h + 1) & (TS - 1);
  }
}

// Zeroes (or copies the prelearned model into) the wave's model slot and
// resets the LDS-resident parts of the model.
__device__ void init_model(const CodecArgs &a, const Task &tk, const WaveMem &M, uint8_t *slot,
                           uint32_t Umax, uint32_t lane) {
  const uint64_t bytes = 7ull << tk.tb;
  const uint64_t n16 = bytes >> 4;
  uint4 *dst = (uint4 *)slot;
  if (tk.flags & kTaskInitModel) {
    const DevInitModel *im = a.init + tk.model;
    const uint4 *src = (const uint4 *)im->states;
    for (uint64_t i = lane; i < n16; i += 64) dst[i] = src[i];
    for (uint32_t i = lane; i < 8 * 256; i += 64) M.stp[i] = im->statep[i >> 8][i & 255];
    for (uint32_t i = lane; i < kMixCtx * kOrders; i += 64) M.w[i] = im->w[i];
    for (uint32_t i = lane; i < 256; i += 64) M.st0[i] = im->st0[i];
  } else {
    const uint4 z = make_uint4(0, 0, 0, 0);
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
  global_sync();
  wave_sync();
}

// Applies the deferred updates of one unit (cucm_compress_v2.cu:794-839,
// cucm_decompress_v2.cu:870-896): mixer deltas from the recorded (pr, y,
// cxt) and unit-start predictions, then every order's StateMap updates in
// bit order, then the distinct slots' final states back to HBM.
__device__ void update_unit(const WaveMem &M, const DevTables *T, uint8_t *slot, uint32_t tb,
                            uint32_t Umax, uint32_t nvalid, uint32_t lane) {
  const uint32_t TS = table_size(Umax);
  for (uint32_t r = lane; r < nvalid; r += 64) {
    const uint32_t prm = M.rprm[r];
    const int32_t err = ((int32_t)((prm & 1u) << 12) - (int32_t)(prm >> 1)) * 7;
    const uint32_t cxt = M.rcxt[r];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t g = M.rg[k * Umax + r];
      const int32_t smp = T->stretch[M.stp[k * 256 + g] >> 20];
      const int32_t d = (int32_t)(((int64_t)err * smp) >> 16);
      atomicAdd(&M.w[cxt * 8 + k], d);
    }
  }
  wave_sync();
  if (lane < 8) {
    const uint32_t k = lane;
    uint32_t *sp = M.stp + k * 256;
    const uint16_t *ri = M.ridx + k * Umax;
    const uint8_t *rgk = M.rg + k * Umax;
    uint16_t *tv = M.tval + (k == 0 ? 0 : (k - 1) * TS);
    for (uint32_t r = 0; r < nvalid; ++r) {
      const uint32_t y = M.rprm[r] & 1u;
      const uint32_t i = ri[r];
      uint32_t s;
      if (k == 0) {
        s = M.st0[i];
      } else {
        const uint32_t v = tv[i];
        s = (v > 255u) ? (uint32_t)rgk[r] : v;
      }
      const uint32_t p = sp[s];
      const uint32_t n = p & 1023u;
      const int32_t e = (int32_t)((y << 22) - (p >> 10));
      const uint32_t adj = ((uint32_t)(e >> 3) * (uint32_t)T->dt[n]) & 0xfffffc00u;
      sp[s] = p + adj + (n < 1023u ? 1u : 0u);
      const uint32_t ns = T->nex[2 * s + y];
      if (k == 0)
        M.st0[i] = (uint8_t)ns;
      else
        tv[i] = (uint16_t)ns;
    }
  }
  wave_sync();
  for (uint32_t e = lane; e < 7 * TS; e += 64) {
    const uint32_t key = M.tkey[e];
    if (key != kEmpty) {
      const uint32_t kk = e / TS;
      slot[((uint64_t)kk << tb) + key] = (uint8_t)M.tval[e];
      M.tkey[e] = kEmpty;
      M.tval[e] = kNoState;
    }
  }
  global_sync();
  wave_sync();
}

// ---------------------------------------------------------------------------
// Encoder
// ---------------------------------------------------------------------------

__device__ void encode_task(const CodecArgs &a, const Task &tk, uint32_t t, const WaveMem &M,
                            const DevTables *T, uint8_t *slot, uint32_t Umax, uint32_t lane) {
  const uint32_t U = tk.U, tb = tk.tb, mask = (1u << tb) - 1u;
  const uint32_t TS = table_size(Umax);
  const uint64_t L = tk.L;
  const uint8_t *src = (const uint8_t *)(uintptr_t)tk.src;
  uint8_t *dst = (uint8_t *)(uintptr_t)tk.dst;
  const uint64_t cap = tk.aux;
  init_model(a, tk, M, slot, Umax, lane);
  const uint64_t vbits = 8 * L;
  const uint64_t coded_end = (tk.flags & kTaskTailPad) ? 8ull * tk.S : vbits;
  uint32_t x1 = 0, x2 = 0xffffffffu, acc = 0;
  uint64_t opos = 0;
  bool ovf = false;
  for (uint64_t base = 0; base < coded_end; base += U) {
    const uint32_t nvalid = base < vbits ? (uint32_t)min((uint64_t)U, vbits - base) : 0u;
    if (nvalid) {
      // Stage the 8 history bytes and the unit's bytes (zero outside the chunk).
      const int64_t b0 = (int64_t)(base >> 3) - 8;
      for (uint32_t i = lane; i < 8 + U / 8; i += 64) {
        const int64_t p = b0 + i;
        M.ub[i] = (p >= 0 && p < (int64_t)L) ? src[p] : (uint8_t)0;
      }
      wave_sync();
      for (uint32_t r = lane; r < nvalid; r += 64) {
        const uint32_t bi = r >> 3, j = r & 7u;
        const uint32_t cur = M.ub[8 + bi];
        const uint32_t h0 = (1u << j) | (j ? (cur >> (8 - j)) : 0u);
        uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
        uint32_t sl[8];
#pragma unroll
        for (int k = 1; k < 8; ++k) {
          H = (H ^ (uint32_t)M.ub[8 + bi - k]) * kFnvPrime;
          sl[k] = H & mask;
        }
        uint32_t g[8];
        g[0] = M.st0[h0];
#pragma unroll
        for (int k = 1; k < 8; ++k) g[k] = slot[((uint64_t)(k - 1) << tb) + sl[k]];
        M.ridx[r] = (uint16_t)h0;
        M.rg[r] = (uint8_t)g[0];
#pragma unroll
        for (int k = 1; k < 8; ++k) {
          M.ridx[k * Umax + r] = (uint16_t)table_insert(M.tkey + (k - 1) * TS, TS, sl[k]);
          M.rg[k * Umax + r] = (uint8_t)g[k];
        }
        int32_t smp[8];
        uint32_t active = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          smp[k] = T->stretch[M.stp[k * 256 + g[k]] >> 20];
          active += (smp[k] != 1);
        }
        const uint32_t cxt = active + (uint32_t)(M.ub[8 + bi - 1] >> 5) * 10u;
        const uint32_t pr = mix_predict(M.w + cxt * 8, smp, T->squash);
        const uint32_t y = (cur >> (7 - j)) & 1u;
        M.rprm[r] = (uint16_t)((pr << 1) | y);
        M.rcxt[r] = (uint8_t)cxt;
      }
      wave_sync();
    }
    if (lane == 0) {
      for (uint32_t r = 0; r < U; ++r) {
        if (base + r >= coded_end) break;
        const uint32_t prm = M.rprm[r];
        const uint32_t xmid = coder_mid(x1, x2, prm >> 1);
        if (prm & 1u)
          x2 = xmid;
        else
          x1 = xmid + 1u;
        while (((x1 ^ x2) & 0xff000000u) == 0) {
          acc |= (x2 >> 24) << (8 * (opos & 3));
          ++opos;
          if ((opos & 3) == 0) {
            if (opos <= cap)
              *(uint32_t *)(dst + opos - 4) = acc;
            else
              ovf = true;
            acc = 0;
          }
          x1 <<= 8;
          x2 = (x2 << 8) | 255u;
        }
      }
    }
    if (nvalid) update_unit(M, T, slot, tb, Umax, nvalid, lane);
    wave_sync();
  }
  if (lane == 0) {
    acc |= (x1 >> 24) << (8 * (opos & 3));
    ++opos;
    const uint64_t wstart = (opos - 1) & ~3ull;
    if (wstart + 4 <= cap)
      *(uint32_t *)(dst + wstart) = acc;
    else
      ovf = true;
    a.result[t] = opos | (ovf ? kResultOverflow : 0ull);
  }
  wave_sync();
}

// ---------------------------------------------------------------------------
// Decoder
// ---------------------------------------------------------------------------

__device__ void decode_task(const CodecArgs &a, const Task &tk, uint32_t t, const WaveMem &M,
                            const DevTables *T, uint8_t *slot, uint32_t Umax, uint32_t lane) {
  const uint32_t U = tk.U, tb = tk.tb, mask = (1u << tb) - 1u;
  const uint32_t TS = table_size(Umax);
  const uint32_t D = a.spec_bits, P = (1u << D) - 1u;
  const uint64_t L = tk.L;
  const uint8_t *pay = (const uint8_t *)(uintptr_t)tk.src;
  const uint64_t plen = tk.aux;
  uint8_t *dst = (uint8_t *)(uintptr_t)tk.dst;
  init_model(a, tk, M, slot, Umax, lane);
  uint64_t winbase = 0;
  for (uint32_t i = lane; i < kWindow; i += 64) M.win[i] = (i < plen) ? pay[i] : (uint8_t)255;
  wave_sync();
  uint32_t x1 = 0, x2 = 0xffffffffu, x = 0;
  uint64_t ip = 4;
  if (lane == 0)
    x = ((uint32_t)M.win[0] << 24) | ((uint32_t)M.win[1] << 16) | ((uint32_t)M.win[2] << 8) |
        (uint32_t)M.win[3];
  const uint64_t vbits = 8 * L;
  for (uint64_t base = 0; base < vbits; base += U) {
    const uint32_t nbits = (uint32_t)min((uint64_t)U, vbits - base);
    for (uint32_t rb = 0; rb < nbits; rb += D) {
      // Keep the coder's next bytes in the LDS window (<= 4 bytes per bit).
      const uint64_t ipu = __shfl(ip, 0);
      if (ipu + 4 * D + 8 > winbase + kWindow) {
        winbase = ipu;
        for (uint32_t i = lane; i < kWindow; i += 64)
          M.win[i] = (winbase + i < plen) ? pay[winbase + i] : (uint8_t)255;
        wave_sync();
      }
      const uint32_t j0 = rb & 7u, bi = rb >> 3;
      const uint32_t known = j0 ? ((uint32_t)M.ub[8 + bi] >> (8 - j0)) : 0u;
      // Speculative predictions: path q fixes the next plen(q) bits to pbits(q).
      for (uint32_t it = lane; it < 8 * P; it += 64) {
        const uint32_t q = it >> 3, k = it & 7u;
        const uint32_t plen_q = 31u - __clz(q + 1);
        if (rb + plen_q >= nbits) continue;
        const uint32_t pbits = q + 1u - (1u << plen_q);
        const uint32_t h0 = (1u << (j0 + plen_q)) | (known << plen_q) | pbits;
        uint32_t g, sl;
        if (k == 0) {
          g = M.st0[h0];
          sl = h0;
        } else {
          uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
          for (uint32_t i = 1; i <= k; ++i) H = (H ^ (uint32_t)M.ub[8 + bi - i]) * kFnvPrime;
          sl = H & mask;
          g = slot[((uint64_t)(k - 1) << tb) + sl];
        }
        M.pg[q * 8 + k] = (uint8_t)g;
        M.pslot[q * 8 + k] = sl;
        M.psmp[q * 8 + k] = T->stretch[M.stp[k * 256 + g] >> 20];
      }
      wave_sync();
      for (uint32_t q = lane; q < P; q += 64) {
        const uint32_t plen_q = 31u - __clz(q + 1);
        if (rb + plen_q >= nbits) continue;
        int32_t smp[8];
        uint32_t active = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          smp[k] = M.psmp[q * 8 + k];
          active += (smp[k] != 1);
        }
        const uint32_t cxt = active + (uint32_t)(M.ub[8 + bi - 1] >> 5) * 10u;
        M.ppr[q] = (uint16_t)mix_predict(M.w + cxt * 8, smp, T->squash);
        M.pcxt[q] = (uint8_t)cxt;
      }
      wave_sync();
      const uint32_t nd = min(D, nbits - rb);
      if (lane == 0) {
        uint32_t pre = 0;
        for (uint32_t i = 0; i < nd; ++i) {
          const uint32_t q = (1u << i) - 1u + pre;
          const uint32_t pr = M.ppr[q];
          const uint32_t xmid = coder_mid(x1, x2, pr);
          const uint32_t y = x <= xmid;
          if (y)
            x2 = xmid;
          else
            x1 = xmid + 1u;
          while (((x1 ^ x2) & 0xff000000u) == 0) {
            x1 <<= 8;
            x2 = (x2 << 8) | 255u;
            x = (x << 8) | (uint32_t)M.win[ip - winbase];
            ++ip;
          }
          pre = (pre << 1) | y;
          M.chosen[i] = (uint8_t)q;
          M.rprm[rb + i] = (uint16_t)((pr << 1) | y);
          M.rcxt[rb + i] = M.pcxt[q];
        }
        M.ub[8 + bi] = (uint8_t)(M.ub[8 + bi] | (pre << (8 - j0 - nd)));
      }
      wave_sync();
      if (lane < 8 * nd) {
        const uint32_t i = lane >> 3, k = lane & 7u, r = rb + i;
        const uint32_t q = M.chosen[i];
        M.rg[k * Umax + r] = M.pg[q * 8 + k];
        const uint32_t sl = M.pslot[q * 8 + k];
        M.ridx[k * Umax + r] =
            (uint16_t)(k == 0 ? sl : table_insert(M.tkey + (k - 1) * TS, TS, sl));
      }
      wave_sync();
    }
    update_unit(M, T, slot, tb, Umax, nbits, lane);
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
  wave_sync();
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__device__ __forceinline__ void stage_tables(const DevTables *g, DevTables *s) {
  const uint32_t n = sizeof(DevTables) / 16;
  const uint4 *src = (const uint4 *)g;
  uint4 *dst = (uint4 *)s;
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
}

template <bool kDecode>
__global__ void __launch_bounds__(512) codec_kernel(CodecArgs a, uint32_t Umax) {
  extern __shared__ __align__(16) uint8_t lds[];
  DevTables *T = (DevTables *)lds;
  stage_tables(a.tabs, T);
  const uint32_t wid = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  const uint32_t P = kDecode ? (1u << a.spec_bits) - 1u : 0u;
  const uint32_t wbytes = wave_lds_bytes(Umax, P, kDecode);
  uint8_t *wbase = lds + align16(sizeof(DevTables)) + wid * wbytes;
  WaveMem M;
  wave_lds_bytes(Umax, P, kDecode, &M, wbase);
  const uint64_t gw = (uint64_t)blockIdx.x * (blockDim.x >> 6) + wid;
  uint8_t *slot = a.pool + gw * a.slot_bytes;
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(a.counter, 1u);
    t = __shfl(t, 0);
    if (t >= a.ntasks) break;
    const Task tk = a.tasks[t];
    if (kDecode)
      decode_task(a, tk, t, M, T, slot, Umax, lane);
    else
      encode_task(a, tk, t, M, T, slot, Umax, lane);
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

hipError_t launch_codec(bool decode, const CodecArgs &a, uint32_t Umax, uint32_t blocks,
                        uint32_t waves_per_block, hipStream_t stream) {
  const uint32_t lds = codec_lds_bytes(Umax, a.spec_bits, decode, waves_per_block);
  if (decode) {
    hipError_t e = hipFuncSetAttribute((const void *)codec_kernel<true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(codec_kernel<true>, dim3(blocks), dim3(64 * waves_per_block), lds, stream,
                       a, Umax);
  } else {
    hipError_t e = hipFuncSetAttribute((const void *)codec_kernel<false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(codec_kernel<false>, dim3(blocks), dim3(64 * waves_per_block), lds, stream,
                       a, Umax);
  }
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
