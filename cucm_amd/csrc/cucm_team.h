// cucm_team.h -- the decoder of one chunk run by a team of kTeamWaves waves
// (one workgroup per chunk), for launches with few chunks per CU.
//
// A single wave per chunk is bound by its own instruction issue (one
// instruction per ~4 cycles per wave) and leaves three of the CU's four SIMDs
// idle when there is one chunk per CU.  The decoder's critical path is the
// chain of speculative rounds: gather the round's path states, mix, decode D
// bits, hash and issue the next round.  The team keeps that chain on the
// leader wave and moves everything else off it:
//
//   wave 0 (leader)  paths_finish, the scalar decode of D bits, the next
//                    round's gathers (from precomputed slots when they are
//                    ready), and publishes the round's bits.
//   wave 1           records the previous byte (distinct-slot inserts, unit
//                    start states, mixer deltas into a delta table) and
//                    resolves its states (nex steps).
//   wave 2           walks the statep recurrences of the byte before that
//                    into a second statep copy.
//   wave 3           hashes the byte's second round for every outcome of its
//                    first (2^D x P slot sets), so the leader only selects
//                    and issues.
//
// Predictions inside a unit read the frozen model of the unit start (statep,
// mixer weights, order-0 and HBM states), exactly as the reference's unit
// schedule requires (cucm_decompress_v2.cu:870-896): the walk writes the copy
// `stpn`, mixer deltas go to `wd`, order-0 resolution to `st0r`, and HBM
// states are written back only at the unit end, when the copies are folded
// into the model.  One workgroup barrier per decoded byte orders the waves'
// LDS traffic (an LDS-only barrier: the leader's gathers stay in flight); the only other cross-wave signal is the candidate-ready word,
// which the leader tests once and never waits on (if wave 3 is late the
// leader hashes the round itself), so no wave ever spins.
#ifndef CUCM_TEAM_H
#define CUCM_TEAM_H
#include "cucm_wave.h"

namespace cucm {

constexpr uint32_t kTeamMaxSpec = 4;  // 2^D x (2^D - 1) candidate slot sets fit LDS for D <= 4

// LDS of a team: the chunk's WaveMem (team_round_buffers rounds of path
// buffers) and:
// A stretch value of the decoder's per-unit table with its mixer-context
// flag (stretch != 1, the orders counted as active) in the high half.
CUCM_HDI uint32_t sm_pack(int32_t st) {
  return ((uint32_t)st & 0xffffu) | ((uint32_t)(st != 1) << 16);
}
CUCM_HDI int32_t sm_stretch(uint32_t v) { return (int32_t)(int16_t)(v & 0xffffu); }

struct TeamMem {
  WaveMem M;
  LDS uint32_t *stpn;  // [8][256] statep being walked (predictions read sm, frozen per unit)
  LDS int32_t *wd;     // [80][8] mixer deltas of the unit
  LDS uint8_t *st0r;   // [256] order-0 states being resolved (M.st0 stays frozen)
  LDS uint32_t *sm;    // [8][256] stretch(statep >> 20) of the frozen statep, packed by sm_pack
  LDS uint32_t *pcb;   // [NB][P] pr | cxt << 16 of every path of a round, per path buffer
  LDS uint32_t *cand;  // [2^D][P][8] next round's slots per outcome of the current round
  LDS uint8_t *ubw;    // encoder: [Umax / 8] the walked unit's bytes
  LDS uint16_t *t0;    // encoder: [256] running order-0 state by h0 (kNoState: unit start)
  LDS uint32_t *ctl;   // [16] 0: task index, 1: candidate round id, 2-7: two published
                       // (history lo, history hi, partial byte) states, by round parity
};

// Path buffers (slots and unit-start states per path) for three bytes of
// rounds: the byte wave 1 records, the byte the leader decodes and the first
// round of the next byte.
CUCM_HDI uint32_t team_round_buffers(uint32_t D) { return 3u * (8u / D); }

struct TeamLayout {
  WaveLayout w;
  uint32_t stpn, wd, st0r, sm, pcb, cand, dummy, ctl, total;
};

CUCM_HDI TeamLayout team_layout(uint32_t Umax, uint32_t D) {
  const uint32_t P = (1u << D) - 1u;
  TeamLayout t;
  t.w = wave_layout(Umax, P, true, team_round_buffers(D), true);
  uint32_t off = t.w.total;
  t.stpn = off;
  off = align16(off + 8 * 256 * 4);
  t.wd = off;
  off = align16(off + kMixCtx * kOrders * 4);
  t.st0r = off;
  off = align16(off + 256);
  t.sm = off;
  off = align16(off + 8 * 256 * 4);
  t.pcb = off;
  off = align16(off + team_round_buffers(D) * P * 4);
  t.cand = off;
  off = align16(off + (1u << D) * P * 8 * 4);
  t.dummy = off;
  off = align16(off + kTeamWaves * 64 * 4);
  t.ctl = off;
  off = align16(off + 16 * 4);
  t.total = off;
  return t;
}

CUCM_HD uint32_t team_lds_bytes(uint32_t Umax, uint32_t D) { return team_layout(Umax, D).total; }

// The team's LDS views for wave `wid` (each wave has its own dummy words).
CUCM_DEV TeamMem carve_team(LDS uint8_t *base, uint32_t Umax, uint32_t D, uint32_t wid) {
  const TeamLayout l = team_layout(Umax, D);
  TeamMem x;
  x.M = carve(base, Umax, (1u << D) - 1u, true, team_round_buffers(D), true);
  x.M.dummy = (LDS uint32_t *)(base + l.dummy) + 64 * wid;
  x.stpn = (LDS uint32_t *)(base + l.stpn);
  x.wd = (LDS int32_t *)(base + l.wd);
  x.st0r = base + l.st0r;
  x.sm = (LDS uint32_t *)(base + l.sm);
  x.pcb = (LDS uint32_t *)(base + l.pcb);
  x.cand = (LDS uint32_t *)(base + l.cand);
  x.ubw = nullptr;
  x.t0 = nullptr;
  x.ctl = (LDS uint32_t *)(base + l.ctl);
  return x;
}

// init_model for a team: thread tid of kTeamWaves * 64; also resets the unit
// copies and the control words.  Followed by a team_sync.
CUCM_DEV void team_init_model(const CodecArgs &a, const Task &tk, const TeamMem &X,
                              const LDS DevTables *T, GLB uint8_t *slot, uint32_t Umax,
                              uint32_t tid) {
  constexpr uint32_t NT = kTeamWaves * 64;
  const WaveMem &M = X.M;
  const uint64_t n16 = (7ull << tk.tb) >> 4;
  GLB u32x4 *dst = (GLB u32x4 *)slot;
  if (tk.flags & kTaskInitModel) {
    const DevInitModel *im = a.init + tk.model;
    const GLB u32x4 *src = (const GLB u32x4 *)im->states;
    for (uint64_t i = tid; i < n16; i += NT) dst[i] = src[i];
    for (uint32_t i = tid; i < 8 * 256; i += NT) {
      const uint32_t v = im->statep[i >> 8][i & 255];
      M.stp[i] = v;
      X.stpn[i] = v;
      X.sm[i] = sm_pack(T->stretch[v >> 20]);
    }
    for (uint32_t i = tid; i < kMixCtx * kOrders; i += NT) M.w[i] = im->w[i];
    for (uint32_t i = tid; i < 256; i += NT) {
      M.st0[i] = im->st0[i];
      X.st0r[i] = im->st0[i];
    }
  } else {
    const u32x4 z = {0u, 0u, 0u, 0u};
    for (uint64_t i = tid; i < n16; i += NT) dst[i] = z;
    for (uint32_t i = tid; i < 8 * 256; i += NT) {
      M.stp[i] = 1u << 31;
      X.stpn[i] = 1u << 31;
      X.sm[i] = sm_pack(T->stretch[(1u << 31) >> 20]);
    }
    for (uint32_t i = tid; i < kMixCtx * kOrders; i += NT) M.w[i] = 0;
    for (uint32_t i = tid; i < 256; i += NT) {
      M.st0[i] = 0;
      X.st0r[i] = 0;
    }
  }
  for (uint32_t i = tid; i < kMixCtx * kOrders; i += NT) X.wd[i] = 0;
  const uint32_t TS = table_size(Umax);
  for (uint32_t i = tid; i < 7 * TS; i += NT) {
    M.tkey[i] = kEmpty;
    M.tval[i] = kNoState;
  }
  for (uint32_t i = tid; i < 8 + Umax / 8 + 8; i += NT) M.ub[i] = 0;
  if (tid < 16) X.ctl[tid] = tid == 1 ? kEmpty : 0u;
}

// init_model for the encoder team (thread tid of kTeamWaves * 64).
CUCM_DEV void team_init_enc(const CodecArgs &a, const Task &tk, const TeamMem &X,
                            GLB uint8_t *slot, uint32_t Umax, uint32_t tid) {
  constexpr uint32_t NT = kTeamWaves * 64;
  const WaveMem &M = X.M;
  const uint64_t n16 = (7ull << tk.tb) >> 4;
  GLB u32x4 *dst = (GLB u32x4 *)slot;
  if (tk.flags & kTaskInitModel) {
    const DevInitModel *im = a.init + tk.model;
    const GLB u32x4 *src = (const GLB u32x4 *)im->states;
    for (uint64_t i = tid; i < n16; i += NT) dst[i] = src[i];
    for (uint32_t i = tid; i < 8 * 256; i += NT) M.stp[i] = im->statep[i >> 8][i & 255];
    for (uint32_t i = tid; i < kMixCtx * kOrders; i += NT) M.w[i] = im->w[i];
    for (uint32_t i = tid; i < 256; i += NT) M.st0[i] = im->st0[i];
  } else {
    const u32x4 z = {0u, 0u, 0u, 0u};
    for (uint64_t i = tid; i < n16; i += NT) dst[i] = z;
    for (uint32_t i = tid; i < 8 * 256; i += NT) M.stp[i] = 1u << 31;
    for (uint32_t i = tid; i < kMixCtx * kOrders; i += NT) M.w[i] = 0;
    for (uint32_t i = tid; i < 256; i += NT) M.st0[i] = 0;
  }
  const uint32_t TS = table_size(Umax);
  for (uint32_t i = tid; i < 7 * TS; i += NT) {
    M.tkey[i] = kEmpty;
    M.tval[i] = kNoState;
  }
  for (uint32_t i = tid; i < Umax; i += NT) M.rprm[i] = 0;
  for (uint32_t i = tid; i < 8 + Umax / 8 + 8; i += NT) M.ub[i] = 0;
  for (uint32_t i = tid; i < 256; i += NT) X.t0[i] = kNoState;
}

// The leader's gathers of a round from slot sets wave 3 precomputed for
// outcome o of the previous round (same loads and pslot record as
// paths_issue).
template <int NM, uint32_t P>
CUCM_DEV void paths_issue_cand(const WaveMem &M, const LDS uint32_t *cand, uint32_t o,
                               const GLB uint8_t *slot, uint32_t tb, uint32_t buf, uint32_t lane,
                               uint32_t g[NM][8]) {
  LDS uint32_t *ps = M.pslot + buf * (P * 8u);
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const uint32_t q = lane + 64u * m;
    const bool ok = q < P;
    const LDS u32x4 *c = (const LDS u32x4 *)(cand + (o * P + (ok ? q : 0u)) * 8u);
    const u32x4 c0 = c[0], c1 = c[1];
    const uint32_t sl[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    g[m][0] = M.st0[sl[0]];
    // lanes without a path load path 0's slots (in bounds, same lines); the
    // offset (k - 1) * 2^tb + slot < 7 * 2^25 fits 32 bits
#pragma unroll
    for (int k = 1; k < 8; ++k) g[m][k] = slot[((uint32_t)(k - 1) << tb) | sl[k]];
    if (ok) {
      LDS u32x4 *d = (LDS u32x4 *)(ps + q * 8u);
      d[0] = c0;
      d[1] = c1;
    }
  }
}

// paths_issue of a byte's first round (j0 = 0) in two parts: the gathers go
// out before the leader's byte barrier, the LDS part (order-0 state, slot
// record) follows it, so the barrier's LDS wait is only the byte's own
// stores.
template <int NM, uint32_t P>
CUCM_DEV void paths_issue_gathers(const GLB uint8_t *slot, uint32_t tb, uint32_t mask,
                                  uint64_t hist, uint32_t lane, uint32_t g[NM][8],
                                  uint32_t sl[NM][8]) {
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    // lanes without a path repeat path 0 (its lines, in bounds)
    const uint32_t q = lane + 64u * m < P ? lane + 64u * m : 0u;
    const uint32_t plen = 31u - clz32(q + 1u);
    const uint32_t pbits = q + 1u - (1u << plen);
    const uint32_t h0 = ((1u << plen) | pbits) & 255u;
    uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
    sl[m][0] = h0;
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      sl[m][k] = fnv_pos(H, (uint32_t)((hist >> (8 * (k - 1))) & 255u), mask);
    }
#pragma unroll
    for (int k = 1; k < 8; ++k) g[m][k] = slot[((uint32_t)(k - 1) << tb) | sl[m][k]];
  }
}
template <int NM, uint32_t P>
CUCM_DEV void paths_issue_record(const WaveMem &M, uint32_t buf, uint32_t lane,
                                 const uint32_t sl[NM][8], uint32_t g[NM][8]) {
  LDS uint32_t *ps = M.pslot + buf * (P * 8u);
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const uint32_t q = lane + 64u * m;
    g[m][0] = M.st0[sl[m][0]];
    if (q < P) {
      LDS u32x4 *d = (LDS u32x4 *)(ps + q * 8u);
      d[0] = u32x4{sl[m][0], sl[m][1], sl[m][2], sl[m][3]};
      d[1] = u32x4{sl[m][4], sl[m][5], sl[m][6], sl[m][7]};
    }
  }
}

// paths_finish for the team, in two parts.  The mix: the stretch of each
// order's statep comes from the unit's sm table in one lookup; pc[m] = pr |
// cxt << 16 of path q.  The record (the round's unit-start states and pc
// into the pg / pcb buffers for wave 1) is left for after the decode and the
// next issue, off the round chain.
template <int NM, uint32_t P>
CUCM_DEV void paths_mix_team(const TeamMem &X, const LDS DevTables *T, uint32_t c1,
                             const uint32_t g[NM][8], uint32_t pc[NM]) {
  const WaveMem &M = X.M;
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    int32_t smp[8];
    uint32_t sv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      sv[k] = X.sm[k * 256 + g[m][k]];
      smp[k] = sm_stretch(sv[k]);
    }
    // the active orders (stretch != 1) are the sum of the high halves
    const uint32_t active = ((sv[0] >> 16) + (sv[1] >> 16)) + ((sv[2] >> 16) + (sv[3] >> 16)) +
                            ((sv[4] >> 16) + (sv[5] >> 16)) + ((sv[6] >> 16) + (sv[7] >> 16));
    const uint32_t cxt = active + c1 * 10u;
    const LDS u32x4 *wr = (const LDS u32x4 *)(M.w + cxt * 8u);
    const u32x4 wa = wr[0], wb = wr[1];
    const int32_t w[8] = {(int32_t)wa.x, (int32_t)wa.y, (int32_t)wa.z, (int32_t)wa.w,
                          (int32_t)wb.x, (int32_t)wb.y, (int32_t)wb.z, (int32_t)wb.w};
    pc[m] = mix_weights(w, smp, T->squash) | (cxt << 16);
  }
}
template <int NM, uint32_t P>
CUCM_DEV void paths_record_team(const TeamMem &X, uint32_t buf, uint32_t lane,
                                const uint32_t g[NM][8], const uint32_t pc[NM]) {
  LDS uint8_t *pgb = X.M.pg + buf * (P * 8u);
#pragma unroll
  for (int m = 0; m < NM; ++m) {
    const uint32_t q = lane + 64u * m;
    if (q < P) {
      const uint32_t lo = g[m][0] | (g[m][1] << 8) | (g[m][2] << 16) | (g[m][3] << 24);
      const uint32_t hi = g[m][4] | (g[m][5] << 8) | (g[m][6] << 16) | (g[m][7] << 24);
      *(LDS uint64_t *)(pgb + q * 8u) = ((uint64_t)hi << 32) | lo;
      X.pcb[buf * P + q] = pc[m];
    }
  }
}

// Wave 3: slot sets of the round after the current one, for every outcome o
// of the current round's D bits (the current round starts at bit j0 of its
// byte with `partial` decided and `hist` the previous bytes).
template <uint32_t D>
CUCM_DEV void team_candidates(LDS uint32_t *cand, uint32_t mask, uint32_t j0, uint32_t partial,
                              uint64_t hist, uint32_t lane) {
  constexpr uint32_t P = (1u << D) - 1u;
  constexpr uint32_t NC = (1u << D) * P;
  for (uint32_t c = lane; c < NC; c += 64) {
    const uint32_t o = c / P, q = c - o * P;
    uint32_t pn = (partial << D) | o;
    uint64_t hn = hist;
    uint32_t jn = j0 + D;
    if (jn == 8) {
      hn = (hist << 8) | pn;
      pn = 0;
      jn = 0;
    }
    const uint32_t plen = 31u - clz32(q + 1u);
    const uint32_t pbits = q + 1u - (1u << plen);
    const uint32_t h0 = ((1u << (jn + plen)) | (pn << plen) | pbits) & 255u;
    uint32_t H = (kFnvBasis ^ h0) * kFnvPrime;
    uint32_t sl[8];
    sl[0] = h0;
#pragma unroll
    for (int k = 1; k < 8; ++k) {
      sl[k] = fnv_pos(H, (uint32_t)((hn >> (8 * (k - 1))) & 255u), mask);
    }
    LDS u32x4 *d = (LDS u32x4 *)(cand + c * 8u);
    d[0] = u32x4{sl[0], sl[1], sl[2], sl[3]};
    d[1] = u32x4{sl[4], sl[5], sl[6], sl[7]};
  }
}

// Wave 1: records the 8 bits of byte b from the leader's round buffers
// (slot and unit-start state per order, distinct-slot insert; lane = bit x 8
// + order) and adds their mixer deltas (cucm_decompress_v2.cu:870-883) to
// wd.  Round r of the unit used path buffer (ubuf + r) mod NB; the path a
// bit took, its y and its (pr, cxt) follow from the decoded byte.
template <uint32_t D>
CUCM_DEV void team_record_byte(const TeamMem &X, const LDS DevTables *T, uint32_t b,
                               uint32_t ubuf, uint32_t Umax, uint32_t lane) {
  constexpr uint32_t P = (1u << D) - 1u;
  constexpr uint32_t NB = 3u * (8u / D);
  const WaveMem &M = X.M;
  const uint32_t TS = table_size(Umax);
  const uint32_t i = lane >> 3, k = lane & 7u;
  const uint32_t r = 8 * b + i;
  const uint32_t jr = i / D, t = i - jr * D;  // round of the byte, bit of the round
  const uint32_t buf = (ubuf + (8 * b) / D + jr) % NB;
  const uint32_t byte = M.ub[8 + b];
  const uint32_t o = (byte >> (8 - (jr + 1) * D)) & P;  // the round's outcome
  const uint32_t q = (1u << t) - 1u + (o >> (D - t));
  const uint32_t y = (o >> (D - 1 - t)) & 1u;
  const uint32_t pcv = X.pcb[buf * P + q];
  const uint32_t gq = M.pg[(buf * P + q) * 8 + k];
  const uint32_t sl = M.pslot[(buf * P + q) * 8 + k];
  const bool ins = k != 0;
  uint32_t h = sl & (TS - 1);
  LDS uint32_t *tkk = M.tkey + (k ? (k - 1) * TS : 0);
  uint32_t old = lds_cas((ins ? &tkk[h] : &M.dummy[lane]), kEmpty, sl);
  if (wave_ballot(ins && old != kEmpty && old != sl)) {
    if (ins)
      while (old != kEmpty && old != sl) {
        h = (h + 1) & (TS - 1);
        old = lds_cas(&tkk[h], kEmpty, sl);
      }
  }
  M.rg[k * Umax + r] = (uint8_t)gq;
  M.ridx[k * Umax + r] = (uint16_t)(k == 0 ? sl : h);
  const int32_t err = ((int32_t)(y << 12) - (int32_t)(pcv & 0xffffu)) * 7;
  const int32_t smp = sm_stretch(X.sm[k * 256 + gq]);
  lds_add(&X.wd[(pcv >> 16) * 8u + k], (int32_t)(((int64_t)err * smp) >> 16));
  wave_sync();
}

// Wave 1: the state each StateMap update of byte b sees (the resolution step
// of update_states for one byte): lane (k, j) = (order, bit of the byte);
// order 0 runs on st0r, orders 1..7 on the distinct-slot tables.
CUCM_DEV void team_resolve_byte(const TeamMem &X, const LDS DevTables *T, uint32_t b,
                                uint32_t Umax, uint32_t lane) {
  const WaveMem &M = X.M;
  const uint32_t TS = table_size(Umax);
  const uint32_t k = lane >> 3, j = lane & 7u;
  const uint32_t r = 8 * b + j;
  const uint32_t y = ((uint32_t)M.ub[8 + b] >> (7 - j)) & 1u;
  const uint32_t i = M.ridx[k * Umax + r];
  const uint32_t g = M.rg[k * Umax + r];
  // two bits of this byte on one slot of an order (a hash collision): the
  // byte is stepped bit by bit
  bool coll = false;
  if (k != 0) {
    const u32x4 v = *(const LDS u32x4 *)(M.ridx + k * Umax + 8 * b);
    const uint32_t e[8] = {v.x & 0xffffu, v.x >> 16, v.y & 0xffffu, v.y >> 16,
                           v.z & 0xffffu, v.z >> 16, v.w & 0xffffu, v.w >> 16};
#pragma unroll
    for (int z = 0; z < 8; ++z) coll |= ((uint32_t)z != j) && (e[z] == i);
  }
  const uint32_t reps = wave_ballot(coll) ? 8u : 1u;
  LDS uint16_t *tv = M.tval + (k ? (k - 1) * TS : 0);
  for (uint32_t q = 0; q < reps; ++q) {
    if (reps == 1u || q == j) {
      uint32_t st;
      if (k == 0) {
        st = X.st0r[i];
      } else {
        const uint32_t v = tv[i];
        st = v > 255u ? g : v;
      }
      const uint32_t ns = T->nex[2 * st + y];
      if (k == 0)
        X.st0r[i] = (uint8_t)ns;
      else
        tv[i] = (uint16_t)ns;
      M.rg[k * Umax + r] = (uint8_t)st;
    }
    wave_sync();
  }
}

// Wave 2: the statep walk of update_states' second part for byte b (its
// bits y MSB first in ybyte) on stpn (lane k < 8 walks order k; callers run
// it on those 8 lanes only).  The walk restarts at every byte: the entries
// of its first two bits are read after the previous byte's writes, so it
// never looks into a byte that is not resolved yet.
#define CUCM_TWALK_BYTE(X, T, b, ybyte, Umax, lane)                           \
  do {                                                                        \
    const uint32_t wk_ = min((lane), 7u);                                     \
    w_sp = (X).stpn + wk_ * 256;                                              \
    w_dummy = (X).M.dummy + (lane);                                           \
    w_dtab = (T)->dt;                                                         \
    w_writer = (lane) < 8;                                                    \
    w_cur = *(const LDS uint64_t *)((X).M.rg + wk_ * (Umax) + 8 * (b));        \
    w_nxt = w_cur;                                                            \
    w_s0 = (uint32_t)w_cur & 255u;                                            \
    w_s1 = (uint32_t)(w_cur >> 8) & 255u;                                     \
    w_sprev = 0x100u;                                                         \
    w_p = w_sp[w_s0];                                                         \
    w_raw1 = w_sp[w_s1];                                                      \
    w_dA = w_dtab[w_p & 1023u];                                               \
    w_dB = w_dC = 0;                                                          \
    w_f1 = true;                                                              \
    w_f2 = false;                                                             \
    w_pnprev = w_dfprev = 0;                                                  \
    const uint32_t wyb_ = (ybyte);                                            \
    CUCM_WALK_STEP(0, (wyb_ >> 7) & 1u);                                      \
    CUCM_WALK_STEP(1, (wyb_ >> 6) & 1u);                                      \
    CUCM_WALK_STEP(2, (wyb_ >> 5) & 1u);                                      \
    CUCM_WALK_STEP(3, (wyb_ >> 4) & 1u);                                      \
    CUCM_WALK_STEP(4, (wyb_ >> 3) & 1u);                                      \
    CUCM_WALK_STEP(5, (wyb_ >> 2) & 1u);                                      \
    CUCM_WALK_STEP(6, (wyb_ >> 1) & 1u);                                      \
    CUCM_WALK_STEP(7, wyb_ & 1u);                                             \
  } while (0)

// The touched slots of up to three orders ks[0..no) (unit bits r < nbits, a
// multiple of 8) written back to HBM from the distinct-slot tables, then
// cleared there: every slot index of the wave goes out together, then every
// table entry, then the stores, so the wave waits for LDS twice in all.
// Lanes past nbits repeat the last bit's loads (in bounds) and store nothing.
CUCM_DEV void writeback_orders(const WaveMem &M, GLB uint8_t *slot, uint32_t tb, uint32_t nbits,
                               uint32_t Umax, const uint32_t ks[3], uint32_t no,
                               uint32_t lane) {
  constexpr int NR = (int)(kMaxUnitBits / 64);
  const uint32_t TS = table_size(Umax);
  uint32_t idx[3][NR], key[3][NR], val[3][NR];
#pragma unroll
  for (int o = 0; o < 3; ++o)
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const uint32_t k = o < (int)no ? ks[o] : ks[0];
      idx[o][q] = (k - 1) * TS + M.ridx[k * Umax + min(lane + 64u * q, nbits - 1)];
    }
#pragma unroll
  for (int o = 0; o < 3; ++o)
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      key[o][q] = M.tkey[idx[o][q]];
      val[o][q] = M.tval[idx[o][q]];
    }
#pragma unroll
  for (int o = 0; o < 3; ++o)
#pragma unroll
    for (int q = 0; q < NR; ++q)
      if (o < (int)no && lane + 64u * q < nbits)
        slot[((ks[o] - 1) << tb) | key[o][q]] = (uint8_t)val[o][q];
  wave_sync();
#pragma unroll
  for (int o = 0; o < 3; ++o)
#pragma unroll
    for (int q = 0; q < NR; ++q)
      if (o < (int)no && lane + 64u * q < nbits) {
        M.tkey[idx[o][q]] = kEmpty;
        M.tval[idx[o][q]] = kNoState;
      }
}

// Unit-end part run by waves 0, 1 and 3 (part 0, 1, 2): the touched slots of
// a third of the orders written back to HBM and the tables cleared, a third
// of the mixer deltas folded into the weights.
CUCM_DEV void team_fold(const TeamMem &X, GLB uint8_t *slot, uint32_t tb, uint32_t nbits,
                        uint32_t Umax, uint32_t part, uint32_t lane) {
  const WaveMem &M = X.M;
  const uint32_t ks[3] = {1 + part, 4 + part, 7 + part};
  writeback_orders(M, slot, tb, nbits, Umax, ks, part == 0 ? 3u : 2u, lane);
  // weights e = part * 64 + lane + 192 q < 640
  int32_t w[4], d[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t e = min(part * 64 + lane + 192u * q, (uint32_t)(kMixCtx * kOrders - 1));
    w[q] = M.w[e];
    d[q] = X.wd[e];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t e = part * 64 + lane + 192u * q;
    if (e < kMixCtx * kOrders) {
      M.w[e] = w[q] + d[q];
      X.wd[e] = 0;
    }
  }
}

// The walked statep becomes the next unit's frozen model (read through sm);
// wave wid does its quarter, all of its loads in flight together.
CUCM_DEV void team_refresh_sm(const TeamMem &X, const LDS DevTables *T, uint32_t wid,
                              uint32_t lane) {
  constexpr int NQ = (int)(8 * 256 / (kTeamWaves * 64));
  uint32_t v[NQ];
  int32_t st[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) v[q] = X.stpn[wid * 64 + lane + kTeamWaves * 64 * q];
#pragma unroll
  for (int q = 0; q < NQ; ++q) st[q] = T->stretch[v[q] >> 20];
#pragma unroll
  for (int q = 0; q < NQ; ++q) X.sm[wid * 64 + lane + kTeamWaves * 64 * q] = sm_pack(st[q]);
}

// Every role runs the same barrier sequence per unit: one LDS barrier after
// each of the nblk + 1 iterations, a full barrier after the unit-end fold
// (write-backs land before the next unit's gathers), an LDS barrier after the
// statep refresh.  The roles are separate code paths so each keeps only its
// own state in registers.

// Wave 0: the speculative round chain (iteration b decodes byte b).
template <int NJ, int D>
CUCM_DEV void team_leader(const CodecArgs &a, const Task &tk, uint32_t t, const TeamMem &X,
                          const LDS DevTables *T, GLB uint8_t *slot, uint32_t Umax,
                          uint32_t lane, uint32_t rid, PhaseClock &clk) {
  constexpr uint32_t P = (1u << D) - 1u;
  constexpr uint32_t RB = 8u / D;
  constexpr uint32_t NB = 3u * RB;
  constexpr int NM = 1;
  const WaveMem &M = X.M;
  const uint32_t U = tk.U, tb = tk.tb, mask = (1u << tb) - 1u;
  const GLB uint8_t *pay = (const GLB uint8_t *)(uintptr_t)tk.src;
  const uint32_t plen = tk.aux;
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  const uint32_t L = tk.L;
  uint32_t winbase = 0;
  uint32_t win = load_window(pay, plen, 0, lane);
  uint32_t x = lane_read(win, 0), x1 = 0, x2 = 0xffffffffu;
  uint32_t ip = 4;
  uint64_t hist = 0;  // the 8 bytes before the current one, latest lowest
  // gathered states, double buffered: round j of a byte mixes g[j & 1] (RB
  // is even) while the next round's gathers land in the other half, so a
  // round's record can follow the next issue
  static_assert(RB % 2 == 0, "rounds per byte");
  uint32_t g[2][NM][8];
  uint32_t ubuf = 0, it = 0;
  if (L) paths_issue<NM, P>(M, slot, tb, mask, 0, 0, hist, ubuf, lane, g[0]);
  for (uint32_t base = 0; base < L; base += U / 8) {
    const uint32_t nblk = min(U / 8, L - base), R = nblk * RB;
    for (uint32_t b = 0; b <= nblk; ++b, ++it) {
      uint32_t sln[NM][8];  // the next byte's first-round slots (recorded after the barrier)
      bool next = false;
      if (b < nblk) {
        uint32_t partial = 0;
#pragma unroll
        for (uint32_t j = 0; j < RB; ++j) {
          const uint32_t r = b * RB + j;
          const uint32_t buf = (ubuf + r) % NB, nbuf = (ubuf + r + 1) % NB;
          if (ip + 4 * D + 8 > winbase + 256) {
            winbase = ip & ~3u;
            win = load_window(pay, plen, winbase, lane);
          }
          uint32_t pc[NM];
          paths_mix_team<NM, P>(X, T, (uint32_t)(hist & 255u) >> 5, g[j & 1], pc);
          clk.mark(9);
          // wave 3's candidate-ready word: loaded before the decode, so its
          // latency overlaps it, and acquired only where it is used
          const uint32_t cflag = (j == 0 && RB > 1) ? flag_peek(&X.ctl[1]) : 0u;
          const uint32_t prv = (pc[0] & 0xffffu) << 20;  // pr << 20 (coder_mid20)
          // q: the path of the next bit, 2^i - 1 + (the round's bits so far)
          uint32_t q = 0;
#pragma unroll
          for (int i = 0; i < D; ++i) {
            const uint32_t pr = lane_read(prv, q);
            const uint32_t xmid = coder_mid20(x1, x2, pr);
            coder_split_path(x, xmid, x1, x2, q);
            // (the refill is the rare case: its code goes out of line)
            if (__builtin_expect(((x1 ^ x2) & 0xff000000u) == 0, 0)) {
              // shift out every leading byte x1 and x2 agree on (at most 4)
              const uint32_t nb = min((uint32_t)clz32(x1 ^ x2) >> 3, 4u);
              const uint32_t sh = 8 * nb;
              x1 = (uint32_t)((uint64_t)x1 << sh);
              x2 = (uint32_t)(((uint64_t)x2 << sh) | ((1ull << sh) - 1ull));
              const uint32_t nxt = window_be32(win, ip - winbase);
              x = (uint32_t)((((uint64_t)x << 32) | nxt) >> (32 - sh));
              ip += nb;
            }
          }
          const uint32_t pre = q - P;  // the round's D bits
          clk.mark(10);
          partial = (partial << D) | pre;
          if (j + 1 == RB) {
            if (lane == 0) M.ub[8 + b] = (uint8_t)partial;
            hist = (hist << 8) | partial;
          }
          if (r + 1 < R) {
            // wave 3's slot sets for this outcome, if they are complete
            if (j == 0 && RB > 1 && lane_first(cflag) == rid + r + 1) {
              flag_acquire();  // the candidates written before the word are visible
              paths_issue_cand<NM, P>(M, X.cand, pre, slot, tb, nbuf, lane, g[(j + 1) & 1]);
            } else if (j + 1 == RB) {
              paths_issue_gathers<NM, P>(slot, tb, mask, hist, lane, g[0], sln);
              next = true;
            } else {
              paths_issue<NM, P>(M, slot, tb, mask, (j + 1) * D, partial, hist, nbuf, lane,
                                 g[(j + 1) & 1]);
            }
          }
          paths_record_team<NM, P>(X, buf, lane, g[j & 1], pc);
          clk.mark(1);
        }
      }
      // the history before the next byte, for wave 3
      if (lane == 0) {
        LDS uint32_t *pub = X.ctl + 2 + 3 * (it & 1u);
        pub[0] = (uint32_t)hist;
        pub[1] = (uint32_t)(hist >> 32);
      }
      if (b == nblk)
        for (uint32_t i = lane; i < nblk; i += 64) dst[base + i] = M.ub[8 + i];
      team_sync_lds();
      if (next) paths_issue_record<NM, P>(M, (ubuf + (b + 1) * RB) % NB, lane, sln, g[0]);
      clk.mark(4);
    }
    team_fold(X, slot, tb, 8 * nblk, Umax, 0, lane);
    for (uint32_t i = lane; i < 256; i += 64) M.st0[i] = X.st0r[i];
    clk.mark(8);
    team_sync();
    team_refresh_sm(X, T, 0, lane);
    rid += R;
    ubuf = (ubuf + R) % NB;
    if (base + U / 8 < L) paths_issue<NM, P>(M, slot, tb, mask, 0, 0, hist, ubuf, lane, g[0]);
    team_sync_lds();
    clk.mark(8);
  }
  if (lane == 0) a.result[t] = ip;
}

// Wave 1: records and resolves byte b - 1 in iteration b.
template <int D>
CUCM_DEV void team_recorder(const Task &tk, const TeamMem &X, const LDS DevTables *T,
                            GLB uint8_t *slot, uint32_t Umax, uint32_t lane, PhaseClock &clk) {
  constexpr uint32_t RB = 8u / D;
  constexpr uint32_t NB = 3u * RB;
  const uint32_t U = tk.U, L = tk.L;
  uint32_t ubuf = 0;
  for (uint32_t base = 0; base < L; base += U / 8) {
    const uint32_t nblk = min(U / 8, L - base);
    for (uint32_t b = 0; b <= nblk; ++b) {
      if (b >= 1) {
        team_record_byte<D>(X, T, b - 1, ubuf, Umax, lane);
        clk.mark(5);
        team_resolve_byte(X, T, b - 1, Umax, lane);
        clk.mark(6);
      }
      team_sync_lds();
      clk.mark(3);
    }
    team_fold(X, slot, tk.tb, 8 * nblk, Umax, 1, lane);
    clk.mark(8);
    team_sync();
    team_refresh_sm(X, T, 1, lane);
    ubuf = (ubuf + nblk * RB) % NB;
    team_sync_lds();
  }
}

// Wave 2: walks byte b - 2 in iteration b, the last byte at the unit end.
template <int D>
CUCM_DEV void team_walker(const Task &tk, const TeamMem &X, const LDS DevTables *T,
                          uint32_t Umax, uint32_t lane, PhaseClock &clk) {
  const uint32_t U = tk.U, L = tk.L;
  CUCM_WALK_STATE;
  (void)w_sv, (void)w_nblk, (void)w_blk;  // (the per-byte walk uses none of these)
  for (uint32_t base = 0; base < L; base += U / 8) {
    const uint32_t nblk = min(U / 8, L - base);
    for (uint32_t b = 0; b <= nblk; ++b) {
      if (b >= 2 && lane < 8) CUCM_TWALK_BYTE(X, T, b - 2, X.M.ub[8 + b - 2], Umax, lane);
      clk.mark(7);
      team_sync_lds();
      clk.mark(3);
    }
    if (lane < 8) CUCM_TWALK_BYTE(X, T, nblk - 1, X.M.ub[8 + nblk - 1], Umax, lane);
    clk.mark(7);
    team_sync();
    team_refresh_sm(X, T, 2, lane);
    team_sync_lds();
  }
}

// Wave 3: in iteration b, the candidates of byte b's second round for every
// outcome of its first, from the history the leader published in the
// previous iteration.
template <int D>
CUCM_DEV void team_hasher(const Task &tk, const TeamMem &X, const LDS DevTables *T,
                          GLB uint8_t *slot, uint32_t Umax, uint32_t lane, uint32_t rid,
                          PhaseClock &clk) {
  constexpr uint32_t RB = 8u / D;
  const uint32_t U = tk.U, L = tk.L, mask = (1u << tk.tb) - 1u;
  uint32_t it = 0;
  for (uint32_t base = 0; base < L; base += U / 8) {
    const uint32_t nblk = min(U / 8, L - base);
    for (uint32_t b = 0; b <= nblk; ++b, ++it) {
      if (RB > 1 && b < nblk) {
        const LDS uint32_t *pub = X.ctl + 2 + 3 * ((it - 1) & 1u);
        const uint64_t h = (uint64_t)pub[0] | ((uint64_t)pub[1] << 32);
        team_candidates<D>(X.cand, mask, 0, 0, h, lane);
        wave_sync();
        if (lane == 0) flag_publish(&X.ctl[1], rid + b * RB + 1);
      }
      clk.mark(2);
      team_sync_lds();
      clk.mark(3);
    }
    team_fold(X, slot, tk.tb, 8 * nblk, Umax, 2, lane);
    clk.mark(8);
    team_sync();
    team_refresh_sm(X, T, 3, lane);
    rid += nblk * RB;
    team_sync_lds();
  }
}

// One chunk decoded by the team (all kTeamWaves waves call this together):
// the leader decodes byte b while wave 1 records and resolves byte b - 1,
// wave 2 walks byte b - 2 and wave 3 hashes the candidates of byte b's
// second round.  `rid` numbers rounds across the workgroup's chunks.
template <int NJ, int D, bool kProf>
CUCM_DEV void decode_team(const CodecArgs &a, const Task &tk, uint32_t t, const TeamMem &X,
                          const LDS DevTables *T, GLB uint8_t *slot, uint32_t Umax, uint32_t wid,
                          uint32_t lane, uint32_t &rid) {
  static_assert(D >= 1 && D <= (int)kTeamMaxSpec && (8 % D) == 0, "team speculation width");
  // (phase clocks only in the profiling instantiation: their registers
  // would otherwise crowd the leader's)
  PhaseClock clk(kProf && a.prof != nullptr);
  team_init_model(a, tk, X, T, slot, Umax, wid * 64 + lane);
  team_sync();
  clk.mark(0);
  if (wid == 0)
    team_leader<NJ, D>(a, tk, t, X, T, slot, Umax, lane, rid, clk);
  else if (wid == 1)
    team_recorder<D>(tk, X, T, slot, Umax, lane, clk);
  else if (wid == 2)
    team_walker<D>(tk, X, T, Umax, lane, clk);
  else
    team_hasher<D>(tk, X, T, slot, Umax, lane, rid, clk);
  // every wave counts the chunk's rounds
  const uint32_t U = tk.U;
  for (uint32_t base = 0; base < tk.L; base += U / 8) rid += min(U / 8, tk.L - base) * (8u / D);
  clk.flush(a.prof, lane);
}

// ---------------------------------------------------------------------------
// Team encoder: one chunk per workgroup of kTeamWaves waves.  Per unit u:
//   P1  waves 0, 1 stage, hash, gather and insert unit u (half the bits
//       each); wave 2 walks unit u - 1's statep recurrences; wave 3 codes
//       the second half of unit u - 1's bits;
//   P2  waves 0, 1 predict unit u, keeping each bit's mixer deltas in
//       registers;
//   P3  waves 0, 1 add the deltas; wave 1 resolves the unit's states byte by
//       byte; wave 2 keeps the unit's bytes for its walk; wave 3 codes the
//       first half of unit u's bits;
//   P4  all four write the touched slots back and clear the tables (a full
//       barrier: the write-backs land before unit u + 1's gathers).
// The model is updated in place: nothing reads statep during P1 but the
// walk, and predictions (P2) come after it.
// ---------------------------------------------------------------------------

struct TeamEncLayout {
  WaveLayout w;
  uint32_t ubw, t0, dummy, ctl, total;
};

CUCM_HDI TeamEncLayout team_enc_layout(uint32_t Umax) {
  TeamEncLayout t;
  t.w = wave_layout(Umax, 0, false, 2, true);
  uint32_t off = t.w.total;
  t.ubw = off;
  off = align16(off + Umax / 8);
  t.t0 = off;
  off = align16(off + 256 * 2);
  t.dummy = off;
  off = align16(off + kTeamWaves * 64 * 4);
  t.ctl = off;
  off = align16(off + 16 * 4);
  t.total = off;
  return t;
}

CUCM_HD uint32_t team_enc_lds_bytes(uint32_t Umax) { return team_enc_layout(Umax).total; }

// The encoder team's views: the walk runs on the model's own statep and the
// resolution on its own order-0 states.
CUCM_DEV TeamMem carve_team_enc(LDS uint8_t *base, uint32_t Umax, uint32_t wid) {
  const TeamEncLayout l = team_enc_layout(Umax);
  TeamMem x;
  x.M = carve(base, Umax, 0, false, 2, true);
  x.M.dummy = (LDS uint32_t *)(base + l.dummy) + 64 * wid;
  x.stpn = x.M.stp;
  x.wd = nullptr;
  x.st0r = x.M.st0;
  x.sm = nullptr;
  x.pcb = nullptr;
  x.cand = nullptr;
  x.ubw = base + l.ubw;
  x.t0 = (LDS uint16_t *)(base + l.t0);
  x.ctl = (LDS uint32_t *)(base + l.ctl);
  return x;
}

// encode_predict for the bits r = r0 + lane + 64j of one issuing wave, with
// each bit's mixer deltas (cucm_compress_v2.cu:794-818: err x stretch of the
// unit-start statep) and context returned in registers.
template <int NJ>
CUCM_DEV void encode_predict_team(const WaveMem &M, const LDS DevTables *T, uint32_t nvalid,
                                  uint32_t Umax, uint32_t lane, uint32_t r0,
                                  const uint32_t sl[NJ][8], const uint32_t g[NJ][8],
                                  const uint32_t idx[NJ][8], int32_t dl[NJ][8],
                                  uint32_t cx[NJ]) {
  int32_t smp[NJ][8];
#pragma unroll
  for (int j = 0; j < NJ; ++j)
#pragma unroll
    for (int k = 0; k < 8; ++k) smp[j][k] = T->stretch[M.stp[k * 256 + g[j][k]] >> 20];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const uint32_t r = r0 + lane + 64u * j;
    const bool ok = r < nvalid;
    const uint32_t rr = min(r, Umax - 1);
    uint32_t active = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) active += (smp[j][k] != 1);
    const uint32_t cxt = active + (uint32_t)(M.ub[8 + (rr >> 3) - 1] >> 5) * 10u;
    const LDS u32x4 *wr = (const LDS u32x4 *)(M.w + cxt * 8u);
    const u32x4 wa = wr[0], wb = wr[1];
    const int32_t w[8] = {(int32_t)wa.x, (int32_t)wa.y, (int32_t)wa.z, (int32_t)wa.w,
                          (int32_t)wb.x, (int32_t)wb.y, (int32_t)wb.z, (int32_t)wb.w};
    const uint32_t pr = mix_weights(w, smp[j], T->squash);
    const uint32_t y = ((uint32_t)M.ub[8 + (rr >> 3)] >> (7 - (rr & 7u))) & 1u;
    const int32_t err = ((int32_t)(y << 12) - (int32_t)pr) * 7;
    cx[j] = cxt;
#pragma unroll
    for (int k = 0; k < 8; ++k) dl[j][k] = ok ? (int32_t)(((int64_t)err * smp[j][k]) >> 16) : 0;
    if (ok) {
      M.rprm[rr] = (uint16_t)((pr << 1) | y);
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
}

// P3 of wave 1: the state each StateMap update of the unit sees (update_states'
// resolution for nbytes <= NB bytes; lane (k, j) = (order, bit of the
// byte)).  The bytes with an in-byte collision are found first, then every
// byte's slot, unit-start state and bit are loaded into registers (one
// packed word per byte, all loads in flight together), so the fully unrolled
// byte loop carries only the dependent state lookup, nex lookup and stores.
template <uint32_t NB>
CUCM_DEV void team_resolve_unit(const TeamMem &X, const LDS DevTables *T, uint32_t nbytes,
                                uint32_t Umax, uint32_t lane) {
  const WaveMem &M = X.M;
  const uint32_t TS = table_size(Umax);
  // bytes where two bits of an order >= 1 hash to one slot
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
    uint64_t bal = wave_ballot(coll);
    while (bal) {
      const uint32_t bit = (uint32_t)__builtin_ctzll(bal);
      bal &= bal - 1;
      cmask |= 1u << ((base + bit) / 7);
    }
  }
  const uint32_t k = lane >> 3, j = lane & 7u;
  // running state per slot: order 0 in t0 (by h0), orders 1..7 in the
  // distinct-slot tables; both read kNoState until the unit touches a slot
  LDS uint16_t *tv = k ? M.tval + (k - 1) * TS : X.t0;
  LDS uint8_t *rgk = M.rg + k * Umax;
  const LDS uint16_t *rik = M.ridx + k * Umax;
  // byte b of the unit: slot index | unit-start state << 16 | bit << 24
  // (read before any of the unit's resolved states overwrite rg)
  uint32_t in[NB];
#pragma unroll
  for (uint32_t b = 0; b < NB; ++b) {
    const uint32_t bb = min(b, nbytes - 1);
    in[b] = (uint32_t)rik[8 * bb + j] | ((uint32_t)rgk[8 * bb + j] << 16) |
            ((((uint32_t)M.ub[8 + bb] >> (7 - j)) & 1u) << 24);
  }
  // all inputs in registers before the chain starts, so no wait inside it
  // also waits for the previous byte's stores
  wait_lgkm();
#pragma unroll
  for (uint32_t b = 0; b < NB; ++b) {
    if (b < nbytes) {
      const uint32_t i = in[b] & 0xffffu, g = (in[b] >> 16) & 255u, y = in[b] >> 24;
      const uint32_t r = 8 * b + j;
      if ((cmask >> b) & 1u) {
        // two bits of one order share a slot: one bit at a time
        for (uint32_t q = 0; q < 8; ++q) {
          if (q == j) {
            const uint32_t v = tv[i];
            const uint32_t st = v > 255u ? g : v;
            tv[i] = (uint16_t)T->nex[2 * st + y];
            rgk[r] = (uint8_t)st;
          }
          lane_step();
        }
      } else {
        const uint32_t v = tv[i];
        const uint32_t st = v > 255u ? g : v;
        tv[i] = (uint16_t)T->nex[2 * st + y];
        rgk[r] = (uint8_t)st;
        lane_step();
      }
    }
  }
  wave_sync();
}

// P4 part of wave `part` (0..3): orders 1 + part and 5 + part written back
// and their tables cleared (wave 3 also folds order 0's states into st0).
CUCM_DEV void team_enc_writeback(const TeamMem &X, GLB uint8_t *slot, uint32_t tb,
                                 uint32_t nvalid, uint32_t Umax, uint32_t part, uint32_t lane) {
  const WaveMem &M = X.M;
  const uint32_t TS = table_size(Umax);
  if (part == 3) {
    // order 0: the final state of every touched h0 back into the model
    for (uint32_t r = lane; r < nvalid; r += 64) {
      const uint32_t h0 = M.ridx[r];
      M.st0[h0] = (uint8_t)X.t0[h0];
    }
    wave_sync();
    for (uint32_t r = lane; r < nvalid; r += 64) X.t0[M.ridx[r]] = kNoState;
  }
#pragma unroll 1
  for (uint32_t k = 1 + part; k < 8; k += 4) {
    GLB uint8_t *tab = slot + ((uint64_t)(k - 1) << tb);
    for (uint32_t r = lane; r < nvalid; r += 64) {
      const uint32_t i = M.ridx[k * Umax + r];
      tab[M.tkey[(k - 1) * TS + i]] = (uint8_t)M.tval[(k - 1) * TS + i];
    }
    wave_sync();
    for (uint32_t r = lane; r < nvalid; r += 64) {
      const uint32_t i = M.ridx[k * Umax + r];
      M.tkey[(k - 1) * TS + i] = kEmpty;
      M.tval[(k - 1) * TS + i] = kNoState;
    }
  }
}

// Unit schedule shared by the roles: units u = 0..NU-1 cover the coded bits
// (the tail-padded ones included); nvalid(u) bits of unit u carry data.
struct EncUnits {
  uint64_t vbits, coded_end;
  uint32_t U, NU;
  CUCM_DEV explicit EncUnits(const Task &tk) {
    U = tk.U;
    vbits = 8ull * tk.L;
    coded_end = (tk.flags & kTaskTailPad) ? 8ull * tk.S : vbits;
    NU = (uint32_t)((coded_end + U - 1) / U);
  }
  CUCM_DEV uint32_t nvalid(uint32_t u) const {
    const uint64_t base = (uint64_t)u * U;
    return (u < NU && base < vbits) ? (uint32_t)min((uint64_t)U, vbits - base) : 0u;
  }
  CUCM_DEV uint32_t ncoded(uint32_t u) const {
    return (uint32_t)min((uint64_t)U, coded_end - (uint64_t)u * U);
  }
};

// Waves 0 and 1: issue, predict, deltas; wave 1 also resolves.
template <int NJ>
CUCM_DEV void team_enc_issuer(const Task &tk, const TeamMem &X, const LDS DevTables *T,
                              GLB uint8_t *slot, uint32_t Umax, uint32_t w, uint32_t lane,
                              PhaseClock &clk) {
  constexpr int NJW = NJ >= 2 ? NJ / 2 : 1;
  const uint32_t r0 = w * 64u * NJW;
  const WaveMem &M = X.M;
  const EncUnits E(tk);
  const uint32_t tb = tk.tb, mask = (1u << tb) - 1u;
  const GLB uint8_t *src = (const GLB uint8_t *)(uintptr_t)tk.src;
  for (uint32_t u = 0; u <= E.NU; ++u) {
    const uint32_t nvalid = E.nvalid(u);
    uint32_t sl[NJW][8], g[NJW][8], idx[NJW][8];
    if (nvalid)
      encode_issue<NJW>(M, src, tk.L, (uint64_t)u * E.U, E.U, nvalid, slot, tb, mask, Umax, lane,
                        sl, g, idx, r0);
    clk.mark(2);
    team_sync_lds();  // P1
    clk.mark(9);
    if (!nvalid) continue;
    int32_t dl[NJW][8];
    uint32_t cx[NJW];
    encode_predict_team<NJW>(M, T, nvalid, Umax, lane, r0, sl, g, idx, dl, cx);
    clk.mark(3);
    team_sync_lds();  // P2
    clk.mark(9);
#pragma unroll
    for (int j = 0; j < NJW; ++j)
      if (r0 + lane + 64u * j < nvalid)
#pragma unroll
        for (int k = 0; k < 8; ++k) lds_add(&M.w[cx[j] * 8 + k], dl[j][k]);
    clk.mark(5);
    if (w == 1) team_resolve_unit<8u * NJ>(X, T, nvalid / 8, Umax, lane);
    clk.mark(6);
    team_sync_lds();  // P3
    clk.mark(9);
    team_enc_writeback(X, slot, tb, nvalid, Umax, w, lane);
    clk.mark(8);
    team_sync();  // P4
    clk.mark(9);
  }
}

// Wave 2: walks unit u - 1 during P1 of unit u (restarting at every byte,
// which measures faster than one continuous walk here).
CUCM_DEV void team_enc_walker(const Task &tk, const TeamMem &X, const LDS DevTables *T,
                              GLB uint8_t *slot, uint32_t Umax, uint32_t lane, PhaseClock &clk) {
  const WaveMem &M = X.M;
  const EncUnits E(tk);
  CUCM_WALK_STATE;
  (void)w_sv, (void)w_nblk, (void)w_blk;  // (the per-byte walk uses none of these)
  for (uint32_t u = 0; u <= E.NU; ++u) {
    const uint32_t prev = u ? E.nvalid(u - 1) : 0u;
    // (only the 8 walking lanes run: the step's dummy store drops out)
    if (lane < 8)
      for (uint32_t b = 0; b < prev / 8; ++b) CUCM_TWALK_BYTE(X, T, b, X.ubw[b], Umax, lane);
    clk.mark(7);
    team_sync_lds();  // P1
    clk.mark(10);
    const uint32_t nvalid = E.nvalid(u);
    if (!nvalid) continue;
    team_sync_lds();  // P2
    for (uint32_t i = lane; i < nvalid / 8; i += 64) X.ubw[i] = M.ub[8 + i];
    team_sync_lds();  // P3
    clk.mark(10);
    team_enc_writeback(X, slot, tk.tb, nvalid, Umax, 2, lane);
    clk.mark(8);
    team_sync();  // P4
    clk.mark(10);
  }
}

// Wave 3: codes unit u - 1 during P1 of unit u (its (pr, y) read into
// registers before unit u's predictions overwrite them), then flushes.
template <int NJ>
CUCM_DEV void team_enc_coder(const CodecArgs &a, const Task &tk, uint32_t t, const TeamMem &X,
                             GLB uint8_t *slot, uint32_t Umax, uint32_t lane, PhaseClock &clk) {
  const WaveMem &M = X.M;
  const EncUnits E(tk);
  GLB uint8_t *dst = (GLB uint8_t *)(uintptr_t)tk.dst;
  const uint64_t cap = tk.aux;
  uint32_t x1 = 0, x2 = 0xffffffffu, acc = 0;
  uint32_t opos = 0, ovf = 0;
  for (uint32_t u = 0; u <= E.NU; ++u) {
    if (u) {
      uint32_t pv[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) pv[j] = prm20(M.rprm[lane + 64u * j]);
      // coded bits are a multiple of 8: blocks of 8 with constant lane offsets
      const uint32_t nc = E.ncoded(u - 1);
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const uint32_t lo = 64u * j, hi = min(nc, lo + 64u);
        for (uint32_t r0 = lo; r0 < hi; r0 += 8) {
#pragma unroll
          for (uint32_t t8 = 0; t8 < 8; ++t8) CUCM_CODE_BIT(lane_read(pv[j], r0 - lo + t8));
        }
      }
    }
    clk.mark(4);
    team_sync_lds();  // P1
    clk.mark(11);
    const uint32_t nvalid = E.nvalid(u);
    if (!nvalid) continue;
    team_sync_lds();  // P2
    team_sync_lds();  // P3
    clk.mark(11);
    team_enc_writeback(X, slot, tk.tb, nvalid, Umax, 3, lane);
    clk.mark(8);
    team_sync();  // P4
    clk.mark(11);
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
}

// One chunk encoded by the team (all kTeamWaves waves call this together).
template <int NJ, bool kProf>
CUCM_DEV void encode_team(const CodecArgs &a, const Task &tk, uint32_t t, const TeamMem &X,
                          const LDS DevTables *T, GLB uint8_t *slot, uint32_t Umax, uint32_t wid,
                          uint32_t lane) {
  PhaseClock clk(kProf && a.prof != nullptr);
  team_init_enc(a, tk, X, slot, Umax, wid * 64 + lane);
  team_sync();
  clk.mark(0);
  if (wid <= 1)
    team_enc_issuer<NJ>(tk, X, T, slot, Umax, wid, lane, clk);
  else if (wid == 2)
    team_enc_walker(tk, X, T, slot, Umax, lane, clk);
  else
    team_enc_coder<NJ>(a, tk, t, X, slot, Umax, lane, clk);
  clk.flush(a.prof, lane);
}

}  // namespace cucm

#endif
