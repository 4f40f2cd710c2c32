// cucm_host_model.cpp -- lookup tables and the host-side prelearn coder.
//
// Prelearn is sequential per bit with immediate updates (the reference runs
// it on the host too: cucm_compress_v2.cu:924-981, cucm_decompress_v2.cu:
// 1056-1182); it trains the model every chunk starts from.  The trained
// model is uploaded once and copied into each chunk's slot on the device.
#include "cucm_host_model.h"

#include <cstring>

namespace cucm {

void build_tables(DevTables *t) {
  // squash: piecewise-linear interpolation of the 33 knots over 128-wide
  // segments of the stretch domain; stretch: its inverse (smallest x whose
  // squash reaches p); dt: StateMap adaptation rates (cucm_compress_v2.cu:681-705).
  int sq[4096];
  for (int k = 0; k < 4095; ++k) {
    const int d = k - 2047;
    const int seg = (d >> 7) + 16, frac = d & 127;
    sq[k] = (kSquashKnots[seg] * (128 - frac) + kSquashKnots[seg + 1] * frac + 64) >> 7;
  }
  sq[4095] = 4095;
  int st[4096];
  int lo = 0;
  for (int x = -2047; x <= 2047; ++x) {
    const int hi = sq[x + 2047];
    for (int p = lo; p <= hi; ++p) st[p] = x;
    lo = hi + 1;
  }
  st[4095] = 2047;
  for (int i = 0; i < 4096; ++i) {
    // every prediction bumps pr < 2048 by one after the squash
    // (cucm_compress_v2.cu:787-788): the table holds the bumped value
    t->squash[i] = (uint16_t)(sq[i] + (sq[i] < 2048 ? 1 : 0));
    t->stretch[i] = (int16_t)st[i];
  }
  for (int n = 0; n < 1024; ++n) t->dt[n] = (uint16_t)(16384 / (2 * n + 3));
  for (int s = 0; s < 256; ++s) {
    t->nex[2 * s] = (uint8_t)(kNextState[s] & 0xff);
    t->nex[2 * s + 1] = (uint8_t)(kNextState[s] >> 8);
  }
}

const DevTables &host_tables() {
  static DevTables t;
  static bool ready = [] {
    build_tables(&t);
    return true;
  }();
  (void)ready;
  return t;
}

HostModel::HostModel(int tb_) : tb(tb_), mask((1u << tb_) - 1u), states((size_t)8 << tb_, 0) {
  for (int k = 0; k < kOrders; ++k)
    for (int s = 0; s < 256; ++s) statep[k][s] = 1u << 31;
  std::memset(w, 0, sizeof w);
}

namespace {

struct BitEval {
  uint32_t slot[kOrders];
  int32_t smp[kOrders];
  uint32_t cxt, pr;
};

void eval(const HostModel &m, const DevTables &T, const uint8_t *buf, uint64_t pos, BitEval &e) {
  const uint64_t byte = pos >> 3;
  const uint32_t j = (uint32_t)(pos & 7);
  uint32_t h = kFnvBasis;
  for (int k = 0; k < kOrders; ++k) {
    uint32_t ctx;
    if (k == 0)
      ctx = (1u << j) | (j ? (uint32_t)buf[byte] >> (8 - j) : 0u);
    else
      ctx = pos >= 8ull * k ? buf[byte - k] : 0u;
    if (k == 0) {
      h = (h ^ ctx) * kFnvPrime;
      e.slot[0] = h & m.mask;
    } else {
      e.slot[k] = fnv_pos(h, ctx, m.mask);  // the device tables' layout
    }
  }
  uint32_t active = 0;
  for (int k = 0; k < kOrders; ++k) {
    const uint8_t s = m.states[((size_t)k << m.tb) + e.slot[k]];
    e.smp[k] = T.stretch[m.statep[k][s] >> 20];
    active += e.smp[k] != 1;
  }
  const uint32_t c1 = pos >= 8 ? buf[byte - 1] : 0u;
  e.cxt = active + (c1 >> 5) * 10u;
  int64_t dot = 0;
  for (int k = 0; k < kOrders; ++k) dot += (int64_t)m.w[e.cxt * 8 + k] * e.smp[k];
  dot >>= 16;
  const int idx = dot < -2047 ? -2047 : (dot > 2048 ? 2048 : (int)dot);
  e.pr = T.squash[idx + 2047];  // the pr bump is in the table
}

void update(HostModel &m, const DevTables &T, const BitEval &e, uint32_t y) {
  const int32_t err = ((int32_t)(y << 12) - (int32_t)e.pr) * 7;
  for (int k = 0; k < kOrders; ++k) {
    const int32_t d = (int32_t)(((int64_t)err * e.smp[k]) >> 16);
    m.w[e.cxt * 8 + k] = (int32_t)((uint32_t)m.w[e.cxt * 8 + k] + (uint32_t)d);
  }
  for (int k = 0; k < kOrders; ++k) {
    uint8_t &cell = m.states[((size_t)k << m.tb) + e.slot[k]];
    const uint32_t s = cell;
    const uint32_t p = m.statep[k][s];
    const uint32_t n = p & 1023u;
    const int32_t er = (int32_t)((y << 22) - (p >> 10));
    const uint32_t adj = ((uint32_t)(er >> 3) * (uint32_t)T.dt[n]) & 0xfffffc00u;
    m.statep[k][s] = p + adj + (n < 1023u ? 1u : 0u);
    cell = T.nex[2 * s + y];
  }
}

inline uint32_t mid(uint32_t x1, uint32_t x2, uint32_t pr) {
  const uint32_t r = x2 - x1;
  return x1 + (r >> 12) * pr + (((r & 0xfffu) * pr) >> 12);
}

}  // namespace

void prelearn_encode(const uint8_t *in, uint64_t P, HostModel &m, std::vector<uint8_t> &out) {
  const DevTables &T = host_tables();
  out.clear();
  uint32_t x1 = 0, x2 = 0xffffffffu;
  BitEval e;
  for (uint64_t pos = 0; pos < 8 * P; ++pos) {
    eval(m, T, in, pos, e);
    const uint32_t y = (in[pos >> 3] >> (7 - (pos & 7))) & 1u;
    update(m, T, e, y);
    const uint32_t xm = mid(x1, x2, e.pr);
    if (y) x2 = xm; else x1 = xm + 1u;
    while (((x1 ^ x2) & 0xff000000u) == 0) {
      out.push_back((uint8_t)(x2 >> 24));
      x1 <<= 8;
      x2 = (x2 << 8) | 255u;
    }
  }
  out.push_back((uint8_t)(x1 >> 24));
}

void prelearn_train(const uint8_t *in, uint64_t P, HostModel &m) {
  const DevTables &T = host_tables();
  BitEval e;
  for (uint64_t pos = 0; pos < 8 * P; ++pos) {
    eval(m, T, in, pos, e);
    update(m, T, e, (in[pos >> 3] >> (7 - (pos & 7))) & 1u);
  }
}

void prelearn_decode(const uint8_t *in, uint64_t inlen, uint64_t P, HostModel &m, uint8_t *out) {
  const DevTables &T = host_tables();
  std::memset(out, 0, P);
  uint64_t ip = 0;
  uint32_t x = 0, x1 = 0, x2 = 0xffffffffu;
  for (int i = 0; i < 4; ++i) x = (x << 8) | (ip < inlen ? in[ip++] : 255u);
  BitEval e;
  for (uint64_t pos = 0; pos < 8 * P; ++pos) {
    eval(m, T, out, pos, e);
    const uint32_t xm = mid(x1, x2, e.pr);
    const uint32_t y = x <= xm;
    if (y) x2 = xm; else x1 = xm + 1u;
    update(m, T, e, y);
    while (((x1 ^ x2) & 0xff000000u) == 0) {
      x1 <<= 8;
      x2 = (x2 << 8) | 255u;
      x = (x << 8) | (ip < inlen ? in[ip++] : 255u);
    }
    out[pos >> 3] |= (uint8_t)(y << (7 - (pos & 7)));
  }
}

void order0_states(const HostModel &m, uint8_t st0[256]) {
  st0[0] = 0;
  for (uint32_t h0 = 1; h0 < 256; ++h0) {
    const uint32_t slot = ((kFnvBasis ^ h0) * kFnvPrime) & m.mask;
    st0[h0] = m.states[slot];
  }
}

}  // namespace cucm
