/*
 * cucm_oracle.c -- CPU restatement of CuCM's chunked CM codec.
 * TEST INFRASTRUCTURE ONLY (see cucm_oracle.h).  Build: oracle/Makefile.
 * Compiled with -fwrapv; all wrap-around arithmetic is written on uint32_t
 * explicitly anyway.
 */
#include "cucm_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* Tables                                                              */
/* ------------------------------------------------------------------ */

/* Bit-history next-state table nex[s][y], 256 rows x 2 bytes, row-major.
 * The values must equal the reference's state_table / d_State_table
 * (cucm_compress_v2.cu:35-285, cucm_decompress_v2.cu:16-265): the coded
 * stream depends on every entry.  Stored here as hex; tests compare it
 * byte-for-byte with the product's copy and, when the reference is present,
 * with all four copies in the reference sources. */
static const char NEX_HEX[] =
  "010203050406070a080c090d0b0e0f13101711181219141b151c161d1a1e1f21"
  "202320232023202322252225222522252225222524272427242724272628292b"
  "2a2d2a2d2c2f2c2f2e312e31303330333234352b36393639383b383b3a3d3a3d"
  "3c3f3c3f3e413e41324243374439443946494649484b484b4a4d4a4d4c4f4c4f"
  "3e513e514052534554475447564956492c3b2c3b3a3d3a3d3c313c314c594c59"
  "4e5b4e5b505c5d455e575e57602d602d30633063586558655066674568576857"
  "6a396a393e6d3e6d586f586f5070715572577257743974393e773e7758795879"
  "5a7a7b557c617c617e397e393e813e81628362835a8485558661866188398839"
  "3e8b3e8b628d628d5a8e8f5f90619061443944393e513e51629362936494955f"
  "966b966b6c976c976498995f9a6b6c9b649c9d5f9e6b6c9f64a0a169a26b6ca3"
  "6ea4a569a67576a76ea8a969aa7576ab6eacad69ae7576af6eb0b169b27576b3"
  "6eb4b573b67576b778b8b973ba7f80bb78bcbd73be7f80bf78c0c173c27f80c3"
  "78c4c573c67f80c778c8c973ca7f80cb78cccd73ce7f80cf78d0d17dd27f80d3"
  "82d4d57dd6898ad782d8d97dda898adb82dcdd7dde898adf82e0e17de2898ae3"
  "82e4e57de6898ae782e8e97dea898aeb82eced7dee898aef82f0f17df2898af3"
  "82f4f587f6898af78cf8f987fa4550fb8cfcf987fa4550fb8cfc000000000000";

static int32_t SQ[4096], ST[4096], DT[1024];
static uint8_t NEX[512];
static int tables_ready = 0;

static int hexval(char c) {
  return (c >= '0' && c <= '9') ? c - '0' : (c >= 'a' && c <= 'f') ? c - 'a' + 10 : -1;
}

/* squash knots and interpolation, stretch as its inverse, dt = 16384/(2i+3):
 * the closed forms of cucm_compress_v2.cu:681-705. */
static void build_tables(void) {
  static const int knots[33] = {1,    2,    3,    6,    10,   16,   27,
                                45,   73,   120,  194,  310,  488,  747,
                                1101, 1546, 2047, 2549, 2994, 3348, 3607,
                                3785, 3901, 3975, 4022, 4050, 4068, 4079,
                                4085, 4089, 4092, 4093, 4094};
  for (int k = 0; k < 4095; ++k) {
    int d = k - 2047;
    int frac = d & 127;
    int seg = (d >> 7) + 16;
    SQ[k] = (knots[seg] * (128 - frac) + knots[seg + 1] * frac + 64) >> 7;
  }
  SQ[4095] = 4095;
  int next = 0;
  for (int x = -2047; x <= 2047; ++x) {
    int v = SQ[x + 2047];
    for (int j = next; j <= v; ++j) ST[j] = x;
    next = v + 1;
  }
  ST[4095] = 2047;
  for (int i = 0; i < 1024; ++i) DT[i] = 16384 / (2 * i + 3);
  for (int i = 0; i < 512; ++i)
    NEX[i] = (uint8_t)(hexval(NEX_HEX[2 * i]) * 16 + hexval(NEX_HEX[2 * i + 1]));
  tables_ready = 1;
}

static inline void ensure_tables(void) {
  if (!tables_ready) {
#ifdef _OPENMP
#pragma omp critical(orc_tables)
#endif
    if (!tables_ready) build_tables();
  }
}

void orc_get_tables(int32_t sq[4096], int32_t st[4096], int32_t dt[1024],
                    uint8_t nex[512]) {
  ensure_tables();
  if (sq) memcpy(sq, SQ, sizeof SQ);
  if (st) memcpy(st, ST, sizeof ST);
  if (dt) memcpy(dt, DT, sizeof DT);
  if (nex) memcpy(nex, NEX, sizeof NEX);
}

/* ------------------------------------------------------------------ */
/* Primitives                                                          */
/* ------------------------------------------------------------------ */

#define FNV_BASIS 2166136261u
#define FNV_PRIME 16777619u

uint32_t orc_fnv1a32(const uint8_t *p, size_t n) {
  uint32_t h = FNV_BASIS;
  for (size_t i = 0; i < n; ++i) { h ^= p[i]; h *= FNV_PRIME; }
  return h;
}

/* Context of chunk-relative bit b: h0 = leading-one partial byte, h_i = the
 * i-th previous byte of the same chunk or 0 (cucm_compress_v2.cu:727-743).
 * Order k hashes (h0..hk) with FNV-1a; the reference keeps the state in a
 * uint64_t whose low 32 bits equal 32-bit FNV-1a. */
void orc_slots(const uint8_t *chunk, uint64_t b, int tb, uint32_t slots[8]) {
  const uint64_t byte = b >> 3;
  const unsigned j = (unsigned)(b & 7);
  const uint32_t mask = (tb >= 32) ? 0xffffffffu : ((1u << tb) - 1u);
  uint32_t h = FNV_BASIS;
  uint32_t h0 = (1u << j) | (j ? ((uint32_t)chunk[byte] >> (8 - j)) : 0u);
  h ^= h0; h *= FNV_PRIME;
  slots[0] = h & mask;
  for (int i = 1; i < 8; ++i) {
    uint32_t hi = (b >= 8u * (uint64_t)i) ? chunk[byte - i] : 0u;
    h ^= hi; h *= FNV_PRIME;
    slots[i] = h & mask;
  }
}

/* Carry-less binary arithmetic coder step, 12-bit probability of a 1
 * (cucm_compress_v2.cu:855-862). */
void orc_coder_step(uint32_t *x1p, uint32_t *x2p, int pr, int y, uint8_t *out,
                    uint64_t *n) {
  uint32_t x1 = *x1p, x2 = *x2p;
  const uint32_t r = x2 - x1;
  const uint32_t p = (uint32_t)pr;
  const uint32_t xmid = x1 + (r >> 12) * p + (((r & 0xfffu) * p) >> 12);
  if (y) x2 = xmid; else x1 = xmid + 1u;
  while (((x1 ^ x2) & 0xff000000u) == 0) {
    if (out) out[*n] = (uint8_t)(x2 >> 24);
    ++*n;
    x1 <<= 8;
    x2 = (x2 << 8) | 255u;
  }
  *x1p = x1; *x2p = x2;
}

struct orc_model {
  int tb;
  uint32_t mask;
  uint8_t *states; /* [8][1 << tb] */
  uint32_t statep[8][256];
  int32_t w[ORC_MIXCTX * ORC_ORDERS];
};

orc_model *orc_model_new(int tb) {
  if (tb < 8 || tb > 28) return NULL;
  ensure_tables();
  orc_model *m = (orc_model *)malloc(sizeof *m);
  if (!m) return NULL;
  m->tb = tb;
  m->mask = (1u << tb) - 1u;
  m->states = (uint8_t *)calloc((size_t)8 << tb, 1);
  if (!m->states) { free(m); return NULL; }
  for (int k = 0; k < 8; ++k)
    for (int s = 0; s < 256; ++s) m->statep[k][s] = 1u << 31;
  memset(m->w, 0, sizeof m->w);
  return m;
}

void orc_model_free(orc_model *m) {
  if (!m) return;
  free(m->states);
  free(m);
}

int orc_model_copy(orc_model *dst, const orc_model *src) {
  if (!dst || !src || dst->tb != src->tb) return ORC_E_PARAM;
  memcpy(dst->states, src->states, (size_t)8 << src->tb);
  memcpy(dst->statep, src->statep, sizeof dst->statep);
  memcpy(dst->w, src->w, sizeof dst->w);
  return ORC_OK;
}

static orc_model *model_clone(const orc_model *src) {
  orc_model *m = orc_model_new(src->tb);
  if (m) orc_model_copy(m, src);
  return m;
}

int orc_model_tb(const orc_model *m) { return m->tb; }
uint8_t orc_model_state(const orc_model *m, int k, uint32_t slot) {
  return m->states[((size_t)k << m->tb) + (slot & m->mask)];
}
uint32_t orc_model_statep(const orc_model *m, int k, int s) { return m->statep[k][s & 255]; }
int32_t orc_model_weight(const orc_model *m, int i) { return m->w[i]; }

int orc_model_predict(const orc_model *m, int k, uint32_t slot) {
  return (int)(m->statep[k][orc_model_state(m, k, slot)] >> 20);
}

/* StateMap::update (cucm_compress_v2.cu:592-601): count-adaptive
 * probability step on statep[s], then the bit-history transition. */
void orc_statemap_update(orc_model *m, int k, uint32_t slot, int y) {
  uint8_t *cell = &m->states[((size_t)k << m->tb) + (slot & m->mask)];
  const int s = *cell;
  const uint32_t p = m->statep[k][s];
  const int n = (int)(p & 1023u);
  const int32_t e = (int32_t)(((uint32_t)y << 22) - (p >> 10));
  const int64_t prod = (int64_t)(e >> 3) * (int64_t)DT[n];
  const uint32_t adj = (uint32_t)prod & 0xfffffc00u;
  m->statep[k][s] = p + adj + (n < 1023 ? 1u : 0u);
  *cell = NEX[2 * s + (y & 1)];
}

/* Mixer::p (cucm_compress_v2.cu:636-644). */
int orc_mixer_predict(const orc_model *m, int cxt, const int32_t smp[8]) {
  int64_t dot = 0;
  for (int i = 0; i < 8; ++i) dot += (int64_t)m->w[cxt * 8 + i] * smp[i];
  dot >>= 16;
  int idx = dot < -2047 ? -2047 : (dot > 2048 ? 2048 : (int)dot);
  return SQ[idx + 2047];
}

/* Mixer::update (cucm_compress_v2.cu:663-669), wrapping int32 adds. */
void orc_mixer_update(orc_model *m, int y, int pr, int cxt, const int32_t smp[8]) {
  const int32_t err = ((y << 12) - pr) * 7;
  for (int i = 0; i < 8; ++i) {
    const int32_t d = (int32_t)(((int64_t)err * smp[i]) >> 16);
    m->w[cxt * 8 + i] = (int32_t)((uint32_t)m->w[cxt * 8 + i] + (uint32_t)d);
  }
}

uint64_t orc_model_digest(const orc_model *m) {
  /* FNV-1a-64 over states, statep (LE) and weights (LE). */
  uint64_t h = 1469598103934665603ull;
  const size_t ns = (size_t)8 << m->tb;
  for (size_t i = 0; i < ns; ++i) { h ^= m->states[i]; h *= 1099511628211ull; }
  for (int k = 0; k < 8; ++k)
    for (int s = 0; s < 256; ++s)
      for (int b = 0; b < 4; ++b) { h ^= (m->statep[k][s] >> (8 * b)) & 255u; h *= 1099511628211ull; }
  for (int i = 0; i < 640; ++i)
    for (int b = 0; b < 4; ++b) { h ^= ((uint32_t)m->w[i] >> (8 * b)) & 255u; h *= 1099511628211ull; }
  return h;
}

/* Per-bit model evaluation: slots, stretched predictions, mixer context and
 * final probability (cucm_compress_v2.cu:711-792). */
typedef struct {
  uint32_t slot[8];
  int32_t smp[8];
  int cxt;
  int pr;
} bitrec;

static void eval_bit(const orc_model *m, const uint8_t *chunk, uint64_t b, bitrec *r) {
  orc_slots(chunk, b, m->tb, r->slot);
  int active = 0;
  for (int k = 0; k < 8; ++k) {
    r->smp[k] = ST[orc_model_predict(m, k, r->slot[k])];
    if (r->smp[k] != 1) ++active;
  }
  const int c1 = (b >= 8) ? chunk[(b >> 3) - 1] : 0;
  r->cxt = active + (c1 >> 5) * 10;
  int pr = orc_mixer_predict(m, r->cxt, r->smp);
  if (pr < 2048) ++pr;
  r->pr = pr;
}

static void apply_bit_updates(orc_model *m, const bitrec *recs, const uint8_t *ys, int cnt) {
  for (int r = 0; r < cnt; ++r) orc_mixer_update(m, ys[r], recs[r].pr, recs[r].cxt, recs[r].smp);
  for (int k = 0; k < 8; ++k)
    for (int r = 0; r < cnt; ++r) orc_statemap_update(m, k, recs[r].slot[k], ys[r]);
}

/* ------------------------------------------------------------------ */
/* Clean per-chunk encoder / decoder                                   */
/* ------------------------------------------------------------------ */

static inline int bit_of(const uint8_t *p, uint64_t b) {
  return (p[b >> 3] >> (7 - (b & 7))) & 1;
}

int64_t orc_encode_chunk(const uint8_t *chunk, uint64_t L, uint64_t S,
                         uint32_t U, const orc_model *init, int tail_pad,
                         uint8_t *out, uint64_t cap) {
  if (!chunk || !init || U == 0 || S == 0 || L == 0 || L > S) return ORC_E_PARAM;
  orc_model *m = model_clone(init);
  bitrec *recs = (bitrec *)malloc(sizeof(bitrec) * U);
  uint8_t *ys = (uint8_t *)malloc(U);
  uint16_t *stale = (uint16_t *)calloc(U, sizeof(uint16_t)); /* pr<<1|y per slot r */
  if (!m || !recs || !ys || !stale) {
    orc_model_free(m); free(recs); free(ys); free(stale);
    return ORC_E_NOMEM;
  }
  uint32_t x1 = 0, x2 = 0xffffffffu;
  uint64_t n = 0;
  const uint64_t vbits = 8 * L, nbits = 8 * S;
  const uint64_t coded_end = tail_pad ? nbits : vbits;
  uint8_t *sink = out;
  for (uint64_t base = 0; base < coded_end; base += U) {
    int cnt = 0;
    for (uint32_t r = 0; r < U; ++r) {
      const uint64_t b = base + r;
      if (b >= vbits) break;
      eval_bit(m, chunk, b, &recs[r]);
      ys[r] = (uint8_t)bit_of(chunk, b);
      stale[r] = (uint16_t)((recs[r].pr << 1) | ys[r]);
      cnt = (int)r + 1;
    }
    for (uint32_t r = 0; r < U; ++r) {
      const uint64_t b = base + r;
      if (b >= coded_end) break;
      if (n + 4 >= cap) sink = NULL; /* count only once the buffer is full */
      orc_coder_step(&x1, &x2, stale[r] >> 1, stale[r] & 1, sink, &n);
    }
    apply_bit_updates(m, recs, ys, cnt);
  }
  int64_t ret;
  if (n + 1 > cap || sink == NULL) ret = ORC_E_SMALL;
  else { out[n++] = (uint8_t)(x1 >> 24); ret = (int64_t)n; }
  orc_model_free(m); free(recs); free(ys); free(stale);
  return ret;
}

int orc_decode_chunk(const uint8_t *pay, uint64_t plen, uint64_t L, uint64_t S,
                     uint32_t U, const orc_model *init, uint8_t *out) {
  if (!pay || !init || !out || U == 0 || L == 0 || L > S) return ORC_E_PARAM;
  orc_model *m = model_clone(init);
  bitrec *recs = (bitrec *)malloc(sizeof(bitrec) * U);
  uint8_t *ys = (uint8_t *)malloc(U);
  if (!m || !recs || !ys) { orc_model_free(m); free(recs); free(ys); return ORC_E_NOMEM; }
  memset(out, 0, L);
  uint64_t ip = 0;
  uint32_t x = 0;
  for (int i = 0; i < 4; ++i) x = (x << 8) | (ip < plen ? pay[ip++] : 255u);
  uint32_t x1 = 0, x2 = 0xffffffffu;
  const uint64_t vbits = 8 * L;
  for (uint64_t base = 0; base < vbits; base += U) {
    int cnt = 0;
    for (uint32_t r = 0; r < U && base + r < vbits; ++r) {
      const uint64_t b = base + r;
      eval_bit(m, out, b, &recs[r]);
      const uint32_t range = x2 - x1, p = (uint32_t)recs[r].pr;
      const uint32_t xmid = x1 + (range >> 12) * p + (((range & 0xfffu) * p) >> 12);
      const int y = x <= xmid;
      if (y) x2 = xmid; else x1 = xmid + 1u;
      while (((x1 ^ x2) & 0xff000000u) == 0) {
        x1 <<= 8;
        x2 = (x2 << 8) | 255u;
        x = (x << 8) | (ip < plen ? pay[ip++] : 255u);
      }
      ys[r] = (uint8_t)y;
      out[b >> 3] |= (uint8_t)(y << (7 - (b & 7)));
      cnt = (int)r + 1;
    }
    apply_bit_updates(m, recs, ys, cnt);
  }
  orc_model_free(m); free(recs); free(ys);
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Frozen model: every bit predicted from the initial model, never     */
/* updated.  The reference has no code for this mode; its per-bit      */
/* model (eval_bit, cucm_compress_v2.cu:711-792) and coder (:841-868)  */
/* are the normal path's, only the updates and tail padding are gone.  */
/* ------------------------------------------------------------------ */

int64_t orc_encode_chunk_frozen(const uint8_t *chunk, uint64_t L, const orc_model *m,
                                uint8_t *out, uint64_t cap) {
  if (!chunk || !m || !out || L == 0) return ORC_E_PARAM;
  uint32_t x1 = 0, x2 = 0xffffffffu;
  uint64_t n = 0;
  uint8_t *sink = out;
  bitrec r;
  for (uint64_t b = 0; b < 8 * L; ++b) {
    eval_bit(m, chunk, b, &r);
    if (n + 4 >= cap) sink = NULL; /* count only once the buffer is full */
    orc_coder_step(&x1, &x2, r.pr, bit_of(chunk, b), sink, &n);
  }
  if (n + 1 > cap || sink == NULL) return ORC_E_SMALL;
  out[n++] = (uint8_t)(x1 >> 24);
  return (int64_t)n;
}

int orc_decode_chunk_frozen(const uint8_t *pay, uint64_t plen, uint64_t L, const orc_model *m,
                            uint8_t *out) {
  if (!pay || !m || !out || L == 0) return ORC_E_PARAM;
  memset(out, 0, L);
  uint64_t ip = 0;
  uint32_t x = 0;
  for (int i = 0; i < 4; ++i) x = (x << 8) | (ip < plen ? pay[ip++] : 255u);
  uint32_t x1 = 0, x2 = 0xffffffffu;
  bitrec r;
  for (uint64_t b = 0; b < 8 * L; ++b) {
    eval_bit(m, out, b, &r);
    const uint32_t range = x2 - x1, p = (uint32_t)r.pr;
    const uint32_t xmid = x1 + (range >> 12) * p + (((range & 0xfffu) * p) >> 12);
    const int y = x <= xmid;
    if (y) x2 = xmid; else x1 = xmid + 1u;
    while (((x1 ^ x2) & 0xff000000u) == 0) {
      x1 <<= 8;
      x2 = (x2 << 8) | 255u;
      x = (x << 8) | (ip < plen ? pay[ip++] : 255u);
    }
    out[b >> 3] |= (uint8_t)(y << (7 - (b & 7)));
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Batch blobs                                                         */
/* ------------------------------------------------------------------ */

static inline void put_be32(uint8_t *p, uint32_t v) {
  p[0] = (uint8_t)(v >> 24); p[1] = (uint8_t)(v >> 16); p[2] = (uint8_t)(v >> 8); p[3] = (uint8_t)v;
}
static inline uint32_t get_be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static inline uint32_t get_be16(const uint8_t *p) { return ((uint32_t)p[0] << 8) | p[1]; }
static inline void put_le32(uint8_t *p, uint32_t v) {
  p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); p[2] = (uint8_t)(v >> 16); p[3] = (uint8_t)(v >> 24);
}
static inline uint32_t get_le32(const uint8_t *p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

/* Strict per-chunk payload bound: every coded bit costs < 12 output bits
 * (pr in [1, 4095]) plus one flush byte and the coder's pending bytes. */
static inline uint64_t chunk_bound(uint64_t S) { return 12 * S + 8; }

uint64_t orc_blob_bound(uint64_t N, uint64_t S) {
  if (S == 0) return 0;
  const uint64_t C = (N + S - 1) / S;
  return 4 * C + C * chunk_bound(S);
}

/* Blob layout (cucm_compress_v2.cu:870-914): chunk c emitted size_c bytes
 * (flush included); E_c = 4C + sum_{j<=c} size_j - 1 is written big-endian
 * at blob[4c]; the payload occupies blob[E_c - size_c + 1 .. E_c]. */
static int assemble_blob(uint8_t **pays, const uint64_t *sizes, uint64_t C,
                         uint8_t *blob, uint64_t cap, uint64_t *blob_len) {
  uint64_t total = 4 * C;
  for (uint64_t c = 0; c < C; ++c) total += sizes[c];
  if (total > cap) return ORC_E_SMALL;
  if (total > 0xffffffffull) return ORC_E_PARAM;
  uint64_t pos = 4 * C;
  for (uint64_t c = 0; c < C; ++c) {
    memcpy(blob + pos, pays[c], sizes[c]);
    pos += sizes[c];
    put_be32(blob + 4 * c, (uint32_t)(pos - 1));
  }
  *blob_len = total;
  return ORC_OK;
}

static int encode_batch_impl(const uint8_t *in, uint64_t N, uint64_t S, uint32_t U,
                             const orc_model *init, int tail_pad, int frozen, int threads,
                             uint8_t *blob, uint64_t cap, uint64_t *blob_len) {
  if (!in || !init || !blob || !blob_len || S == 0 || U == 0 || N == 0) return ORC_E_PARAM;
  const uint64_t C = (N + S - 1) / S;
  uint8_t **pays = (uint8_t **)calloc(C, sizeof(uint8_t *));
  uint64_t *sizes = (uint64_t *)calloc(C, sizeof(uint64_t));
  if (!pays || !sizes) { free(pays); free(sizes); return ORC_E_NOMEM; }
  int err = ORC_OK;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int64_t c = 0; c < (int64_t)C; ++c) {
    const uint64_t L = (c == (int64_t)C - 1) ? N - (uint64_t)c * S : S;
    const uint64_t bcap = chunk_bound(S);
    pays[c] = (uint8_t *)malloc(bcap);
    int64_t r = !pays[c] ? ORC_E_NOMEM
                : frozen ? orc_encode_chunk_frozen(in + (uint64_t)c * S, L, init, pays[c], bcap)
                         : orc_encode_chunk(in + (uint64_t)c * S, L, S, U, init,
                                            tail_pad && (c == (int64_t)C - 1), pays[c], bcap);
    if (r < 0) {
#ifdef _OPENMP
#pragma omp critical(orc_err)
#endif
      err = (int)r;
    } else sizes[c] = (uint64_t)r;
  }
  (void)threads;
  if (err == ORC_OK) err = assemble_blob(pays, sizes, C, blob, cap, blob_len);
  for (uint64_t c = 0; c < C; ++c) free(pays[c]);
  free(pays); free(sizes);
  return err;
}

int orc_encode_batch(const uint8_t *in, uint64_t N, uint64_t S, uint32_t U,
                     const orc_model *init, int tail_pad, int threads,
                     uint8_t *blob, uint64_t cap, uint64_t *blob_len) {
  return encode_batch_impl(in, N, S, U, init, tail_pad, 0, threads, blob, cap, blob_len);
}

int orc_encode_batch_frozen(const uint8_t *in, uint64_t N, uint64_t S, const orc_model *init,
                            int threads, uint8_t *blob, uint64_t cap, uint64_t *blob_len) {
  return encode_batch_impl(in, N, S, 32, init, 0, 1, threads, blob, cap, blob_len);
}

/* Literal restatement of compress_batch (cucm_compress_v2.cu:990-1115):
 * every phase runs over all chunks before the next phase, with the
 * reference's per-chunk staging buffers (pred, smhashes, prm, cxts) that
 * persist across unit iterations and the reference's guards:
 *   predict  : skip if g >= 8N                   (:721-722)
 *   mix      : skip if g >= 8N or b >= 8S        (:764-765)
 *   mixer upd: skip if g >= 8N                   (:803-804)
 *   SM update: len = min(U, 8N - (int)(poss + S*c))  (:832-834, bits/bytes mixed)
 *   encode   : for r < U while b < 8S, no g < 8N check (:850-851)
 * smhashes is never initialised by the reference (cudaMalloc); entries that
 * are read before being written are taken as 0 here. */
int orc_encode_batch_literal(const uint8_t *in, uint64_t N, uint64_t S,
                             uint32_t U, const orc_model *init, uint8_t *blob,
                             uint64_t cap, uint64_t *blob_len) {
  if (!in || !init || !blob || !blob_len || S == 0 || U == 0 || N == 0) return ORC_E_PARAM;
  const uint64_t C = (N + S - 1) / S;
  orc_model **ms = (orc_model **)calloc(C, sizeof *ms);
  uint64_t *poss = (uint64_t *)calloc(C, sizeof(uint64_t));
  uint32_t *x1s = (uint32_t *)calloc(C, sizeof(uint32_t));
  uint32_t *x2s = (uint32_t *)malloc(C * sizeof(uint32_t));
  uint16_t *pred = (uint16_t *)calloc(C * 8 * U, sizeof(uint16_t));
  uint32_t *smh = (uint32_t *)calloc(C * 8 * U, sizeof(uint32_t));
  uint16_t *prm = (uint16_t *)calloc(C * U, sizeof(uint16_t));
  uint8_t *cxts = (uint8_t *)calloc(C * U, 1);
  uint8_t **pays = (uint8_t **)calloc(C, sizeof(uint8_t *));
  uint64_t *sizes = (uint64_t *)calloc(C, sizeof(uint64_t));
  int err = ORC_OK;
  if (!ms || !poss || !x1s || !x2s || !pred || !smh || !prm || !cxts || !pays || !sizes) {
    err = ORC_E_NOMEM;
    goto done;
  }
  for (uint64_t c = 0; c < C; ++c) {
    ms[c] = model_clone(init);
    pays[c] = (uint8_t *)malloc(chunk_bound(S));
    if (!ms[c] || !pays[c]) { err = ORC_E_NOMEM; goto done; }
    x2s[c] = 0xffffffffu;
  }
  const uint64_t nb = 8 * N;
  for (uint64_t it = 0; it < S * 8 / U + 1; ++it) {
    /* cudapredictUnit */
    for (uint64_t c = 0; c < C; ++c)
      for (int k = 0; k < 8; ++k)
        for (uint32_t r = 0; r < U; ++r) {
          const uint64_t pos = poss[c] + r + 8 * S * c;
          if (pos >= nb) continue;
          uint32_t h[8] = {0};
          for (int i = 1; i <= k; ++i)
            if (poss[c] + r >= 8u * (uint64_t)i) h[i] = in[pos / 8 - i];
          const unsigned j = (unsigned)(pos % 8);
          h[0] = (1u << j) | (j ? ((uint32_t)in[pos / 8] >> (8 - j)) : 0u);
          uint64_t hash = 2166136261ull;
          for (int i = 0; i <= k; ++i) { hash ^= h[i]; hash *= 16777619ull; }
          hash &= ms[c]->mask;
          pred[(c * 8 + k) * U + r] = (uint16_t)orc_model_predict(ms[c], k, (uint32_t)hash);
          smh[(c * 8 + k) * U + r] = (uint32_t)hash;
        }
    /* cudamixUnit */
    for (uint64_t c = 0; c < C; ++c)
      for (uint32_t r = 0; r < U; ++r) {
        const uint64_t pos = poss[c] + r + 8 * S * c;
        if (pos >= nb || poss[c] + r >= 8 * S) continue;
        int32_t smp[8];
        int active = 0;
        for (int k = 0; k < 8; ++k) {
          smp[k] = ST[pred[(c * 8 + k) * U + r]];
          if (smp[k] != 1) ++active;
        }
        int cxt = (poss[c] + r >= 8) ? in[pos / 8 - 1] : 0;
        cxt = active + (cxt >> 5) * 10;
        int pr = orc_mixer_predict(ms[c], cxt, smp);
        if (pr < 2048) ++pr;
        const int y = (in[pos / 8] >> (7 - pos % 8)) & 1;
        prm[c * U + r] = (uint16_t)((pr << 1) | y);
        cxts[c * U + r] = (uint8_t)cxt;
      }
    /* cudaupdateMixer: deltas from recorded pr, cxt and current pred */
    for (uint64_t c = 0; c < C; ++c) {
      /* all deltas of a unit are computed from the weights' frozen copy in
       * the reference only through pr, so sequential application is
       * identical to the atomicAdd schedule. */
      for (uint32_t r = 0; r < U; ++r) {
        const uint64_t pos = poss[c] + r + 8 * S * c;
        if (pos >= nb) continue;
        int32_t smp[8];
        for (int k = 0; k < 8; ++k) smp[k] = ST[pred[(c * 8 + k) * U + r]];
        orc_mixer_update(ms[c], prm[c * U + r] & 1, prm[c * U + r] >> 1, cxts[c * U + r], smp);
      }
    }
    /* cudaStateMapUpdate */
    for (uint64_t c = 0; c < C; ++c) {
      const int32_t pos = (int32_t)(uint32_t)(poss[c] + S * c);
      const uint64_t rem = nb - (uint64_t)(int64_t)pos;
      const int32_t len = (int32_t)((uint64_t)U < rem ? (uint64_t)U : rem);
      for (int k = 0; k < 8; ++k)
        for (int32_t i = 0; i < len; ++i)
          orc_statemap_update(ms[c], k, smh[(c * 8 + k) * U + i], prm[c * U + i] & 1);
    }
    /* cudaencodeUnit */
    for (uint64_t c = 0; c < C; ++c) {
      for (uint32_t i = 0; i < U; ++i) {
        if (poss[c] + i >= S * 8) break;
        orc_coder_step(&x1s[c], &x2s[c], prm[c * U + i] >> 1, prm[c * U + i] & 1, pays[c], &sizes[c]);
      }
      poss[c] += U;
    }
  }
  /* cudaCalEncodeSize + scan + cudaencodefinal: flush byte x1>>24 */
  for (uint64_t c = 0; c < C; ++c) pays[c][sizes[c]++] = (uint8_t)(x1s[c] >> 24);
  err = assemble_blob(pays, sizes, C, blob, cap, blob_len);
done:
  if (ms) for (uint64_t c = 0; c < C; ++c) orc_model_free(ms[c]);
  if (pays) for (uint64_t c = 0; c < C; ++c) free(pays[c]);
  free(ms); free(poss); free(x1s); free(x2s); free(pred); free(smh); free(prm);
  free(cxts); free(pays); free(sizes);
  return err;
}

/* Parses the blob's offset table into payload-relative [start, end). */
static int parse_blob(const uint8_t *blob, uint64_t blen, uint64_t C, uint64_t *starts) {
  if (blen < 4 * C) return ORC_E_CORRUPT;
  uint64_t prev = 4 * C - 1;
  starts[0] = 0;
  for (uint64_t c = 0; c < C; ++c) {
    const uint64_t e = get_be32(blob + 4 * c);
    if (e <= prev || e >= blen) return ORC_E_CORRUPT;
    starts[c + 1] = e + 1 - 4 * C;
    prev = e;
  }
  if (prev != blen - 1) return ORC_E_CORRUPT;
  return ORC_OK;
}

static int decode_batch_impl(const uint8_t *blob, uint64_t blen, uint64_t N, uint64_t S,
                             uint32_t U, const orc_model *init, int frozen, int threads,
                             uint8_t *out) {
  if (!blob || !init || !out || S == 0 || U == 0 || N == 0) return ORC_E_PARAM;
  const uint64_t C = (N + S - 1) / S;
  uint64_t *st = (uint64_t *)malloc((C + 1) * sizeof(uint64_t));
  if (!st) return ORC_E_NOMEM;
  int err = parse_blob(blob, blen, C, st);
  if (err) { free(st); return err; }
  const uint8_t *payload = blob + 4 * C;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int64_t c = 0; c < (int64_t)C; ++c) {
    const uint64_t L = (c == (int64_t)C - 1) ? N - (uint64_t)c * S : S;
    int r = frozen ? orc_decode_chunk_frozen(payload + st[c], st[c + 1] - st[c], L, init,
                                             out + (uint64_t)c * S)
                   : orc_decode_chunk(payload + st[c], st[c + 1] - st[c], L, S, U, init,
                                      out + (uint64_t)c * S);
    if (r) {
#ifdef _OPENMP
#pragma omp critical(orc_err)
#endif
      err = r;
    }
  }
  (void)threads;
  free(st);
  return err;
}

int orc_decode_batch(const uint8_t *blob, uint64_t blen, uint64_t N, uint64_t S,
                     uint32_t U, const orc_model *init, int threads, uint8_t *out) {
  return decode_batch_impl(blob, blen, N, S, U, init, 0, threads, out);
}

int orc_decode_batch_frozen(const uint8_t *blob, uint64_t blen, uint64_t N, uint64_t S,
                            const orc_model *init, int threads, uint8_t *out) {
  return decode_batch_impl(blob, blen, N, S, 32, init, 1, threads, out);
}

/* Literal restatement of decompressbatch (cucm_decompress_v2.cu:898-1054)
 * and its kernels predict (:698-770), mix (:774-809), decode (:813-868),
 * updateStateMap (:870-883) and updateMixer (:885-896).  Chunks advance in
 * lock step; the previous-byte guard is global (realposs >= 8i) and reads
 * the shared output buffer; x is initialised from 4 payload bytes without a
 * boundary check; per-unit records are applied for all U slots. */
static int ilog2u(uint64_t v) { int l = 0; while (v >>= 1) ++l; return l; }

int orc_decode_batch_literal(const uint8_t *blob, uint64_t blen, uint64_t N,
                             uint64_t S, uint32_t U, const orc_model *init,
                             int m, uint8_t *out, uint32_t *flags) {
  if (!blob || !init || !out || S == 0 || U == 0 || N == 0 || m < 0 || m > 12) return ORC_E_PARAM;
  const uint64_t C = (N + S - 1) / S;
  const uint64_t P = (1ull << (m + 1)) - 1;
  uint64_t *cs = (uint64_t *)malloc((C + 1) * sizeof(uint64_t));
  if (!cs) return ORC_E_NOMEM;
  int err = parse_blob(blob, blen, C, cs);
  if (err) { free(cs); return err; }
  const uint8_t *input = blob + 4 * C;
  const uint64_t input_size = blen - 4 * C;
  uint32_t fl = 0;
  orc_model **ms = (orc_model **)calloc(C, sizeof *ms);
  uint64_t *poss = (uint64_t *)calloc(C, sizeof(uint64_t));
  uint32_t *x1s = (uint32_t *)calloc(C, 4), *x2s = (uint32_t *)malloc(C * 4), *xs = (uint32_t *)calloc(C, 4);
  uint64_t *bnd = (uint64_t *)malloc(C * 8), *ioff = (uint64_t *)malloc(C * 8);
  int32_t *pred = (int32_t *)calloc(C * P * 8, 4), *smh = (int32_t *)calloc(C * P * 8, 4);
  int32_t *cxt = (int32_t *)calloc(C * P, 4), *prs = (int32_t *)calloc(C * P, 4);
  int32_t *rpred = (int32_t *)calloc(C * 8 * U, 4), *rsmh = (int32_t *)calloc(C * 8 * U, 4);
  int32_t *rcxt = (int32_t *)calloc(C * U, 4), *rprs = (int32_t *)calloc(C * U, 4);
  if (!ms || !poss || !x1s || !x2s || !xs || !bnd || !ioff || !pred || !smh || !cxt || !prs ||
      !rpred || !rsmh || !rcxt || !rprs) { err = ORC_E_NOMEM; goto done; }
  memset(out, 0, N);
  for (uint64_t c = 0; c < C; ++c) {
    ms[c] = model_clone(init);
    if (!ms[c]) { err = ORC_E_NOMEM; goto done; }
    x2s[c] = 0xffffffffu;
    bnd[c] = cs[c + 1];
    uint32_t x = 0;
    for (int i = 0; i < 4; ++i) {
      const uint64_t at = cs[c]++;
      if (at >= input_size) fl |= 1u;
      x = (x << 8) + (at < input_size ? input[at] : 0u);
    }
    xs[c] = x;
    ioff[c] = cs[c];
  }
  const uint64_t lastchunk = (N % S) ? (N % S) : S;
  for (uint64_t unit = 0; unit < S * 8 / U + 1; ++unit) {
    for (uint64_t rnd = 0; rnd < (U + m) / (uint64_t)(m + 1); ++rnd) {
      /* predict */
      for (uint64_t c = 0; c < C; ++c)
        for (int k = 0; k < 8; ++k)
          for (uint64_t pid = 0; pid < P; ++pid) {
            int plen = ilog2u(pid + 1);
            int pcx = (int)(pid - ((1ull << plen) - 1));
            const uint64_t piu = poss[c] % U;
            const uint64_t rp = poss[c] + c * S * 8 + (uint64_t)plen;
            if (piu + plen >= U || poss[c] + plen >= S * 8 || rp >= N * 8) continue;
            uint64_t h[9] = {0};
            for (int i = 1; i <= k; ++i)
              if (rp >= 8u * (uint64_t)i) h[i] = out[rp / 8 - i];
            const int cbl = (int)(rp % 8);
            if (cbl == 0) {
              h[0] = 0;
            } else if (plen >= cbl) {
              h[0] = (uint64_t)(pcx & ((1 << cbl) - 1));
              plen -= cbl;
              pcx >>= cbl;
            } else {
              h[0] = (uint64_t)(out[rp / 8] >> (8 - cbl));
              h[0] += (uint64_t)pcx;
              plen = 0;
            }
            h[0] = (1ull << cbl) | h[0];
            int tgt = 1;
            while (plen > 0) { h[tgt++] += (uint64_t)(pcx & 0xff); pcx >>= 8; plen -= 8; }
            uint64_t hash = 2166136261ull;
            for (int i = 0; i <= k; ++i) { hash ^= h[i]; hash *= 16777619ull; }
            hash &= ms[c]->mask;
            pred[(c * P + pid) * 8 + k] = orc_model_predict(ms[c], k, (uint32_t)hash);
            smh[(c * P + pid) * 8 + k] = (int32_t)hash;
          }
      /* mix */
      for (uint64_t c = 0; c < C; ++c)
        for (uint64_t pid = 0; pid < P; ++pid) {
          int plen = ilog2u(pid + 1);
          int pcx = (int)(pid - ((1ull << plen) - 1));
          const uint64_t rb = poss[c] + (uint64_t)plen;
          const uint64_t rp = c * S * 8 + rb;
          if (poss[c] % U + plen >= U || rb >= S * 8 || rp >= N * 8) continue;
          int32_t smp[8];
          for (int i = 0; i < 8; ++i) smp[i] = ST[pred[(c * P + pid) * 8 + i]];
          int cx = 0;
          if (rb >= 8) {
            const int sh = (int)(rp % 8);
            pcx >>= sh;
            cx = ((pcx & 0xff) + out[rp / 8 - 1]) >> 5;
          }
          cx *= 10;
          for (int i = 0; i < 8; ++i) if (smp[i] != 1) ++cx;
          int pr = orc_mixer_predict(ms[c], cx, smp);
          if (pr < 2048) ++pr;
          cxt[c * P + pid] = cx;
          prs[c * P + pid] = pr;
        }
      /* decode */
      for (uint64_t c = 0; c < C; ++c) {
        uint32_t x1 = x1s[c], x2 = x2s[c], x = xs[c];
        uint64_t io = ioff[c];
        int pcx = 0;
        for (int i = 0; i <= m; ++i) {
          const uint64_t unitid = (poss[c] + i) % U;
          if (poss[c] + i >= S * 8 || poss[c] % U + i >= U) break;
          if (c == C - 1 && poss[c] + i >= lastchunk * 8) break;
          const uint64_t pid = i ? ((1ull << i) - 1 + (uint64_t)pcx) : 0;
          const uint32_t p = (uint32_t)prs[c * P + pid];
          const uint32_t xmid = x1 + ((x2 - x1) >> 12) * p + ((((x2 - x1) & 0xfffu) * p) >> 12);
          const int y = x <= xmid;
          if (y) x2 = xmid; else x1 = xmid + 1u;
          while (((x1 ^ x2) & 0xff000000u) == 0) {
            x1 <<= 8;
            x2 = (x2 << 8) + 255u;
            if (io < input_size && io < bnd[c]) x = (x << 8) + input[io++];
            else x = (x << 8) + 255u;
          }
          for (int k = 0; k < 8; ++k) {
            rpred[(c * 8 + k) * U + unitid] = pred[(c * P + pid) * 8 + k];
            rsmh[(c * 8 + k) * U + unitid] = smh[(c * P + pid) * 8 + k];
          }
          rprs[c * U + unitid] = (prs[c * P + pid] << 1) | y;
          rcxt[c * U + unitid] = cxt[c * P + pid];
          pcx = (pcx << 1) | y;
          out[c * S + (poss[c] + i) / 8] |= (uint8_t)(y << (7 - (poss[c] + i) % 8));
        }
        x1s[c] = x1; x2s[c] = x2; xs[c] = x; ioff[c] = io;
      }
      for (uint64_t c = 0; c < C; ++c) {
        const uint64_t pos = poss[c];
        const uint64_t left = U - pos % U;
        poss[c] = pos + (left < (uint64_t)(m + 1) ? left : (uint64_t)(m + 1));
      }
    }
    /* updateStateMap over all U records, then updateMixer */
    for (uint64_t c = 0; c < C; ++c) {
      for (int k = 0; k < 8; ++k)
        for (uint32_t i = 0; i < U; ++i)
          orc_statemap_update(ms[c], k, (uint32_t)rsmh[(c * 8 + k) * U + i], rprs[c * U + i] & 1);
      for (uint32_t i = 0; i < U; ++i) {
        int32_t smp[8];
        for (int k = 0; k < 8; ++k) smp[k] = ST[rpred[(c * 8 + k) * U + i]];
        orc_mixer_update(ms[c], rprs[c * U + i] & 1, rprs[c * U + i] >> 1, rcxt[c * U + i], smp);
      }
    }
  }
done:
  if (ms) for (uint64_t c = 0; c < C; ++c) orc_model_free(ms[c]);
  free(ms); free(poss); free(x1s); free(x2s); free(xs); free(bnd); free(ioff);
  free(pred); free(smh); free(cxt); free(prs); free(rpred); free(rsmh); free(rcxt); free(rprs);
  free(cs);
  if (flags) *flags = fl;
  return err;
}

/* ------------------------------------------------------------------ */
/* Prelearn: per-bit coder with immediate updates                      */
/* (cucm_compress_v2.cu:924-981, cucm_decompress_v2.cu:1056-1182)     */
/* ------------------------------------------------------------------ */

static void prelearn_bit(orc_model *m, const uint8_t *buf, uint64_t pos, bitrec *r) {
  eval_bit(m, buf, pos, r);
}

int64_t orc_prelearn_encode(const uint8_t *in, uint64_t P, orc_model *m,
                            uint8_t *out, uint64_t cap) {
  uint32_t x1 = 0, x2 = 0xffffffffu;
  uint64_t n = 0;
  uint8_t *sink = out;
  for (uint64_t pos = 0; pos < 8 * P; ++pos) {
    bitrec r;
    prelearn_bit(m, in, pos, &r);
    const uint8_t y = (uint8_t)bit_of(in, pos);
    apply_bit_updates(m, &r, &y, 1);
    if (n + 4 >= cap) sink = NULL;
    orc_coder_step(&x1, &x2, r.pr, y, sink, &n);
  }
  if (sink == NULL || n + 1 > cap) return ORC_E_SMALL;
  out[n++] = (uint8_t)(x1 >> 24);
  return (int64_t)n;
}

void orc_prelearn_train(const uint8_t *in, uint64_t P, orc_model *m) {
  for (uint64_t pos = 0; pos < 8 * P; ++pos) {
    bitrec r;
    prelearn_bit(m, in, pos, &r);
    const uint8_t y = (uint8_t)bit_of(in, pos);
    apply_bit_updates(m, &r, &y, 1);
  }
}

int orc_prelearn_decode(const uint8_t *in, uint64_t inlen, uint64_t P,
                        orc_model *m, uint8_t *out) {
  memset(out, 0, P);
  uint64_t ip = 0;
  uint32_t x = 0, x1 = 0, x2 = 0xffffffffu;
  for (int i = 0; i < 4; ++i) x = (x << 8) | (ip < inlen ? in[ip++] : 255u);
  for (uint64_t pos = 0; pos < 8 * P; ++pos) {
    bitrec r;
    prelearn_bit(m, out, pos, &r);
    const uint32_t range = x2 - x1, p = (uint32_t)r.pr;
    const uint32_t xmid = x1 + (range >> 12) * p + (((range & 0xfffu) * p) >> 12);
    const uint8_t y = (uint8_t)(x <= xmid);
    if (y) x2 = xmid; else x1 = xmid + 1u;
    apply_bit_updates(m, &r, &y, 1);
    while (((x1 ^ x2) & 0xff000000u) == 0) {
      x1 <<= 8;
      x2 = (x2 << 8) | 255u;
      x = (x << 8) | (ip < inlen ? in[ip++] : 255u);
    }
    out[pos >> 3] |= (uint8_t)(y << (7 - (pos & 7)));
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Container (cucm_compress_v2.cu:1188-1248, :1420-1443;              */
/*            cucm_decompress_v2.cu:1189-1247, :1297-1301)            */
/* ------------------------------------------------------------------ */
/* Header (21 bytes, big-endian): N u32 | batch_num u16 | U u16 | S u32 |
 * Pc u32 | P u32 | flag u8; then the prelearn segment (raw P bytes if
 * flag == 0, else Pc coded bytes); then per batch a host-order (LE) u32 blob
 * length and the blob.  Reference U values are multiples of 32, so the low
 * five bits of the U field are free: 0 means table_bits 25 (the reference's
 * fixed 2^25 slots), any other value v means table_bits v. */

uint64_t orc_compress_bound(uint64_t N, uint32_t S, uint32_t P, uint32_t batch_num) {
  if (S == 0 || batch_num == 0 || P > N) return 0;
  const uint64_t C = (N - P + S - 1) / S;
  const uint64_t batches = (C + batch_num - 1) / batch_num;
  return ORC_HEADER + (uint64_t)P + 4 * batches + 4 * C + C * chunk_bound(S);
}

static int compress_impl(const uint8_t *in, uint64_t N, uint32_t S, uint32_t U,
                         uint32_t batch_num, uint32_t P, int tb, int tail_pad, int frozen,
                         int threads, uint8_t *out, uint64_t cap, uint64_t *out_len) {
  if (!out || !out_len || (N && !in)) return ORC_E_PARAM;
  if (N > 0xffffffffull || S == 0 || U == 0 || U % 32 || U > 65504 || batch_num == 0 ||
      batch_num > 65535 || P > N || tb < 8 || tb > 25)
    return ORC_E_PARAM;
  if (cap < ORC_HEADER) return ORC_E_SMALL;
  orc_model *m = orc_model_new(tb);
  if (!m) return ORC_E_NOMEM;
  uint64_t Pc = 0;
  uint8_t *pseg = NULL;
  int flag = 0;
  if (P > 0) {
    const uint64_t pcap = 12ull * P + 8;
    pseg = (uint8_t *)malloc(pcap);
    if (!pseg) { orc_model_free(m); return ORC_E_NOMEM; }
    int64_t r = orc_prelearn_encode(in, P, m, pseg, pcap);
    if (r < 0) { orc_model_free(m); free(pseg); return (int)r; }
    Pc = (uint64_t)r;
    flag = (P <= Pc) ? 0 : 1;
  }
  const uint64_t seg = flag ? Pc : P;
  if (cap < ORC_HEADER + seg) { orc_model_free(m); free(pseg); return ORC_E_SMALL; }
  put_be32(out, (uint32_t)N);
  out[4] = (uint8_t)(batch_num >> 8); out[5] = (uint8_t)batch_num;
  const uint32_t ufield = U | (tb == 25 ? 0u : (uint32_t)tb);
  out[6] = (uint8_t)(ufield >> 8); out[7] = (uint8_t)ufield;
  put_be32(out + 8, S);
  put_be32(out + 12, (uint32_t)Pc);
  put_be32(out + 16, P);
  out[20] = (uint8_t)(flag | (frozen ? ORC_FLAG_FROZEN : 0));
  if (seg) memcpy(out + ORC_HEADER, flag ? pseg : in, seg);
  free(pseg);
  uint64_t pos = ORC_HEADER + seg;
  const uint64_t body = N - P;
  const uint64_t C = (body + S - 1) / S;
  const uint64_t bbytes = (uint64_t)S * batch_num;
  int err = ORC_OK;
  for (uint64_t b0 = 0; b0 < C; b0 += batch_num) {
    const uint64_t off = (b0 / batch_num) * bbytes;
    const uint64_t nb = (body - off < bbytes) ? body - off : bbytes;
    if (pos + 4 > cap) { err = ORC_E_SMALL; break; }
    uint64_t blen = 0;
    err = encode_batch_impl(in + P + off, nb, S, U, m, tail_pad, frozen, threads,
                            out + pos + 4, cap - pos - 4, &blen);
    if (err) break;
    put_le32(out + pos, (uint32_t)blen);
    pos += 4 + blen;
  }
  orc_model_free(m);
  if (err) return err;
  *out_len = pos;
  return ORC_OK;
}

int orc_compress(const uint8_t *in, uint64_t N, uint32_t S, uint32_t U,
                 uint32_t batch_num, uint32_t P, int tb, int tail_pad,
                 int threads, uint8_t *out, uint64_t cap, uint64_t *out_len) {
  return compress_impl(in, N, S, U, batch_num, P, tb, tail_pad, 0, threads, out, cap, out_len);
}

int orc_compress_frozen(const uint8_t *in, uint64_t N, uint32_t S, uint32_t U,
                        uint32_t batch_num, uint32_t P, int tb, int threads, uint8_t *out,
                        uint64_t cap, uint64_t *out_len) {
  return compress_impl(in, N, S, U, batch_num, P, tb, 0, 1, threads, out, cap, out_len);
}

int orc_decompress(const uint8_t *in, uint64_t len, int threads, uint8_t *out,
                   uint64_t cap, uint64_t *out_len) {
  if (!in || !out_len || len < ORC_HEADER) return ORC_E_CORRUPT;
  const uint64_t N = get_be32(in);
  const uint32_t batch_num = get_be16(in + 4);
  const uint32_t ufield = get_be16(in + 6);
  const uint64_t S = get_be32(in + 8);
  const uint64_t Pc = get_be32(in + 12);
  const uint64_t P = get_be32(in + 16);
  const int flag = in[20] & 1, frozen = (in[20] & ORC_FLAG_FROZEN) != 0;
  const uint32_t U = ufield & ~31u;
  const int tb = (ufield & 31u) ? (int)(ufield & 31u) : 25;
  if (N > cap) return ORC_E_SMALL;
  if (P > N || in[20] > (1 | ORC_FLAG_FROZEN) || (N > P && (S == 0 || U == 0 || batch_num == 0)) || tb < 8)
    return ORC_E_CORRUPT;
  const uint64_t seg = flag ? Pc : P;
  if (ORC_HEADER + seg > len) return ORC_E_CORRUPT;
  if (N == 0) { *out_len = 0; return ORC_OK; }
  orc_model *m = orc_model_new(tb);
  if (!m) return ORC_E_NOMEM;
  if (P > 0) {
    if (flag == 0) {
      memcpy(out, in + ORC_HEADER, P);
      orc_prelearn_train(out, P, m);
    } else {
      orc_prelearn_decode(in + ORC_HEADER, Pc, P, m, out);
    }
  }
  uint64_t pos = ORC_HEADER + seg;
  const uint64_t body = N - P;
  const uint64_t C = (body + S - 1) / S;
  int err = ORC_OK;
  for (uint64_t b0 = 0; b0 < C; b0 += batch_num) {
    const uint64_t off = (b0 / batch_num) * S * batch_num;
    const uint64_t nb = (body - off < S * batch_num) ? body - off : S * batch_num;
    if (pos + 4 > len) { err = ORC_E_CORRUPT; break; }
    const uint64_t blen = get_le32(in + pos);
    if (pos + 4 + blen > len) { err = ORC_E_CORRUPT; break; }
    err = decode_batch_impl(in + pos + 4, blen, nb, S, U, m, frozen, threads, out + P + off);
    if (err) break;
    pos += 4 + blen;
  }
  orc_model_free(m);
  if (err) return err;
  *out_len = N;
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Many independent files, one file per task                           */
/* ------------------------------------------------------------------ */

/* Compresses nfiles files of the arena (file i at offs[i], lens[i] bytes)
 * into one container each, in parallel over files (OpenMP dynamic, one file
 * per task, each file coded single-threaded), with chunk size
 * min(lens[i], chunk_max) as the batched GPU path cuts them.  Containers go
 * to out + ooffs[i] (capacity ocaps[i]); their sizes to out_lens[i].  This
 * is compress_file (cucm_compress_v2.cu:1151-1250) run once per file. */
int orc_compress_files(const uint8_t *arena, const uint64_t *offs, const uint64_t *lens,
                       uint32_t nfiles, uint32_t chunk_max, uint32_t U, uint32_t batch_num,
                       int tb, int threads, uint8_t *out, const uint64_t *ooffs,
                       const uint64_t *ocaps, uint64_t *out_lens) {
  if (!arena || !offs || !lens || !out || !ooffs || !ocaps || !out_lens || chunk_max == 0)
    return ORC_E_PARAM;
  int err = ORC_OK;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int64_t i = 0; i < (int64_t)nfiles; ++i) {
    const uint64_t n = lens[i];
    const uint32_t S = (uint32_t)(n < chunk_max ? (n ? n : 1) : chunk_max);
    const int rc = orc_compress(arena + offs[i], n, S, U, batch_num, 0, tb, 1, 1,
                                out + ooffs[i], ocaps[i], &out_lens[i]);
    if (rc) {
#ifdef _OPENMP
#pragma omp critical(orc_err)
#endif
      err = rc;
    }
  }
  (void)threads;
  return err;
}

/* Decompresses nfiles containers (container i at in + ioffs[i], ilens[i]
 * bytes) into out + ooffs[i] (capacity ocaps[i]), in parallel over files. */
int orc_decompress_files(const uint8_t *in, const uint64_t *ioffs, const uint64_t *ilens,
                         uint32_t nfiles, int threads, uint8_t *out, const uint64_t *ooffs,
                         const uint64_t *ocaps, uint64_t *out_lens) {
  if (!in || !ioffs || !ilens || !out || !ooffs || !ocaps || !out_lens) return ORC_E_PARAM;
  int err = ORC_OK;
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int64_t i = 0; i < (int64_t)nfiles; ++i) {
    const int rc = orc_decompress(in + ioffs[i], ilens[i], 1, out + ooffs[i], ocaps[i],
                                  &out_lens[i]);
    if (rc) {
#ifdef _OPENMP
#pragma omp critical(orc_err)
#endif
      err = rc;
    }
  }
  (void)threads;
  return err;
}
