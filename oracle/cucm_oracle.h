/*
 * cucm_oracle.h -- CPU restatement of CuCM's chunked context-mixing codec.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product (cucm_amd/) never
 * links, calls or falls back to it.
 *
 * What it restates (reference file:line):
 *   tables            cucm_compress_v2.cu:35-285 (nex), :681-705 (squash/stretch/dt)
 *   StateMap          cucm_compress_v2.cu:558-601
 *   Mixer             cucm_compress_v2.cu:621-679
 *   context hash      cucm_compress_v2.cu:727-743
 *   unit encoder      cucm_compress_v2.cu:707-914, 1067-1115
 *   container         cucm_compress_v2.cu:1151-1250 (writer), cucm_decompress_v2.cu:1184-1248 (reader)
 *   prelearn          cucm_compress_v2.cu:924-981, cucm_decompress_v2.cu:1056-1182
 *   decoder           cucm_decompress_v2.cu:685-1054
 *
 * Two independent encoders are provided: a "clean" per-chunk sequential one
 * and a "literal" one that walks the reference's kernel sequence over all
 * chunks of a batch with the reference's per-slot staging buffers and guards.
 * Two decoders likewise: a clean sequential one and a literal restatement of
 * the reference's lock-step speculative decoder.
 *
 * Parity pinning (none of the reference's own tests exist):
 *   squash/stretch/dt  closed-form known answers + stretch(squash(x)) == x
 *   nex                checksum against all four copies in the reference
 *   FNV hash           textbook FNV-1a-32 vectors
 *   StateMap/Mixer     closed-form first-update / zero-weight known answers
 *   coder              first-step known answer + invertibility
 *   unit schedule, tail padding, zeroed staging:  closed forms of the zero
 *                      model (one-unit chunks code their complement; zeroed
 *                      slots step x1 += 1), a hand-derived two-unit known
 *                      answer, the stale-slot rule on per-bit traces, and a
 *                      restatement of the schedule from the pinned primitives
 *                      (tests/test_oracle_schedule.py)
 *   framing            big-endian offsets/lengths checked field by field;
 *                      clean vs literal encoders and decoders; round trips
 */
#ifndef CUCM_ORACLE_H
#define CUCM_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_ORDERS 8
#define ORC_MIXCTX 80
#define ORC_HEADER 21

enum {
  ORC_OK = 0,
  ORC_E_PARAM = -1,
  ORC_E_SMALL = -2,
  ORC_E_CORRUPT = -3,
  ORC_E_NOMEM = -4,
};

/* ---- tables ---------------------------------------------------------- */
void orc_get_tables(int32_t sq[4096], int32_t st[4096], int32_t dt[1024],
                    uint8_t nex[512]);

/* ---- primitives (exposed for known-answer tests) --------------------- */
uint32_t orc_fnv1a32(const uint8_t *p, size_t n);
/* Slots of the 8 orders for chunk-relative bit b of `chunk`. */
void orc_slots(const uint8_t *chunk, uint64_t b, int tb, uint32_t slots[8]);
void orc_coder_step(uint32_t *x1, uint32_t *x2, int pr, int y, uint8_t *out,
                    uint64_t *n);

typedef struct orc_model orc_model;
orc_model *orc_model_new(int tb);
void orc_model_free(orc_model *m);
int orc_model_copy(orc_model *dst, const orc_model *src);
int orc_model_tb(const orc_model *m);
uint8_t orc_model_state(const orc_model *m, int k, uint32_t slot);
uint32_t orc_model_statep(const orc_model *m, int k, int s);
int32_t orc_model_weight(const orc_model *m, int i);
int orc_model_predict(const orc_model *m, int k, uint32_t slot);
void orc_statemap_update(orc_model *m, int k, uint32_t slot, int y);
int orc_mixer_predict(const orc_model *m, int cxt, const int32_t smp[8]);
void orc_mixer_update(orc_model *m, int y, int pr, int cxt,
                      const int32_t smp[8]);
/* 64-bit digest of a model (states, statep, weights), for equality tests. */
uint64_t orc_model_digest(const orc_model *m);

/* ---- chunk codec (clean) -------------------------------------------- */
/* Encodes one chunk of L valid bytes with nominal chunk size S.  With
 * tail_pad != 0 bits [8L, 8S) are coded with the stale per-slot (pr, y)
 * values exactly as the reference does.  Returns payload length (including
 * the one flush byte) or a negative error. */
int64_t orc_encode_chunk(const uint8_t *chunk, uint64_t L, uint64_t S,
                         uint32_t U, const orc_model *init, int tail_pad,
                         uint8_t *out, uint64_t cap);
int orc_decode_chunk(const uint8_t *pay, uint64_t plen, uint64_t L,
                     uint64_t S, uint32_t U, const orc_model *init,
                     uint8_t *out);

/* ---- frozen model --------------------------------------------------- */
/* The reference has no code for this mode (the paper's "frozen model":
 * no runtime updates after prelearn).  This build's definition: every bit
 * of the chunk is predicted from `m` exactly as the normal path predicts
 * from its model (slots, statep, stretch, mixer context, mix, pr bump), the
 * model is never updated, and exactly 8L bits are coded (no tail padding),
 * then the one flush byte.  The container marks it with bit 1 of the flag
 * byte (offset 20). */
#define ORC_FLAG_FROZEN 2
int64_t orc_encode_chunk_frozen(const uint8_t *chunk, uint64_t L,
                                const orc_model *m, uint8_t *out,
                                uint64_t cap);
int orc_decode_chunk_frozen(const uint8_t *pay, uint64_t plen, uint64_t L,
                            const orc_model *m, uint8_t *out);
int orc_encode_batch_frozen(const uint8_t *in, uint64_t N, uint64_t S,
                            const orc_model *init, int threads, uint8_t *blob,
                            uint64_t cap, uint64_t *blob_len);
int orc_decode_batch_frozen(const uint8_t *blob, uint64_t blen, uint64_t N,
                            uint64_t S, const orc_model *init, int threads,
                            uint8_t *out);
/* ---- batch blob: 4C big-endian end offsets + payloads ---------------- */
uint64_t orc_blob_bound(uint64_t N, uint64_t S);
int orc_encode_batch(const uint8_t *in, uint64_t N, uint64_t S, uint32_t U,
                     const orc_model *init, int tail_pad, int threads,
                     uint8_t *blob, uint64_t cap, uint64_t *blob_len);
int orc_encode_batch_literal(const uint8_t *in, uint64_t N, uint64_t S,
                             uint32_t U, const orc_model *init, uint8_t *blob,
                             uint64_t cap, uint64_t *blob_len);
int orc_decode_batch(const uint8_t *blob, uint64_t blen, uint64_t N,
                     uint64_t S, uint32_t U, const orc_model *init,
                     int threads, uint8_t *out);
/* Restatement of the reference decoder (lock-step chunks, speculation depth
 * m+1 bits, global previous-byte guard, 4-byte x initialisation).  `flags`
 * receives bit 0 if the reference would have read past its payload buffer. */
int orc_decode_batch_literal(const uint8_t *blob, uint64_t blen, uint64_t N,
                             uint64_t S, uint32_t U, const orc_model *init,
                             int m, uint8_t *out, uint32_t *flags);

/* ---- prelearn (per-bit sequential coder, immediate updates) ---------- */
int64_t orc_prelearn_encode(const uint8_t *in, uint64_t P, orc_model *m,
                            uint8_t *out, uint64_t cap);
void orc_prelearn_train(const uint8_t *in, uint64_t P, orc_model *m);
int orc_prelearn_decode(const uint8_t *in, uint64_t inlen, uint64_t P,
                        orc_model *m, uint8_t *out);

/* ---- container ------------------------------------------------------ */
uint64_t orc_compress_bound(uint64_t N, uint32_t S, uint32_t P,
                            uint32_t batch_num);
int orc_compress(const uint8_t *in, uint64_t N, uint32_t S, uint32_t U,
                 uint32_t batch_num, uint32_t P, int tb, int tail_pad,
                 int threads, uint8_t *out, uint64_t cap, uint64_t *out_len);
/* orc_compress with the frozen model: the header's flag byte carries
 * ORC_FLAG_FROZEN and orc_decompress decodes it accordingly. */
int orc_compress_frozen(const uint8_t *in, uint64_t N, uint32_t S, uint32_t U,
                        uint32_t batch_num, uint32_t P, int tb, int threads,
                        uint8_t *out, uint64_t cap, uint64_t *out_len);
int orc_decompress(const uint8_t *in, uint64_t len, int threads, uint8_t *out,
                   uint64_t cap, uint64_t *out_len);

/* ---- many files, one file per task (OpenMP over files) --------------- */
int orc_compress_files(const uint8_t *arena, const uint64_t *offs, const uint64_t *lens,
                       uint32_t nfiles, uint32_t chunk_max, uint32_t U, uint32_t batch_num,
                       int tb, int threads, uint8_t *out, const uint64_t *ooffs,
                       const uint64_t *ocaps, uint64_t *out_lens);
int orc_decompress_files(const uint8_t *in, const uint64_t *ioffs, const uint64_t *ilens,
                         uint32_t nfiles, int threads, uint8_t *out, const uint64_t *ooffs,
                         const uint64_t *ocaps, uint64_t *out_lens);

#ifdef __cplusplus
}
#endif
#endif
