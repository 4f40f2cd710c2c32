// cucm_stream.cpp -- streaming container I/O.
//
// The reference reads a whole file into host memory and pushes it through
// the GPU batch by batch (compress_file, cucm_compress_v2.cu:1151-1250;
// decompress, cucm_decompress_v2.cu:1184-1248).  A stream does the same
// with bounded memory: the header and prelearn segment are produced or
// consumed once (the prelearned model is uploaded to the device once), then
// the body goes through in segments of whole batch records, each coded in
// one persistent launch.  Containers are byte-identical to one-shot
// cucm_compress output, so inputs far larger than the HBM budget (up to the
// container's 32-bit size fields) stream through a small device footprint.
#include <cstring>
#include <memory>
#include <vector>

#include "cucm_engine.h"
#include "cucm_host_model.h"

using namespace cucm;

namespace {

constexpr uint64_t kHead = 21;

uint32_t be32(const uint8_t *p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

}  // namespace

struct cucm_stream {
  cucm_ctx *ctx = nullptr;
  bool enc = false;
  cucm_params p{};
  uint64_t N = 0, P = 0, body = 0, done = 0;
  uint32_t C = 0, next_chunk = 0, spec = 0;
  bool has_model = false;
  DevBuf states;
  DevInitModel dm{};
};

extern "C" {

int cucm_container_head_len(const uint8_t *in, uint64_t in_len, uint64_t *head_len) {
  if (!in || !head_len) return CUCM_E_PARAM;
  if (in_len < kHead) return CUCM_E_CORRUPT;
  const uint32_t flag = in[20];
  if (flag > (1u | kHeaderFrozen)) return CUCM_E_CORRUPT;
  *head_len = kHead + ((flag & 1u) ? be32(in + 12) : be32(in + 16));
  return CUCM_OK;
}

int cucm_stream_compress_begin(cucm_ctx *ctx, uint64_t total_len, const cucm_params *p,
                               const uint8_t *prelearn, uint8_t *head, uint64_t head_cap,
                               uint64_t *head_len, cucm_stream **out) {
  if (!ctx || !p || !head || !head_len || !out) return CUCM_E_PARAM;
  *out = nullptr;
  const int rc = validate_params(*p, total_len);
  if (rc) return rc;
  if (p->prelearn_size && !prelearn) return CUCM_E_PARAM;
  auto s = std::make_unique<cucm_stream>();
  s->ctx = ctx;
  s->enc = true;
  s->p = *p;
  s->N = total_len;
  s->P = p->prelearn_size;
  s->body = total_len - s->P;
  s->C = (uint32_t)((s->body + p->chunk_size - 1) / p->chunk_size);
  std::vector<uint8_t> seg;
  uint32_t Pc = 0, flag = 0;
  if (s->P) {
    HostModel m(p->table_bits);
    prelearn_encode(prelearn, s->P, m, seg);
    Pc = (uint32_t)seg.size();
    flag = s->P <= Pc ? 0u : 1u;
    if (!flag) seg.assign(prelearn, prelearn + s->P);
    if (s->body) {
      const int urc = upload_init_model(ctx, m, s->states, s->dm);
      if (urc) return urc;
      s->has_model = true;
    }
  }
  *head_len = kHead + seg.size();
  if (*head_len > head_cap) return CUCM_E_SMALL;
  write_header(head, total_len, *p, Pc, flag);
  if (!seg.empty()) std::memcpy(head + kHead, seg.data(), seg.size());
  *out = s.release();
  return CUCM_OK;
}

int cucm_stream_compress_next(cucm_stream *s, const uint8_t *in, uint64_t len, uint8_t *out,
                              uint64_t out_cap, uint64_t *out_len) {
  if (!s || !s->enc || !out_len || (len && (!in || !out))) return CUCM_E_PARAM;
  *out_len = 0;
  if (len == 0) return CUCM_OK;
  const uint64_t per_batch = (uint64_t)s->p.batch_chunks * s->p.chunk_size;
  if (s->done + len > s->body) return CUCM_E_PARAM;
  if (s->done + len < s->body && len % per_batch) return CUCM_E_PARAM;
  cucm_params q = s->p;
  q.prelearn_size = 0;
  EncJob rec;
  rec.records_only = true;
  rec.ext_model = s->has_model ? &s->dm : nullptr;
  int32_t st = 0;
  const int rc = compress_host(s->ctx, 1, &in, &len, &q, &out, &out_cap, out_len, &st, &rec);
  if (rc) return rc;
  if (st) return st;
  s->done += len;
  return CUCM_OK;
}

int cucm_stream_decompress_begin(cucm_ctx *ctx, const uint8_t *head, uint64_t head_len,
                                 uint32_t spec_bits, uint64_t *orig_len, cucm_params *p,
                                 uint8_t *prelearn_out, uint64_t prelearn_cap,
                                 cucm_stream **out) {
  if (!ctx || !head || !out) return CUCM_E_PARAM;
  *out = nullptr;
  uint64_t need = 0;
  int rc = cucm_container_head_len(head, head_len, &need);
  if (rc) return rc;
  if (head_len < need) return CUCM_E_CORRUPT;
  cucm_params hp;
  uint64_t N = 0;
  rc = cucm_peek_header(head, head_len, &N, &hp);
  if (rc) return rc;
  const uint64_t P = hp.prelearn_size;
  if (P > N || hp.table_bits < 8 || hp.table_bits > 25) return CUCM_E_CORRUPT;
  if (N - P > 0 && (hp.chunk_size == 0 || hp.unit_bits == 0 || hp.batch_chunks == 0))
    return CUCM_E_CORRUPT;
  if (hp.unit_bits > kMaxBigUnitBits) return CUCM_E_CORRUPT;
  if (P > prelearn_cap || (P && !prelearn_out)) return CUCM_E_SMALL;
  auto s = std::make_unique<cucm_stream>();
  s->ctx = ctx;
  s->p = hp;
  s->N = N;
  s->P = P;
  s->body = N - P;
  s->C = s->body ? (uint32_t)((s->body + hp.chunk_size - 1) / hp.chunk_size) : 0;
  s->spec = spec_bits;
  if (P) {
    HostModel m(hp.table_bits);
    const uint8_t *seg = head + kHead;
    if (head[20] & 1u)
      prelearn_decode(seg, need - kHead, P, m, prelearn_out);
    else {
      std::memcpy(prelearn_out, seg, P);
      prelearn_train(prelearn_out, P, m);
    }
    if (s->body) {
      rc = upload_init_model(ctx, m, s->states, s->dm);
      if (rc) return rc;
      s->has_model = true;
    }
  }
  if (orig_len) *orig_len = N;
  if (p) *p = hp;
  *out = s.release();
  return CUCM_OK;
}

int cucm_stream_decompress_next(cucm_stream *s, const uint8_t *records, uint64_t len, uint8_t *out,
                                uint64_t out_cap, uint64_t *out_len) {
  if (!s || s->enc || !out_len || (len && !records)) return CUCM_E_PARAM;
  *out_len = 0;
  if (len == 0) return CUCM_OK;
  // whole records only: LE u32 blob length, then the blob
  uint64_t pos = 0, nrec = 0;
  while (pos < len) {
    if (pos + 4 > len) return CUCM_E_CORRUPT;
    const uint64_t blen = (uint64_t)records[pos] | (uint64_t)records[pos + 1] << 8 |
                          (uint64_t)records[pos + 2] << 16 | (uint64_t)records[pos + 3] << 24;
    pos += 4 + blen;
    if (pos > len) return CUCM_E_CORRUPT;
    ++nrec;
  }
  const uint64_t left = s->C - s->next_chunk;
  const uint64_t chunks = std::min<uint64_t>(nrec * s->p.batch_chunks, left);
  if (chunks == 0 || (chunks + s->p.batch_chunks - 1) / s->p.batch_chunks != nrec)
    return CUCM_E_CORRUPT;
  const uint64_t S = s->p.chunk_size;
  const uint64_t first = s->next_chunk;
  const uint64_t bytes = (first + chunks == s->C) ? s->body - first * S : chunks * S;
  if (bytes > out_cap || !out) return CUCM_E_SMALL;
  DecJob rec;
  rec.records_only = true;
  rec.body = bytes;
  rec.S = s->p.chunk_size;
  rec.U = s->p.unit_bits;
  rec.tb = s->p.table_bits;
  rec.batch = s->p.batch_chunks;
  rec.frozen = (s->p.flags & CUCM_FLAG_FROZEN) != 0;
  rec.ext_model = s->has_model ? &s->dm : nullptr;
  int32_t st = 0;
  const int rc = decompress_host(s->ctx, 1, &records, &len, &out, &out_cap, out_len, &st, s->spec,
                                 &rec);
  if (rc) return rc;
  if (st) return st;
  s->next_chunk += (uint32_t)chunks;
  s->done += bytes;
  return CUCM_OK;
}

void cucm_stream_end(cucm_stream *s) { delete s; }

}  // extern "C"
