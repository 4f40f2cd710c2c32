// cucm_host_model.h -- host-side tables and prelearn coder.
#ifndef CUCM_HOST_MODEL_H
#define CUCM_HOST_MODEL_H
#include <cstdint>
#include <vector>

#include "cucm_model.h"

namespace cucm {

void build_tables(DevTables *t);
const DevTables &host_tables();

// Full model on the host: 8 order tables of 2^tb one-byte states, statep and
// mixer weights (StateMap/Mixer, cucm_compress_v2.cu:558-679).  Order 0 is
// indexed by slot; orders 1..7 are laid out as on the device (fnv_pos), so
// they upload as one copy.
struct HostModel {
  explicit HostModel(int tb);
  int tb;
  uint32_t mask;
  std::vector<uint8_t> states;  // [8][1 << tb]
  uint32_t statep[kOrders][256];
  int32_t w[kMixCtx * kOrders];
};

// Per-bit sequential coder with immediate updates over the first P bytes
// (cucm_compress_v2.cu:924-981); trains m and writes the coded segment.
void prelearn_encode(const uint8_t *in, uint64_t P, HostModel &m, std::vector<uint8_t> &out);
// Trains m on P raw bytes (cucm_decompress_v2.cu:1134-1182).
void prelearn_train(const uint8_t *in, uint64_t P, HostModel &m);
// Decodes P bytes from a coded segment while training m
// (cucm_decompress_v2.cu:1056-1132).
void prelearn_decode(const uint8_t *in, uint64_t inlen, uint64_t P, HostModel &m, uint8_t *out);
// Order-0 states indexed by the partial byte h0 (the device keeps order 0 in LDS).
void order0_states(const HostModel &m, uint8_t st0[256]);

}  // namespace cucm
#endif
