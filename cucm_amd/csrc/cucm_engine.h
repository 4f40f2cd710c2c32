// cucm_engine.h -- internal engine types shared by the engine and the
// serving loop.
#ifndef CUCM_ENGINE_H
#define CUCM_ENGINE_H
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

#include "cucm.h"
#include "cucm_model.h"

struct DevBuf {
  void *p = nullptr;
  uint64_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf();
  int reserve(uint64_t n);
  void release();
};

struct PinnedBuf {
  void *p = nullptr;
  uint64_t cap = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  ~PinnedBuf();
  int reserve(uint64_t n);
};

struct Server;

struct cucm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  uint32_t cus = 0, lds_cu = 0, lds_block = 0;
  uint32_t max_waves = 0;
  uint32_t waves_per_block = 0;  // 0: as many as LDS allows, spread over the CUs
  uint32_t team_mode = 0;        // 0: teams when every chunk fits at once, 1: never, 2: always
  bool profile = false;
  uint64_t phase[cucm::kPhases] = {};
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  DevBuf tabs, pool, bigrec, staging, overflow, tasks, tasks2, results, results2, taskdst, asmjobs,
      headers, joblen, counter, initm, ptrs, hdrs, parsejobs, status, arena_in, arena_out, prof;
  PinnedBuf pin;
  std::vector<cucm::Task> last_tasks;
  std::vector<uint64_t> last_results;
  cucm_stats last{};
  std::mutex mu;
  Server *server = nullptr;  // serving loop, owned
  ~cucm_ctx();
};

// Bit 1 of the container's flag byte: chunks coded with the frozen model.
constexpr uint32_t kHeaderFrozen = 2;

struct EncJob {
  const uint8_t *din = nullptr;      // device input
  const uint8_t *host_in = nullptr;  // host copy of the input, if any
  uint64_t n = 0;
  cucm_params p{};
  uint8_t *dout = nullptr;
  uint64_t cap = 0;
  bool blob_only = false;
  // streaming: batch records only (no header), coded from this already
  // uploaded initial model instead of a fresh one (p.prelearn_size must be 0)
  bool records_only = false;
  const cucm::DevInitModel *ext_model = nullptr;
  int32_t status = 0;
  uint64_t out_len = 0;
  uint32_t C = 0, first = 0;
  int model = -1;
  std::vector<uint8_t> header;
};

struct DecJob {
  const uint8_t *din = nullptr;
  const uint8_t *host_in = nullptr;
  uint64_t len = 0;
  uint8_t *dout = nullptr;
  uint8_t *host_out = nullptr;
  uint64_t cap = 0;
  int32_t status = 0;
  uint64_t out_len = 0;
  // streaming: din holds whole batch records (no header) that decode to
  // `body` bytes with these parameters and initial model
  bool records_only = false;
  uint64_t body = 0;
  uint32_t S = 0, U = 0, tb = 0, batch = 0;
  bool frozen = false;
  const cucm::DevInitModel *ext_model = nullptr;
};

// Waves and launch shape for a persistent launch of `ntasks` chunks: either
// single-wave chunks (wpb waves per workgroup sharing the tables, one model
// slot per wave) or teams (one chunk and one model slot per workgroup of
// kTeamWaves waves).
struct Shape {
  uint32_t blocks = 0, wpb = 0;
  bool team = false;
  uint32_t waves() const { return blocks * wpb; }
  uint32_t slots() const { return team ? blocks : blocks * wpb; }
};
// Shape of the unit codec's launch within mem_left bytes of model slots of
// slot_bytes each (no slot limit when slot_bytes is 0); one_wave rules out
// teams.
Shape pick_shape(const cucm_ctx *c, uint64_t ntasks, uint32_t Umax, uint32_t spec, bool dec,
                 uint64_t slot_bytes, uint64_t mem_left, bool one_wave = false);
// Shape of the frozen-model codec's launch (no model slots).
Shape frozen_shape(const cucm_ctx *c, uint64_t ntasks);
// Tasks of one codec pass by kind: the unit codec's (U <= 256), the frozen
// codec's and the large-unit codec's, with the largest U of the first and
// last kinds and the largest table bits.
struct PassCounts {
  uint64_t nmain = 0, nfrozen = 0, nbig = 0;
  uint32_t Umax = 32, Ubig = 0, tbmax = 8;
  void add(uint64_t n, bool frozen, bool big, uint32_t U, uint32_t tb) {
    if (frozen)
      nfrozen += n;
    else if (big) {
      nbig += n;
      Ubig = U > Ubig ? U : Ubig;
    } else {
      nmain += n;
      Umax = U > Umax ? U : Umax;
    }
    tbmax = tb > tbmax ? tb : tbmax;
  }
};
// Launch shapes of a pass and the buffers they need.
struct PassPlan {
  Shape sh, fz, bg;  // unit codec, frozen codec, large-unit codec
  uint32_t Umax = 32, Ubig = 0;
  uint32_t spec = 0;  // decoder bits per round (spec 0 resolved with the shape)
  uint64_t slot_bytes = 0, big_bytes = 0;
};
Shape big_shape(const cucm_ctx *c, uint64_t ntasks, uint64_t per_wave, uint64_t mem_left);
// Shapes within mem_left bytes, model slots and large-unit records reserved;
// a decoder pass with spec 0 also chooses its speculation width (pl.spec).
int plan_pass(cucm_ctx *c, const PassCounts &n, uint32_t spec, bool dec, uint64_t mem_left,
              PassPlan &pl, cucm_stats &st);
// The plan's buffers in a's pool / slot / large-unit fields.
void plan_args(const cucm_ctx *c, const PassPlan &pl, cucm::CodecArgs &a);
// One codec pass: the unit codec, the frozen codec and the large-unit codec
// back to back on the context's stream, each pulling from its own queue head
// (a.counter[0..2], cleared here).  With `timed`, events 0 and 1 bracket the
// launches.
int launch_pass(cucm_ctx *c, const cucm::CodecArgs &a, bool decode, const PassPlan &pl,
                bool timed, cucm_stats &st);
// Staging capacity of one chunk's payload before assembly (real data stays
// below ~1.01 S; an overflow is flagged in the chunk's result word).
inline uint64_t staging_cap(uint64_t S) { return (S + S / 2 + 256 + 15) / 16 * 16; }

int ctx_init(cucm_ctx *c, int device, uint64_t budget);
// Looks up the allocation-free entry point's device scratch on `device`
// (once per device; the device must be current).
int noalloc_setup(int device);
int encode_run(cucm_ctx *c, std::vector<EncJob> &jobs, cucm_stats &st);
int encode_assemble(cucm_ctx *c, std::vector<EncJob> &jobs, cucm_stats &st);
int decode_run(cucm_ctx *c, std::vector<DecJob> &jobs, uint32_t spec, cucm_stats &st);
void serve_shutdown(cucm_ctx *c);
void write_header(uint8_t *h, uint64_t N, const cucm_params &p, uint32_t Pc, uint32_t flag);
int validate_params(const cucm_params &p, uint64_t n);
int compress_host(cucm_ctx *ctx, uint32_t n, const uint8_t *const in[], const uint64_t in_len[],
                  const cucm_params params[], uint8_t *const out[], const uint64_t out_cap[],
                  uint64_t out_len[], int32_t status[], const EncJob *rec);
int decompress_host(cucm_ctx *ctx, uint32_t n, const uint8_t *const in[], const uint64_t in_len[],
                    uint8_t *const out[], const uint64_t out_cap[], uint64_t out_len[],
                    int32_t status[], uint32_t spec_bits, const DecJob *rec);
namespace cucm {
struct HostModel;
}
int upload_init_model(cucm_ctx *c, const cucm::HostModel &m, DevBuf &states,
                      cucm::DevInitModel &d);

#endif
