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
  bool profile = false;
  uint64_t phase[cucm::kPhases] = {};
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  DevBuf tabs, pool, staging, overflow, tasks, tasks2, results, results2, taskdst, asmjobs,
      headers, joblen, counter, initm, ptrs, hdrs, parsejobs, status, arena_in, arena_out, prof;
  PinnedBuf pin;
  std::vector<cucm::Task> last_tasks;
  std::vector<uint64_t> last_results;
  cucm_stats last{};
  std::mutex mu;
  Server *server = nullptr;  // serving loop, owned
  ~cucm_ctx();
};

struct EncJob {
  const uint8_t *din = nullptr;      // device input
  const uint8_t *host_in = nullptr;  // host copy of the input, if any
  uint64_t n = 0;
  cucm_params p{};
  uint8_t *dout = nullptr;
  uint64_t cap = 0;
  bool blob_only = false;
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
};

int ctx_init(cucm_ctx *c, int device, uint64_t budget);
int encode_run(cucm_ctx *c, std::vector<EncJob> &jobs, cucm_stats &st);
int encode_assemble(cucm_ctx *c, std::vector<EncJob> &jobs, cucm_stats &st);
int decode_run(cucm_ctx *c, std::vector<DecJob> &jobs, uint32_t spec, cucm_stats &st);
void serve_shutdown(cucm_ctx *c);

#endif
