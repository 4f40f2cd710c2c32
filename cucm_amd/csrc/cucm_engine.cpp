// cucm_engine.cpp -- host engine and C ABI of the MI355X CuCM codec.
//
// The engine turns files into chunk tasks, runs them through one persistent
// launch of the codec kernel (a device work queue over all chunks of all
// files), and assembles the reference container on the device:
//   header (21 B + prelearn segment) | per batch: LE u32 length, 4C BE end
//   offsets, payloads (cucm_compress_v2.cu:1188-1248).
// Decompression parses container headers and batch records on the device
// and decodes all chunks of all containers in one launch.
#define CUCM_WITH_LAUNCHERS
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "cucm.h"
#include "cucm_engine.h"
#include "cucm_host_model.h"
#include "cucm_model.h"

using namespace cucm;

namespace {

constexpr uint32_t kHeader = 21;

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

inline void put_be32(uint8_t *p, uint32_t v) {
  p[0] = (uint8_t)(v >> 24);
  p[1] = (uint8_t)(v >> 16);
  p[2] = (uint8_t)(v >> 8);
  p[3] = (uint8_t)v;
}
inline uint32_t get_be32(const uint8_t *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

// Staging capacity of one chunk.  Real data stays below ~1.01 S; a chunk
// that overflows is re-run with the strict bound (the carry-less coder can
// emit at most 4 bytes per coded bit).
inline uint64_t strict_cap(uint64_t S) { return align_up(32 * S + 16, 16); }

}  // namespace

// ---------------------------------------------------------------------------
// Device buffers
// ---------------------------------------------------------------------------

DevBuf::~DevBuf() {
  if (p) (void)hipFree(p);
}

int DevBuf::reserve(uint64_t n) {
  if (n <= cap && p) return CUCM_OK;
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
  const uint64_t want = std::max<uint64_t>(align_up(n + n / 4, 4096), 4096);
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    if (hipMalloc(&p, std::max<uint64_t>(n, 16)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return CUCM_E_NOMEM;
    }
    cap = std::max<uint64_t>(n, 16);
    return CUCM_OK;
  }
  cap = want;
  return CUCM_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

PinnedBuf::~PinnedBuf() {
  if (p) (void)hipHostFree(p);
}

int PinnedBuf::reserve(uint64_t n) {
  if (n <= cap && p) return CUCM_OK;
  if (p) (void)hipHostFree(p);
  p = nullptr;
  cap = 0;
  const uint64_t want = std::max<uint64_t>(align_up(n + n / 4, 4096), 4096);
  if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    return CUCM_E_NOMEM;
  }
  cap = want;
  return CUCM_OK;
}

#define HIP_TRY(expr)                      \
  do {                                     \
    if ((expr) != hipSuccess) {            \
      (void)hipGetLastError();             \
      return CUCM_E_DEVICE;                \
    }                                      \
  } while (0)

#define CU_TRY(expr)          \
  do {                        \
    int rc_ = (expr);         \
    if (rc_ != CUCM_OK) return rc_; \
  } while (0)

// The 21-byte container header (cucm_compress_v2.cu:1172-1224, format
// comment :1420-1443): BE u32 N, BE u16 batch chunks, BE u16 U with the
// table bits in its low 5 bits (0 = 25), BE u32 S, Pc, P, the flag byte
// (bit 0: the prelearn segment is coded; bit 1, kHeaderFrozen: frozen model).
void write_header(uint8_t *h, uint64_t N, const cucm_params &p, uint32_t Pc, uint32_t flag) {
  put_be32(h, (uint32_t)N);
  h[4] = (uint8_t)(p.batch_chunks >> 8);
  h[5] = (uint8_t)p.batch_chunks;
  const uint32_t uf = p.unit_bits | (p.table_bits == 25 ? 0u : p.table_bits);
  h[6] = (uint8_t)(uf >> 8);
  h[7] = (uint8_t)uf;
  put_be32(h + 8, p.chunk_size);
  put_be32(h + 12, Pc);
  put_be32(h + 16, p.prelearn_size);
  h[20] = (uint8_t)(flag | ((p.flags & CUCM_FLAG_FROZEN) ? kHeaderFrozen : 0u));
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------

int ctx_init(cucm_ctx *c, int device, uint64_t budget) {
  c->device = device;
  HIP_TRY(hipSetDevice(device));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  c->cus = (uint32_t)prop.multiProcessorCount;
  int lds_cu = 0, lds_optin = 0;
  if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) !=
          hipSuccess ||
      lds_cu <= 0)
    lds_cu = 64 * 1024;
  if (hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, device) !=
          hipSuccess ||
      lds_optin <= 0)
    lds_optin = (int)prop.sharedMemPerBlock;
  (void)hipGetLastError();
  c->lds_cu = (uint32_t)lds_cu;
  c->lds_block = (uint32_t)std::min(lds_optin, lds_cu);
  size_t freeb = 0, total = 0;
  HIP_TRY(hipMemGetInfo(&freeb, &total));
  c->budget = budget ? budget : (uint64_t)(freeb * 0.9);
  // the allocation-free entry point's scratch address on this device; if
  // the lookup fails only that entry point reports CUCM_E_DEVICE
  (void)noalloc_setup(device);
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto &e : c->ev) HIP_TRY(hipEventCreate(&e));
  CU_TRY(c->tabs.reserve(sizeof(DevTables)));
  HIP_TRY(hipMemcpy(c->tabs.p, &host_tables(), sizeof(DevTables), hipMemcpyHostToDevice));
  return CUCM_OK;
}

cucm_ctx::~cucm_ctx() {
  serve_shutdown(this);
  (void)hipSetDevice(device);
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto &e : ev)
    if (e) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

}  // namespace

int validate_params(const cucm_params &p, uint64_t n) {
  if (n > 0xffffffffull) return CUCM_E_PARAM;
  if (p.chunk_size == 0 || p.unit_bits == 0 || p.unit_bits % 32 || p.batch_chunks == 0)
    return CUCM_E_PARAM;
  if (p.table_bits < 8 || p.table_bits > 25) return CUCM_E_PARAM;
  if (p.flags & ~(CUCM_FLAG_NO_TAIL_PAD | CUCM_FLAG_FROZEN)) return CUCM_E_PARAM;
  if (p.prelearn_size > n) return CUCM_E_PARAM;
  if (p.unit_bits > kMaxBigUnitBits) return CUCM_E_PARAM;
  return CUCM_OK;
}

namespace {

// Team workgroups the launch can hold at once (0 if a team does not fit).
uint64_t team_capacity(const cucm_ctx *c, uint32_t Umax, uint32_t spec, bool dec) {
  const uint32_t lds = team_codec_lds_bytes(Umax, spec, dec);
  if (lds > c->lds_block) return 0;
  return (uint64_t)c->cus * std::max<uint32_t>(1, c->lds_cu / lds);
}

// Teams run when forced on, or (automatic) when every chunk of the launch
// can be resident as a team at once -- then a chunk's latency bounds the
// launch and the team's extra waves shorten it -- or when the chunks fill at
// least two full passes of resident teams (measured: 4096 x 256 KiB chunks
// code 6% faster as teams than as single waves; 1024 chunks, 1.33 passes of
// teams, 12% slower because of the ragged last pass).
bool use_team(const cucm_ctx *c, uint64_t ntasks, uint32_t Umax, uint32_t spec, bool dec,
              bool one_wave = false) {
  if (one_wave || c->team_mode == 1 || !team_supported(dec, spec)) return false;
  const uint64_t cap = team_capacity(c, Umax, spec, dec);
  if (cap == 0) return false;
  return c->team_mode == 2 || ntasks <= cap || ntasks >= 2 * cap;
}

}  // namespace

Shape pick_shape(const cucm_ctx *c, uint64_t ntasks, uint32_t Umax, uint32_t spec, bool dec,
                 uint64_t slot_bytes, uint64_t mem_left, bool one_wave) {
  Shape s;
  if (use_team(c, ntasks, Umax, spec, dec, one_wave)) {
    const uint64_t cap = team_capacity(c, Umax, spec, dec);
    uint64_t blocks = std::min<uint64_t>(ntasks, cap);
    if (c->max_waves) {
      blocks = std::min<uint64_t>(blocks, std::max<uint32_t>(1, c->max_waves / kTeamWaves));
    } else if (!dec && ntasks >= 4 * cap) {
      // Encoder teams over four or more full passes: at most 9 per 4 CUs
      // resident, fewer chunks in flight than LDS allows serve the memory
      // system better (measured, one box each: the 10,000-file batch at
      // 2^18 codes in 3,650 ms with 576 teams on 256 CUs against 3,770 with
      // 640 and ~3,980 with all 768; 4096 x 256 KiB chunks at 2^22 at
      // 432 MB/s with 512-576 against 426 with 768).
      blocks = std::min<uint64_t>(blocks, std::max<uint64_t>(1, (uint64_t)c->cus * 9 / 4));
    }
    if (slot_bytes) blocks = std::min<uint64_t>(blocks, mem_left / slot_bytes);
    s.team = true;
    s.wpb = kTeamWaves;
    s.blocks = (uint32_t)blocks;
    return s;
  }
  // Decoding with model slots of 8 MiB or more (2^20 slots and up), at most
  // six waves per CU: the seventh that LDS allows at U = 128 speeds the
  // 10,000-file batch up (2^18: 229.9 -> 243.9 MB/s) but slows 4096 x 256
  // KiB chunks at 2^22 down (218.4 -> 207.8 MB/s), where the memory system
  // saturates with fewer chunks in flight.
  const uint32_t wcap = (dec && slot_bytes >= (8ull << 20)) ? 6u : 8u;
  uint32_t wmax = 1;
  for (uint32_t w = wcap; w >= 1; --w)
    if (codec_lds_bytes(Umax, spec, dec, w) <= c->lds_block) {
      wmax = w;
      break;
    }
  // resident capacity for the workgroup size the launch uses (a fixed
  // waves_per_block packs differently into a CU's LDS than wmax does)
  const uint32_t wfit = c->waves_per_block ? std::min<uint32_t>(c->waves_per_block, wmax) : wmax;
  uint64_t waves = ntasks;
  const uint64_t per_cu_blocks = std::max<uint64_t>(1, c->lds_cu / codec_lds_bytes(Umax, spec, dec, wfit));
  waves = std::min<uint64_t>(waves, (uint64_t)c->cus * per_cu_blocks * wfit);
  if (c->max_waves) waves = std::min<uint64_t>(waves, c->max_waves);
  if (slot_bytes) waves = std::min<uint64_t>(waves, mem_left / slot_bytes);
  if (waves == 0) return s;
  uint32_t w = (uint32_t)std::min<uint64_t>(wmax, (waves + c->cus - 1) / c->cus);
  if (c->waves_per_block) w = wfit;
  s.wpb = std::max<uint32_t>(w, 1);
  s.blocks = (uint32_t)((waves + s.wpb - 1) / s.wpb);
  if ((uint64_t)s.blocks * s.wpb > waves && s.blocks > 1) {
    // keep the pool within budget: round down instead of up
    s.blocks = (uint32_t)(waves / s.wpb);
  }
  return s;
}

// Frozen-model tasks: one wave per chunk, workgroups of up to 8 waves
// sharing the tables, as many resident as LDS allows; no model slots (the
// read-only model is read in place).
Shape frozen_shape(const cucm_ctx *c, uint64_t ntasks) {
  Shape s;
  uint32_t wmax = 8;
  while (wmax > 1 && frozen_lds_bytes(wmax) > c->lds_block) --wmax;
  const uint32_t wfit = c->waves_per_block ? std::min<uint32_t>(c->waves_per_block, wmax) : wmax;
  const uint64_t per_cu = std::max<uint64_t>(1, c->lds_cu / frozen_lds_bytes(wfit));
  uint64_t waves = std::min<uint64_t>(ntasks, (uint64_t)c->cus * per_cu * wfit);
  if (c->max_waves) waves = std::min<uint64_t>(waves, c->max_waves);
  if (waves == 0) return s;
  uint32_t w = (uint32_t)std::min<uint64_t>(wmax, (waves + c->cus - 1) / c->cus);
  if (c->waves_per_block) w = wfit;
  s.wpb = std::max<uint32_t>(w, 1);
  s.blocks = (uint32_t)((waves + s.wpb - 1) / s.wpb);
  return s;
}

// Large-unit tasks: one wave per chunk, up to 4 waves per workgroup sharing
// the tables, each wave with a model slot and a region of unit records
// (per_wave bytes together), within mem_left.
Shape big_shape(const cucm_ctx *c, uint64_t ntasks, uint64_t per_wave, uint64_t mem_left) {
  Shape s;
  uint32_t wmax = 4;
  while (wmax > 1 && big_lds_bytes(wmax) > c->lds_block) --wmax;
  const uint64_t per_cu = std::max<uint64_t>(1, c->lds_cu / big_lds_bytes(wmax));
  uint64_t waves = std::min<uint64_t>(ntasks, (uint64_t)c->cus * per_cu * wmax);
  if (c->max_waves) waves = std::min<uint64_t>(waves, c->max_waves);
  if (per_wave) waves = std::min<uint64_t>(waves, mem_left / per_wave);
  if (waves == 0) return s;
  s.wpb = (uint32_t)std::min<uint64_t>(wmax, waves);
  s.blocks = (uint32_t)(waves / s.wpb);
  return s;
}

int plan_pass(cucm_ctx *c, const PassCounts &n, uint32_t spec, bool dec, uint64_t mem_left,
              PassPlan &pl, cucm_stats &st) {
  pl = PassPlan{};
  pl.Umax = n.Umax;
  pl.slot_bytes = align_up(7ull << n.tbmax, 256);
  pl.big_bytes = n.nbig ? big_record_bytes(n.Ubig) : 0;
  pl.Ubig = n.Ubig;
  // Automatic speculation width (spec 0, decoder): 4-bit rounds as teams
  // while every chunk can be resident as a team (the launch is latency-bound
  // and the team's helpers shorten each chunk's chain); with more chunks the
  // launch is bound by memory transactions, and one wave per chunk with
  // 2-bit rounds keeps the most chunks in flight and gathers the fewest lines
  // per bit (measured: 4096 x 256 KiB chunks at 2^22 decode at 218 MB/s
  // against 165 MB/s as 4-bit teams; the 10,000-file batch at 229.5 against
  // 176 MB/s).  Team mode 1 or 2 still forces the choice of execution.
  bool one_wave = false;
  if (dec && spec == 0) {
    spec = 4;
    if (n.nmain && c->team_mode != 2) {
      const uint64_t cap = team_capacity(c, n.Umax, 4, true);
      if (c->team_mode == 1 || cap == 0 || n.nmain > cap) {
        spec = 2;
        one_wave = true;
      }
    }
  }
  pl.spec = spec;
  uint64_t slots = 0;
  if (n.nmain) {
    pl.sh = pick_shape(c, n.nmain, n.Umax, spec, dec, pl.slot_bytes, mem_left, one_wave);
    if (pl.sh.slots() == 0) return CUCM_E_BUDGET;
    slots = pl.sh.slots();
    if (pl.sh.team) st.teams = std::max<uint64_t>(st.teams, pl.sh.blocks);
  }
  if (n.nfrozen) pl.fz = frozen_shape(c, n.nfrozen);
  if (n.nbig) {
    // the large-unit launch runs after the others on the same stream and
    // reuses their model slots
    pl.bg = big_shape(c, n.nbig, pl.slot_bytes + pl.big_bytes, mem_left);
    if (pl.bg.waves() == 0) return CUCM_E_BUDGET;
    slots = std::max<uint64_t>(slots, pl.bg.waves());
    CU_TRY(c->bigrec.reserve(pl.big_bytes * pl.bg.waves()));
  }
  if (slots) CU_TRY(c->pool.reserve(pl.slot_bytes * slots));
  st.model_bytes = std::max<uint64_t>(st.model_bytes, pl.slot_bytes * slots);
  st.waves = std::max<uint64_t>(st.waves, pl.sh.waves() + pl.fz.waves() + pl.bg.waves());
  return CUCM_OK;
}

void plan_args(const cucm_ctx *c, const PassPlan &pl, CodecArgs &a) {
  a.pool = (uint8_t *)c->pool.p;
  a.slot_bytes = pl.slot_bytes;
  a.big = (uint8_t *)c->bigrec.p;
  a.big_bytes = pl.big_bytes;
  a.big_umax = pl.Ubig;
}

// One codec pass over the tasks on the device: the unit codec (shape sh)
// for the tasks without kTaskFrozen or kTaskBigUnit, the frozen codec (fz)
// and the large-unit codec (bg) for those, back to back on the context's
// stream, each kernel pulling from its own queue head (counter words 0, 1
// and 2) and skipping the others' tasks.  With `timed`, events 0 and 1
// bracket the launches.
int launch_pass(cucm_ctx *c, const CodecArgs &a, bool decode, const PassPlan &pl, bool timed,
                cucm_stats &st) {
  const Shape &sh = pl.sh, &fz = pl.fz, &bg = pl.bg;
  const uint32_t Umax = pl.Umax;
  HIP_TRY(hipMemsetAsync(a.counter, 0, 16, c->stream));
  if (timed) HIP_TRY(hipEventRecord(c->ev[0], c->stream));
  if (sh.blocks) {
    HIP_TRY(launch_codec(decode, a, Umax, sh.blocks, sh.wpb, sh.team, c->stream));
    st.launches += 1;
  }
  if (fz.blocks) {
    CodecArgs f = a;
    f.counter = a.counter + 1;
    HIP_TRY(launch_frozen(decode, f, fz.blocks, fz.wpb, c->stream));
    st.launches += 1;
  }
  if (bg.blocks) {
    CodecArgs b = a;
    b.counter = a.counter + 2;
    b.prof = nullptr;
    HIP_TRY(launch_big(decode, b, bg.blocks, bg.wpb, c->stream));
    st.launches += 1;
  }
  if (timed) HIP_TRY(hipEventRecord(c->ev[1], c->stream));
  return CUCM_OK;
}

namespace {

struct InitModels {
  std::vector<DevInitModel> hostside;
  std::vector<std::unique_ptr<DevBuf>> states;
};

int upload_model(cucm_ctx *c, const HostModel &m, InitModels &im) {
  DevInitModel d{};
  auto buf = std::make_unique<DevBuf>();
  const uint64_t bytes = 7ull << m.tb;
  CU_TRY(buf->reserve(bytes));
  HIP_TRY(hipMemcpyAsync(buf->p, m.states.data() + ((size_t)1 << m.tb), bytes,
                         hipMemcpyHostToDevice, c->stream));
  d.states = (const uint8_t *)buf->p;
  std::memcpy(d.statep, m.statep, sizeof d.statep);
  std::memcpy(d.w, m.w, sizeof d.w);
  order0_states(m, d.st0);
  d.tb = (uint32_t)m.tb;
  im.hostside.push_back(d);
  im.states.push_back(std::move(buf));
  // the host model must stay alive until the copy completes
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUCM_OK;
}

}  // namespace

// Arms the optional per-phase cycle counters for one codec launch.
static int prof_arm(cucm_ctx *c, CodecArgs &a) {
  a.prof = nullptr;
  if (!c->profile) return CUCM_OK;
  CU_TRY(c->prof.reserve(8 * kPhases));
  HIP_TRY(hipMemsetAsync(c->prof.p, 0, 8 * kPhases, c->stream));
  a.prof = (uint64_t *)c->prof.p;
  return CUCM_OK;
}

static int prof_collect(cucm_ctx *c, const CodecArgs &a) {
  if (!a.prof) return CUCM_OK;
  HIP_TRY(hipMemcpyAsync(c->phase, a.prof, 8 * kPhases, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUCM_OK;
}

// ---------------------------------------------------------------------------
// Encode
// ---------------------------------------------------------------------------

int encode_run(cucm_ctx *c, std::vector<EncJob> &jobs, cucm_stats &st) {
  InitModels im;
  std::vector<Task> tasks;
  PassCounts cnt;
  uint64_t staging_total = 0;
  for (auto &J : jobs) {
    J.status = validate_params(J.p, J.n);
    J.C = 0;
    J.model = -1;
    J.header.clear();
    if (J.status) continue;
    const uint64_t P = J.p.prelearn_size, S = J.p.chunk_size;
    const uint64_t body = J.n - P;
    J.C = (uint32_t)((body + S - 1) / S);
    // header + prelearn segment
    std::vector<uint8_t> seg;
    uint32_t Pc = 0, flag = 0;
    if (J.ext_model) {
      if (P) {
        J.status = CUCM_E_PARAM;
        continue;
      }
      im.hostside.push_back(*J.ext_model);
      J.model = (int)im.hostside.size() - 1;
    } else if (P > 0) {
      std::vector<uint8_t> raw(P);
      if (J.host_in)
        std::memcpy(raw.data(), J.host_in, P);
      else
        HIP_TRY(hipMemcpy(raw.data(), J.din, P, hipMemcpyDeviceToHost));
      HostModel m(J.p.table_bits);
      std::vector<uint8_t> coded;
      prelearn_encode(raw.data(), P, m, coded);
      Pc = (uint32_t)coded.size();
      flag = P <= Pc ? 0u : 1u;
      seg = flag ? coded : raw;
      CU_TRY(upload_model(c, m, im));
      J.model = (int)im.hostside.size() - 1;
    }
    if (!J.blob_only && !J.records_only) {
      J.header.resize(kHeader + seg.size());
      write_header(J.header.data(), J.n, J.p, Pc, flag);
      if (!seg.empty()) std::memcpy(J.header.data() + kHeader, seg.data(), seg.size());
    }
    J.first = (uint32_t)tasks.size();
    const bool frozen = (J.p.flags & CUCM_FLAG_FROZEN) != 0;
    const bool pad = !(J.p.flags & CUCM_FLAG_NO_TAIL_PAD) && !frozen;
    const bool big = !frozen && J.p.unit_bits > kMaxUnitBits;
    for (uint32_t ci = 0; ci < J.C; ++ci) {
      Task t{};
      t.src = (uint64_t)(uintptr_t)(J.din + P + (uint64_t)ci * S);
      t.L = (ci == J.C - 1) ? (uint32_t)(body - (uint64_t)ci * S) : (uint32_t)S;
      t.S = (uint32_t)S;
      t.U = J.p.unit_bits;
      t.tb = J.p.table_bits;
      t.flags = (uint8_t)(((pad && ci == J.C - 1) ? kTaskTailPad : 0) |
                          (J.model >= 0 ? kTaskInitModel : 0) | (frozen ? kTaskFrozen : 0) |
                          (big ? kTaskBigUnit : 0));
      t.model = J.model >= 0 ? (uint32_t)J.model : 0;
      t.dst = staging_total;  // offset until the staging buffer is placed
      const uint64_t cap = staging_cap(S);
      t.aux = (uint32_t)std::min<uint64_t>(cap, 0xfffffff0ull);
      staging_total += cap;
      tasks.push_back(t);
    }
    cnt.add(J.C, frozen, big, J.p.unit_bits, J.p.table_bits);
  }
  const uint32_t ntasks = (uint32_t)tasks.size();
  st.tasks += ntasks;
  c->last_tasks = tasks;
  if (ntasks == 0) {
    for (auto &J : jobs)
      if (!J.status) J.out_len = J.header.size();
    c->last_results.clear();
    return CUCM_OK;
  }
  // device buffers
  CU_TRY(c->staging.reserve(staging_total));
  for (auto &t : tasks) t.dst += (uint64_t)(uintptr_t)c->staging.p;
  CU_TRY(c->tasks.reserve(sizeof(Task) * ntasks));
  CU_TRY(c->results.reserve(8ull * ntasks));
  CU_TRY(c->counter.reserve(16));
  CU_TRY(c->initm.reserve(sizeof(DevInitModel) * std::max<size_t>(im.hostside.size(), 1)));
  if (!im.hostside.empty())
    HIP_TRY(hipMemcpyAsync(c->initm.p, im.hostside.data(), sizeof(DevInitModel) * im.hostside.size(),
                           hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->tasks.p, tasks.data(), sizeof(Task) * ntasks, hipMemcpyHostToDevice,
                         c->stream));
  const uint64_t used = c->staging.cap + c->tasks.cap + c->results.cap + c->arena_in.cap +
                        c->arena_out.cap;
  const uint64_t mem_left = c->budget > used ? c->budget - used : 0;
  PassPlan pl;
  CU_TRY(plan_pass(c, cnt, 0, false, mem_left, pl, st));
  CodecArgs a{};
  a.tasks = (const Task *)c->tasks.p;
  a.ntasks = ntasks;
  a.spec_bits = 0;
  a.counter = (uint32_t *)c->counter.p;
  a.result = (uint64_t *)c->results.p;
  a.tabs = (const DevTables *)c->tabs.p;
  a.init = (const DevInitModel *)c->initm.p;
  plan_args(c, pl, a);
  CU_TRY(prof_arm(c, a));
  CU_TRY(launch_pass(c, a, false, pl, true, st));
  CU_TRY(prof_collect(c, a));
  std::vector<uint64_t> res(ntasks);
  HIP_TRY(hipMemcpyAsync(res.data(), c->results.p, 8ull * ntasks, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  st.codec_ms += ms;
  // Overflowed chunks: re-run with the strict capacity.
  std::vector<uint32_t> redo;
  for (uint32_t t = 0; t < ntasks; ++t)
    if (res[t] & kResultOverflow) redo.push_back(t);
  if (!redo.empty()) {
    uint64_t extra = 0;
    for (uint32_t t : redo) extra += strict_cap(tasks[t].S);
    CU_TRY(c->overflow.reserve(extra));
    std::vector<Task> rt;
    uint64_t off = 0;
    for (uint32_t t : redo) {
      Task x = tasks[t];
      x.dst = (uint64_t)(uintptr_t)c->overflow.p + off;
      const uint64_t cap = strict_cap(x.S);
      if (cap > 0xfffffff0ull) return CUCM_E_UNSUPPORTED;
      x.aux = (uint32_t)cap;
      off += cap;
      tasks[t] = x;
      rt.push_back(x);
    }
    CU_TRY(c->tasks2.reserve(sizeof(Task) * rt.size()));
    CU_TRY(c->results2.reserve(8 * rt.size()));
    HIP_TRY(hipMemcpyAsync(c->tasks2.p, rt.data(), sizeof(Task) * rt.size(), hipMemcpyHostToDevice,
                           c->stream));
    CodecArgs b = a;
    b.tasks = (const Task *)c->tasks2.p;
    b.ntasks = (uint32_t)rt.size();
    b.result = (uint64_t *)c->results2.p;
    b.prof = nullptr;
    PassCounts rc2;
    for (const Task &x : rt)
      rc2.add(1, (x.flags & kTaskFrozen) != 0, (x.flags & kTaskBigUnit) != 0, x.U, x.tb);
    PassPlan p2;
    CU_TRY(plan_pass(c, rc2, 0, false, mem_left, p2, st));
    plan_args(c, p2, b);
    CU_TRY(launch_pass(c, b, false, p2, false, st));
    std::vector<uint64_t> r2(rt.size());
    HIP_TRY(hipMemcpyAsync(r2.data(), c->results2.p, 8 * rt.size(), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < redo.size(); ++i) {
      if (r2[i] & kResultOverflow) return CUCM_E_UNSUPPORTED;
      res[redo[i]] = r2[i];
    }
    HIP_TRY(hipMemcpyAsync(c->tasks.p, tasks.data(), sizeof(Task) * ntasks, hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipMemcpyAsync(c->results.p, res.data(), 8ull * ntasks, hipMemcpyHostToDevice,
                           c->stream));
  }
  // Container lengths (the same arithmetic the layout kernel performs).
  for (auto &J : jobs) {
    if (J.status) continue;
    uint64_t len = J.header.size();
    for (uint32_t c0 = 0; c0 < J.C; c0 += J.p.batch_chunks) {
      const uint32_t Cb = std::min<uint32_t>(J.p.batch_chunks, J.C - c0);
      uint64_t blob = 4ull * Cb;
      for (uint32_t k = 0; k < Cb; ++k) blob += res[J.first + c0 + k];
      if (blob > 0xffffffffull) {
        J.status = CUCM_E_UNSUPPORTED;
        break;
      }
      len += (J.blob_only ? 0 : 4) + blob;
    }
    J.out_len = len;
  }
  c->last_results = std::move(res);
  return CUCM_OK;
}

int encode_assemble(cucm_ctx *c, std::vector<EncJob> &jobs, cucm_stats &st) {
  std::vector<AsmJob> aj;
  std::vector<uint8_t> headers;
  for (auto &J : jobs) {
    if (J.status) continue;
    if (J.out_len > J.cap) {
      J.status = CUCM_E_SMALL;
      continue;
    }
    AsmJob a{};
    a.out_base = (uint64_t)(uintptr_t)J.dout;
    a.header_off = headers.size();
    a.header_len = (uint32_t)J.header.size();
    a.first_task = J.first;
    a.ntasks = J.C;
    a.batch_chunks = J.blob_only ? std::max<uint32_t>(J.C, 1) : J.p.batch_chunks;
    a.blob_only = J.blob_only ? 1 : 0;
    headers.insert(headers.end(), J.header.begin(), J.header.end());
    aj.push_back(a);
  }
  if (aj.empty()) return CUCM_OK;
  const uint32_t ntasks = (uint32_t)c->last_tasks.size();
  CU_TRY(c->asmjobs.reserve(sizeof(AsmJob) * aj.size()));
  CU_TRY(c->headers.reserve(std::max<size_t>(headers.size(), 16)));
  CU_TRY(c->taskdst.reserve(8ull * std::max<uint32_t>(ntasks, 1)));
  CU_TRY(c->joblen.reserve(8 * aj.size()));
  HIP_TRY(hipMemcpyAsync(c->asmjobs.p, aj.data(), sizeof(AsmJob) * aj.size(),
                         hipMemcpyHostToDevice, c->stream));
  if (!headers.empty())
    HIP_TRY(hipMemcpyAsync(c->headers.p, headers.data(), headers.size(), hipMemcpyHostToDevice,
                           c->stream));
  HIP_TRY(hipEventRecord(c->ev[2], c->stream));
  // task_dst == 0 marks tasks of jobs that are not assembled (copy skips them)
  HIP_TRY(hipMemsetAsync(c->taskdst.p, 0, 8ull * ntasks, c->stream));
  HIP_TRY(launch_assemble((const AsmJob *)c->asmjobs.p, (uint32_t)aj.size(),
                          (const uint8_t *)c->headers.p, (const Task *)c->tasks.p, ntasks,
                          (const uint64_t *)c->results.p, (uint64_t *)c->taskdst.p,
                          (uint64_t *)c->joblen.p, c->stream));
  HIP_TRY(hipEventRecord(c->ev[3], c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev[2], c->ev[3]));
  st.assemble_ms += ms;
  return CUCM_OK;
}

// ---------------------------------------------------------------------------
// Decode
// ---------------------------------------------------------------------------

int decode_run(cucm_ctx *c, std::vector<DecJob> &jobs, uint32_t spec, cucm_stats &st) {
  // (spec 0: chosen with the launch shape, plan_pass)
  if (spec != 0 && spec != 1 && spec != 2 && spec != 4 && spec != 8) return CUCM_E_PARAM;
  const uint32_t n = (uint32_t)jobs.size();
  if (n == 0) return CUCM_OK;
  // 1. headers
  std::vector<uint8_t> hdr(21ull * n);
  {
    std::vector<uint64_t> src(n), len(n);
    for (uint32_t j = 0; j < n; ++j) {
      src[j] = (uint64_t)(uintptr_t)jobs[j].din;
      len[j] = jobs[j].len;
    }
    CU_TRY(c->ptrs.reserve(16ull * n));
    CU_TRY(c->hdrs.reserve(21ull * n));
    HIP_TRY(hipMemcpyAsync(c->ptrs.p, src.data(), 8ull * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync((uint8_t *)c->ptrs.p + 8ull * n, len.data(), 8ull * n,
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_read_headers((const uint64_t *)c->ptrs.p, (const uint64_t *)c->ptrs.p + n, n,
                                (uint8_t *)c->hdrs.p, c->stream));
    HIP_TRY(hipMemcpyAsync(hdr.data(), c->hdrs.p, 21ull * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
  }
  // 2. parse and size
  InitModels im;
  std::vector<ParseJob> pj;
  std::vector<uint32_t> pj_of;
  // Chunk counts come from untrusted headers: each chunk needs at least a
  // 4-byte end offset and one payload byte, so a count above (len - seg) / 5
  // is corrupt; totals are kept in 64 bits and refused above 2^32 - 1.
  uint64_t ntasks = 0;
  PassCounts cnt;
  for (uint32_t j = 0; j < n; ++j) {
    DecJob &J = jobs[j];
    J.status = CUCM_OK;
    J.out_len = 0;
    if (J.records_only) {
      if (J.body > J.cap) {
        J.status = CUCM_E_SMALL;
        continue;
      }
      if (J.body > 0 && (J.S == 0 || J.U == 0 || J.batch == 0 || J.U > kMaxBigUnitBits)) {
        J.status = CUCM_E_PARAM;
        continue;
      }
      J.out_len = J.body;
      if (J.body == 0) continue;
      int model = -1;
      if (J.ext_model) {
        im.hostside.push_back(*J.ext_model);
        model = (int)im.hostside.size() - 1;
      }
      const uint64_t C = (J.body + J.S - 1) / J.S;
      if (C > J.len / 5 || ntasks + C > 0xffffffffull) {
        J.status = CUCM_E_CORRUPT;
        J.out_len = 0;
        continue;
      }
      ParseJob p{};
      p.src = (uint64_t)(uintptr_t)J.din;
      p.len = J.len;
      p.dst = (uint64_t)(uintptr_t)J.dout;
      p.body = J.body;
      p.first_task = (uint32_t)ntasks;
      p.ntasks = (uint32_t)C;
      p.seg = 0;
      p.S = J.S;
      p.U = J.U;
      p.tb = J.tb;
      p.batch = J.batch;
      const bool big = !J.frozen && J.U > (uint32_t)kMaxUnitBits;
      p.flags = (model >= 0 ? kTaskInitModel : 0) | (J.frozen ? kTaskFrozen : 0) |
                (big ? kTaskBigUnit : 0);
      p.model = model >= 0 ? (uint32_t)model : 0;
      ntasks += p.ntasks;
      cnt.add(p.ntasks, J.frozen, big, J.U, J.tb);
      pj.push_back(p);
      pj_of.push_back(j);
      continue;
    }
    if (J.len < kHeader) {
      J.status = CUCM_E_CORRUPT;
      continue;
    }
    const uint8_t *h = hdr.data() + 21ull * j;
    const uint64_t N = get_be32(h);
    const uint32_t batch = ((uint32_t)h[4] << 8) | h[5];
    const uint32_t uf = ((uint32_t)h[6] << 8) | h[7];
    const uint64_t S = get_be32(h + 8), Pc = get_be32(h + 12), P = get_be32(h + 16);
    const uint32_t flag = h[20] & 1u;
    const bool frozen = (h[20] & kHeaderFrozen) != 0;
    const uint32_t U = uf & ~31u;
    const uint32_t tb = (uf & 31u) ? (uf & 31u) : 25u;
    const uint64_t segb = flag ? Pc : P;
    if (P > N || h[20] > (1u | kHeaderFrozen) || tb < 8 || tb > 25 || kHeader + segb > J.len) {
      J.status = CUCM_E_CORRUPT;
      continue;
    }
    const uint64_t body = N - P;
    if (body > 0 && (S == 0 || U == 0 || batch == 0)) {
      J.status = CUCM_E_CORRUPT;
      continue;
    }
    if (N > J.cap) {
      J.status = CUCM_E_SMALL;
      continue;
    }
    const uint64_t C = body ? (body + S - 1) / S : 0;
    if (C > (J.len - kHeader - segb) / 5 || ntasks + C > 0xffffffffull) {
      J.status = CUCM_E_CORRUPT;
      continue;
    }
    int model = -1;
    if (P > 0) {
      std::vector<uint8_t> segbuf(segb), out(P);
      if (J.host_in)
        std::memcpy(segbuf.data(), J.host_in + kHeader, segb);
      else
        HIP_TRY(hipMemcpy(segbuf.data(), J.din + kHeader, segb, hipMemcpyDeviceToHost));
      HostModel m((int)tb);
      if (flag) {
        prelearn_decode(segbuf.data(), segb, P, m, out.data());
      } else {
        out = segbuf;
        prelearn_train(out.data(), P, m);
      }
      if (J.host_out)
        std::memcpy(J.host_out, out.data(), P);
      else
        HIP_TRY(hipMemcpy(J.dout, out.data(), P, hipMemcpyHostToDevice));
      if (body > 0) {
        CU_TRY(upload_model(c, m, im));
        model = (int)im.hostside.size() - 1;
      }
    }
    J.out_len = N;
    if (body == 0) continue;
    ParseJob p{};
    p.src = (uint64_t)(uintptr_t)J.din;
    p.len = J.len;
    p.dst = (uint64_t)(uintptr_t)(J.dout + P);
    p.body = body;
    p.first_task = (uint32_t)ntasks;
    p.ntasks = (uint32_t)C;
    p.seg = (uint32_t)(kHeader + segb);
    p.S = (uint32_t)S;
    p.U = U;
    p.tb = tb;
    p.batch = batch;
    const bool big = !frozen && U > (uint32_t)kMaxUnitBits;
    p.flags = (model >= 0 ? kTaskInitModel : 0) | (frozen ? kTaskFrozen : 0) |
              (big ? kTaskBigUnit : 0);
    p.model = model >= 0 ? (uint32_t)model : 0;
    ntasks += p.ntasks;
    cnt.add(p.ntasks, frozen, big, U, tb);
    pj.push_back(p);
    pj_of.push_back(j);
  }
  st.tasks += ntasks;
  if (ntasks == 0) return CUCM_OK;
  // 3. tasks on the device
  CU_TRY(c->parsejobs.reserve(sizeof(ParseJob) * pj.size()));
  CU_TRY(c->status.reserve(4 * pj.size()));
  CU_TRY(c->tasks.reserve(sizeof(Task) * ntasks));
  CU_TRY(c->results.reserve(8ull * ntasks));
  CU_TRY(c->counter.reserve(16));
  CU_TRY(c->initm.reserve(sizeof(DevInitModel) * std::max<size_t>(im.hostside.size(), 1)));
  if (!im.hostside.empty())
    HIP_TRY(hipMemcpyAsync(c->initm.p, im.hostside.data(), sizeof(DevInitModel) * im.hostside.size(),
                           hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->parsejobs.p, pj.data(), sizeof(ParseJob) * pj.size(),
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(launch_build_decode_tasks((const ParseJob *)c->parsejobs.p, (uint32_t)pj.size(),
                                    (Task *)c->tasks.p, (int32_t *)c->status.p, c->stream));
  const uint64_t used = c->staging.cap + c->tasks.cap + c->results.cap + c->arena_in.cap +
                        c->arena_out.cap;
  const uint64_t mem_left = c->budget > used ? c->budget - used : 0;
  PassPlan pl;
  CU_TRY(plan_pass(c, cnt, spec, true, mem_left, pl, st));
  st.spec_bits = pl.spec;
  CodecArgs a{};
  a.tasks = (const Task *)c->tasks.p;
  a.ntasks = (uint32_t)ntasks;
  a.spec_bits = pl.spec;
  a.counter = (uint32_t *)c->counter.p;
  a.result = (uint64_t *)c->results.p;
  a.tabs = (const DevTables *)c->tabs.p;
  a.init = (const DevInitModel *)c->initm.p;
  plan_args(c, pl, a);
  CU_TRY(prof_arm(c, a));
  CU_TRY(launch_pass(c, a, true, pl, true, st));
  CU_TRY(prof_collect(c, a));
  std::vector<int32_t> stv(pj.size());
  HIP_TRY(hipMemcpyAsync(stv.data(), c->status.p, 4 * pj.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  st.codec_ms += ms;
  for (size_t i = 0; i < pj.size(); ++i)
    if (stv[i]) {
      jobs[pj_of[i]].status = CUCM_E_CORRUPT;
      jobs[pj_of[i]].out_len = 0;
    }
  return CUCM_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

namespace {

std::mutex g_default_mu;
cucm_ctx *g_default = nullptr;

int default_ctx(cucm_ctx **out) {
  if (!g_default) {
    const char *dev = std::getenv("CUCM_DEVICE");
    int rc = cucm_ctx_create(&g_default, dev ? std::atoi(dev) : 0, 0);
    if (rc) {
      g_default = nullptr;
      return rc;
    }
  }
  *out = g_default;
  return CUCM_OK;
}

struct CallTimer {
  cucm_ctx *c;
  cucm_stats st{};
  uint64_t t0;
  explicit CallTimer(cucm_ctx *cc) : c(cc), t0(now_ns()) {}
  ~CallTimer() {
    st.wall_ms = (now_ns() - t0) / 1e6;
    c->last = st;
  }
};

}  // namespace

extern "C" {

void cucm_default_params(cucm_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->chunk_size = 1u << 20;
  p->unit_bits = 128;
  p->batch_chunks = 100;
  p->prelearn_size = 0;
  p->table_bits = 25;
  p->spec_bits = 0;  // automatic
  p->flags = 0;
}

const char *cucm_strerror(int code) {
  switch (code) {
    case CUCM_OK: return "ok";
    case CUCM_E_PARAM: return "invalid parameter";
    case CUCM_E_CORRUPT: return "malformed container";
    case CUCM_E_SMALL: return "output buffer too small";
    case CUCM_E_NOMEM: return "allocation failed";
    case CUCM_E_BUDGET: return "job exceeds the HBM budget";
    case CUCM_E_DEVICE: return "HIP runtime error";
    case CUCM_E_UNSUPPORTED: return "unsupported by this build";
    case CUCM_E_NOTREADY: return "job not finished";
    default: return "unknown error";
  }
}

uint64_t cucm_compress_bound(uint64_t n, const cucm_params *p) {
  if (!p || p->chunk_size == 0 || p->batch_chunks == 0 || p->prelearn_size > n) return 0;
  const uint64_t P = p->prelearn_size, S = p->chunk_size;
  const uint64_t C = (n - P + S - 1) / S;
  const uint64_t B = (C + p->batch_chunks - 1) / p->batch_chunks;
  // prelearn segment <= P raw bytes; per chunk <= 4 bytes per coded bit + flush
  return kHeader + P + 4 * B + 4 * C + C * (32 * S + 1);
}

int cucm_peek_header(const uint8_t *in, uint64_t len, uint64_t *orig, cucm_params *out) {
  if (!in || len < kHeader) return CUCM_E_CORRUPT;
  const uint32_t uf = ((uint32_t)in[6] << 8) | in[7];
  if (orig) *orig = get_be32(in);
  if (out) {
    cucm_default_params(out);
    out->batch_chunks = (uint16_t)(((uint32_t)in[4] << 8) | in[5]);
    out->unit_bits = (uint16_t)(uf & ~31u);
    out->table_bits = (uint8_t)((uf & 31u) ? (uf & 31u) : 25u);
    out->chunk_size = get_be32(in + 8);
    out->prelearn_size = get_be32(in + 16);
    if (in[20] & kHeaderFrozen) out->flags |= CUCM_FLAG_FROZEN;
  }
  return CUCM_OK;
}

int cucm_ctx_create(cucm_ctx **ctx, int device, uint64_t budget) {
  if (!ctx) return CUCM_E_PARAM;
  *ctx = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return CUCM_E_DEVICE;
  }
  if (device < 0 || device >= ndev) return CUCM_E_PARAM;
  cucm_ctx *c = new (std::nothrow) cucm_ctx();
  if (!c) return CUCM_E_NOMEM;
  int rc = ctx_init(c, device, budget);
  if (rc) {
    delete c;
    return rc;
  }
  *ctx = c;
  return CUCM_OK;
}

void cucm_ctx_destroy(cucm_ctx *ctx) { delete ctx; }

int cucm_ctx_set_max_waves(cucm_ctx *ctx, uint32_t waves) {
  if (!ctx) return CUCM_E_PARAM;
  ctx->max_waves = waves;
  return CUCM_OK;
}

int cucm_ctx_set_team(cucm_ctx *ctx, uint32_t mode) {
  if (!ctx || mode > 2) return CUCM_E_PARAM;
  ctx->team_mode = mode;
  return CUCM_OK;
}

int cucm_ctx_set_waves_per_block(cucm_ctx *ctx, uint32_t waves) {
  if (!ctx || waves > 8) return CUCM_E_PARAM;
  ctx->waves_per_block = waves;
  return CUCM_OK;
}

int cucm_ctx_set_phase_profile(cucm_ctx *ctx, int enable) {
  if (!ctx) return CUCM_E_PARAM;
  ctx->profile = enable != 0;
  return CUCM_OK;
}

int cucm_ctx_phase_cycles(cucm_ctx *ctx, uint64_t cycles[CUCM_PHASES]) {
  if (!ctx || !cycles) return CUCM_E_PARAM;
  for (int i = 0; i < CUCM_PHASES; ++i) cycles[i] = ctx->phase[i];
  return CUCM_OK;
}

int cucm_ctx_last_stats(cucm_ctx *ctx, cucm_stats *st) {
  if (!ctx || !st) return CUCM_E_PARAM;
  *st = ctx->last;
  return CUCM_OK;
}

int cucm_ctx_malloc(cucm_ctx *ctx, void **p, uint64_t bytes) {
  if (!ctx || !p) return CUCM_E_PARAM;
  HIP_TRY(hipSetDevice(ctx->device));
  if (hipMalloc(p, std::max<uint64_t>(bytes, 1)) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return CUCM_E_NOMEM;
  }
  return CUCM_OK;
}

int cucm_ctx_free(cucm_ctx *ctx, void *p) {
  if (!ctx) return CUCM_E_PARAM;
  HIP_TRY(hipSetDevice(ctx->device));
  if (p) HIP_TRY(hipFree(p));
  return CUCM_OK;
}

int cucm_ctx_memcpy_h2d(cucm_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx) return CUCM_E_PARAM;
  HIP_TRY(hipSetDevice(ctx->device));
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return CUCM_OK;
}

int cucm_ctx_memcpy_d2h(cucm_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx) return CUCM_E_PARAM;
  HIP_TRY(hipSetDevice(ctx->device));
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return CUCM_OK;
}

int cucm_ctx_synchronize(cucm_ctx *ctx) {
  if (!ctx) return CUCM_E_PARAM;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  HIP_TRY(hipDeviceSynchronize());
  return CUCM_OK;
}

int cucm_compress_many_device(cucm_ctx *ctx, uint32_t n, const uint8_t *const d_in[],
                              const uint64_t in_len[], const cucm_params params[],
                              uint8_t *const d_out[], const uint64_t out_cap[],
                              uint64_t out_len[], int32_t status[]) {
  if (!ctx || (n && (!d_in || !in_len || !params || !d_out || !out_cap || !out_len)))
    return CUCM_E_PARAM;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  CallTimer tm(ctx);
  std::vector<EncJob> jobs(n);
  for (uint32_t i = 0; i < n; ++i) {
    jobs[i].din = d_in[i];
    jobs[i].n = in_len[i];
    jobs[i].p = params[i];
    jobs[i].dout = d_out[i];
    jobs[i].cap = out_cap[i];
  }
  CU_TRY(encode_run(ctx, jobs, tm.st));
  CU_TRY(encode_assemble(ctx, jobs, tm.st));
  int rc = CUCM_OK;
  for (uint32_t i = 0; i < n; ++i) {
    out_len[i] = jobs[i].status ? 0 : jobs[i].out_len;
    if (status) status[i] = jobs[i].status;
    if (jobs[i].status && rc == CUCM_OK) rc = jobs[i].status;
    tm.st.in_bytes += in_len[i];
    tm.st.out_bytes += out_len[i];
  }
  return status ? CUCM_OK : rc;
}

int cucm_decompress_many_device(cucm_ctx *ctx, uint32_t n, const uint8_t *const d_in[],
                                const uint64_t in_len[], uint8_t *const d_out[],
                                const uint64_t out_cap[], uint64_t out_len[], int32_t status[],
                                uint32_t spec_bits) {
  if (!ctx || (n && (!d_in || !in_len || !d_out || !out_cap || !out_len))) return CUCM_E_PARAM;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  CallTimer tm(ctx);
  std::vector<DecJob> jobs(n);
  for (uint32_t i = 0; i < n; ++i) {
    jobs[i].din = d_in[i];
    jobs[i].len = in_len[i];
    jobs[i].dout = d_out[i];
    jobs[i].cap = out_cap[i];
  }
  CU_TRY(decode_run(ctx, jobs, spec_bits, tm.st));
  int rc = CUCM_OK;
  for (uint32_t i = 0; i < n; ++i) {
    out_len[i] = jobs[i].status ? 0 : jobs[i].out_len;
    if (status) status[i] = jobs[i].status;
    if (jobs[i].status && rc == CUCM_OK) rc = jobs[i].status;
    tm.st.in_bytes += in_len[i];
    tm.st.out_bytes += out_len[i];
  }
  return status ? CUCM_OK : rc;
}

int cucm_compress_plain_device(cucm_ctx *ctx, uint32_t n, const uint8_t *const d_in[],
                               const uint64_t in_len[], const cucm_params params[],
                               uint8_t *const d_out[], const uint64_t out_cap[],
                               uint64_t out_len[]) {
  if (!ctx) return CUCM_E_PARAM;
  cucm_stats acc{};
  const uint64_t t0 = now_ns();
  for (uint32_t i = 0; i < n; ++i) {
    int32_t st = 0;
    int rc = cucm_compress_many_device(ctx, 1, d_in + i, in_len + i, params + i, d_out + i,
                                       out_cap + i, out_len + i, &st);
    if (rc) return rc;
    if (st) return st;
    acc.codec_ms += ctx->last.codec_ms;
    acc.assemble_ms += ctx->last.assemble_ms;
    acc.tasks += ctx->last.tasks;
    acc.launches += ctx->last.launches;
    acc.waves = std::max(acc.waves, ctx->last.waves);
    acc.teams = std::max(acc.teams, ctx->last.teams);
    acc.in_bytes += ctx->last.in_bytes;
    acc.out_bytes += ctx->last.out_bytes;
    acc.model_bytes = std::max(acc.model_bytes, ctx->last.model_bytes);
  }
  acc.wall_ms = (now_ns() - t0) / 1e6;
  ctx->last = acc;
  return CUCM_OK;
}

int cucm_decompress_plain_device(cucm_ctx *ctx, uint32_t n, const uint8_t *const d_in[],
                                 const uint64_t in_len[], uint8_t *const d_out[],
                                 const uint64_t out_cap[], uint64_t out_len[],
                                 uint32_t spec_bits) {
  if (!ctx) return CUCM_E_PARAM;
  cucm_stats acc{};
  const uint64_t t0 = now_ns();
  for (uint32_t i = 0; i < n; ++i) {
    int32_t st = 0;
    int rc = cucm_decompress_many_device(ctx, 1, d_in + i, in_len + i, d_out + i, out_cap + i,
                                         out_len + i, &st, spec_bits);
    if (rc) return rc;
    if (st) return st;
    acc.codec_ms += ctx->last.codec_ms;
    acc.spec_bits = std::max(acc.spec_bits, ctx->last.spec_bits);
    acc.tasks += ctx->last.tasks;
    acc.launches += ctx->last.launches;
    acc.waves = std::max(acc.waves, ctx->last.waves);
    acc.teams = std::max(acc.teams, ctx->last.teams);
    acc.in_bytes += ctx->last.in_bytes;
    acc.out_bytes += ctx->last.out_bytes;
    acc.model_bytes = std::max(acc.model_bytes, ctx->last.model_bytes);
  }
  acc.wall_ms = (now_ns() - t0) / 1e6;
  ctx->last = acc;
  return CUCM_OK;
}

int cucm_compress_many(cucm_ctx *ctx, uint32_t n, const uint8_t *const in[],
                       const uint64_t in_len[], const cucm_params params[], uint8_t *const out[],
                       const uint64_t out_cap[], uint64_t out_len[], int32_t status[]) {
  return compress_host(ctx, n, in, in_len, params, out, out_cap, out_len, status, nullptr);
}

int cucm_decompress_many(cucm_ctx *ctx, uint32_t n, const uint8_t *const in[],
                         const uint64_t in_len[], uint8_t *const out[], const uint64_t out_cap[],
                         uint64_t out_len[], int32_t status[], uint32_t spec_bits) {
  return decompress_host(ctx, n, in, in_len, out, out_cap, out_len, status, spec_bits, nullptr);
}

}  // extern "C"

// Host-buffer compression through the context's pinned staging arena.  With
// rec set, every job is coded as batch records only (no header) from the
// initial model rec->ext_model.
int compress_host(cucm_ctx *ctx, uint32_t n, const uint8_t *const in[], const uint64_t in_len[],
                  const cucm_params params[], uint8_t *const out[], const uint64_t out_cap[],
                  uint64_t out_len[], int32_t status[], const EncJob *rec) {
  if (!ctx || (n && (!in || !in_len || !params || !out || !out_cap || !out_len)))
    return CUCM_E_PARAM;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  CallTimer tm(ctx);
  // inputs: one pinned staging buffer, one H2D copy
  std::vector<uint64_t> off(n);
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    off[i] = total;
    total += align_up(in_len[i], 256);
  }
  CU_TRY(ctx->arena_in.reserve(std::max<uint64_t>(total, 256)));
  CU_TRY(ctx->pin.reserve(std::max<uint64_t>(total, 256)));
  for (uint32_t i = 0; i < n; ++i)
    if (in_len[i]) std::memcpy((uint8_t *)ctx->pin.p + off[i], in[i], in_len[i]);
  const uint64_t th0 = now_ns();
  if (total)
    HIP_TRY(hipMemcpyAsync(ctx->arena_in.p, ctx->pin.p, total, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  tm.st.h2d_ms = (now_ns() - th0) / 1e6;
  std::vector<EncJob> jobs(n);
  for (uint32_t i = 0; i < n; ++i) {
    jobs[i].din = (const uint8_t *)ctx->arena_in.p + off[i];
    jobs[i].host_in = in[i];
    jobs[i].n = in_len[i];
    jobs[i].p = params[i];
    jobs[i].cap = UINT64_MAX;
    if (rec) {
      jobs[i].records_only = true;
      jobs[i].ext_model = rec->ext_model;
    }
  }
  CU_TRY(encode_run(ctx, jobs, tm.st));
  std::vector<uint64_t> ooff(n);
  uint64_t ototal = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (jobs[i].status == CUCM_OK && jobs[i].out_len > out_cap[i]) jobs[i].status = CUCM_E_SMALL;
    ooff[i] = ototal;
    if (jobs[i].status == CUCM_OK) ototal += align_up(jobs[i].out_len, 256);
  }
  CU_TRY(ctx->arena_out.reserve(std::max<uint64_t>(ototal, 256)));
  for (uint32_t i = 0; i < n; ++i) {
    jobs[i].dout = (uint8_t *)ctx->arena_out.p + ooff[i];
    jobs[i].cap = jobs[i].out_len;
  }
  CU_TRY(encode_assemble(ctx, jobs, tm.st));
  CU_TRY(ctx->pin.reserve(std::max<uint64_t>(ototal, 256)));
  const uint64_t td0 = now_ns();
  if (ototal)
    HIP_TRY(hipMemcpyAsync(ctx->pin.p, ctx->arena_out.p, ototal, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  tm.st.d2h_ms = (now_ns() - td0) / 1e6;
  int rc = CUCM_OK;
  for (uint32_t i = 0; i < n; ++i) {
    if (jobs[i].status == CUCM_OK) {
      std::memcpy(out[i], (uint8_t *)ctx->pin.p + ooff[i], jobs[i].out_len);
      out_len[i] = jobs[i].out_len;
    } else {
      out_len[i] = 0;
      if (rc == CUCM_OK) rc = jobs[i].status;
    }
    if (status) status[i] = jobs[i].status;
    tm.st.in_bytes += in_len[i];
    tm.st.out_bytes += out_len[i];
  }
  return status ? CUCM_OK : rc;
}

// Host-buffer decompression through the pinned arena.  With rec set, every
// input is a run of whole batch records decoding to rec->body bytes with
// rec's parameters and initial model (n must then be 1).
int decompress_host(cucm_ctx *ctx, uint32_t n, const uint8_t *const in[], const uint64_t in_len[],
                    uint8_t *const out[], const uint64_t out_cap[], uint64_t out_len[],
                    int32_t status[], uint32_t spec_bits, const DecJob *rec) {
  if (!ctx || (n && (!in || !in_len || !out || !out_cap || !out_len))) return CUCM_E_PARAM;
  if (rec && n != 1) return CUCM_E_PARAM;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  CallTimer tm(ctx);
  std::vector<uint64_t> off(n), ooff(n), N(n, 0);
  uint64_t total = 0, ototal = 0;
  for (uint32_t i = 0; i < n; ++i) {
    off[i] = total;
    total += align_up(in_len[i], 256);
    N[i] = rec ? rec->body : (in_len[i] >= kHeader ? get_be32(in[i]) : 0);
    ooff[i] = ototal;
    ototal += align_up(std::min<uint64_t>(N[i], out_cap[i]), 256);
  }
  CU_TRY(ctx->arena_in.reserve(std::max<uint64_t>(total, 256)));
  CU_TRY(ctx->arena_out.reserve(std::max<uint64_t>(ototal, 256)));
  CU_TRY(ctx->pin.reserve(std::max<uint64_t>(std::max(total, ototal), 256)));
  for (uint32_t i = 0; i < n; ++i)
    if (in_len[i]) std::memcpy((uint8_t *)ctx->pin.p + off[i], in[i], in_len[i]);
  const uint64_t th0 = now_ns();
  if (total)
    HIP_TRY(hipMemcpyAsync(ctx->arena_in.p, ctx->pin.p, total, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  tm.st.h2d_ms = (now_ns() - th0) / 1e6;
  std::vector<DecJob> jobs(n);
  for (uint32_t i = 0; i < n; ++i) {
    jobs[i].din = (const uint8_t *)ctx->arena_in.p + off[i];
    jobs[i].host_in = in[i];
    jobs[i].len = in_len[i];
    jobs[i].dout = (uint8_t *)ctx->arena_out.p + ooff[i];
    jobs[i].host_out = out[i];
    jobs[i].cap = out_cap[i];
    if (rec) {
      jobs[i].records_only = true;
      jobs[i].body = rec->body;
      jobs[i].S = rec->S;
      jobs[i].U = rec->U;
      jobs[i].tb = rec->tb;
      jobs[i].batch = rec->batch;
      jobs[i].frozen = rec->frozen;
      jobs[i].ext_model = rec->ext_model;
    }
  }
  CU_TRY(decode_run(ctx, jobs, spec_bits, tm.st));
  const uint64_t td0 = now_ns();
  if (ototal)
    HIP_TRY(hipMemcpyAsync(ctx->pin.p, ctx->arena_out.p, ototal, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  tm.st.d2h_ms = (now_ns() - td0) / 1e6;
  int rc = CUCM_OK;
  for (uint32_t i = 0; i < n; ++i) {
    if (jobs[i].status == CUCM_OK) {
      // the prelearn bytes were written to host_out directly by the host coder
      cucm_params hp{};
      uint64_t Nn = 0;
      if (!rec) cucm_peek_header(in[i], in_len[i], &Nn, &hp);
      const uint64_t P = rec ? 0 : hp.prelearn_size;
      if (jobs[i].out_len > P)
        std::memcpy(out[i] + P, (uint8_t *)ctx->pin.p + ooff[i] + P, jobs[i].out_len - P);
      out_len[i] = jobs[i].out_len;
    } else {
      out_len[i] = 0;
      if (rc == CUCM_OK) rc = jobs[i].status;
    }
    if (status) status[i] = jobs[i].status;
    tm.st.in_bytes += in_len[i];
    tm.st.out_bytes += out_len[i];
  }
  return status ? CUCM_OK : rc;
}

// Uploads a trained host model once; the caller owns the device copy.
int upload_init_model(cucm_ctx *c, const HostModel &m, DevBuf &states, DevInitModel &d) {
  std::lock_guard<std::mutex> lk(c->mu);
  HIP_TRY(hipSetDevice(c->device));
  d = DevInitModel{};
  const uint64_t bytes = 7ull << m.tb;
  CU_TRY(states.reserve(bytes));
  HIP_TRY(hipMemcpyAsync(states.p, m.states.data() + ((size_t)1 << m.tb), bytes,
                         hipMemcpyHostToDevice, c->stream));
  d.states = (const uint8_t *)states.p;
  std::memcpy(d.statep, m.statep, sizeof d.statep);
  std::memcpy(d.w, m.w, sizeof d.w);
  order0_states(m, d.st0);
  d.tb = (uint32_t)m.tb;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return CUCM_OK;
}

extern "C" {

int cucm_ctx_compress(cucm_ctx *ctx, const uint8_t *in, uint64_t in_len, const cucm_params *p,
                      uint8_t *out, uint64_t out_cap, uint64_t *out_len) {
  if (!ctx || !p || !out || !out_len || (in_len && !in)) return CUCM_E_PARAM;
  const uint8_t *ins[1] = {in};
  uint8_t *outs[1] = {out};
  int32_t st = 0;
  int rc = cucm_compress_many(ctx, 1, ins, &in_len, p, outs, &out_cap, out_len, &st);
  return rc ? rc : st;
}

int cucm_ctx_decompress(cucm_ctx *ctx, const uint8_t *in, uint64_t in_len, uint8_t *out,
                        uint64_t out_cap, uint64_t *out_len) {
  if (!ctx || !in || !out_len) return CUCM_E_PARAM;
  const uint8_t *ins[1] = {in};
  uint8_t *outs[1] = {out};
  int32_t st = 0;
  int rc = cucm_decompress_many(ctx, 1, ins, &in_len, outs, &out_cap, out_len, &st, 0);
  return rc ? rc : st;
}

int cucm_compress(const uint8_t *in, uint64_t in_len, const cucm_params *p, uint8_t *out,
                  uint64_t out_cap, uint64_t *out_len) {
  std::lock_guard<std::mutex> lk(g_default_mu);
  cucm_ctx *c = nullptr;
  CU_TRY(default_ctx(&c));
  return cucm_ctx_compress(c, in, in_len, p, out, out_cap, out_len);
}

int cucm_decompress(const uint8_t *in, uint64_t in_len, uint8_t *out, uint64_t out_cap,
                    uint64_t *out_len) {
  std::lock_guard<std::mutex> lk(g_default_mu);
  cucm_ctx *c = nullptr;
  CU_TRY(default_ctx(&c));
  return cucm_ctx_decompress(c, in, in_len, out, out_cap, out_len);
}

int cucm_compress_batch(cucm_ctx *ctx, const uint8_t *in, uint64_t n, uint32_t chunk_size,
                        uint32_t unit_bits, uint32_t table_bits, uint8_t *blob, uint64_t cap,
                        uint64_t *blob_len) {
  if (!ctx || !blob || !blob_len || (n && !in) || n == 0) return CUCM_E_PARAM;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  CallTimer tm(ctx);
  CU_TRY(ctx->arena_in.reserve(n));
  HIP_TRY(hipMemcpy(ctx->arena_in.p, in, n, hipMemcpyHostToDevice));
  std::vector<EncJob> jobs(1);
  EncJob &J = jobs[0];
  J.din = (const uint8_t *)ctx->arena_in.p;
  J.host_in = in;
  J.n = n;
  cucm_default_params(&J.p);
  J.p.chunk_size = chunk_size;
  J.p.unit_bits = (uint16_t)unit_bits;
  J.p.table_bits = (uint8_t)table_bits;
  J.p.batch_chunks = 65535;
  J.blob_only = true;
  J.cap = UINT64_MAX;
  if (unit_bits > 65535) return CUCM_E_PARAM;
  CU_TRY(encode_run(ctx, jobs, tm.st));
  if (J.status) return J.status;
  if ((uint64_t)J.C > 65535) return CUCM_E_UNSUPPORTED;
  if (J.out_len > cap) return CUCM_E_SMALL;
  CU_TRY(ctx->arena_out.reserve(J.out_len));
  J.dout = (uint8_t *)ctx->arena_out.p;
  J.cap = J.out_len;
  CU_TRY(encode_assemble(ctx, jobs, tm.st));
  if (J.status) return J.status;
  HIP_TRY(hipMemcpy(blob, ctx->arena_out.p, J.out_len, hipMemcpyDeviceToHost));
  *blob_len = J.out_len;
  tm.st.in_bytes = n;
  tm.st.out_bytes = J.out_len;
  return CUCM_OK;
}

int cucm_decompress_batch(cucm_ctx *ctx, const uint8_t *blob, uint64_t blob_len, uint8_t *out,
                          uint64_t out_len, uint32_t chunk_size, uint32_t unit_bits,
                          uint32_t table_bits, uint32_t spec_bits) {
  if (!ctx || !blob || !out || out_len == 0 || chunk_size == 0 || unit_bits == 0 ||
      unit_bits % 32 || unit_bits > 65535 || table_bits < 8 || table_bits > 25)
    return CUCM_E_PARAM;
  const uint64_t C = (out_len + chunk_size - 1) / chunk_size;
  if (out_len > 0xffffffffull || C > 65535 || blob_len + 4 > 0xffffffffull) return CUCM_E_PARAM;
  // Wrap the blob in a one-batch container and decode it.
  std::vector<uint8_t> box(kHeader + 4 + blob_len);
  uint8_t *h = box.data();
  put_be32(h, (uint32_t)out_len);
  h[4] = (uint8_t)(C >> 8);
  h[5] = (uint8_t)C;
  const uint32_t uf = unit_bits | (table_bits == 25 ? 0u : table_bits);
  h[6] = (uint8_t)(uf >> 8);
  h[7] = (uint8_t)uf;
  put_be32(h + 8, chunk_size);
  put_be32(h + 12, 0);
  put_be32(h + 16, 0);
  h[20] = 0;
  const uint32_t bl = (uint32_t)blob_len;
  std::memcpy(h + kHeader, &bl, 4);
  std::memcpy(h + kHeader + 4, blob, blob_len);
  const uint8_t *ins[1] = {box.data()};
  const uint64_t lens[1] = {box.size()};
  uint8_t *outs[1] = {out};
  uint64_t got = 0;
  int32_t st = 0;
  int rc = cucm_decompress_many(ctx, 1, ins, lens, outs, &out_len, &got, &st, spec_bits);
  return rc ? rc : st;
}

}  // extern "C"
