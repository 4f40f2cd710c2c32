// cucm_serve.cpp -- serving loop: continuous batching into a running kernel.
//
// Requests (compress or decompress of caller-owned host buffers) enter a FIFO
// through cucm_submit.  A scheduler thread admits them, oldest first, as soon
// as a job slot, room in the task ring and arena memory are free: it copies
// the request into pinned host memory and from there into device memory,
// builds its chunk tasks on the host (headers, prelearned models and batch
// records are parsed there) and publishes them by advancing the ring's
// tail.  One persistent
// launch of serve_kernel (cucm_kernels.hip) takes the tasks of every admitted
// request, compress and decompress mixed, as its waves free up, so a request
// that arrives while others are on the GPU starts on the next free wave
// instead of waiting for a launch to end.  The wave that finishes a job's
// last task assembles the container and raises the job's done word, which
// the scheduler polls; it copies the result to the caller's buffer and frees
// the job's memory.  The kernel's waves leave after a stretch with no work
// (the scheduler stops it 20 ms after the last job, so an idle device is not
// held) and the scheduler relaunches it when work arrives.
//
// Requests the serving kernel does not take (frozen model, units above 128
// bits, tables above 2^20 slots, anything larger than the arenas, or a chunk
// that overflowed its staging region) go to fallback lanes that run them as
// static batches through the one-launch-per-call path.  Started with
// CUCM_SERVE_STATIC_LANES, every request takes the lanes (the design before
// the serving kernel: each lane takes the oldest pending requests of one kind
// and runs them as one batch to the end).
//
// Each request's output depends only on its own input and parameters, never
// on what it ran beside.
#define CUCM_WITH_LAUNCHERS
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>

#include "cucm_engine.h"
#include "cucm_host_model.h"

using namespace cucm;

namespace {

constexpr uint32_t kServeUmax = 128;      // units the serving kernel takes
constexpr uint32_t kServeTbMax = 20;      // table bits its model slots hold
constexpr uint32_t kRing = 1u << 16;      // task ring entries
constexpr uint64_t kIdleTicks = 200000000;  // a wave leaves after 2 s without a task
constexpr uint64_t kStopAfterNs = 20000000;   // the kernel is stopped 20 ms after the last job
constexpr uint64_t kPollTicks = 200;          // host words read at most every 2 us

uint64_t clock_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

uint32_t be32(const uint8_t *p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}
uint32_t le32(const uint8_t *p) {
  return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24;
}
inline uint64_t up256(uint64_t x) { return (x + 255) & ~255ull; }

// First-fit ranges of one buffer (offsets, 256-byte granules).
struct Arena {
  uint64_t cap = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> bytes
  void init(uint64_t c) {
    cap = c;
    free_.clear();
    if (c) free_[0] = c;
  }
  bool alloc(uint64_t n, uint64_t &off) {
    n = up256(n ? n : 1);
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= n) {
        off = it->first;
        const uint64_t rest = it->second - n;
        free_.erase(it);
        if (rest) free_[off + n] = rest;
        return true;
      }
    return false;
  }
  void release(uint64_t off, uint64_t n) {
    n = up256(n ? n : 1);
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + n == next->first) {
      n += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += n;
        return;
      }
    }
    free_[off] = n;
  }
};

// Host memory shared with the device (coherent, mapped): p for the host, d
// the device's address of the same bytes.
struct Shared {
  void *p = nullptr;
  void *d = nullptr;
  ~Shared() {
    if (p) (void)hipHostFree(p);
  }
  int alloc(uint64_t n) {
    if (hipHostMalloc(&p, n, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return CUCM_E_NOMEM;
    }
    if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess || !d) {
      (void)hipGetLastError();
      return CUCM_E_DEVICE;
    }
    std::memset(p, 0, n);
    return CUCM_OK;
  }
  template <class T>
  T *dev(const void *host) const {
    return (T *)((uint8_t *)d + ((const uint8_t *)host - (const uint8_t *)p));
  }
};

// What a request's admission computes from the request alone: its
// parameters checked; for compression the container header and prelearn
// segment, for decompression the header, the batch records' chunk payloads
// and the prelearned model.  Made once per request by the scheduler thread
// outside the server lock (the prelearn coder is serial), and kept while the
// request waits for a job slot, ring room or arena memory.
struct AdmitPrep {
  int rc = CUCM_OK;
  std::vector<uint8_t> head_bytes;                   // compress: header + prelearn segment
  std::unique_ptr<HostModel> model;                  // prelearned model (P > 0)
  uint64_t S = 0, C = 0, P = 0, N = 0;
  uint32_t U = 0, tb = 0, batch = 0;
  std::vector<std::pair<uint64_t, uint32_t>> pays;  // decompress: (offset, length) per chunk
  std::vector<uint8_t> pre;                          // decompress: the prelearn bytes
};

struct SJob {
  cucm_job job{};
  std::shared_ptr<AdmitPrep> prep;  // serving kernel: admission inputs, once made
  uint64_t t_submit = 0, t_start = 0, t_done = 0;
  int32_t status = 0;
  uint32_t batch = 0;
  uint64_t out_len = 0;
  bool done = false;
  // serving kernel
  int slot = -1;
  uint32_t gen = 0, ntasks = 0;
  uint64_t h_off = 0, h_len = 0, d_off = 0, d_len = 0, i_off = 0, i_len = 0;
  uint64_t out_off = 0;  // container / decoded bytes in the shared arena
};

// Device bytes a request holds while it runs on a fallback lane: its input,
// output and (for compression) the per-chunk staging of 1.5 S + 256 bytes.
uint64_t footprint(const SJob &j) {
  if (j.job.op == CUCM_OP_COMPRESS) {
    const uint64_t n = j.job.in_len;
    const uint64_t S = j.job.params.chunk_size ? j.job.params.chunk_size : 1;
    return n + (n + n / 2) + 256 * (n / S + 1) + j.job.out_cap;
  }
  const uint64_t N = j.job.in_len >= 4 ? be32(j.job.in) : 0;
  return j.job.in_len + N;
}

// The serving kernel's control block, in shared host memory.
struct Ctl {
  uint32_t tail;
  uint32_t stop;
};

}  // namespace

struct Server {
  cucm_ctx *parent = nullptr;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done, cv_work;
  std::unordered_map<uint64_t, SJob> jobs;
  uint64_t next_id = 1;
  bool stop = false;
  uint64_t launches = 0;
  // scheduler and serving kernel
  std::deque<uint64_t> pending;       // not yet admitted
  std::vector<uint64_t> slot_job;     // job id per slot (0: free)
  std::vector<uint32_t> free_slots;
  uint32_t J = 0;
  uint32_t tail = 0, inflight_tasks = 0, inflight_jobs = 0, gen = 0;
  hipStream_t ks = nullptr;
  Shared ctl_mem, ring_mem, jobs_mem, done_mem, lens_mem, models_mem, harena_mem;
  Ctl *ctl = nullptr;
  Task *ring = nullptr;
  ServeJob *sjobs = nullptr;
  uint32_t *done_w = nullptr;
  uint64_t *lens = nullptr;
  DevInitModel *models = nullptr;
  uint8_t *hbase = nullptr;
  DevBuf pool, results, head, jobcount, staging, dinput;
  Arena harena, darena, iarena;
  hipStream_t cs = nullptr;  // copies of admitted requests' inputs to the device
  Shape shape;
  uint64_t slot_bytes = 0;
  bool engine_ok = false;
  std::thread sched;
  // fallback lanes
  std::deque<uint64_t> fb_pending;
  std::vector<cucm_ctx *> lanes;
  std::vector<std::thread> lane_threads;
  uint64_t lane_budget = 0;
  uint32_t fb_inflight = 0;

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_sched.notify_all();
    cv_work.notify_all();
    if (sched.joinable()) sched.join();
    for (auto &t : lane_threads)
      if (t.joinable()) t.join();
    for (auto *c : lanes) delete c;
    if (ctl) {
      __atomic_store_n(&ctl->stop, 1u, __ATOMIC_RELEASE);
      if (ks) (void)hipStreamSynchronize(ks);
    }

    if (ks) (void)hipStreamDestroy(ks);
    if (cs) (void)hipStreamDestroy(cs);
  }

  int init_engine(uint32_t max_jobs);
  void sched_loop();
  void lane_loop(cucm_ctx *c);
  bool eligible(const SJob &j) const;
  int admit(uint64_t id, SJob &j);  // CUCM_OK, CUCM_E_NOTREADY (retry later) or an error
  void publish(SJob &j, uint32_t first, uint32_t n);
  void finish(uint64_t id, SJob &j, uint32_t word);
  void release(SJob &j);
  void complete(SJob &j, int32_t status, uint64_t out_len);
  int launch_kernel();
};

int Server::init_engine(uint32_t max_jobs) {
  cucm_ctx *c = parent;
  J = max_jobs;
  if (hipSetDevice(c->device) != hipSuccess) return CUCM_E_DEVICE;
  if (hipStreamCreateWithFlags(&ks, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
    return CUCM_E_DEVICE;
  int rc;
  if ((rc = ctl_mem.alloc(sizeof(Ctl))) || (rc = ring_mem.alloc(sizeof(Task) * kRing)) ||
      (rc = jobs_mem.alloc(sizeof(ServeJob) * J)) || (rc = done_mem.alloc(4ull * J)) ||
      (rc = lens_mem.alloc(8ull * J)) || (rc = models_mem.alloc(sizeof(DevInitModel) * J)))
    return rc;
  ctl = (Ctl *)ctl_mem.p;
  ring = (Task *)ring_mem.p;
  sjobs = (ServeJob *)jobs_mem.p;
  done_w = (uint32_t *)done_mem.p;
  lens = (uint64_t *)lens_mem.p;
  models = (DevInitModel *)models_mem.p;
  // arenas: a quarter of the budget for request bytes in shared host memory
  // (at most 2 GiB) and as much again for payload staging on the device
  const uint64_t arena = std::min<uint64_t>(c->budget / 4, 2ull << 30);
  if ((rc = harena_mem.alloc(arena))) return rc;
  hbase = (uint8_t *)harena_mem.p;
  harena.init(arena);
  if ((rc = staging.reserve(arena))) return rc;
  darena.init(arena);
  // and as much again for the requests' inputs (and prelearned models) on
  // the device, copied there at admission: read across the host link, a
  // decoder's payload window and an encoder's unit bytes stall every wave
  if ((rc = dinput.reserve(arena))) return rc;
  iarena.init(arena);
  // the kernel's resident waves: as LDS allows, within the rest of the budget
  slot_bytes = up256(7ull << kServeTbMax);
  // as many waves per CU as LDS holds, in one workgroup (seven at U = 128):
  // measured on the serving mix against four per CU, 296 against 216 MB/s
  // saturated and p99 169 against 202 ms at the static lanes' saturated
  // arrival rate (cucm_ctx_set_waves_per_block sets fewer)
  uint32_t wmax = c->waves_per_block ? std::min<uint32_t>(c->waves_per_block, 8u) : 8u;
  while (wmax > 1 && serve_lds_bytes(wmax) > c->lds_block) --wmax;
  const uint64_t per_cu = std::max<uint64_t>(1, c->lds_cu / serve_lds_bytes(wmax));
  uint64_t waves = (uint64_t)c->cus * per_cu * wmax;
  const uint64_t left = c->budget > 2 * arena ? c->budget - 2 * arena : 0;
  waves = std::min<uint64_t>(waves, left / slot_bytes);
  if (c->max_waves) waves = std::min<uint64_t>(waves, c->max_waves);
  if (waves == 0) return CUCM_E_BUDGET;
  const uint32_t wpb = (uint32_t)std::min<uint64_t>(wmax, (waves + c->cus - 1) / c->cus);
  shape.wpb = wpb;
  shape.blocks = (uint32_t)std::max<uint64_t>(1, waves / wpb);
  if ((rc = pool.reserve(slot_bytes * shape.waves())) || (rc = results.reserve(8ull * kRing)) ||
      (rc = head.reserve(256)) || (rc = jobcount.reserve(4ull * J)))
    return rc;
  if (hipMemset(head.p, 0, 256) != hipSuccess || hipMemset(jobcount.p, 0, 4ull * J) != hipSuccess)
    return CUCM_E_DEVICE;
  slot_job.assign(J, 0);
  for (uint32_t s = J; s-- > 0;) free_slots.push_back(s);
  engine_ok = true;
  return CUCM_OK;
}

bool Server::eligible(const SJob &j) const {
  if (!engine_ok) return false;
  if (j.job.op == CUCM_OP_COMPRESS) {
    const cucm_params &p = j.job.params;
    return p.unit_bits <= kServeUmax && p.table_bits <= kServeTbMax &&
           !(p.flags & CUCM_FLAG_FROZEN);
  }
  if (j.job.in_len < 21) return true;  // reported corrupt by admit
  const uint8_t *h = j.job.in;
  const uint32_t uf = ((uint32_t)h[6] << 8) | h[7];
  const uint32_t U = uf & ~31u, tb = (uf & 31u) ? (uf & 31u) : 25u;
  return U <= kServeUmax && tb <= kServeTbMax && !(h[20] & kHeaderFrozen);
}

void Server::complete(SJob &j, int32_t status, uint64_t out_len) {
  j.status = status;
  j.out_len = status ? 0 : out_len;
  j.t_done = clock_ns();
  j.done = true;
  cv_done.notify_all();
}

void Server::release(SJob &j) {
  if (j.h_len) harena.release(j.h_off, j.h_len);
  if (j.d_len) darena.release(j.d_off, j.d_len);
  if (j.i_len) iarena.release(j.i_off, j.i_len);
  j.h_len = j.d_len = j.i_len = 0;
  if (j.slot >= 0) {
    slot_job[j.slot] = 0;
    free_slots.push_back((uint32_t)j.slot);
    j.slot = -1;
  }
}

void Server::publish(SJob &j, uint32_t first, uint32_t n) {
  (void)first;
  // the ring entries and the job record land before the new tail
  std::atomic_thread_fence(std::memory_order_release);
  tail += n;
  __atomic_store_n(&ctl->tail, tail, __ATOMIC_RELEASE);
  inflight_tasks += n;
  ++inflight_jobs;
  j.batch = inflight_jobs;
}

// Admission of one request: CUCM_E_NOTREADY when a slot, ring room or arena
// memory is not free yet.  On success the request's tasks are published (or,
// with no chunks, it is complete).
// The request-only part of admission (see AdmitPrep).
static void prepare_request(const cucm_job &q, AdmitPrep &a) {
  uint64_t seg = 0;
  if (q.op == CUCM_OP_COMPRESS) {
    const cucm_params &p = q.params;
    if ((a.rc = validate_params(p, q.in_len))) return;
    a.S = p.chunk_size;
    a.P = p.prelearn_size;
    a.U = p.unit_bits;
    a.tb = p.table_bits;
    a.batch = p.batch_chunks;
    a.C = (q.in_len - a.P + a.S - 1) / a.S;
    uint32_t Pc = 0, flag = 0;
    std::vector<uint8_t> segb;
    if (a.P) {
      a.model.reset(new HostModel((int)a.tb));
      std::vector<uint8_t> coded;
      prelearn_encode(q.in, a.P, *a.model, coded);
      Pc = (uint32_t)coded.size();
      flag = a.P <= Pc ? 0u : 1u;
      if (flag)
        segb = coded;
      else
        segb.assign(q.in, q.in + a.P);
    }
    a.head_bytes.resize(21 + segb.size());
    write_header(a.head_bytes.data(), q.in_len, p, Pc, flag);
    if (!segb.empty()) std::memcpy(a.head_bytes.data() + 21, segb.data(), segb.size());
    return;
  }
  if (q.in_len < 21) {
    a.rc = CUCM_E_CORRUPT;
    return;
  }
  const uint8_t *h = q.in;
  a.N = be32(h);
  a.batch = ((uint32_t)h[4] << 8) | h[5];
  const uint32_t uf = ((uint32_t)h[6] << 8) | h[7];
  a.S = be32(h + 8);
  const uint64_t Pc = be32(h + 12);
  a.P = be32(h + 16);
  const uint32_t flag = h[20] & 1u;
  a.U = uf & ~31u;
  a.tb = (uf & 31u) ? (uf & 31u) : 25u;
  seg = flag ? Pc : a.P;
  if (a.P > a.N || h[20] > (1u | kHeaderFrozen) || a.tb < 8 || 21 + seg > q.in_len) {
    a.rc = CUCM_E_CORRUPT;
    return;
  }
  const uint64_t body = a.N - a.P;
  if (body > 0 && (a.S == 0 || a.U == 0 || a.batch == 0)) {
    a.rc = CUCM_E_CORRUPT;
    return;
  }
  if (a.N > q.out_cap) {
    a.rc = CUCM_E_SMALL;
    return;
  }
  a.C = body ? (body + a.S - 1) / a.S : 0;
  if (a.C > (q.in_len - 21 - seg) / 5) {
    a.rc = CUCM_E_CORRUPT;
    return;
  }
  // batch records, as build_decode_tasks_kernel walks them
  uint64_t pos = 21 + seg;
  for (uint64_t c0 = 0; c0 < a.C; c0 += a.batch) {
    const uint64_t Cb = std::min<uint64_t>(a.batch, a.C - c0);
    const uint64_t blen = pos + 4 <= q.in_len ? le32(q.in + pos) : 0;
    if (pos + 4 > q.in_len || pos + 4 + blen > q.in_len || blen < 5 * Cb) {
      a.rc = CUCM_E_CORRUPT;
      return;
    }
    const uint8_t *blob = q.in + pos + 4;
    uint64_t prev = 4 * Cb - 1;
    for (uint64_t c = 0; c < Cb; ++c) {
      const uint64_t E = be32(blob + 4 * c);
      if (E <= prev || E >= blen) {
        a.rc = CUCM_E_CORRUPT;
        return;
      }
      a.pays.push_back({pos + 4 + prev + 1, (uint32_t)(E - prev)});
      prev = E;
    }
    if (prev != blen - 1) {
      a.rc = CUCM_E_CORRUPT;
      return;
    }
    pos += 4 + blen;
  }
  if (a.P) {
    a.model.reset(new HostModel((int)a.tb));
    a.pre.resize(a.P);
    if (flag) {
      prelearn_decode(q.in + 21, seg, a.P, *a.model, a.pre.data());
    } else {
      std::memcpy(a.pre.data(), q.in + 21, a.P);
      prelearn_train(a.pre.data(), a.P, *a.model);
    }
  }
}

int Server::admit(uint64_t id, SJob &j) {
  if (free_slots.empty() || inflight_tasks >= kRing) return CUCM_E_NOTREADY;
  const cucm_job &q = j.job;
  const AdmitPrep &a = *j.prep;
  if (a.rc) return a.rc;
  const uint64_t S = a.S, C = a.C, P = a.P, N = a.N;
  const uint32_t U = a.U, tb = a.tb, batch = a.batch;
  const HostModel *model = a.model.get();
  const std::vector<uint8_t> &head_bytes = a.head_bytes;
  const std::vector<std::pair<uint64_t, uint32_t>> &pays = a.pays;
  const std::vector<uint8_t> &pre = a.pre;
  if (C > kRing) return CUCM_E_UNSUPPORTED;  // more chunks than the ring: a fallback lane
  if (inflight_tasks + C > kRing) return CUCM_E_NOTREADY;
  // memory: input | output | model states in the shared arena; staging on the device
  const uint64_t cap = staging_cap(S ? S : 1);
  const uint64_t out_bytes = q.op == CUCM_OP_COMPRESS
                                 ? head_bytes.size() + 4 * ((C + batch - 1) / std::max<uint32_t>(batch, 1)) +
                                       C * (4 + cap)
                                 : N;
  const uint64_t states = model ? (7ull << tb) : 0;
  const uint64_t in_off = 0, out_off = up256(q.in_len), st_off = out_off + up256(out_bytes);
  const uint64_t hneed = st_off + up256(states);
  // device staging: compress payloads, or decompress chunks' decoded bytes
  const uint64_t dneed = q.op == CUCM_OP_COMPRESS ? C * cap : C * S;
  // the input and the model's states, on the device
  const uint64_t ineed = up256(q.in_len) + up256(states);
  if (hneed > harena.cap || dneed > darena.cap || ineed > iarena.cap) return CUCM_E_UNSUPPORTED;
  uint64_t hoff = 0, doff = 0, ioff = 0;
  if (!harena.alloc(hneed, hoff)) return CUCM_E_NOTREADY;
  if (dneed && !darena.alloc(dneed, doff)) {
    harena.release(hoff, hneed);
    return CUCM_E_NOTREADY;
  }
  if (!iarena.alloc(ineed, ioff)) {
    harena.release(hoff, hneed);
    if (dneed) darena.release(doff, dneed);
    return CUCM_E_NOTREADY;
  }
  j.h_off = hoff;
  j.h_len = hneed;
  j.d_off = doff;
  j.d_len = dneed;
  j.i_off = ioff;
  j.i_len = ineed;
  uint8_t *dib = (uint8_t *)dinput.p + ioff;
  j.slot = (int)free_slots.back();
  free_slots.pop_back();
  slot_job[j.slot] = id;
  j.gen = (++gen) & 0x7fffffffu;
  if (j.gen == 0) j.gen = (++gen) & 0x7fffffffu;
  uint8_t *hb = hbase + hoff;
  if (q.in_len) std::memcpy(hb + in_off, q.in, q.in_len);
  bool copy_ok = !q.in_len || hipMemcpyAsync(dib, hb + in_off, q.in_len, hipMemcpyHostToDevice,
                                             cs) == hipSuccess;
  uint8_t *out = hb + out_off;
  j.out_off = hoff + out_off;
  uint32_t mflag = 0;
  if (model) {
    std::memcpy(hb + st_off, model->states.data() + ((size_t)1 << tb), states);
    copy_ok = copy_ok && hipMemcpyAsync(dib + up256(q.in_len), hb + st_off, states,
                                        hipMemcpyHostToDevice, cs) == hipSuccess;
    DevInitModel &d = models[j.slot];
    d.states = dib + up256(q.in_len);
    std::memcpy(d.statep, model->statep, sizeof d.statep);
    std::memcpy(d.w, model->w, sizeof d.w);
    order0_states(*model, d.st0);
    d.tb = tb;
    mflag = kTaskInitModel;
  }
  // the tasks read the device copy: it must be complete before they are published
  if (copy_ok && hipStreamSynchronize(cs) != hipSuccess) copy_ok = false;
  if (!copy_ok) {
    (void)hipGetLastError();
    release(j);
    return CUCM_E_DEVICE;
  }
  j.t_start = clock_ns();
  if (q.op == CUCM_OP_COMPRESS) {
    std::memcpy(out, head_bytes.data(), head_bytes.size());
  } else if (P) {
    std::memcpy(out, pre.data(), P);
  }
  j.ntasks = (uint32_t)C;
  if (C == 0) {
    const uint64_t len = q.op == CUCM_OP_COMPRESS ? head_bytes.size() : N;
    int32_t st = CUCM_OK;
    if (len > q.out_cap)
      st = CUCM_E_SMALL;
    else if (len)
      std::memcpy(q.out, out, len);
    release(j);
    complete(j, st, len);
    return CUCM_OK;
  }
  const uint32_t first = tail;
  const bool pad = q.op == CUCM_OP_COMPRESS && !(q.params.flags & CUCM_FLAG_NO_TAIL_PAD);
  for (uint64_t c = 0; c < C; ++c) {
    Task t{};
    const uint64_t body = (q.op == CUCM_OP_COMPRESS ? q.in_len : N) - P;
    t.L = (c == C - 1) ? (uint32_t)(body - c * S) : (uint32_t)S;
    t.S = (uint32_t)S;
    t.U = (uint16_t)U;
    t.tb = (uint8_t)tb;
    t.model = (uint32_t)j.slot;
    t.job = (uint32_t)j.slot;
    if (q.op == CUCM_OP_COMPRESS) {
      t.src = (uint64_t)(uintptr_t)(dib + in_off + P + c * S);
      t.dst = (uint64_t)(uintptr_t)((uint8_t *)staging.p + doff + c * cap);
      t.aux = (uint32_t)cap;
      t.flags = (uint8_t)(((pad && c == C - 1) ? kTaskTailPad : 0) | mflag);
    } else {
      t.src = (uint64_t)(uintptr_t)(dib + in_off + pays[c].first);
      t.aux = pays[c].second;
      t.dst = (uint64_t)(uintptr_t)((uint8_t *)staging.p + doff + c * S);
      t.flags = (uint8_t)(kTaskDecode | mflag);
    }
    ring[(first + c) & (kRing - 1)] = t;
  }
  ServeJob &sj = sjobs[j.slot];
  sj.out_base = (uint64_t)(uintptr_t)harena_mem.dev<uint8_t>(out);
  sj.first = first;
  sj.ntasks = (uint32_t)C;
  // compress: the header's length; decompress: where chunk 0's bytes go
  sj.header_len = (uint32_t)(q.op == CUCM_OP_COMPRESS ? head_bytes.size() : P);
  sj.batch = batch;
  sj.id = j.gen;
  publish(j, first, (uint32_t)C);
  return CUCM_OK;
}

void Server::finish(uint64_t id, SJob &j, uint32_t word) {
  (void)id;
  inflight_tasks -= j.ntasks;
  --inflight_jobs;
  const cucm_job &q = j.job;
  if (word & 0x80000000u) {
    // a chunk overflowed its staging region: the fallback lane re-runs the
    // request with the strict bound
    release(j);
    fb_pending.push_back(id);
    cv_work.notify_one();
    return;
  }
  const uint8_t *out = hbase + j.out_off;
  uint64_t len;
  if (q.op == CUCM_OP_COMPRESS)
    len = __atomic_load_n(&lens[j.slot], __ATOMIC_ACQUIRE);
  else
    len = be32(q.in);
  int32_t st = CUCM_OK;
  if (len > q.out_cap)
    st = CUCM_E_SMALL;
  else
    std::memcpy(q.out, out, len);
  release(j);
  complete(j, st, len);
}

int Server::launch_kernel() {
  ServeArgs s{};
  s.a.tasks = ring_mem.dev<Task>(ring);
  s.a.ntasks = 0;
  s.a.spec_bits = 2;
  s.a.counter = (uint32_t *)head.p;
  s.a.pool = (uint8_t *)pool.p;
  s.a.slot_bytes = slot_bytes;
  s.a.result = (uint64_t *)results.p;
  s.a.prof = nullptr;
  s.a.tabs = (const DevTables *)parent->tabs.p;
  s.a.init = models_mem.dev<DevInitModel>(models);
  s.tail = ctl_mem.dev<uint32_t>(&ctl->tail);
  s.stop = ctl_mem.dev<uint32_t>(&ctl->stop);
  s.jobs = jobs_mem.dev<ServeJob>(sjobs);
  s.jobcount = (uint32_t *)jobcount.p;
  s.done = done_mem.dev<uint32_t>(done_w);
  s.lens = lens_mem.dev<uint64_t>(lens);
  s.mirror = (uint32_t *)((uint8_t *)head.p + 64);  // a line of its own, past the claim head
  s.ring_mask = kRing - 1;
  s.job_mask = J - 1;
  s.idle_ticks = kIdleTicks;
  s.poll_ticks = kPollTicks;
  // the mirror's stop word still holds the last launch's stop
  if (hipMemsetAsync(s.mirror + 1, 0, 4, ks) != hipSuccess ||
      launch_serve(s, shape.blocks, shape.wpb, ks) != hipSuccess) {
    (void)hipGetLastError();
    return CUCM_E_DEVICE;
  }
  ++launches;
  return CUCM_OK;
}

void Server::sched_loop() {
  (void)hipSetDevice(parent->device);
  std::unique_lock<std::mutex> lk(mu);
  bool kernel_live = false, stop_sent = false;
  uint64_t idle_since = 0;
  for (;;) {
    bool progress = false;
    // admission, oldest first
    while (!pending.empty()) {
      const uint64_t id = pending.front();
      SJob &j = jobs[id];
      if (!eligible(j)) {
        pending.pop_front();
        fb_pending.push_back(id);
        cv_work.notify_one();
        progress = true;
        continue;
      }
      if (!j.prep) {
        // the request-only work (a serial prelearn coder among it) runs
        // once, without the lock, so submit and poll are not held up; only
        // this thread admits, so the request stays at the queue's front
        const cucm_job q = j.job;
        auto prep = std::make_shared<AdmitPrep>();
        lk.unlock();
        prepare_request(q, *prep);
        lk.lock();
        jobs[id].prep = std::move(prep);
      }
      SJob &ja = jobs[id];
      const int rc = admit(id, ja);
      if (rc == CUCM_E_NOTREADY) break;
      ja.prep.reset();
      pending.pop_front();
      progress = true;
      if (rc == CUCM_E_UNSUPPORTED) {
        fb_pending.push_back(id);
        cv_work.notify_one();
      } else if (rc) {
        ja.t_start = clock_ns();
        complete(ja, rc, 0);
      }
    }
    // completions
    if (inflight_jobs) {
      for (uint32_t s = 0; s < J; ++s) {
        const uint64_t id = slot_job[s];
        if (!id) continue;
        SJob &j = jobs[id];
        const uint32_t w = __atomic_load_n(&done_w[s], __ATOMIC_ACQUIRE);
        if ((w & 0x7fffffffu) == j.gen && j.ntasks) {
          finish(id, j, w);
          progress = true;
        }
      }
    }
    // the kernel runs while there is work, and is stopped once idle
    if (kernel_live && hipStreamQuery(ks) == hipSuccess) {
      kernel_live = false;
      if (stop_sent) {
        __atomic_store_n(&ctl->stop, 0u, __ATOMIC_RELEASE);
        stop_sent = false;
      }
    }
    if (inflight_jobs || !pending.empty()) {
      idle_since = 0;
    } else if (kernel_live && !stop_sent) {
      const uint64_t now = clock_ns();
      if (!idle_since) idle_since = now;
      if (now - idle_since > kStopAfterNs) {
        __atomic_store_n(&ctl->stop, 1u, __ATOMIC_RELEASE);
        stop_sent = true;
      }
    }
    if (inflight_jobs && !kernel_live && !stop_sent) {
      if (launch_kernel() == CUCM_OK) {
        kernel_live = true;
      } else {
        // the device refused the launch: every job in flight fails
        for (uint32_t s = 0; s < J; ++s)
          if (slot_job[s]) {
            SJob &j = jobs[slot_job[s]];
            inflight_tasks -= j.ntasks;
            --inflight_jobs;
            release(j);
            complete(j, CUCM_E_DEVICE, 0);
          }
      }
      progress = true;
    }
    if (stop && pending.empty() && inflight_jobs == 0) break;
    if (!progress)
      cv_sched.wait_for(lk, std::chrono::microseconds(inflight_jobs || kernel_live ? 50 : 2000));
  }
}

void Server::lane_loop(cucm_ctx *c) {
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_work.wait(lk, [&] { return stop || !fb_pending.empty(); });
    if (stop && fb_pending.empty()) return;
    // the oldest fallback request's kind, FIFO, within the lane's memory share
    const int32_t op = jobs[fb_pending.front()].job.op;
    std::vector<uint64_t> take;
    uint64_t mem = 0;
    for (auto it = fb_pending.begin(); it != fb_pending.end();) {
      SJob &j = jobs[*it];
      if (j.job.op != op) {
        ++it;
        continue;
      }
      const uint64_t need = footprint(j);
      if (!take.empty() && mem + need > lane_budget) break;
      mem += need;
      take.push_back(*it);
      it = fb_pending.erase(it);
    }
    fb_inflight += (uint32_t)take.size();
    ++launches;
    const uint32_t n = (uint32_t)take.size();
    std::vector<const uint8_t *> in(n);
    std::vector<uint8_t *> out(n);
    std::vector<uint64_t> in_len(n), out_cap(n), out_len(n, 0);
    std::vector<cucm_params> params(n);
    std::vector<int32_t> status(n, 0);
    uint32_t spec = 0;
    const uint64_t t_start = clock_ns();
    for (uint32_t i = 0; i < n; ++i) {
      SJob &j = jobs[take[i]];
      if (!j.t_start) j.t_start = t_start;
      in[i] = j.job.in;
      in_len[i] = j.job.in_len;
      out[i] = j.job.out;
      out_cap[i] = j.job.out_cap;
      params[i] = j.job.params;
      spec = std::max<uint32_t>(spec, j.job.params.spec_bits);
    }
    lk.unlock();
    int rc;
    if (op == CUCM_OP_COMPRESS)
      rc = cucm_compress_many(c, n, in.data(), in_len.data(), params.data(), out.data(),
                              out_cap.data(), out_len.data(), status.data());
    else
      rc = cucm_decompress_many(c, n, in.data(), in_len.data(), out.data(), out_cap.data(),
                                out_len.data(), status.data(), spec);
    lk.lock();
    for (uint32_t i = 0; i < n; ++i) {
      SJob &j = jobs[take[i]];
      j.batch = n;
      complete(j, rc ? rc : status[i], out_len[i]);
    }
    fb_inflight -= n;
    cv_sched.notify_all();
  }
}

void serve_shutdown(cucm_ctx *c) {
  delete c->server;
  c->server = nullptr;
}

extern "C" {

int cucm_serve_start_ex(cucm_ctx *ctx, uint32_t lanes, uint32_t max_inflight, uint32_t flags) {
  if (!ctx || lanes > 16 || max_inflight > 4096 || (flags & ~CUCM_SERVE_STATIC_LANES))
    return CUCM_E_PARAM;
  if (ctx->server) return CUCM_E_PARAM;
  const bool static_lanes = (flags & CUCM_SERVE_STATIC_LANES) != 0;
  if (lanes == 0) lanes = static_lanes ? 3 : 1;
  uint32_t J = 256;
  if (max_inflight) {
    J = 1;
    while (J < max_inflight) J <<= 1;
  }
  auto *s = new (std::nothrow) Server();
  if (!s) return CUCM_E_NOMEM;
  s->parent = ctx;
  const int rc = static_lanes ? CUCM_E_UNSUPPORTED : s->init_engine(J);
  if (rc) {
    // without the serving kernel every request runs on the lanes
    s->engine_ok = false;
    if (rc == CUCM_E_DEVICE || rc == CUCM_E_NOMEM) {
      delete s;
      return rc;
    }
  }
  // the lanes share the budget; beside the serving kernel they get a quarter
  s->lane_budget = ctx->budget / ((static_lanes ? 1 : 4) * lanes);
  for (uint32_t i = 0; i < lanes; ++i) {
    auto *c = new (std::nothrow) cucm_ctx();
    if (!c) {
      delete s;
      return CUCM_E_NOMEM;
    }
    const int lrc = ctx_init(c, ctx->device, s->lane_budget);
    if (lrc) {
      delete c;
      delete s;
      return lrc;
    }
    c->max_waves = ctx->max_waves;
    s->lanes.push_back(c);
  }
  for (auto *c : s->lanes) s->lane_threads.emplace_back([s, c] { s->lane_loop(c); });
  s->sched = std::thread([s] { s->sched_loop(); });
  ctx->server = s;
  return CUCM_OK;
}

int cucm_serve_start(cucm_ctx *ctx, uint32_t lanes, uint32_t max_inflight) {
  return cucm_serve_start_ex(ctx, lanes, max_inflight, 0);
}

int cucm_serve_stop(cucm_ctx *ctx) {
  if (!ctx) return CUCM_E_PARAM;
  Server *s = ctx->server;
  if (!s) return CUCM_OK;
  {
    // drain: every submitted request completes before the loop exits
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
      return s->pending.empty() && s->inflight_jobs == 0 && s->fb_pending.empty() &&
             s->fb_inflight == 0;
    });
  }
  serve_shutdown(ctx);
  return CUCM_OK;
}

int cucm_submit(cucm_ctx *ctx, const cucm_job *job, uint64_t *job_id) {
  if (!ctx || !job || !job_id) return CUCM_E_PARAM;
  Server *s = ctx->server;
  if (!s) return CUCM_E_PARAM;
  if (job->op != CUCM_OP_COMPRESS && job->op != CUCM_OP_DECOMPRESS) return CUCM_E_PARAM;
  if ((job->in_len && !job->in) || (job->out_cap && !job->out)) return CUCM_E_PARAM;
  SJob j;
  j.job = *job;
  j.t_submit = clock_ns();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return CUCM_E_PARAM;
    const uint64_t id = s->next_id++;
    s->jobs.emplace(id, j);
    s->pending.push_back(id);
    *job_id = id;
  }
  s->cv_sched.notify_one();
  return CUCM_OK;
}

int cucm_poll(cucm_ctx *ctx, uint64_t job_id, int wait, cucm_result *res) {
  if (!ctx || !res) return CUCM_E_PARAM;
  Server *s = ctx->server;
  if (!s) return CUCM_E_PARAM;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return CUCM_E_PARAM;
  if (!it->second.done) {
    if (!wait) return CUCM_E_NOTREADY;
    // look the job up again after every wake-up: submissions made while
    // this thread waited may have rehashed the table.  Each id has one
    // waiter; if another thread reported the job meanwhile it is gone.
    s->cv_done.wait(lk, [&] {
      auto f = s->jobs.find(job_id);
      return f == s->jobs.end() || f->second.done;
    });
    it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return CUCM_E_PARAM;
  }
  const SJob &j = it->second;
  res->status = j.status;
  res->batch_jobs = (int32_t)j.batch;
  res->out_len = j.out_len;
  res->queue_ms = j.t_start >= j.t_submit ? (j.t_start - j.t_submit) / 1e6 : 0.0;
  res->latency_ms = (j.t_done - j.t_submit) / 1e6;
  s->jobs.erase(it);
  return CUCM_OK;
}

int cucm_serve_launches(cucm_ctx *ctx, uint64_t *launches) {
  if (!ctx || !launches) return CUCM_E_PARAM;
  Server *s = ctx->server;
  if (!s) return CUCM_E_PARAM;
  std::lock_guard<std::mutex> lk(s->mu);
  *launches = s->launches;
  return CUCM_OK;
}

}  // extern "C"
