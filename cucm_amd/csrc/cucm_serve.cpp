// cucm_serve.cpp -- continuous-batching serving loop.
//
// Requests (compress or decompress of caller-owned host buffers) enter a FIFO
// through cucm_submit.  The loop runs `lanes` worker threads; each owns an
// internal context on the same device (its own HIP stream, device buffers and
// an equal share of the HBM budget).  A lane that falls idle takes the oldest
// pending request and every other pending request of the same kind that fits
// its memory share and the in-flight limit, and runs them as one persistent
// launch (one wave per chunk, device work queue).  Requests that arrive while
// launches are on the GPU are admitted by the next lane that frees up, so
// their work overlaps on the device with the launches already running: the
// GPU places a new launch's workgroups on the CUs a draining launch leaves.
//
// Each request's output depends only on its own input and parameters, never
// on what it was batched with.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>

#include "cucm_engine.h"

namespace {

uint64_t clock_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

uint32_t be32(const uint8_t *p) {
  return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

struct SJob {
  cucm_job job{};
  uint64_t t_submit = 0, t_start = 0, t_done = 0;
  int32_t status = 0;
  uint32_t batch = 0;
  uint64_t out_len = 0;
  bool done = false;
};

// Device bytes a request holds while it runs: its input, output and (for
// compression) the per-chunk staging of 1.5 S + 256 bytes.
uint64_t footprint(const SJob &j) {
  if (j.job.op == CUCM_OP_COMPRESS) {
    const uint64_t n = j.job.in_len;
    const uint64_t S = j.job.params.chunk_size ? j.job.params.chunk_size : 1;
    return n + (n + n / 2) + 256 * (n / S + 1) + j.job.out_cap;
  }
  const uint64_t N = j.job.in_len >= 4 ? be32(j.job.in) : 0;
  return j.job.in_len + N;
}

}  // namespace

struct Server {
  cucm_ctx *parent = nullptr;
  std::vector<cucm_ctx *> lanes;
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> pending;
  std::unordered_map<uint64_t, SJob> jobs;
  uint64_t next_id = 1;
  uint32_t max_inflight = 0;
  uint32_t inflight = 0;    // requests taken by a lane and not finished
  uint64_t lane_budget = 0;
  uint64_t launches = 0;
  bool stop = false;

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv_work.notify_all();
    for (auto &t : threads)
      if (t.joinable()) t.join();
    for (auto *c : lanes) delete c;
  }

  void lane_loop(cucm_ctx *c);
};

void Server::lane_loop(cucm_ctx *c) {
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_work.wait(lk, [&] { return stop || (!pending.empty() && inflight < max_inflight); });
    // Stopping without a drain (context destroyed while serving) abandons
    // the requests still pending; cucm_serve_stop drains first.
    if (stop) return;
    // Admission: the oldest request's kind, FIFO, within the lane's memory
    // share (the first request is always taken) and the in-flight limit.
    const int32_t op = jobs[pending.front()].job.op;
    std::vector<uint64_t> take;
    uint64_t mem = 0;
    for (auto it = pending.begin(); it != pending.end() && inflight + take.size() < max_inflight;) {
      SJob &j = jobs[*it];
      if (j.job.op != op) {
        ++it;
        continue;
      }
      const uint64_t need = footprint(j);
      if (!take.empty() && mem + need > lane_budget) break;
      mem += need;
      take.push_back(*it);
      it = pending.erase(it);
    }
    inflight += (uint32_t)take.size();
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
      j.t_start = t_start;
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
    const uint64_t t_done = clock_ns();
    lk.lock();
    for (uint32_t i = 0; i < n; ++i) {
      SJob &j = jobs[take[i]];
      j.status = rc ? rc : status[i];
      j.out_len = j.status ? 0 : out_len[i];
      j.t_done = t_done;
      j.batch = n;
      j.done = true;
    }
    inflight -= n;
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

void serve_shutdown(cucm_ctx *c) {
  delete c->server;
  c->server = nullptr;
}

extern "C" {

int cucm_serve_start(cucm_ctx *ctx, uint32_t lanes, uint32_t max_inflight) {
  if (!ctx || lanes > 16) return CUCM_E_PARAM;
  if (ctx->server) return CUCM_E_PARAM;
  if (lanes == 0) lanes = 3;
  auto *s = new (std::nothrow) Server();
  if (!s) return CUCM_E_NOMEM;
  s->parent = ctx;
  s->max_inflight = max_inflight ? max_inflight : 4096;
  s->lane_budget = ctx->budget / lanes;
  for (uint32_t i = 0; i < lanes; ++i) {
    auto *c = new (std::nothrow) cucm_ctx();
    if (!c) {
      delete s;
      return CUCM_E_NOMEM;
    }
    const int rc = ctx_init(c, ctx->device, s->lane_budget);
    if (rc) {
      delete c;
      delete s;
      return rc;
    }
    c->max_waves = ctx->max_waves;
    s->lanes.push_back(c);
  }
  for (auto *c : s->lanes) s->threads.emplace_back([s, c] { s->lane_loop(c); });
  ctx->server = s;
  return CUCM_OK;
}

int cucm_serve_stop(cucm_ctx *ctx) {
  if (!ctx) return CUCM_E_PARAM;
  Server *s = ctx->server;
  if (!s) return CUCM_OK;
  {
    // drain: every submitted request completes before the lanes exit
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->pending.empty() && s->inflight == 0; });
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
  s->cv_work.notify_one();
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
    // this thread waited may have rehashed the table
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
  res->queue_ms = (j.t_start - j.t_submit) / 1e6;
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
