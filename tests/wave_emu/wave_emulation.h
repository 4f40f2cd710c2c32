// wave_emulation.h -- CPU stand-ins for the cross-lane primitives of
// cucm_wave.h (test infrastructure).
//
// The 64 lanes of a wave run as fibers on one host thread.  Each lane runs
// until it reaches a cross-lane operation (readlane, readfirstlane, ballot,
// wave_sync); when all 64 are there, the operation is resolved and every
// lane continues.  Cross-lane LDS traffic is therefore ordered only at those
// points, so device code that relies on lock step without a wave_sync shows
// up as a mismatch here.  A workgroup of several waves (run_team) steps its
// waves in a seeded random order, one cross-lane operation at a time, and
// holds each wave at a workgroup barrier (team_sync) until every unfinished
// wave is there, so LDS traffic between waves that is not ordered by a
// barrier or a flag varies from seed to seed.  LDS and every global buffer
// are separate heap blocks of their exact sizes, so AddressSanitizer reports
// any access out of bounds (the GPU would silently drop an out-of-range LDS
// access).
#ifndef CUCM_WAVE_EMULATION_H
#define CUCM_WAVE_EMULATION_H
#include <sanitizer/common_interface_defs.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <vector>

#define LDS
#define GLB
#define CUCM_DEV inline
#define CUCM_HD inline
#define CUCM_HDI inline

template <class T>
inline T min(T a, T b) {
  return a < b ? a : b;
}
template <class T>
inline T max(T a, T b) {
  return a > b ? a : b;
}

namespace emu {

// Minimal x86-64 fiber switch: saves the callee-saved registers and the
// stack pointer of the running fiber in *from, resumes the fiber whose
// stack pointer is to.
extern "C" void emu_fiber_swap(void **from, void *to);
asm(R"(
.text
.globl emu_fiber_swap
.type emu_fiber_swap,@function
emu_fiber_swap:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
.size emu_fiber_swap, .-emu_fiber_swap
)");

constexpr int kLanes = 64;
constexpr size_t kStack = 1 << 20;
enum Kind { kNone, kRead, kFirst, kBallot, kSync, kBarrier };

struct Wave {
  void *sched_sp = nullptr;
  void *sp[kLanes];
  std::vector<char> stack[kLanes];
  bool done[kLanes];
  Kind kind[kLanes];
  uint64_t val[kLanes];
  int line[kLanes];
  uint64_t snap[kLanes];  // values of the resolved operation
  uint64_t ballot = 0;
  int cur = 0;
  const void *sched_bottom = nullptr;
  size_t sched_size = 0;
  std::function<void(uint32_t)> body;
};

inline Wave *&current() {
  static Wave *w = nullptr;
  return w;
}

// Switches from a lane back to the scheduler (AddressSanitizer is told
// which stack becomes current).
inline void to_sched(Wave &w, int l) {
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, w.sched_bottom, w.sched_size);
  emu_fiber_swap(&w.sp[l], w.sched_sp);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
}

inline void to_lane(Wave &w, int l) {
  void *fake = nullptr;
  w.cur = l;
  __sanitizer_start_switch_fiber(&fake, w.stack[l].data(), w.stack[l].size());
  emu_fiber_swap(&w.sched_sp, w.sp[l]);
  __sanitizer_finish_switch_fiber(fake, &w.sched_bottom, &w.sched_size);
}

extern "C" inline void emu_lane_entry() {
  __sanitizer_finish_switch_fiber(nullptr, nullptr, nullptr);
  Wave &w = *current();
  const int l = w.cur;
  w.body((uint32_t)l);
  w.done[l] = true;
  w.kind[l] = kNone;
  to_sched(w, l);
  abort();  // a finished lane is never resumed
}

// Lane side of a cross-lane operation: publish, yield, read the result.
inline uint64_t meet(Kind k, uint64_t v, uint32_t arg, int line) {
  Wave &w = *current();
  const int l = w.cur;
  w.kind[l] = k;
  w.val[l] = v;
  w.line[l] = line;
  to_sched(w, l);
  switch (k) {
    case kRead: return w.snap[arg & 63u];
    case kFirst: return w.snap[0];
    case kBallot: return w.ballot;
    default: return 0;
  }
}

// Creates the 64 lane fibers of a wave that will run body(lane).
inline Wave *make_wave(const std::function<void(uint32_t)> &body) {
  Wave *w = new Wave();
  w->body = body;
  for (int l = 0; l < kLanes; ++l) {
    w->stack[l].assign(kStack, 0);
    w->done[l] = false;
    // initial frame: six zero registers, then the entry as return address
    uintptr_t top = ((uintptr_t)(w->stack[l].data() + kStack) & ~(uintptr_t)15) - 64;
    void **f = (void **)top;
    for (int i = 0; i < 6; ++i) f[i] = nullptr;
    f[6] = (void *)&emu_lane_entry;
    f[7] = nullptr;
    w->sp[l] = (void *)top;
  }
  return w;
}

// Runs every lane of the wave up to its next cross-lane operation.  Returns
// the operation they all wait at (kNone once the wave has finished); wave
// operations are resolved, a workgroup barrier is left pending.
inline Kind step_wave(Wave *w) {
  current() = w;
  for (int l = 0; l < kLanes; ++l)
    if (!w->done[l]) to_lane(*w, l);
  current() = nullptr;
  int ndone = 0;
  for (int l = 0; l < kLanes; ++l) ndone += w->done[l];
  if (ndone == kLanes) return kNone;
  if (ndone) {
    int at = 0;
    while (w->done[at]) ++at;
    fprintf(stderr, "wave emulation: %d lanes finished while others wait at line %d\n", ndone,
            w->line[at]);
    abort();
  }
  for (int l = 1; l < kLanes; ++l)
    if (w->kind[l] != w->kind[0] || w->line[l] != w->line[0]) {
      fprintf(stderr,
              "wave emulation: divergent cross-lane operation: lane 0 at line %d, lane %d at "
              "line %d\n",
              w->line[0], l, w->line[l]);
      abort();
    }
  if (w->kind[0] != kBarrier) {
    w->ballot = 0;
    for (int l = 0; l < kLanes; ++l) {
      w->snap[l] = w->val[l];
      if (w->val[l]) w->ballot |= 1ull << l;
    }
  }
  return w->kind[0];
}

// Runs body(lane) for the 64 lanes of one wave to completion.
inline void run_wave(const std::function<void(uint32_t)> &body) {
  Wave *w = make_wave(body);
  for (;;) {
    const Kind k = step_wave(w);
    if (k == kNone) break;
    if (k == kBarrier) {
      fprintf(stderr, "wave emulation: workgroup barrier in a one-wave run (line %d)\n", w->line[0]);
      abort();
    }
  }
  delete w;
}

// Runs body(wave, lane) for a workgroup of `waves` waves to completion.  The
// next wave to step is drawn from a generator seeded with `seed`; a wave at
// a barrier waits until every unfinished wave has reached one.
inline void run_team(uint32_t waves, uint64_t seed,
                     const std::function<void(uint32_t, uint32_t)> &body) {
  std::vector<Wave *> ws;
  for (uint32_t i = 0; i < waves; ++i)
    ws.push_back(make_wave([&body, i](uint32_t lane) { body(i, lane); }));
  std::vector<int> state(waves, 0);  // 0 runnable, 1 at a barrier, 2 finished
  uint64_t rng = seed * 0x9e3779b97f4a7c15ull + 1;
  for (;;) {
    std::vector<uint32_t> run;
    for (uint32_t i = 0; i < waves; ++i)
      if (state[i] == 0) run.push_back(i);
    if (run.empty()) {
      bool any = false;
      for (uint32_t i = 0; i < waves; ++i)
        if (state[i] == 1) {
          state[i] = 0;
          any = true;
        }
      if (!any) break;
      continue;
    }
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    const uint32_t i = run[rng % run.size()];
    const Kind k = step_wave(ws[i]);
    state[i] = k == kNone ? 2 : (k == kBarrier ? 1 : 0);
  }
  for (Wave *w : ws) delete w;
}

}  // namespace emu

namespace cucm {
#define CUCM_EMU_MEET(k, v, a) emu::meet(k, v, a, __LINE__)
inline uint32_t lane_read_at(uint32_t v, uint32_t lane, int line) {
  return (uint32_t)emu::meet(emu::kRead, v, lane, line);
}
#define lane_read(v, l) lane_read_at((v), (l), __LINE__)
inline uint32_t lane_first_at(uint32_t v, int line) {
  return (uint32_t)emu::meet(emu::kFirst, v, 0, line);
}
#define lane_first(v) lane_first_at((v), __LINE__)
inline uint64_t wave_ballot_at(bool p, int line) { return emu::meet(emu::kBallot, p ? 1 : 0, 0, line); }
#define wave_ballot(p) wave_ballot_at((p), __LINE__)
inline void wave_sync_at(int line) { emu::meet(emu::kSync, 0, 0, line); }
#define wave_sync() wave_sync_at(__LINE__)
inline void team_sync_at(int line) { emu::meet(emu::kBarrier, 0, 0, line); }
#define team_sync() team_sync_at(__LINE__)
#define team_sync_lds() team_sync_at(__LINE__)
#define vm_drain() wave_sync_at(__LINE__)
#define lane_step() wave_sync_at(__LINE__)
inline uint32_t clz32(uint32_t x) { return x ? (uint32_t)__builtin_clz(x) : 32u; }
inline uint64_t shader_clock() { return 0; }
inline void wait_vm() {}
inline void sched_fence() {}
inline void wait_lgkm() {}
inline uint32_t coder_split(uint32_t x, uint32_t xmid, uint32_t &x1, uint32_t &x2) {
  const uint32_t y = x <= xmid ? 1u : 0u;
  if (y)
    x2 = xmid;
  else
    x1 = xmid + 1u;
  return y;
}
inline void coder_encode_split(uint32_t prm, uint32_t xmid, uint32_t &x1, uint32_t &x2) {
  if (prm & 1u)
    x2 = xmid;
  else
    x1 = xmid + 1u;
}
inline void coder_split_path(uint32_t x, uint32_t xmid, uint32_t &x1, uint32_t &x2, uint32_t &q) {
  q = 2 * q + 1 + coder_split(x, xmid, x1, x2);
}
inline void flag_publish(uint32_t *p, uint32_t v) { *p = v; }
inline uint32_t flag_read(const uint32_t *p) { return *p; }
inline uint32_t flag_peek(const uint32_t *p) { return *p; }
inline void flag_acquire() {}
inline uint32_t mul_i24(int32_t a, int32_t b) {
  if (a < -(1 << 23) || a >= (1 << 23) || b < -(1 << 23) || b >= (1 << 23)) {
    fprintf(stderr, "wave emulation: mul_i24 operand out of range (%d, %d)\n", a, b);
    abort();
  }
  return (uint32_t)((int64_t)a * (int64_t)b);
}
inline uint32_t lds_cas(uint32_t *p, uint32_t cmp, uint32_t v) {
  const uint32_t o = *p;
  if (o == cmp) *p = v;
  return o;
}
inline void lds_add(int32_t *p, int32_t v) { *p = (int32_t)((uint32_t)*p + (uint32_t)v); }
inline void glb_add64(uint64_t *p, uint64_t v) { *p += v; }
}  // namespace cucm

#endif
