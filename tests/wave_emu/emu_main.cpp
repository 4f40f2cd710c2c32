// emu_main.cpp -- runs the codec's per-wave device code (cucm_wave.h) on
// the CPU, one chunk at a time, under the fiber wave emulation of
// wave_emulation.h (test infrastructure; the product never runs this).
//
//   wave_emu enc <input> <S> <U> <tb> <tail_pad> <output>
//       output: per chunk, a u32 LE payload length and the payload
//   wave_emu dec <payloads> <N> <S> <U> <tb> <spec_bits> <output>
//       payloads in the format enc writes; output: the N decoded bytes
//   wave_emu tenc <input> <S> <U> <tb> <tail_pad> <seed> <output>
//       enc through the team encoder (cucm_team.h): every chunk in turn on
//       one workgroup of kTeamWaves waves and one LDS
//   wave_emu tdec <payloads> <N> <S> <U> <tb> <spec_bits> <seed> <output>
//       the same through the team decoder (cucm_team.h): every chunk in turn
//       on one workgroup of kTeamWaves waves and one LDS, waves stepped in
//       an order drawn from seed
//   wave_emu fenc <input> <S> <tb> <prelearn|-> <output>
//       every chunk through the frozen-model encoder (cucm_frozen.h) on one
//       wave, with the zero model or the model trained on the prelearn file
//   wave_emu fdec <payloads> <N> <S> <tb> <spec_bits> <prelearn|-> <output>
//       the same through the frozen-model decoder
//   wave_emu benc <input> <S> <U> <tb> <tail_pad> <output>
//       enc through the large-unit codec (cucm_bigunit.h), any U, one wave
//       whose model slot and unit records persist across chunks
//   wave_emu bdec <payloads> <N> <S> <U> <tb> <output>
//       the same through the large-unit decoder
#define CUCM_WAVE_EMULATION 1
#include "wave_emulation.h"
// (the device code follows the emulation shims)
#include "cucm_bigunit.h"
#include "cucm_frozen.h"
#include "cucm_team.h"
#include "cucm_host_model.h"

#include <cstring>
#include <string>

using namespace cucm;

static std::vector<uint8_t> read_file(const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) {
    perror(path);
    exit(2);
  }
  std::vector<uint8_t> v;
  uint8_t buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) v.insert(v.end(), buf, buf + n);
  fclose(f);
  return v;
}

static void write_file(const char *path, const std::vector<uint8_t> &v) {
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(v.data(), 1, v.size(), f) != v.size()) {
    perror(path);
    exit(2);
  }
  fclose(f);
}

// LDS of a one-wave workgroup: the tables, then the wave's carve-out.
struct Lds {
  uint8_t *base;
  const DevTables *T;
  WaveMem M;
  Lds(uint32_t Umax, uint32_t P, bool dec) {
    const size_t tab = align16(sizeof(DevTables));
    const size_t bytes = tab + wave_lds_bytes(Umax, P, dec);
    base = (uint8_t *)aligned_alloc(16, (bytes + 15) & ~(size_t)15);
    memset(base, 0xa5, bytes);  // stale contents, as a reused workgroup sees
    memcpy(base, &host_tables(), sizeof(DevTables));
    T = (const DevTables *)base;
    M = carve(base + tab, Umax, P, dec);
  }
  ~Lds() { free(base); }
};

template <int NJ>
static std::vector<uint8_t> encode_one(const uint8_t *chunk, uint32_t L, uint32_t S, uint32_t U,
                                       uint32_t tb, bool pad, uint8_t *slot) {
  Lds lds(U, 0, false);
  uint8_t *src = (uint8_t *)malloc(L ? L : 1);
  memcpy(src, chunk, L);
  const uint32_t cap = align16(S + S / 2 + 256);
  uint8_t *dst = (uint8_t *)malloc(cap);
  Task tk{};
  tk.src = (uint64_t)(uintptr_t)src;
  tk.dst = (uint64_t)(uintptr_t)dst;
  tk.L = L;
  tk.S = S;
  tk.aux = cap;
  tk.U = (uint16_t)U;
  tk.tb = (uint8_t)tb;
  tk.flags = pad ? kTaskTailPad : 0;
  uint64_t result = 0;
  CodecArgs a{};
  a.tasks = &tk;
  a.ntasks = 1;
  a.result = &result;
  emu::run_wave([&](uint32_t lane) { encode_task<NJ, true>(a, tk, 0, lds.M, lds.T, slot, U, lane); });
  if (result & kResultOverflow) {
    fprintf(stderr, "staging overflow\n");
    exit(3);
  }
  std::vector<uint8_t> out(dst, dst + result);
  free(src);
  free(dst);
  return out;
}

template <int NJ, int D>
static void decode_one(const std::vector<uint8_t> &pay, uint8_t *out, uint32_t L, uint32_t S,
                       uint32_t U, uint32_t tb, uint8_t *slot) {
  constexpr uint32_t P = (1u << D) - 1u;
  Lds lds(U, P, true);
  uint8_t *src = (uint8_t *)malloc(pay.size() ? pay.size() : 1);
  memcpy(src, pay.data(), pay.size());
  uint8_t *dst = (uint8_t *)malloc(L ? L : 1);
  Task tk{};
  tk.src = (uint64_t)(uintptr_t)src;
  tk.dst = (uint64_t)(uintptr_t)dst;
  tk.L = L;
  tk.S = S;
  tk.aux = (uint32_t)pay.size();
  tk.U = (uint16_t)U;
  tk.tb = (uint8_t)tb;
  uint64_t result = 0;
  CodecArgs a{};
  a.tasks = &tk;
  a.ntasks = 1;
  a.spec_bits = D;
  a.result = &result;
  emu::run_wave([&](uint32_t lane) { decode_task<NJ, D, true>(a, tk, 0, lds.M, lds.T, slot, U, lane); });
  memcpy(out, dst, L);
  free(src);
  free(dst);
}

template <int NJ>
static void decode_d(int spec, const std::vector<uint8_t> &pay, uint8_t *out, uint32_t L,
                     uint32_t S, uint32_t U, uint32_t tb, uint8_t *slot) {
  switch (spec) {
    case 1: decode_one<NJ, 1>(pay, out, L, S, U, tb, slot); break;
    case 2: decode_one<NJ, 2>(pay, out, L, S, U, tb, slot); break;
    case 4: decode_one<NJ, 4>(pay, out, L, S, U, tb, slot); break;
    case 8: decode_one<NJ, 8>(pay, out, L, S, U, tb, slot); break;
    default: fprintf(stderr, "spec bits\n"); exit(2);
  }
}

// All chunks through one encoder team, as the persistent team kernel runs them.
template <int NJ>
static std::vector<uint8_t> encode_team_all(const std::vector<uint8_t> &in, uint32_t S,
                                            uint32_t U, uint32_t tb, bool pad, uint64_t seed) {
  const uint64_t N = in.size();
  const uint64_t C = (N + S - 1) / S;
  const uint32_t Umax = 64u * NJ;
  std::vector<Task> tasks(C);
  std::vector<uint8_t *> srcs(C), dsts(C);
  const uint32_t cap = align16(S + S / 2 + 256);
  for (uint64_t c = 0; c < C; ++c) {
    const uint32_t L = (uint32_t)std::min<uint64_t>(S, N - c * S);
    srcs[c] = (uint8_t *)malloc(L ? L : 1);
    memcpy(srcs[c], in.data() + c * S, L);
    dsts[c] = (uint8_t *)malloc(cap);
    Task &tk = tasks[c];
    tk = Task{};
    tk.src = (uint64_t)(uintptr_t)srcs[c];
    tk.dst = (uint64_t)(uintptr_t)dsts[c];
    tk.L = L;
    tk.S = S;
    tk.aux = cap;
    tk.U = (uint16_t)U;
    tk.tb = (uint8_t)tb;
    tk.flags = (pad && c == C - 1) ? kTaskTailPad : 0;
  }
  const size_t tab = align16(sizeof(DevTables));
  const size_t bytes = tab + team_enc_lds_bytes(Umax);
  uint8_t *base = (uint8_t *)aligned_alloc(16, (bytes + 15) & ~(size_t)15);
  memset(base, 0xa5, bytes);
  memcpy(base, &host_tables(), sizeof(DevTables));
  const DevTables *T = (const DevTables *)base;
  uint8_t *slot = (uint8_t *)malloc((size_t)7 << tb);
  std::vector<uint64_t> result(C, 0);
  CodecArgs a{};
  a.tasks = tasks.data();
  a.ntasks = (uint32_t)C;
  a.result = result.data();
  emu::run_team(kTeamWaves, seed, [&](uint32_t wid, uint32_t lane) {
    const TeamMem X = carve_team_enc(base + tab, Umax, wid);
    for (uint64_t c = 0; c < C; ++c) encode_team<NJ, true>(a, tasks[c], (uint32_t)c, X, T, slot, Umax, wid, lane);
  });
  std::vector<uint8_t> out;
  for (uint64_t c = 0; c < C; ++c) {
    if (result[c] & kResultOverflow) {
      fprintf(stderr, "staging overflow\n");
      exit(3);
    }
    const uint32_t n = (uint32_t)result[c];
    for (int i = 0; i < 4; ++i) out.push_back((uint8_t)(n >> (8 * i)));
    out.insert(out.end(), dsts[c], dsts[c] + n);
    free(srcs[c]);
    free(dsts[c]);
  }
  free(slot);
  free(base);
  return out;
}

// All chunks through one team, as the persistent team kernel runs them.
template <int NJ, int D>
static void decode_team_all(const std::vector<uint8_t> &in, uint8_t *out, uint64_t N, uint32_t S,
                            uint32_t U, uint32_t tb, uint64_t seed) {
  const uint64_t C = (N + S - 1) / S;
  std::vector<Task> tasks(C);
  std::vector<uint8_t *> srcs(C), dsts(C);
  size_t pos = 0;
  for (uint64_t c = 0; c < C; ++c) {
    const uint32_t L = (uint32_t)std::min<uint64_t>(S, N - c * S);
    uint32_t n = 0;
    for (int i = 0; i < 4; ++i) n |= (uint32_t)in[pos + i] << (8 * i);
    pos += 4;
    srcs[c] = (uint8_t *)malloc(n ? n : 1);
    memcpy(srcs[c], in.data() + pos, n);
    pos += n;
    dsts[c] = (uint8_t *)malloc(L ? L : 1);
    Task &tk = tasks[c];
    tk = Task{};
    tk.src = (uint64_t)(uintptr_t)srcs[c];
    tk.dst = (uint64_t)(uintptr_t)dsts[c];
    tk.L = L;
    tk.S = S;
    tk.aux = n;
    tk.U = (uint16_t)U;
    tk.tb = (uint8_t)tb;
  }
  const size_t tab = align16(sizeof(DevTables));
  const size_t bytes = tab + team_lds_bytes(U, D);
  uint8_t *base = (uint8_t *)aligned_alloc(16, (bytes + 15) & ~(size_t)15);
  memset(base, 0xa5, bytes);
  memcpy(base, &host_tables(), sizeof(DevTables));
  const DevTables *T = (const DevTables *)base;
  uint8_t *slot = (uint8_t *)malloc((size_t)7 << tb);
  std::vector<uint64_t> result(C, 0);
  CodecArgs a{};
  a.tasks = tasks.data();
  a.ntasks = (uint32_t)C;
  a.spec_bits = D;
  a.result = result.data();
  emu::run_team(kTeamWaves, seed, [&](uint32_t wid, uint32_t lane) {
    const TeamMem X = carve_team(base + tab, U, D, wid);
    uint32_t rid = 0;
    for (uint64_t c = 0; c < C; ++c)
      decode_team<NJ, D, true>(a, tasks[c], (uint32_t)c, X, T, slot, U, wid, lane, rid);
  });
  for (uint64_t c = 0; c < C; ++c) {
    memcpy(out + c * S, dsts[c], tasks[c].L);
    free(srcs[c]);
    free(dsts[c]);
  }
  free(slot);
  free(base);
}

template <int NJ>
static void decode_team_d(int spec, const std::vector<uint8_t> &in, uint8_t *out, uint64_t N,
                          uint32_t S, uint32_t U, uint32_t tb, uint64_t seed) {
  switch (spec) {
    case 1: decode_team_all<NJ, 1>(in, out, N, S, U, tb, seed); break;
    case 2: decode_team_all<NJ, 2>(in, out, N, S, U, tb, seed); break;
    case 4: decode_team_all<NJ, 4>(in, out, N, S, U, tb, seed); break;
    default: fprintf(stderr, "spec bits\n"); exit(2);
  }
}

// The frozen codec's model: zero (no file) or trained on the prelearn file.
struct FrozenSetup {
  HostModel m;
  DevInitModel d{};
  bool init = false;
  FrozenSetup(uint32_t tb, const char *prelearn) : m((int)tb) {
    if (std::string(prelearn) != "-") {
      const std::vector<uint8_t> pre = read_file(prelearn);
      prelearn_train(pre.data(), pre.size(), m);
      d.states = m.states.data() + ((size_t)1 << tb);  // orders 1..7
      std::memcpy(d.statep, m.statep, sizeof d.statep);
      std::memcpy(d.w, m.w, sizeof d.w);
      order0_states(m, d.st0);
      d.tb = tb;
      init = true;
    }
  }
};

// Every chunk through one frozen wave (its model stays loaded across
// chunks, as in the persistent kernel).
static std::vector<uint8_t> frozen_encode_all(const std::vector<uint8_t> &in, uint32_t S,
                                              uint32_t tb, const FrozenSetup &fs) {
  const uint64_t N = in.size();
  const uint64_t C = (N + S - 1) / S;
  std::vector<Task> tasks(C);
  std::vector<uint8_t *> srcs(C), dsts(C);
  const uint32_t cap = align16(S + S / 2 + 256);
  for (uint64_t c = 0; c < C; ++c) {
    const uint32_t L = (uint32_t)std::min<uint64_t>(S, N - c * S);
    srcs[c] = (uint8_t *)malloc(L);
    memcpy(srcs[c], in.data() + c * S, L);
    dsts[c] = (uint8_t *)malloc(cap);
    Task &tk = tasks[c];
    tk = Task{};
    tk.src = (uint64_t)(uintptr_t)srcs[c];
    tk.dst = (uint64_t)(uintptr_t)dsts[c];
    tk.L = L;
    tk.S = S;
    tk.aux = cap;
    tk.U = 32;
    tk.tb = (uint8_t)tb;
    tk.flags = kTaskFrozen | (fs.init ? kTaskInitModel : 0);
  }
  const size_t tab = align16(sizeof(DevTables));
  const size_t bytes = tab + frozen_wave_bytes();
  uint8_t *base = (uint8_t *)aligned_alloc(16, bytes);
  memset(base, 0xa5, bytes);
  memcpy(base, &host_tables(), sizeof(DevTables));
  const DevTables *T = (const DevTables *)base;
  std::vector<uint64_t> result(C, 0);
  CodecArgs a{};
  a.tasks = tasks.data();
  a.ntasks = (uint32_t)C;
  a.result = result.data();
  a.init = &fs.d;
  emu::run_wave([&](uint32_t lane) {
    const FrozenMem F = carve_frozen(base + tab);
    uint32_t loaded = kNoModel;
    for (uint64_t c = 0; c < C; ++c) {
      const GLB uint8_t *states = frozen_model(a, tasks[c], F, T, loaded, lane);
      frozen_encode_task(a, tasks[c], (uint32_t)c, F, T, states, lane);
    }
  });
  std::vector<uint8_t> out;
  for (uint64_t c = 0; c < C; ++c) {
    if (result[c] & kResultOverflow) {
      fprintf(stderr, "staging overflow\n");
      exit(3);
    }
    const uint32_t n = (uint32_t)result[c];
    for (int i = 0; i < 4; ++i) out.push_back((uint8_t)(n >> (8 * i)));
    out.insert(out.end(), dsts[c], dsts[c] + n);
    free(srcs[c]);
    free(dsts[c]);
  }
  free(base);
  return out;
}

template <int D>
static void frozen_decode_all(const std::vector<uint8_t> &in, uint8_t *out, uint64_t N,
                              uint32_t S, uint32_t tb, const FrozenSetup &fs) {
  const uint64_t C = (N + S - 1) / S;
  std::vector<Task> tasks(C);
  std::vector<uint8_t *> srcs(C), dsts(C);
  size_t pos = 0;
  for (uint64_t c = 0; c < C; ++c) {
    const uint32_t L = (uint32_t)std::min<uint64_t>(S, N - c * S);
    uint32_t n = 0;
    for (int i = 0; i < 4; ++i) n |= (uint32_t)in[pos + i] << (8 * i);
    pos += 4;
    srcs[c] = (uint8_t *)malloc(n ? n : 1);
    memcpy(srcs[c], in.data() + pos, n);
    pos += n;
    dsts[c] = (uint8_t *)malloc(L);
    Task &tk = tasks[c];
    tk = Task{};
    tk.src = (uint64_t)(uintptr_t)srcs[c];
    tk.dst = (uint64_t)(uintptr_t)dsts[c];
    tk.L = L;
    tk.S = S;
    tk.aux = n;
    tk.U = 32;
    tk.tb = (uint8_t)tb;
    tk.flags = kTaskFrozen | (fs.init ? kTaskInitModel : 0);
  }
  const size_t tab = align16(sizeof(DevTables));
  const size_t bytes = tab + frozen_wave_bytes();
  uint8_t *base = (uint8_t *)aligned_alloc(16, bytes);
  memset(base, 0xa5, bytes);
  memcpy(base, &host_tables(), sizeof(DevTables));
  const DevTables *T = (const DevTables *)base;
  std::vector<uint64_t> result(C, 0);
  CodecArgs a{};
  a.tasks = tasks.data();
  a.ntasks = (uint32_t)C;
  a.spec_bits = D;
  a.result = result.data();
  a.init = &fs.d;
  emu::run_wave([&](uint32_t lane) {
    const FrozenMem F = carve_frozen(base + tab);
    uint32_t loaded = kNoModel;
    for (uint64_t c = 0; c < C; ++c) {
      const GLB uint8_t *states = frozen_model(a, tasks[c], F, T, loaded, lane);
      frozen_decode_task<D>(a, tasks[c], (uint32_t)c, F, T, states, lane);
    }
  });
  for (uint64_t c = 0; c < C; ++c) {
    memcpy(out + c * S, dsts[c], tasks[c].L);
    free(srcs[c]);
    free(dsts[c]);
  }
  free(base);
}

// Every chunk through one large-unit wave (exact-size LDS, slot and records).
static std::vector<uint8_t> big_run(bool decode, const std::vector<uint8_t> &in, uint64_t N,
                                    uint32_t S, uint32_t U, uint32_t tb, bool pad) {
  const size_t tab = align16(sizeof(DevTables));
  const size_t bytes = tab + big_wave_bytes();
  uint8_t *base = (uint8_t *)aligned_alloc(16, (bytes + 15) & ~(size_t)15);
  memset(base, 0xa5, bytes);
  memcpy(base, &host_tables(), sizeof(DevTables));
  const DevTables *T = (const DevTables *)base;
  const BigMem M = carve_big(base + tab);
  uint8_t *slot = (uint8_t *)malloc((size_t)7 << tb);
  const uint64_t rb = big_record_bytes(U);
  uint8_t *recs = (uint8_t *)malloc(rb);
  memset(recs, 0x5a, rb);
  const BigRec R = big_records(recs, U);
  std::vector<uint8_t> out;
  if (decode) out.resize(N);
  const uint64_t C = (N + S - 1) / S;
  size_t pos = 0;
  for (uint64_t c = 0; c < C; ++c) {
    const uint32_t L = (uint32_t)std::min<uint64_t>(S, N - c * S);
    Task tk{};
    std::vector<uint8_t> buf;
    uint8_t *dstb;
    uint32_t cap = 0;
    if (decode) {
      uint32_t n = 0;
      for (int i = 0; i < 4; ++i) n |= (uint32_t)in[pos + i] << (8 * i);
      pos += 4;
      buf.assign(in.begin() + pos, in.begin() + pos + n);
      pos += n;
      dstb = (uint8_t *)malloc(L);
      tk.aux = n;
    } else {
      buf.assign(in.begin() + c * S, in.begin() + c * S + L);
      cap = align16(S + S / 2 + 256);
      dstb = (uint8_t *)malloc(cap);
      tk.aux = cap;
    }
    uint8_t *src = (uint8_t *)malloc(buf.size() ? buf.size() : 1);
    memcpy(src, buf.data(), buf.size());
    tk.src = (uint64_t)(uintptr_t)src;
    tk.dst = (uint64_t)(uintptr_t)dstb;
    tk.L = L;
    tk.S = S;
    tk.U = (uint16_t)U;
    tk.tb = (uint8_t)tb;
    tk.flags = (uint8_t)(kTaskBigUnit | ((pad && c == C - 1) ? kTaskTailPad : 0));
    uint64_t result = 0;
    CodecArgs a{};
    a.tasks = &tk;
    a.ntasks = 1;
    a.result = &result;
    emu::run_wave([&](uint32_t lane) {
      if (decode)
        big_decode_task(a, tk, 0, M, T, slot, R, lane);
      else
        big_encode_task(a, tk, 0, M, T, slot, R, lane);
    });
    if (decode) {
      memcpy(out.data() + c * S, dstb, L);
    } else {
      if (result & kResultOverflow) {
        fprintf(stderr, "staging overflow\n");
        exit(3);
      }
      const uint32_t n = (uint32_t)result;
      for (int i = 0; i < 4; ++i) out.push_back((uint8_t)(n >> (8 * i)));
      out.insert(out.end(), dstb, dstb + n);
    }
    free(src);
    free(dstb);
  }
  free(recs);
  free(slot);
  free(base);
  return out;
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  const std::string mode = argv[1];
  if (mode == "enc" && argc == 8) {
    const std::vector<uint8_t> in = read_file(argv[2]);
    const uint32_t S = atoi(argv[3]), U = atoi(argv[4]), tb = atoi(argv[5]);
    const bool pad = atoi(argv[6]) != 0;
    uint8_t *slot = (uint8_t *)malloc((size_t)7 << tb);
    std::vector<uint8_t> out;
    const uint64_t N = in.size();
    const uint64_t C = (N + S - 1) / S;
    for (uint64_t c = 0; c < C; ++c) {
      const uint32_t L = (uint32_t)std::min<uint64_t>(S, N - c * S);
      std::vector<uint8_t> p;
      if (U <= 64)
        p = encode_one<1>(in.data() + c * S, L, S, U, tb, pad && c == C - 1, slot);
      else if (U <= 128)
        p = encode_one<2>(in.data() + c * S, L, S, U, tb, pad && c == C - 1, slot);
      else
        p = encode_one<4>(in.data() + c * S, L, S, U, tb, pad && c == C - 1, slot);
      const uint32_t n = (uint32_t)p.size();
      for (int i = 0; i < 4; ++i) out.push_back((uint8_t)(n >> (8 * i)));
      out.insert(out.end(), p.begin(), p.end());
    }
    free(slot);
    write_file(argv[7], out);
    return 0;
  }
  if (mode == "dec" && argc == 9) {
    const std::vector<uint8_t> in = read_file(argv[2]);
    const uint64_t N = strtoull(argv[3], nullptr, 10);
    const uint32_t S = atoi(argv[4]), U = atoi(argv[5]), tb = atoi(argv[6]);
    const int spec = atoi(argv[7]);
    uint8_t *slot = (uint8_t *)malloc((size_t)7 << tb);
    std::vector<uint8_t> out(N);
    size_t pos = 0;
    const uint64_t C = (N + S - 1) / S;
    for (uint64_t c = 0; c < C; ++c) {
      const uint32_t L = (uint32_t)std::min<uint64_t>(S, N - c * S);
      uint32_t n = 0;
      for (int i = 0; i < 4; ++i) n |= (uint32_t)in[pos + i] << (8 * i);
      pos += 4;
      std::vector<uint8_t> p(in.begin() + pos, in.begin() + pos + n);
      pos += n;
      if (U <= 64)
        decode_d<1>(spec, p, out.data() + c * S, L, S, U, tb, slot);
      else if (U <= 128)
        decode_d<2>(spec, p, out.data() + c * S, L, S, U, tb, slot);
      else
        decode_d<4>(spec, p, out.data() + c * S, L, S, U, tb, slot);
    }
    free(slot);
    write_file(argv[8], out);
    return 0;
  }
  if (mode == "tenc" && argc == 9) {
    const std::vector<uint8_t> in = read_file(argv[2]);
    const uint32_t S = atoi(argv[3]), U = atoi(argv[4]), tb = atoi(argv[5]);
    const bool pad = atoi(argv[6]) != 0;
    const uint64_t seed = strtoull(argv[7], nullptr, 10);
    std::vector<uint8_t> out;
    if (U <= 64)
      out = encode_team_all<1>(in, S, U, tb, pad, seed);
    else if (U <= 128)
      out = encode_team_all<2>(in, S, U, tb, pad, seed);
    else
      out = encode_team_all<4>(in, S, U, tb, pad, seed);
    write_file(argv[8], out);
    return 0;
  }
  if (mode == "tdec" && argc == 10) {
    const std::vector<uint8_t> in = read_file(argv[2]);
    const uint64_t N = strtoull(argv[3], nullptr, 10);
    const uint32_t S = atoi(argv[4]), U = atoi(argv[5]), tb = atoi(argv[6]);
    const int spec = atoi(argv[7]);
    const uint64_t seed = strtoull(argv[8], nullptr, 10);
    std::vector<uint8_t> out(N);
    if (U <= 64)
      decode_team_d<1>(spec, in, out.data(), N, S, U, tb, seed);
    else if (U <= 128)
      decode_team_d<2>(spec, in, out.data(), N, S, U, tb, seed);
    else
      decode_team_d<4>(spec, in, out.data(), N, S, U, tb, seed);
    write_file(argv[9], out);
    return 0;
  }
  if (mode == "fenc" && argc == 7) {
    const std::vector<uint8_t> in = read_file(argv[2]);
    const uint32_t S = atoi(argv[3]), tb = atoi(argv[4]);
    const FrozenSetup fs(tb, argv[5]);
    write_file(argv[6], frozen_encode_all(in, S, tb, fs));
    return 0;
  }
  if (mode == "fdec" && argc == 9) {
    const std::vector<uint8_t> in = read_file(argv[2]);
    const uint64_t N = strtoull(argv[3], nullptr, 10);
    const uint32_t S = atoi(argv[4]), tb = atoi(argv[5]);
    const int spec = atoi(argv[6]);
    const FrozenSetup fs(tb, argv[7]);
    std::vector<uint8_t> out(N);
    switch (spec) {
      case 1: frozen_decode_all<1>(in, out.data(), N, S, tb, fs); break;
      case 2: frozen_decode_all<2>(in, out.data(), N, S, tb, fs); break;
      case 4: frozen_decode_all<4>(in, out.data(), N, S, tb, fs); break;
      case 8: frozen_decode_all<8>(in, out.data(), N, S, tb, fs); break;
      default: fprintf(stderr, "spec bits\n"); return 2;
    }
    write_file(argv[8], out);
    return 0;
  }
  if (mode == "benc" && argc == 8) {
    const std::vector<uint8_t> in = read_file(argv[2]);
    const uint32_t S = atoi(argv[3]), U = atoi(argv[4]), tb = atoi(argv[5]);
    write_file(argv[7], big_run(false, in, in.size(), S, U, tb, atoi(argv[6]) != 0));
    return 0;
  }
  if (mode == "bdec" && argc == 8) {
    const std::vector<uint8_t> in = read_file(argv[2]);
    const uint64_t N = strtoull(argv[3], nullptr, 10);
    const uint32_t S = atoi(argv[4]), U = atoi(argv[5]), tb = atoi(argv[6]);
    write_file(argv[7], big_run(true, in, N, S, U, tb, false));
    return 0;
  }
  fprintf(stderr, "usage: wave_emu enc|dec|tenc|tdec|fenc|fdec|benc|bdec ...\n");
  return 2;
}
