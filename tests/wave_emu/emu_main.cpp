// emu_main.cpp -- runs the codec's per-wave device code (cucm_wave.h) on
// the CPU, one chunk at a time, under the fiber wave emulation of
// wave_emulation.h (test infrastructure; the product never runs this).
//
//   wave_emu enc <input> <S> <U> <tb> <tail_pad> <output>
//       output: per chunk, a u32 LE payload length and the payload
//   wave_emu dec <payloads> <N> <S> <U> <tb> <spec_bits> <output>
//       payloads in the format enc writes; output: the N decoded bytes
#define CUCM_WAVE_EMULATION 1
#include "wave_emulation.h"
// (the device code follows the emulation shims)
#include "cucm_wave.h"
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
  emu::run_wave([&](uint32_t lane) { encode_task<NJ>(a, tk, 0, lds.M, lds.T, slot, U, lane); });
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
  emu::run_wave([&](uint32_t lane) { decode_task<NJ, D>(a, tk, 0, lds.M, lds.T, slot, U, lane); });
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
  fprintf(stderr, "usage: wave_emu enc|dec ...\n");
  return 2;
}
