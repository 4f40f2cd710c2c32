// cucm_decompress -- drop-in for the reference's decompressor command line
// (cucm_decompress_v2.cu:1250-1313): -i input, -o output, -a
// maxassumed_length (the reference decides maxassumed_length + 1 bits per
// speculative round, 3 + 1 by default; here -a m selects the largest width
// of 1, 2, 4, 8 not above m + 1, and without -a the library chooses the
// width with the launch shape -- neither ever changes the output).  Batch
// records stream through the C ABI, about 256 MiB of output at a time.
// Extra: -d device.
#include <chrono>
#include <cstring>
#include <iostream>
#include <vector>

#include "cli_common.h"
#include "cucm.h"

static void print_help(const char *prog) {
  std::cout << "Usage: " << prog << " -i <input_file> -o <output_file> [-a <maxassumed_length>]\n";
  std::cout << "  -i <input_file>        Input compressed file\n";
  std::cout << "  -o <output_file>       Output decompressed file\n";
  std::cout << "  -a <N>                 Set maxassumed_length (default: the library's automatic\n"
               "                         speculation width; the output is the same for every N)\n";
  std::cout << "  -d <N>                 GPU to use (default: 0)\n";
}

int main(int argc, char *argv[]) {
  const char *input_file = nullptr, *output_file = nullptr;
  int maxassumed_length = -1, device = 0;  // -1: automatic width
  for (int i = 1; i < argc; ++i) {
    if (std::strcmp(argv[i], "-i") == 0 && i + 1 < argc) {
      input_file = argv[++i];
    } else if (std::strcmp(argv[i], "-o") == 0 && i + 1 < argc) {
      output_file = argv[++i];
    } else if (std::strcmp(argv[i], "-a") == 0 && i + 1 < argc) {
      maxassumed_length = std::atoi(argv[++i]);
    } else if (std::strcmp(argv[i], "-d") == 0 && i + 1 < argc) {
      device = std::atoi(argv[++i]);
    } else {
      print_help(argv[0]);
      return 1;
    }
  }
  if (!input_file || !output_file) {
    print_help(argv[0]);
    return 1;
  }
  FILE *in = std::fopen(input_file, "rb");
  if (!in) {
    std::cerr << "Error: cannot open input file " << input_file << "\n";
    return 1;
  }
  FILE *out = std::fopen(output_file, "wb");
  if (!out) {
    std::cerr << "Error: cannot open output file " << output_file << "\n";
    return 1;
  }
  const uint32_t want = (uint32_t)std::max(1, maxassumed_length + 1);
  const uint32_t spec =
      maxassumed_length < 0 ? 0 : want >= 8 ? 8 : want >= 4 ? 4 : want >= 2 ? 2 : 1;
  const auto t0 = std::chrono::high_resolution_clock::now();
  std::vector<uint8_t> head(21);
  uint64_t head_len = 0;
  if (!read_exact(in, head.data(), 21) ||
      cucm_container_head_len(head.data(), 21, &head_len) != CUCM_OK) {
    std::cerr << "Error: not a container\n";
    return 1;
  }
  head.resize(head_len);
  if (!read_exact(in, head.data() + 21, head_len - 21)) {
    std::cerr << "Error: truncated container\n";
    return 1;
  }
  cucm_ctx *ctx = nullptr;
  int rc = cucm_ctx_create(&ctx, device, 0);
  if (rc) {
    std::cerr << "Error: " << cucm_strerror(rc) << "\n";
    return 1;
  }
  uint64_t N = 0;
  cucm_params p;
  if (cucm_peek_header(head.data(), head.size(), &N, &p) != CUCM_OK) {
    std::cerr << "Error: not a container\n";
    return 1;
  }
  std::vector<uint8_t> pre(p.prelearn_size);  // the leading bytes the model learned from
  cucm_stream *s = nullptr;
  rc = cucm_stream_decompress_begin(ctx, head.data(), head.size(), spec, &N, &p, pre.data(),
                                    pre.size(), &s);
  if (rc || !write_exact(out, pre.data(), pre.size())) {
    std::cerr << "Error: " << (rc ? cucm_strerror(rc) : "write failed") << "\n";
    return 1;
  }
  // groups of whole batch records, about 256 MiB of output each
  const uint64_t per_batch = (uint64_t)p.chunk_size * p.batch_chunks;
  const uint64_t group = std::max<uint64_t>(1, per_batch ? (256ull << 20) / per_batch : 1);
  uint64_t left = N - p.prelearn_size;
  std::vector<uint8_t> recs, obuf;
  while (left > 0) {
    recs.clear();
    uint64_t covered = 0;
    for (uint64_t r = 0; r < group && covered < left; ++r) {
      uint8_t lb[4];
      if (!read_exact(in, lb, 4)) {
        std::cerr << "Error: truncated container\n";
        return 1;
      }
      const uint64_t blen = (uint64_t)lb[0] | (uint64_t)lb[1] << 8 | (uint64_t)lb[2] << 16 |
                            (uint64_t)lb[3] << 24;
      const size_t at = recs.size();
      recs.resize(at + 4 + blen);
      std::memcpy(recs.data() + at, lb, 4);
      if (!read_exact(in, recs.data() + at + 4, blen)) {
        std::cerr << "Error: truncated container\n";
        return 1;
      }
      covered += per_batch;
    }
    const uint64_t n = std::min(covered, left);
    obuf.resize(n);
    uint64_t olen = 0;
    rc = cucm_stream_decompress_next(s, recs.data(), recs.size(), obuf.data(), obuf.size(), &olen);
    if (rc || !write_exact(out, obuf.data(), olen)) {
      std::cerr << "Error: " << (rc ? cucm_strerror(rc) : "write failed") << "\n";
      return 1;
    }
    left -= olen;
  }
  cucm_stream_end(s);
  cucm_ctx_destroy(ctx);
  std::fclose(in);
  std::fclose(out);
  const std::chrono::duration<double> el = std::chrono::high_resolution_clock::now() - t0;
  std::cout << "Decompression finished: " << input_file << " -> " << output_file << " (" << N
            << " bytes, maxassumed_length=";
  if (maxassumed_length < 0)
    std::cout << "automatic";
  else
    std::cout << maxassumed_length;
  std::cout << ")\n";
  std::cout << "Total time: " << el.count() << " sec" << std::endl;
  return 0;
}
