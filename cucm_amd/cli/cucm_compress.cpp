// cucm_compress -- drop-in for the reference's compressor command line
// (cucm_compress_v2.cu:1289-1418): same options, defaults and messages; the
// input goes through the streaming C ABI one segment of batch records at a
// time, so files larger than the GPU's memory compress with a bounded
// footprint.  Extra options: -t/--table-bits (default 25 = reference
// containers), -d/--device, --no-tail-pad, --frozen.
#include <chrono>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "cli_common.h"
#include "cucm.h"

static void print_help(const char *prog) {
  std::cout << "Usage: " << prog << " [options]\n";
  std::cout << "Options:\n";
  std::cout << "  -c, --chunknum <N>    Number of chunks (default: file_size/chunksize + 1)\n";
  std::cout << "  -s, --chunksize <SZ>  Size of each chunk, supports K/M/G suffix (default: 1M)\n";
  std::cout << "  -u, --unit <SZ>       Unit size, supports K/M/G suffix (default: 128)\n";
  std::cout << "  -f, --file <path>     Input file to compress (default: ./enwik8)\n";
  std::cout << "  -o, --output <path>   Output compressed file (default: ./compressed.bin)\n";
  std::cout << "  -p, --prelearn <SZ>   Prelearn size, supports K/M/G suffix (default: 0 = disabled)\n";
  std::cout << "  -b, --batch <NUM>     Chunks per batch record (default: 100)\n";
  std::cout << "  -t, --table-bits <N>  log2 slots per order table, 8..25 (default: 25)\n";
  std::cout << "  -d, --device <N>      GPU to use (default: 0)\n";
  std::cout << "      --no-tail-pad     Code the last chunk only up to its end\n";
  std::cout << "      --frozen          Frozen model: the prelearned (or zero) model codes every\n"
               "                        chunk and is never updated\n";
  std::cout << "  -h, --help            Show this help message\n";
}

static bool is(const char *a, const char *s, const char *l) {
  return std::strcmp(a, s) == 0 || (l && std::strcmp(a, l) == 0);
}

int main(int argc, char *argv[]) {
  if (argc == 1) {
    print_help(argv[0]);
    return 0;
  }
  int unit_size = 128, chunknum = -1, batch = 100, table_bits = 25, device = 0;
  uint64_t chunk_size = 1024 * 1024, prelearn_size = 0;
  bool tail_pad = true, frozen = false;
  std::string filename = "./enwik8", output_filename = "./compressed.bin";
  for (int i = 1; i < argc; ++i) {
    const bool more = i + 1 < argc;
    if (is(argv[i], "-c", "--chunknum") && more) {
      chunknum = std::atoi(argv[++i]);
      if (chunknum <= 0) {
        std::cerr << "Invalid chunknum: " << chunknum << std::endl;
        return 1;
      }
    } else if (is(argv[i], "-s", "--chunksize") && more) {
      chunk_size = cli_parse_size(argv[++i]);
      if (chunk_size == 0) {
        std::cerr << "Invalid chunk size: " << argv[i] << std::endl;
        return 1;
      }
    } else if (is(argv[i], "-f", "--file") && more) {
      filename = argv[++i];
    } else if (is(argv[i], "-o", "--output") && more) {
      output_filename = argv[++i];
    } else if (is(argv[i], "-p", "--prelearn") && more) {
      prelearn_size = cli_parse_size(argv[++i]);
      if (prelearn_size == 0) std::cerr << "Warning: prelearn disabled (size=0)" << std::endl;
    } else if (is(argv[i], "-u", "--unit") && more) {
      unit_size = std::atoi(argv[++i]);
      if (unit_size <= 0) {
        std::cerr << "Invalid unit size: " << unit_size << std::endl;
        return 1;
      }
    } else if (is(argv[i], "-b", "--batch") && more) {
      batch = std::atoi(argv[++i]);
      if (batch <= 0) {
        std::cerr << "Invalid batch number: " << batch << std::endl;
        return 1;
      }
    } else if (is(argv[i], "-t", "--table-bits") && more) {
      table_bits = std::atoi(argv[++i]);
    } else if (is(argv[i], "-d", "--device") && more) {
      device = std::atoi(argv[++i]);
    } else if (is(argv[i], "--no-tail-pad", nullptr)) {
      tail_pad = false;
    } else if (is(argv[i], "--frozen", nullptr)) {
      frozen = true;
    } else if (is(argv[i], "-h", "--help")) {
      print_help(argv[0]);
      return 0;
    } else {
      std::cerr << "Unknown option: " << argv[i] << std::endl;
      print_help(argv[0]);
      return 1;
    }
  }
  FILE *in = std::fopen(filename.c_str(), "rb");
  if (!in) {
    std::cerr << "Failed to open input file: " << filename << std::endl;
    return 1;
  }
  std::fseek(in, 0, SEEK_END);
  uint64_t file_size = (uint64_t)std::ftell(in);
  std::fseek(in, 0, SEEK_SET);
  if (chunknum == -1) chunknum = (int)((file_size + chunk_size - 1) / chunk_size);
  std::cout << "Compressing file: " << filename << ", filesize: " << file_size
            << ", chunksize: " << chunk_size << ", chunknum: " << chunknum
            << ", unit: " << unit_size << ", prelearn: " << prelearn_size << ", batch: " << batch
            << ", output: " << output_filename << std::endl;
  FILE *out = std::fopen(output_filename.c_str(), "wb");
  if (!out) {
    std::cerr << "Failed to open output file: " << output_filename << std::endl;
    return 1;
  }
  // -c truncates the input to chunknum chunks (cucm_compress_v2.cu:1396-1398)
  if (chunk_size * (uint64_t)chunknum < file_size) file_size = chunk_size * (uint64_t)chunknum;
  const auto t0 = std::chrono::high_resolution_clock::now();

  cucm_params p;
  cucm_default_params(&p);
  if (chunk_size > 0xffffffffull || unit_size > 0xffff || batch > 0xffff ||
      prelearn_size > 0xffffffffull || file_size > 0xffffffffull) {
    std::cerr << "Error: sizes exceed the container's 32-bit fields" << std::endl;
    return 1;
  }
  p.chunk_size = (uint32_t)chunk_size;
  p.unit_bits = (uint16_t)unit_size;
  p.batch_chunks = (uint16_t)batch;
  p.prelearn_size = (uint32_t)prelearn_size;
  p.table_bits = (uint8_t)table_bits;
  p.flags = (tail_pad ? 0u : CUCM_FLAG_NO_TAIL_PAD) | (frozen ? CUCM_FLAG_FROZEN : 0u);
  cucm_ctx *ctx = nullptr;
  int rc = cucm_ctx_create(&ctx, device, 0);
  if (rc) {
    std::cerr << "Error: " << cucm_strerror(rc) << std::endl;
    return 1;
  }
  std::vector<uint8_t> pre(prelearn_size);
  if (!read_exact(in, pre.data(), pre.size())) {
    std::cerr << "Error: short read" << std::endl;
    return 1;
  }
  // header + prelearn segment: the coded segment is kept only when shorter
  // than the raw bytes, so it never exceeds 21 + P bytes
  std::vector<uint8_t> head(21 + prelearn_size);
  uint64_t head_len = 0, total = 0;
  cucm_stream *s = nullptr;
  rc = cucm_stream_compress_begin(ctx, file_size, &p, pre.data(), head.data(), head.size(),
                                  &head_len, &s);
  if (rc == CUCM_E_SMALL) {
    head.resize(head_len);
    rc = cucm_stream_compress_begin(ctx, file_size, &p, pre.data(), head.data(), head.size(),
                                    &head_len, &s);
  }
  if (rc || !write_exact(out, head.data(), head_len)) {
    std::cerr << "Error: " << (rc ? cucm_strerror(rc) : "write failed") << std::endl;
    return 1;
  }
  total += head_len;
  // segments of whole batch records, about 256 MiB of input each
  const uint64_t per_batch = chunk_size * (uint64_t)batch;
  const uint64_t seg = per_batch * std::max<uint64_t>(1, (256ull << 20) / per_batch);
  std::vector<uint8_t> buf, obuf;
  for (uint64_t left = file_size - prelearn_size; left > 0;) {
    const uint64_t n = std::min(seg, left);
    buf.resize(n);
    if (!read_exact(in, buf.data(), n)) {
      std::cerr << "Error: short read" << std::endl;
      return 1;
    }
    const uint64_t chunks = (n + chunk_size - 1) / chunk_size;
    obuf.resize(n + n / 2 + chunks * chunk_size / 2 + 4096 * (chunks + 2));
    uint64_t olen = 0;
    rc = cucm_stream_compress_next(s, buf.data(), n, obuf.data(), obuf.size(), &olen);
    if (rc == CUCM_E_SMALL) {  // incompressible beyond 1.5x: use the strict bound
      cucm_params q = p;
      q.prelearn_size = 0;
      obuf.resize(cucm_compress_bound(n, &q));
      rc = cucm_stream_compress_next(s, buf.data(), n, obuf.data(), obuf.size(), &olen);
    }
    if (rc || !write_exact(out, obuf.data(), olen)) {
      std::cerr << "Error: " << (rc ? cucm_strerror(rc) : "write failed") << std::endl;
      return 1;
    }
    total += olen;
    left -= n;
  }
  cucm_stream_end(s);
  cucm_ctx_destroy(ctx);
  std::fclose(in);
  std::fclose(out);
  const std::chrono::duration<double> el = std::chrono::high_resolution_clock::now() - t0;
  std::cout << "Original size  : " << file_size << " bytes\n";
  std::cout << "Compressed size: " << total << " bytes\n";
  std::cout << "Compression ratio: " << (double)file_size / (double)total << "\n";
  std::cout << "Total time: " << el.count() << " sec" << std::endl;
  return 0;
}
