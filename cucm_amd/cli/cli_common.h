// cli_common.h -- helpers shared by the cucm_compress / cucm_decompress
// command-line tools.
#ifndef CUCM_CLI_COMMON_H
#define CUCM_CLI_COMMON_H
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <string>

// Size with an optional K/M/G suffix (binary units), as the reference's
// parse_size (cucm_compress_v2.cu:1252-1287) reads it: leading blanks, the
// digits, optional blanks, then one unit letter (case-insensitive); anything
// else after the digits is ignored; no digits means 0.
inline uint64_t cli_parse_size(const std::string &s) {
  size_t i = 0;
  while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  size_t j = i;
  while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j;
  if (j == i) return 0;
  uint64_t v = std::stoull(s.substr(i, j - i));
  while (j < s.size() && std::isspace((unsigned char)s[j])) ++j;
  if (j < s.size()) {
    switch (std::toupper((unsigned char)s[j])) {
      case 'K': v *= 1024ull; break;
      case 'M': v *= 1024ull * 1024; break;
      case 'G': v *= 1024ull * 1024 * 1024; break;
      default: break;
    }
  }
  return v;
}

inline bool read_exact(FILE *f, void *buf, size_t n) {
  return n == 0 || fread(buf, 1, n, f) == n;
}

inline bool write_exact(FILE *f, const void *buf, size_t n) {
  return n == 0 || fwrite(buf, 1, n, f) == n;
}

#endif
