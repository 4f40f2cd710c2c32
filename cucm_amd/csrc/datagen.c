/*
 * datagen.c -- seeded synthetic inputs for the CuCM benchmark workloads.
 *
 * PRNG: splitmix64 seeding xoshiro256** (public-domain algorithms of
 * Vigna/Blackman).  Every generator is a pure function of (n, seed) and is
 * OpenMP-parallel over independent regions, so any host regenerates the same
 * bytes quickly.
 *
 *   markov2     order-2 Markov source over 64 symbols (bytes 0x20+s); each of
 *               the 4096 (a, b) rows is Zipf(s) over ranks 1..64 mapped
 *               through a per-row random permutation.
 *   mixed       4 KiB slots: 15% uniform random; the rest order-3 Markov text
 *               over 96 printable symbols (Zipf 1.2 rows, a space forced with
 *               probability 0.18) interleaved with LZ copies (length
 *               Geom(1/64), offset U[1, min(1 MiB, pos))) at a 60:25 byte ratio.
 *   structured  4 MiB regions cycling int32 LE random walks (steps
 *               round(N(0, 32^2))), float32 1024x1024 smooth fields (8
 *               sinusoids + N(0, 1e-3) noise) and 16-byte records of four
 *               Zipf(1.0) categorical u32 fields over 256 levels.
 *   lognormal   file sizes clip(round(median * exp(sigma Z)), lo, hi).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed) {
  uint64_t x = seed;
  for (int i = 0; i < 4; ++i) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r) {
  uint64_t *s = r->s;
  const uint64_t res = rotl(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return res;
}

static inline double rng_unit(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

static inline uint64_t rng_below(rng_t *r, uint64_t n) { return n ? rng_next(r) % n : 0; }

static double rng_normal(rng_t *r) {
  double u1 = rng_unit(r), u2 = rng_unit(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Geometric number of trials until success with probability p (>= 1). */
static uint64_t rng_geom(rng_t *r, double p) {
  double u = rng_unit(r);
  if (u < 1e-300) u = 1e-300;
  return 1 + (uint64_t)floor(log(u) / log(1.0 - p));
}

static void zipf_cdf(double *cdf, int k, double s) {
  double acc = 0;
  for (int i = 0; i < k; ++i) { acc += pow((double)(i + 1), -s); cdf[i] = acc; }
  for (int i = 0; i < k; ++i) cdf[i] /= acc;
}

static inline int zipf_draw(rng_t *r, const double *cdf, int k) {
  const double u = rng_unit(r);
  int lo = 0, hi = k - 1;
  while (lo < hi) { int mid = (lo + hi) >> 1; if (cdf[mid] < u) lo = mid + 1; else hi = mid; }
  return lo;
}

static void shuffle(uint8_t *p, int k, uint64_t seed) {
  rng_t r;
  rng_seed(&r, seed);
  for (int i = 0; i < k; ++i) p[i] = (uint8_t)i;
  for (int i = k - 1; i > 0; --i) {
    int j = (int)rng_below(&r, (uint64_t)i + 1);
    uint8_t t = p[i]; p[i] = p[j]; p[j] = t;
  }
}

/* ---- order-2 Markov, 64 symbols ------------------------------------- */
void cucm_gen_markov2(uint8_t *out, uint64_t n, uint64_t seed, double s) {
  enum { K = 64 };
  double cdf[K];
  zipf_cdf(cdf, K, s);
  uint8_t *perm = (uint8_t *)malloc((size_t)K * K * K);
  for (int row = 0; row < K * K; ++row)
    shuffle(perm + (size_t)row * K, K, seed * 0x9e3779b97f4a7c15ull ^ (uint64_t)(row + 1) * 0xd1b54a32d192ed03ull);
  rng_t r;
  rng_seed(&r, seed);
  int a = 0, b = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const int sym = perm[(size_t)(a * K + b) * K + zipf_draw(&r, cdf, K)];
    out[i] = (uint8_t)(0x20 + sym);
    a = b;
    b = sym;
  }
  free(perm);
}

/* Many markov2 files (config 3 content): file i seeded with seed0 + i. */
void cucm_gen_markov2_files(uint8_t *out, const uint64_t *offsets, const uint64_t *sizes,
                            uint32_t nfiles, uint64_t seed0, double s, int threads) {
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int64_t i = 0; i < (int64_t)nfiles; ++i)
    cucm_gen_markov2(out + offsets[i], sizes[i], seed0 + (uint64_t)i, s);
  (void)threads;
}

/* ---- mixed text / LZ / random stream --------------------------------- */
#define SLOT 4096
typedef struct { uint64_t start, len, off; } copy_seg;

static inline uint8_t text_symbol(rng_t *r, const double *cdf, uint32_t ctx, uint64_t seed) {
  /* per-context permutation of 96 ranks: rank -> (mult * rank + add) mod 96
   * with mult coprime to 96, both drawn from a hash of the context. */
  static const uint8_t mults[32] = {1, 5, 7, 11, 13, 17, 19, 23, 25, 29, 31, 35, 37, 41, 43, 47,
                                    49, 53, 55, 59, 61, 65, 67, 71, 73, 77, 79, 83, 85, 89, 91, 95};
  uint64_t h = (uint64_t)ctx * 0x9e3779b97f4a7c15ull ^ seed;
  h ^= h >> 29; h *= 0xbf58476d1ce4e5b9ull; h ^= h >> 32;
  const int mult = mults[h & 31], add = (int)((h >> 8) % 96);
  const int rank = zipf_draw(r, cdf, 96);
  return (uint8_t)(0x20 + (mult * rank + add) % 96);
}

void cucm_gen_mixed(uint8_t *out, uint64_t n, uint64_t seed, int threads) {
  double cdf[96];
  zipf_cdf(cdf, 96, 1.2);
  const uint64_t nslots = (n + SLOT - 1) / SLOT;
  /* up to SLOT/8 copy segments per slot */
  const uint64_t maxseg = SLOT / 8;
  copy_seg *segs = (copy_seg *)malloc(sizeof(copy_seg) * nslots * maxseg);
  uint32_t *nseg = (uint32_t *)calloc(nslots, sizeof(uint32_t));
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
#endif
  for (int64_t si = 0; si < (int64_t)nslots; ++si) {
    rng_t r;
    rng_seed(&r, seed ^ ((uint64_t)si + 1) * 0x2545f4914f6cdd1dull);
    const uint64_t s0 = (uint64_t)si * SLOT;
    const uint64_t s1 = s0 + SLOT < n ? s0 + SLOT : n;
    if (rng_unit(&r) < 0.15) {
      for (uint64_t i = s0; i < s1; ++i) out[i] = (uint8_t)rng_next(&r);
      continue;
    }
    uint32_t ctx = (uint32_t)rng_next(&r) & 0xffffff;
    uint64_t i = s0;
    int in_text = 1;
    while (i < s1) {
      if (in_text) {
        uint64_t len = rng_geom(&r, 1.0 / 154.0);
        for (; len && i < s1; --len, ++i) {
          uint8_t c = rng_unit(&r) < 0.18 ? (uint8_t)' ' : text_symbol(&r, cdf, ctx, seed);
          out[i] = c;
          ctx = ((ctx << 8) | c) & 0xffffff;
        }
      } else {
        uint64_t len = rng_geom(&r, 1.0 / 64.0);
        if (len > s1 - i) len = s1 - i;
        const uint64_t span = i < (1u << 20) ? i : (1u << 20);
        if (span >= 2 && nseg[si] < maxseg) {
          copy_seg *cs = &segs[si * maxseg + nseg[si]++];
          cs->start = i;
          cs->len = len;
          cs->off = 1 + rng_below(&r, span - 1);
          for (uint64_t k = 0; k < len; ++k) out[i + k] = 0; /* filled in pass 2 */
          i += len;
        }
      }
      in_text = !in_text;
    }
  }
  /* pass 2: copies in position order (sources are always earlier bytes) */
  for (uint64_t si = 0; si < nslots; ++si)
    for (uint32_t k = 0; k < nseg[si]; ++k) {
      const copy_seg *cs = &segs[si * maxseg + k];
      for (uint64_t j = 0; j < cs->len; ++j) out[cs->start + j] = out[cs->start + j - cs->off];
    }
  free(segs);
  free(nseg);
  (void)threads;
}

/* ---- structured binary ------------------------------------------------ */
#define REGION (4u << 20)

void cucm_gen_structured(uint8_t *out, uint64_t n, uint64_t seed, int threads) {
  const uint64_t nreg = (n + REGION - 1) / REGION;
  double zcdf[256];
  zipf_cdf(zcdf, 256, 1.0);
#ifdef _OPENMP
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
#endif
  for (int64_t ri = 0; ri < (int64_t)nreg; ++ri) {
    rng_t r;
    rng_seed(&r, seed ^ ((uint64_t)ri + 7) * 0x9e3779b97f4a7c15ull);
    uint8_t *reg = (uint8_t *)malloc(REGION);
    const int kind = (int)(ri % 3);
    if (kind == 0) {
      int32_t v = (int32_t)(rng_next(&r) & 0xffff) - 32768;
      for (uint32_t i = 0; i < REGION / 4; ++i) {
        v += (int32_t)lrint(32.0 * rng_normal(&r));
        memcpy(reg + 4 * i, &v, 4);
      }
    } else if (kind == 1) {
      double A[8], phi[8];
      int u[8], w[8];
      for (int t = 0; t < 8; ++t) {
        A[t] = 0.5 + 1.5 * rng_unit(&r);
        u[t] = 1 + (int)rng_below(&r, 8);
        w[t] = 1 + (int)rng_below(&r, 8);
        phi[t] = 6.283185307179586 * rng_unit(&r);
      }
      for (uint32_t i = 0; i < 1024; ++i)
        for (uint32_t j = 0; j < 1024; ++j) {
          double acc = 0;
          for (int t = 0; t < 8; ++t)
            acc += A[t] * sin(6.283185307179586 * (u[t] * (double)i + w[t] * (double)j) / 1024.0 + phi[t]);
          const float f = (float)(acc + 1e-3 * rng_normal(&r));
          memcpy(reg + 4 * (i * 1024 + j), &f, 4);
        }
    } else {
      uint32_t dict[4][256];
      for (int f = 0; f < 4; ++f)
        for (int l = 0; l < 256; ++l) dict[f][l] = (uint32_t)rng_next(&r);
      for (uint32_t i = 0; i < REGION / 16; ++i)
        for (int f = 0; f < 4; ++f) {
          const uint32_t v = dict[f][zipf_draw(&r, zcdf, 256)];
          memcpy(reg + 16 * i + 4 * f, &v, 4);
        }
    }
    const uint64_t o = (uint64_t)ri * REGION;
    const uint64_t m = n - o < REGION ? n - o : REGION;
    memcpy(out + o, reg, m);
    free(reg);
  }
  (void)threads;
}

/* ---- file sizes --------------------------------------------------------- */
void cucm_gen_lognormal_sizes(uint64_t *sizes, uint32_t count, double median, double sigma,
                              uint64_t lo, uint64_t hi, uint64_t seed) {
  rng_t r;
  rng_seed(&r, seed);
  for (uint32_t i = 0; i < count; ++i) {
    double v = floor(median * exp(sigma * rng_normal(&r)) + 0.5);
    if (v < (double)lo) v = (double)lo;
    if (v > (double)hi) v = (double)hi;
    sizes[i] = (uint64_t)v;
  }
}

/* Poisson arrival times (seconds) at rate lambda. */
void cucm_gen_poisson_arrivals(double *t, uint32_t count, double lambda, uint64_t seed) {
  rng_t r;
  rng_seed(&r, seed);
  double acc = 0;
  for (uint32_t i = 0; i < count; ++i) {
    double u = rng_unit(&r);
    if (u < 1e-300) u = 1e-300;
    acc += -log(u) / lambda;
    t[i] = acc;
  }
}
