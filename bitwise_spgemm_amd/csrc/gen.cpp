// gen.cpp — host-side deterministic input generators (no GPU).
//
// xoshiro256** (Blackman & Vigna, 2018) seeded through splitmix64, the
// published algorithms.  They replace the reference's unseeded glibc rand()
// fill (fill_random, source/src/utils.h:14-22) so that every configuration is
// reproducible from a seed, in this library, in the tests and in the bench.
#include <math.h>
#include <string.h>

#include <vector>

#include "bws.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix(x);
  }
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform double in [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [0, n) (Lemire's multiply-shift with rejection)
  uint64_t below(uint64_t n) {
    uint64_t x = next();
    __uint128_t m = (__uint128_t)x * n;
    uint64_t l = (uint64_t)m;
    if (l < n) {
      uint64_t t = (0 - n) % n;
      while (l < t) {
        x = next();
        m = (__uint128_t)x * n;
        l = (uint64_t)m;
      }
    }
    return (uint64_t)(m >> 64);
  }
  // standard normal (Box–Muller, one value per call)
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
};

inline uint32_t wbit(int64_t i) { return 0x80000000u >> (i & 31); }

}  // namespace

extern "C" {

bws_status bws_gen_bitmap_iid(uint64_t seed, int64_t rows, int64_t cols, double p, uint32_t* W,
                              int64_t ldw) {
  if (rows < 0 || cols < 0 || ldw < (cols + 31) / 32 || !(p >= 0.0 && p <= 1.0)) return BWS_EINVAL;
  if (rows > 0 && ldw > 0 && !W) return BWS_EINVAL;
  Rng g(seed);
  for (int64_t r = 0; r < rows; ++r) {
    uint32_t* w = W + r * ldw;
    memset(w, 0, (size_t)ldw * 4);
    for (int64_t c = 0; c < cols; ++c)
      if (g.uniform() < p) w[c >> 5] |= wbit(c);
  }
  return BWS_OK;
}

bws_status bws_gen_bitmap_clustered(uint64_t seed, int64_t rows, int64_t cols, int32_t n_proto,
                                    double p_proto, double flip, uint32_t* W, int64_t ldw) {
  if (rows < 0 || cols < 0 || ldw < (cols + 31) / 32 || n_proto < 1) return BWS_EINVAL;
  if (!(p_proto >= 0.0 && p_proto <= 1.0) || !(flip >= 0.0 && flip <= 1.0)) return BWS_EINVAL;
  if (rows > 0 && ldw > 0 && !W) return BWS_EINVAL;
  Rng g(seed);
  const int64_t wpr = (cols + 31) / 32;
  std::vector<uint32_t> proto((size_t)n_proto * (size_t)wpr, 0u);
  for (int32_t c = 0; c < n_proto; ++c)
    for (int64_t k = 0; k < cols; ++k)
      if (g.uniform() < p_proto) proto[(size_t)c * wpr + (k >> 5)] |= wbit(k);
  std::vector<uint32_t> tmp((size_t)rows * (size_t)wpr, 0u);
  for (int64_t r = 0; r < rows; ++r) {
    const uint64_t c = g.below((uint64_t)n_proto);
    uint32_t* dst = &tmp[(size_t)r * wpr];
    memcpy(dst, &proto[(size_t)c * wpr], (size_t)wpr * 4);
    for (int64_t k = 0; k < cols; ++k)
      if (g.uniform() < flip) dst[k >> 5] ^= wbit(k);
  }
  // Fisher–Yates shuffle of the row order
  std::vector<int64_t> perm((size_t)rows);
  for (int64_t r = 0; r < rows; ++r) perm[(size_t)r] = r;
  for (int64_t r = rows - 1; r > 0; --r) {
    const int64_t j = (int64_t)g.below((uint64_t)(r + 1));
    int64_t t = perm[(size_t)r];
    perm[(size_t)r] = perm[(size_t)j];
    perm[(size_t)j] = t;
  }
  for (int64_t r = 0; r < rows; ++r) {
    uint32_t* w = W + r * ldw;
    memset(w, 0, (size_t)ldw * 4);
    memcpy(w, &tmp[(size_t)perm[(size_t)r] * wpr], (size_t)wpr * 4);
  }
  return BWS_OK;
}

bws_status bws_gen_f32(uint64_t seed, int64_t n, int32_t kind, double a, double b, float* out) {
  if (n < 0 || (n > 0 && !out) || kind < 0 || kind > 2) return BWS_EINVAL;
  Rng g(seed);
  for (int64_t i = 0; i < n; ++i) {
    double v;
    if (kind == 0) {
      v = a + b * g.normal();
    } else if (kind == 1) {
      v = a + (b - a) * g.uniform();
    } else {
      const double e = a + (b - a) * g.uniform();
      v = pow(10.0, e) * ((g.next() >> 63) ? -1.0 : 1.0);
    }
    out[i] = (float)v;
  }
  return BWS_OK;
}

}  // extern "C"
