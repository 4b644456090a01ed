// gg_sa_common.h -- small host helpers shared by the SA drivers and graph
// builders (RNG, edge hash set, wall clock).
#ifndef GG_SA_COMMON_H
#define GG_SA_COMMON_H
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <vector>

namespace gg {

// xoshiro256** seeded through splitmix64: the generators' own RNG (the SA
// drivers use libc random() instead, to follow the reference's draws).
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (int i = 0; i < 4; ++i) {
      seed += 0x9E3779B97F4A7C15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
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
  // Uniform in [0, m) (Lemire's multiply-shift; bias < 2^-32 for our m).
  uint32_t below(uint32_t m) { return (uint32_t)(((next() >> 32) * (uint64_t)m) >> 32); }
};

inline uint64_t ekey(int a, int b) {
  return a < b ? ((uint64_t)(uint32_t)a << 32) | (uint32_t)b : ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
}

struct EdgeSet {  // open-addressing set of unordered edges
  std::vector<uint64_t> slot;
  uint64_t mask;
  explicit EdgeSet(size_t n_edges) {
    size_t cap = 16;
    while (cap < n_edges * 4) cap <<= 1;
    slot.assign(cap, ~0ull);
    mask = cap - 1;
  }
  static uint64_t h(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return k;
  }
  bool has(uint64_t k) const {
    for (uint64_t i = h(k) & mask;; i = (i + 1) & mask) {
      if (slot[i] == k) return true;
      if (slot[i] == ~0ull) return false;
    }
  }
  void add(uint64_t k) {
    uint64_t i = h(k) & mask;
    while (slot[i] != ~0ull && slot[i] != ~1ull) i = (i + 1) & mask;
    slot[i] = k;
  }
  void remove(uint64_t k) {
    for (uint64_t i = h(k) & mask;; i = (i + 1) & mask) {
      if (slot[i] == k) {
        slot[i] = ~1ull;  // tombstone
        return;
      }
      if (slot[i] == ~0ull) return;
    }
  }
};

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}


// libc random() draws as the reference makes them: getRandom
// (general/common.c:44-47) and uniform_rand (general/sa.c:36-39).
inline int ref_random_below(int max) { return (int)(random() * ((double)max) / (1.0 + RAND_MAX)); }
inline double ref_uniform() { return ((double)random() + 1.0) / ((double)RAND_MAX + 2.0); }

// Moore bound of diameter and ASPL (lower_bound_of_diam_aspl,
// general/main.c:244-265).
inline void moore_bound(int nodes, int degree, int *low_diam, double *low_aspl) {
  int diam = -1, n = 1, r = 1;
  double aspl = 0.0;
  while (true) {
    const int tmp = n + degree * (int)std::pow(degree - 1, r - 1);
    if (tmp >= nodes) break;
    n = tmp;
    aspl += r * degree * std::pow(degree - 1, r - 1);
    diam = r++;
  }
  diam++;
  aspl += diam * (nodes - n);
  aspl /= (nodes - 1);
  *low_diam = diam;
  *low_aspl = aspl;
}

}  // namespace gg
#endif
