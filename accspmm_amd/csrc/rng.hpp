// rng.hpp — xoshiro256** (Blackman & Vigna, "Scrambled linear pseudorandom
// number generators", 2018) seeded through splitmix64; identical streams on
// every x86-64 host.
#pragma once
#include <cstdint>

namespace accspmm {

inline uint64_t splitmix64(uint64_t* x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct Xoshiro256 {
  uint64_t s[4];
  explicit Xoshiro256(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s[i] = splitmix64(&x);
  }
  static inline uint64_t rotl(uint64_t v, int k) {
    return (v << k) | (v >> (64 - k));
  }
  inline uint64_t next() {
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  // U(0,1] on a 2^-24 grid: ((x >> 40) + 1) * 2^-24
  inline float pos_unit() {
    return (float)((next() >> 40) + 1) * (1.0f / 16777216.0f);
  }
  // U[-1,1) on a 2^-23 grid: (x >> 40) * 2^-23 - 1
  inline float sym_unit() {
    return (float)(next() >> 40) * (1.0f / 8388608.0f) - 1.0f;
  }
  // U(0,1) in double: ((x >> 11) + 0.5) * 2^-53
  inline double open_unit() {
    return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
};

}  // namespace accspmm
