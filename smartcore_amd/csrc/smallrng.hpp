// Restatement of the random stream KMeans::kmeans_plus_plus draws
// (kmeans.rs:355-396 via rand_custom.rs:8-33, default features): rand
// 0.8.5's SmallRng, which on 64-bit targets is xoshiro256++.  rand 0.8.5's
// SmallRng implements only from_seed and from_rng, so seed_from_u64 is
// rand_core 0.6's trait default: eight PCG32 (XSH-RR) outputs, each taken
// after advancing the LCG state (multiplier 6364136223846793005, increment
// 11634580027462260723), fill the 32-byte seed little-endian; xoshiro256++'s
// from_seed reads it as four little-endian u64 words and maps an all-zero
// seed to its own SplitMix64 seed_from_u64(0).  Draws: `gen_range(0..n)`
// for usize (widening-multiply rejection sampling) and `gen::<f64>()` (53
// high bits times 2^-53).  The rand crates are not part of the reference
// tree; this follows their published source.
#pragma once

#include <cstdint>

namespace sc {

class SmallRng {
 public:
  // SmallRng::seed_from_u64 (rand_core 0.6's default, then from_seed).
  explicit SmallRng(uint64_t seed) {
    constexpr uint64_t MUL = 6364136223846793005ULL;
    constexpr uint64_t INC = 11634580027462260723ULL;
    uint64_t st = seed;
    uint64_t words[4] = {0, 0, 0, 0};
    for (int c = 0; c < 8; ++c) {
      st = st * MUL + INC;
      const uint32_t xorshifted = static_cast<uint32_t>(((st >> 18) ^ st) >> 27);
      const uint32_t rot = static_cast<uint32_t>(st >> 59);
      const uint32_t x = (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
      // bytes 4c .. 4c+3 of the seed, little-endian, are word c/2's low or
      // high half
      words[c / 2] |= static_cast<uint64_t>(x) << (32 * (c % 2));
    }
    if ((words[0] | words[1] | words[2] | words[3]) == 0) {
      splitmix_fill(0);
    } else {
      for (int i = 0; i < 4; ++i) s_[i] = words[i];
    }
  }

  uint64_t next_u64() {
    const uint64_t result = rotl(s_[0] + s_[3], 23) + s_[0];
    const uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return result;
  }

  // Rng::gen_range(0..n) for usize: UniformInt::sample_single on the
  // inclusive range [0, n-1] with zone = (range << lz(range)) - 1.
  uint64_t gen_range(uint64_t n) {
    const uint64_t range = n;  // (n - 1) - 0 + 1
    if (range == 0) return next_u64();
    const uint64_t zone = (range << __builtin_clzll(range)) - 1;
    for (;;) {
      const uint64_t v = next_u64();
      const unsigned __int128 m = static_cast<unsigned __int128>(v) * range;
      const uint64_t hi = static_cast<uint64_t>(m >> 64), lo = static_cast<uint64_t>(m);
      if (lo <= zone) return hi;
    }
  }

  // Standard distribution for f64: 53 most significant bits * 2^-53.
  double gen_f64() {
    const double scale = 1.0 / static_cast<double>(uint64_t(1) << 53);
    return scale * static_cast<double>(next_u64() >> 11);
  }

 private:
  // xoshiro256++'s own seed_from_u64: four SplitMix64 outputs.
  void splitmix_fill(uint64_t state) {
    constexpr uint64_t PHI = 0x9e3779b97f4a7c15ULL;
    for (int i = 0; i < 4; ++i) {
      state += PHI;
      uint64_t z = state;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      s_[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t s_[4];
};

}  // namespace sc
