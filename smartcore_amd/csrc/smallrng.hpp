// Restatement of the random stream KMeans::kmeans_plus_plus draws
// (kmeans.rs:355-396 via rand_custom.rs:8-33): rand 0.8.5's SmallRng, which
// on 64-bit targets is xoshiro256++ seeded through SplitMix64, with
// `gen_range(0..n)` for usize (widening-multiply rejection sampling) and
// `gen::<f64>()` (53 high bits times 2^-53).  The rand crate is not part of
// the reference tree; this follows its published algorithms.
#pragma once

#include <cstdint>

namespace sc {

class SmallRng {
 public:
  // SeedableRng::seed_from_u64 of xoshiro256++: four SplitMix64 outputs.
  explicit SmallRng(uint64_t state) {
    constexpr uint64_t PHI = 0x9e3779b97f4a7c15ULL;
    for (int i = 0; i < 4; ++i) {
      state += PHI;
      uint64_t z = state;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      s_[i] = z ^ (z >> 31);
    }
    if ((s_[0] | s_[1] | s_[2] | s_[3]) == 0) *this = SmallRng(0);
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
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t s_[4];
};

}  // namespace sc
