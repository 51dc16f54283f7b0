// Workload generator: the Graph500-style R-MAT tuple stream of
// scripts/prepare_publication_datasets.py:80-129 (quadrant walk :94-111,
// A/B/C/D = .57/.19/.19/.05, no vertex permutation, self-loops and
// duplicates kept in the stream).
//
// The script draws from Python's random.Random(seed).random(); this file
// restates that generator so the tuples match the script's output exactly:
// MT19937 (Matsumoto & Nishimura 1998) seeded with init_by_array over the
// seed's 32-bit words, and random() = genrand_res53 (53-bit double from two
// draws, a >> 5 and b >> 6), as documented for CPython's _random module.
#include <cstdint>
#include <string>
#include <vector>

#include "vgx.h"

namespace vgx {
vgx_status set_error(vgx_status s, const std::string& msg);

namespace {

class Mt19937 {
 public:
  explicit Mt19937(uint64_t seed) {
    // CPython: the absolute value of an int seed, split into 32-bit words
    // (least significant first); zero becomes the single word 0.
    std::vector<uint32_t> key;
    if (seed == 0) key.push_back(0);
    while (seed) {
      key.push_back(static_cast<uint32_t>(seed & 0xFFFFFFFFu));
      seed >>= 32;
    }
    init_by_array(key);
  }

  uint32_t next_u32() {
    if (idx_ >= kN) twist();
    uint32_t y = mt_[idx_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }

  double random() {
    uint32_t a = next_u32() >> 5, b = next_u32() >> 6;
    return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
  }

 private:
  static constexpr int kN = 624, kM = 397;

  void init_genrand(uint32_t s) {
    mt_[0] = s;
    for (int i = 1; i < kN; ++i) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + static_cast<uint32_t>(i);
    idx_ = kN;
  }

  void init_by_array(const std::vector<uint32_t>& key) {
    init_genrand(19650218u);
    int i = 1, j = 0;
    const int len = static_cast<int>(key.size());
    for (int k = kN > len ? kN : len; k; --k) {
      mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1664525u)) + key[j] + static_cast<uint32_t>(j);
      ++i;
      ++j;
      if (i >= kN) {
        mt_[0] = mt_[kN - 1];
        i = 1;
      }
      if (j >= len) j = 0;
    }
    for (int k = kN - 1; k; --k) {
      mt_[i] = (mt_[i] ^ ((mt_[i - 1] ^ (mt_[i - 1] >> 30)) * 1566083941u)) - static_cast<uint32_t>(i);
      ++i;
      if (i >= kN) {
        mt_[0] = mt_[kN - 1];
        i = 1;
      }
    }
    mt_[0] = 0x80000000u;
    idx_ = kN;
  }

  static uint32_t mix(uint32_t a, uint32_t b, uint32_t c) {
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return c ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  }

  void twist() {
    int i = 0;
    for (; i < kN - kM; ++i) mt_[i] = mix(mt_[i], mt_[i + 1], mt_[i + kM]);
    for (; i < kN - 1; ++i) mt_[i] = mix(mt_[i], mt_[i + 1], mt_[i + kM - kN]);
    mt_[kN - 1] = mix(mt_[kN - 1], mt_[0], mt_[kM - 1]);
    idx_ = 0;
  }

  uint32_t mt_[kN];
  int idx_ = kN;
};

}  // namespace
}  // namespace vgx

extern "C" vgx_status vgx_rmat_generate(uint32_t scale, uint32_t edgefactor, uint64_t seed, uint32_t* src,
                                        uint32_t* dst) {
  if (scale == 0 || scale >= 32) return vgx::set_error(VGX_INVALID_ARGUMENT, "scale must be in [1, 31]");
  if (edgefactor == 0) return vgx::set_error(VGX_INVALID_ARGUMENT, "edgefactor must be positive");
  if (!src || !dst) return vgx::set_error(VGX_INVALID_ARGUMENT, "null output");
  vgx::Mt19937 rng(seed);
  const uint64_t n = 1ull << scale;
  const uint64_t m = static_cast<uint64_t>(edgefactor) * n;
  // Same left-to-right double sums as the script's probability tests.
  const double p0 = 0.57, p1 = 0.19, p2 = 0.19;
  const double t0 = p0, t1 = p0 + p1, t2 = p0 + p1 + p2;
  for (uint64_t e = 0; e < m; ++e) {
    uint64_t u = 0, v = 0;
    for (uint64_t step = n >> 1; step; step >>= 1) {
      // quadrant A: none, B (t0 <= r < t1): v, C (t1 <= r < t2): u, D: both;
      // written branch-free because the outcome is unpredictable.
      const double r = rng.random();
      const uint64_t bu = r >= t1;
      const uint64_t bv = static_cast<uint64_t>(r >= t0) & static_cast<uint64_t>(r < t1 || r >= t2);
      u += step & (0 - bu);
      v += step & (0 - bv);
    }
    src[e] = static_cast<uint32_t>(u);
    dst[e] = static_cast<uint32_t>(v);
  }
  return VGX_OK;
}
