// Host-side helpers shared by the library's C++ translation units:
// error reporting across the C-ABI, a std::thread parallel-for, the
// Marsaglia-Tsang gamma sampler on the counter-based stream, and the
// held-out/test membership hash layout shared with the device.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "philox.h"

namespace ammsb {

// ---- error reporting (thread-local message behind ammsb_last_error) -------
void set_error(const std::string& msg);
const char* get_error();

// ---- parallel for over [0, n) in contiguous chunks ------------------------
inline int host_threads() {
  unsigned hc = std::thread::hardware_concurrency();
  if (hc == 0) hc = 1;
  return (int)std::min<unsigned>(hc, 64);
}

template <typename Fn>
void parallel_for(int64_t n, Fn fn, int64_t min_chunk = 1024) {
  int nt = host_threads();
  if (n <= min_chunk || nt <= 1) {
    fn((int64_t)0, n, 0);
    return;
  }
  nt = (int)std::min<int64_t>(nt, (n + min_chunk - 1) / min_chunk);
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; ++t) {
    const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    th.emplace_back([=]() { fn(lo, hi, t); });
  }
  for (auto& x : th) x.join();
}

// ---- Gamma(shape, 1) on the counter-based stream ---------------------------
// Marsaglia & Tsang, "A simple method for generating gamma variables",
// ACM TOMS 26(3), 2000.  Attempt j uses Philox blocks (2j, 2j+1) of the
// stream (id, iter, tag); for shape < 1 the boost U^(1/shape) uses block
// 0xFFFFFFFF.  Returns log(G) so tiny shapes do not underflow.
inline double log_gamma_draw(Key key, double shape, uint32_t id, uint32_t iter,
                             uint32_t tag) {
  double boost_log = 0.0;
  double a = shape;
  if (a < 1.0) {
    const u32x4 r = draw4(key, 0xFFFFFFFFu, id, iter, tag);
    const double u = u01_from_u64((uint64_t)r.x | ((uint64_t)r.y << 32));
    boost_log = std::log(u) / a;
    a += 1.0;
  }
  const double d = a - 1.0 / 3.0;
  const double c = 1.0 / std::sqrt(9.0 * d);
  for (uint32_t j = 0;; ++j) {
    double x, unused;
    normal_pair(draw4(key, 2 * j, id, iter, tag), &x, &unused);
    double v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    const u32x4 r = draw4(key, 2 * j + 1, id, iter, tag);
    const double u = u01_from_u64((uint64_t)r.x | ((uint64_t)r.y << 32));
    const double x2 = x * x;
    if (u < 1.0 - 0.0331 * x2 * x2 ||
        std::log(u) < 0.5 * x2 + d * (1.0 - v + std::log(v))) {
      return std::log(d * v) + boost_log;
    }
  }
}

// ---- held-out ∪ test membership hash (device layout) -----------------------
// Open addressing, linear probing, power-of-two capacity, load <= 1/2.
// key = (uint64)a << 32 | b with a < b; empty slot = ~0.
// vals: bit0 held-out, bit1 test, bit2 label (link).
constexpr uint64_t kHashEmpty = ~0ull;
AMMSB_HD uint64_t edge_key(uint32_t a, uint32_t b) {
  return ((uint64_t)a << 32) | (uint64_t)b;
}
AMMSB_HD uint64_t mix64(uint64_t z) {
  // splitmix64 finaliser (Steele, Lea, Flood, OOPSLA 2014)
  z ^= z >> 30;
  z *= 0xBF58476D1CE4E5B9ull;
  z ^= z >> 27;
  z *= 0x94D049BB133111EBull;
  z ^= z >> 31;
  return z;
}

struct EdgeHashHost {
  std::vector<uint64_t> keys;
  std::vector<uint8_t> vals;
  uint64_t mask = 0;
  void init(int64_t count) {
    uint64_t cap = 1024;
    while (cap < (uint64_t)(2 * count + 1)) cap <<= 1;
    keys.assign(cap, kHashEmpty);
    vals.assign(cap, 0);
    mask = cap - 1;
  }
  // returns false if already present (vals OR-ed in either case)
  bool insert(uint64_t key, uint8_t v) {
    uint64_t s = mix64(key) & mask;
    while (true) {
      if (keys[s] == kHashEmpty) {
        keys[s] = key;
        vals[s] = v;
        return true;
      }
      if (keys[s] == key) {
        vals[s] |= v;
        return false;
      }
      s = (s + 1) & mask;
    }
  }
  int find(uint64_t key) const {
    uint64_t s = mix64(key) & mask;
    while (true) {
      if (keys[s] == kHashEmpty) return -1;
      if (keys[s] == key) return vals[s];
      s = (s + 1) & mask;
    }
  }
};

// Blocked Bloom filter over the held-out ∪ test keys, in front of the hash:
// one 64-bit word per key with two bits set in it, about 8 bits per key
// (false-positive rate about 5%), so that most membership tests of random
// pairs (all of which miss) read one word of a table a sixteenth of the
// hash's size.
AMMSB_HD uint64_t bloom_hash(uint64_t key) { return mix64(key ^ 0x9E3779B97F4A7C15ull); }
AMMSB_HD uint64_t bloom_bits(uint64_t h) {
  return (1ull << ((h >> 40) & 63)) | (1ull << ((h >> 50) & 63));
}
struct BloomHost {
  std::vector<uint64_t> words;
  uint64_t mask = 0;
  void init(int64_t count) {
    uint64_t n = 64;
    while (n * 8 < (uint64_t)count) n <<= 1;
    words.assign(n, 0);
    mask = n - 1;
  }
  void insert(uint64_t key) {
    const uint64_t h = bloom_hash(key);
    words[h & mask] |= bloom_bits(h);
  }
};

// Linked test on the symmetric CSR: binary search of b in row a.
inline bool csr_linked(const int64_t* row_ptr, const int32_t* cols, int32_t a,
                       int32_t b) {
  const int32_t* lo = cols + row_ptr[a];
  const int32_t* hi = cols + row_ptr[a + 1];
  const int32_t* it = std::lower_bound(lo, hi, b);
  return it != hi && *it == b;
}

}  // namespace ammsb
