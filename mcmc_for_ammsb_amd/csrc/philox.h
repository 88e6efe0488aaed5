// Counter-based random stream shared by the device samplers and the host-side
// pieces of the library (graph generator, held-out split, theta init).
//
// Philox4x32-10 as published in Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11: a 4x32-bit counter, a 2x32-bit key,
// ten rounds of two 32x32->64 multiplies with a Weyl-sequence key schedule.
//
// This replaces the reference's sequential xorshift128+ stream
// (c++/src/mcmc/random.cc:136-142). The reference's *semantics* are kept:
//   randint(0, n-1) = (64-bit draw) % n        (random.cc:174-176)
// but each draw is addressed by a counter instead of a position in a stream,
// so every draw of every iteration can be produced independently on the GPU
// and replayed bit-exactly on the host.
//
// Counter layout used everywhere in this library:
//   ctr = { idx, id, iter, tag }
//     idx  : draw index inside one (tag, id, iter) stream (attempt number,
//            element pair index, ...)
//     id   : node id (neighbour draws, phi noise) or 0 / source index
//     iter : step_count of the SGRLD iteration (1-based, as the reference)
//     tag  : which stream (AMMSB_TAG_*)
//   key = { low 32 bits of seed, high 32 bits of seed }
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AMMSB_HD __host__ __device__ __forceinline__
#else
#define AMMSB_HD inline
#endif

namespace ammsb {

enum : uint32_t {
  TAG_FLAG = 1,        // links / non-links coin of the stratified sampler
  TAG_SOURCE = 2,      // source node(s) of the stratified sampler
  TAG_NONLINK = 3,     // non-link partner candidates
  TAG_NEIGHBOR = 4,    // neighbour candidates of update_phi
  TAG_PHI_NOISE = 5,   // Gaussian noise of the phi step
  TAG_THETA_NOISE = 6, // Gaussian noise of the theta step
  TAG_INIT_PHI = 7,    // phi ~ Gamma(1,1) at init
  TAG_INIT_THETA = 8,  // theta ~ Gamma(eta0, eta1) at init
  TAG_SPLIT = 9,       // held-out / test split
  TAG_GEN = 10,        // synthetic graph generators
  TAG_RANDOM_EDGE = 11 // random-edge minibatch strategy
};

struct u32x4 {
  uint32_t x, y, z, w;
};

AMMSB_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIPCC__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

struct Key {
  uint32_t k0, k1;
};

AMMSB_HD Key make_key(uint64_t seed) {
  Key k;
  k.k0 = (uint32_t)seed;
  k.k1 = (uint32_t)(seed >> 32);
  return k;
}

AMMSB_HD u32x4 draw4(Key key, uint32_t idx, uint32_t id, uint32_t iter,
                     uint32_t tag) {
  u32x4 c;
  c.x = idx;
  c.y = id;
  c.z = iter;
  c.w = tag;
  return philox4x32_10(c, key.k0, key.k1);
}

// One 64-bit draw per counter: words (x, y) -> x | y << 32.
AMMSB_HD uint64_t draw_u64(Key key, uint32_t idx, uint32_t id, uint32_t iter,
                           uint32_t tag) {
  const u32x4 r = draw4(key, idx, id, iter, tag);
  return (uint64_t)r.x | ((uint64_t)r.y << 32);
}

// randint(0, n-1) with the reference's modulo semantics.
AMMSB_HD uint32_t randint_mod(uint64_t u, uint32_t n) {
  return (uint32_t)(u % (uint64_t)n);
}

// The same u % n without a 64-bit division (Barrett reduction with the
// precomputed r = floor((2^64-1)/n), n >= 1): q = floor(u*r / 2^64) is
// floor(u/n) or one less, because u*r/2^64 > u/n - (1 + (2^64-1) mod n)/n
// >= u/n - 1; so u - q*n < 2n and one conditional subtraction is exact.
AMMSB_HD uint64_t fastmod_recip(uint32_t n) { return ~0ull / (uint64_t)n; }
AMMSB_HD uint32_t randint_fast(uint64_t u, uint32_t n, uint64_t r) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t q = __umul64hi(u, r);
#else
  const uint64_t q = (uint64_t)(((unsigned __int128)u * r) >> 64);
#endif
  uint64_t rem = u - q * (uint64_t)n;
  if (rem >= (uint64_t)n) rem -= (uint64_t)n;
  return (uint32_t)rem;
}

// Uniform doubles in the open interval (0, 1): 53 high bits + 1/2 ulp.
AMMSB_HD double u01_from_u64(uint64_t u) {
  return ((double)(u >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Uniform floats in (0, 1): 24 high bits + 1/2 ulp.
AMMSB_HD float u01f_from_u32(uint32_t u) {
  return ((float)(u >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

}  // namespace ammsb

#include <math.h>

namespace ammsb {

// Box-Muller pair from one Philox block (Box & Muller 1958).
// FP64: u1 from words (x,y), u2 from words (z,w), 53-bit uniforms.
// FP32: u1 from word y, u2 from word w, 24-bit uniforms: these are the high
// bits of the FP64 uniforms of the same block, so both precisions see the
// same Gaussian up to float rounding.
// The angle is 2 pi u2 with u2 in (0, 1).  The device evaluates its sine and
// cosine as sincospi(2 u2): 2 u2 is exact, so there is no rounded angle and
// no general range reduction.  The host evaluates cos and sin of the rounded
// angle; the two differ by about an ulp of the angle (1e-16 absolute in
// FP64, 1e-7 in FP32).
AMMSB_HD void normal_pair(u32x4 r, double* n0, double* n1) {
  const double u1 = u01_from_u64((uint64_t)r.x | ((uint64_t)r.y << 32));
  const double u2 = u01_from_u64((uint64_t)r.z | ((uint64_t)r.w << 32));
  const double rad = sqrt(-2.0 * log(u1));
#if defined(__HIP_DEVICE_COMPILE__)
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  *n0 = rad * cs;
  *n1 = rad * sn;
#else
  const double th = 6.283185307179586476925286766559 * u2;
  *n0 = rad * cos(th);
  *n1 = rad * sin(th);
#endif
}

AMMSB_HD void normal_pair(u32x4 r, float* n0, float* n1) {
  const float u1 = u01f_from_u32(r.y);
  const float u2 = u01f_from_u32(r.w);
  const float rad = sqrtf(-2.0f * logf(u1));
#if defined(__HIP_DEVICE_COMPILE__)
  float sn, cs;
  sincospif(2.0f * u2, &sn, &cs);
  *n0 = rad * cs;
  *n1 = rad * sn;
#else
  const float th = 6.2831853071795864769f * u2;
  *n0 = rad * cosf(th);
  *n1 = rad * sinf(th);
#endif
}

}  // namespace ammsb
