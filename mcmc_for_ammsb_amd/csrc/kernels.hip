// gfx950 kernels of the SGRLD a-MMSB iteration.
//
// Hot path of the reference, MCMCSamplerStochastic::run()
// (c++/src/mcmc/learning/mcmc_sampler_stochastic.cc:131-174), as device
// kernels over HBM-resident state:
//   k_minibatch   Network::sample_mini_batch, stratified random node
//                 (network.cc:274-421) + nodes_in_batch (:315-327)
//   k_neighbors   sample_neighbor_nodes (mcmc_sampler_stochastic.h:87-108)
//   k_update_phi  update_phi (mcmc_sampler_stochastic.cc:269-313) in the
//                 pi ++ phi_sum form of update_phi_node
//                 (mcmc_sampler_stochastic_distr.cc:1158-1227) with the
//                 update_pi renormalisation (pi_from_phi, :743-751) fused as
//                 its epilogue, written to a staging row
//   k_scatter     publishes the staged rows after all phi updates (the
//                 reference refreshes pi only after the node loop, :162-166)
//   k_beta        update_beta gradient (:200-240) as a block reduction
//   k_theta       theta step + beta = theta1/(theta0+theta1) (:242-266)
//   k_ppx         Learner::cal_perplexity (learner.cc:138-183) with
//                 cal_edge_likelihood (learner.h:100-139)
//
// Data layout: pi is [N][ld] with ld a multiple of 128 bytes and slot K of
// each row holding phi_sum (the reference's DKV record pi[0..K) ++ phi_sum).
// A row is processed by a "row group" of LPR lanes; lane lg of the group
// owns vectors c*LPR + lg (c < CH) of 16 bytes each, so one group reads a
// row with fully coalesced 16-byte loads.  Small K packs 64/LPR rows into
// one wavefront.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <limits.h>
#include <math.h>
#include <type_traits>

#include "host_util.h"
#include "kernels.h"

namespace ammsb {

namespace {

template <typename F>
struct VT;
template <>
struct VT<double> {
  typedef double2 type;
  static constexpr int V = 2;
};
template <>
struct VT<float> {
  typedef float4 type;
  static constexpr int V = 4;
};

template <typename F>
__device__ __forceinline__ void load_vec(const F* p, F* out);
template <>
__device__ __forceinline__ void load_vec<double>(const double* p, double* out) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  out[0] = v.x;
  out[1] = v.y;
}
template <>
__device__ __forceinline__ void load_vec<float>(const float* p, float* out) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
  out[3] = v.w;
}

template <typename F>
__device__ __forceinline__ void store_vec(F* p, const F* in);
template <>
__device__ __forceinline__ void store_vec<double>(double* p, const double* in) {
  *reinterpret_cast<double2*>(p) = make_double2(in[0], in[1]);
}
template <>
__device__ __forceinline__ void store_vec<float>(float* p, const float* in) {
  *reinterpret_cast<float4*>(p) = make_float4(in[0], in[1], in[2], in[3]);
}
typedef double st_d2 __attribute__((ext_vector_type(2)));
typedef float st_f4 __attribute__((ext_vector_type(4)));
// 16-byte vector types the nontemporal builtins accept
template <typename F>
struct NTV;
template <>
struct NTV<double> {
  typedef st_d2 type;
};
template <>
struct NTV<float> {
  typedef st_f4 type;
};

// Accesses through pointers that address global memory.  A pointer read
// from memory (the batched kernels' per-chain arguments) is generic to the
// compiler, which then issues flat instructions: they take 64-bit addresses
// and also count against the LDS/scalar counter, so a wait for one row
// becomes a wait for every load in flight.  Only for pointers into device
// memory (never LDS).
template <typename T>
using gptr = T __attribute__((address_space(1)))*;
template <typename T>
__device__ __forceinline__ T ld_g(const T* p) {
  return *(gptr<const T>)p;
}
template <typename F>
__device__ __forceinline__ void load_vec_g(const F* p, F* out) {
  typedef typename NTV<F>::type T;
  const T v = *(gptr<const T>)p;
#pragma unroll
  for (int i = 0; i < VT<F>::V; ++i) out[i] = v[i];
}

// 16-byte non-temporal store (p in device memory, 16-byte aligned).
template <typename F>
__device__ __forceinline__ void store_vec_nt(F* p, const F* in) {
  typedef typename NTV<F>::type T;
  T v;
#pragma unroll
  for (int i = 0; i < VT<F>::V; ++i) v[i] = in[i];
  __builtin_nontemporal_store(v, (gptr<T>)p);
}

// beta[k] in Float, 0 past the row (K >= 1).  The load is unconditional,
// from a clamped index: a load under `k < K ? beta[k] : 0` becomes a branch
// with its own wait, and a row's worth of them run one after another.
template <typename F>
__device__ __forceinline__ F beta_at(const double* beta, int k, int K) {
  const double b = ld_g(beta + (k < K ? k : K - 1));
  return k < K ? (F)b : (F)0;
}

// Streaming (non-temporal) 16-byte loads, for rows read once per use.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));
template <typename F>
__device__ __forceinline__ void load_vec_nt(const F* p, F* out);
template <>
__device__ __forceinline__ void load_vec_nt<double>(const double* p, double* out) {
  const nt_d2 v = __builtin_nontemporal_load((gptr<const nt_d2>)p);
  out[0] = v.x;
  out[1] = v.y;
}
template <>
__device__ __forceinline__ void load_vec_nt<float>(const float* p, float* out) {
  const nt_f4 v = __builtin_nontemporal_load((gptr<const nt_f4>)p);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
  out[3] = v.w;
}

// Whole rows read by one wavefront through a buffer resource: the row's
// address in scalar registers (it is the same for every lane), each lane's
// byte offsets in vector registers, computed once (4 bytes per vector
// instead of a 64-bit address per vector and row).  bytes bounds the
// offsets (the hardware returns 0 past it).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* row, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(row), (short)0, (int)bytes,
                                           0x00020000);
}
// 16 bytes at byte offset off; NT: a non-temporal load.
template <typename F, bool NT>
__device__ __forceinline__ void load_vec_buf(__amdgpu_buffer_rsrc_t r, uint32_t off, F* out) {
  typedef typename NTV<F>::type T;
  const T v = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, NT ? 2 : 0));
#pragma unroll
  for (int i = 0; i < VT<F>::V; ++i) out[i] = v[i];
}

// Row of K elements into registers: vector vi = c*LPR + lg.  The loads are
// unconditional so that every vector of a row (and of several rows) is in
// flight before the first wait: a vector past the end of the row re-reads
// vector 0 of the same row (in bounds) and is zeroed by mask_row().
template <typename F, int LPR, int CH, bool NT = false>
__device__ __forceinline__ void issue_row(const F* row, int K, int lg, F* out) {
  constexpr int V = VT<F>::V;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k0 = (c * LPR + lg) * V;
    if constexpr (NT)
      load_vec_nt<F>(row + (k0 < K ? k0 : 0), out + c * V);
    else
      load_vec_g<F>(row + (k0 < K ? k0 : 0), out + c * V);
  }
}
// Zero the elements k >= K (selects, no branches around memory operations).
template <typename F, int LPR, int CH>
__device__ __forceinline__ void mask_row(int K, int lg, F* out) {
  constexpr int V = VT<F>::V;
  if (CH * LPR * V <= K) return;  // every lane's vectors lie inside the row
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = (c * LPR + lg) * V + v;
      out[c * V + v] = k < K ? out[c * V + v] : (F)0;
    }
}
template <typename F, int LPR, int CH>
__device__ __forceinline__ void load_row(const F* row, int K, int lg, F* out) {
  issue_row<F, LPR, CH>(row, K, lg, out);
  mask_row<F, LPR, CH>(K, lg, out);
}

// 1/x: hardware reciprocal + two Newton-Raphson steps (FP64), or the
// hardware reciprocal + one step (FP32); error within an ulp or two, against
// the reference's exact divide.
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float recip(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  const float e = fmaf(-x, r, 1.0f);
  return fmaf(r, e, r);
}

// Cross-lane moves through DPP (data-parallel primitives of the VALU, no
// LDS round trip).  Controls: quad_perm [1,0,3,2] = 0xB1 (lane ^ 1),
// quad_perm [2,3,0,1] = 0x4E (lane ^ 2), row_half_mirror = 0x141 (lane i of
// an 8-lane half row <- lane 7-i), row_mirror = 0x140 (lane i of a 16-lane
// row <- lane 15-i).
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xFFFFFFFFll), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float read_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double read_lane(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xFFFFFFFFll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the LPR lanes of a row group (LPR in {4,8,16,32,64}), result in
// every lane of the group.  Butterfly inside 16-lane rows with DPP: after
// the ^1 and ^2 steps a quad is uniform, so the half-row mirror adds the
// other quad and the row mirror the other half row.  Rows are combined with
// readlane (uniform scalars).  The summation order is fixed.
template <int LPR, typename T>
__device__ __forceinline__ T group_sum(T v) {
  static_assert(LPR == 4 || LPR == 8 || LPR == 16 || LPR == 32 || LPR == 64, "row group");
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  if (LPR >= 8) v += dpp_mov<0x141>(v);
  if (LPR >= 16) v += dpp_mov<0x140>(v);
  if (LPR == 32) {
    const T r0 = read_lane(v, 0), r1 = read_lane(v, 16), r2 = read_lane(v, 32),
            r3 = read_lane(v, 48);
    v = ((int)(threadIdx.x & 63) < 32) ? r0 + r1 : r2 + r3;
  } else if (LPR == 64) {
    const T r0 = read_lane(v, 0), r1 = read_lane(v, 16), r2 = read_lane(v, 32),
            r3 = read_lane(v, 48);
    v = (r0 + r1) + (r2 + r3);
  }
  return v;
}

// Sum over the lpr lanes of a row group with lpr a uniform runtime value
// (the same butterfly and order as group_sum<LPR>).
template <typename T>
__device__ __forceinline__ T group_sum_rt(T v, int lpr) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  if (lpr >= 8) v += dpp_mov<0x141>(v);
  if (lpr >= 16) v += dpp_mov<0x140>(v);
  if (lpr >= 32) {
    const T r0 = read_lane(v, 0), r1 = read_lane(v, 16), r2 = read_lane(v, 32),
            r3 = read_lane(v, 48);
    if (lpr == 32)
      v = ((int)(threadIdx.x & 63) < 32) ? r0 + r1 : r2 + r3;
    else
      v = (r0 + r1) + (r2 + r3);
  }
  return v;
}

// Sum over the 64/LPR row groups of a wavefront (lanes with equal lg).
template <int LPR, typename T>
__device__ __forceinline__ T cross_group_sum(T v) {
#pragma unroll
  for (int off = 32; off >= LPR; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double eps_of_step(const DevHyper& hp, int64_t t) {
  // get_eps_t (mcmc_sampler_stochastic.h:112-115)
  return hp.a * pow(1.0 + (double)t / hp.b, -hp.c);
}

__device__ __forceinline__ int hash_find(const DevGraph& g, uint32_t lo, uint32_t hi) {
  const uint64_t key = edge_key(lo, hi);
  uint64_t s = mix64(key) & g.hmask;
  while (true) {
    const uint64_t k = ld_g(g.hkeys + s);
    if (k == key) return ld_g(g.hvals + s);
    if (k == kHashEmpty) return -1;
    s = (s + 1) & g.hmask;
  }
}

// (lo, hi) in held-out ∪ test: the Bloom word first (nearly every random
// pair stops there), the hash only for possible members.
__device__ __forceinline__ bool pair_held_out(const DevGraph& g, uint32_t lo, uint32_t hi) {
  const uint64_t h = bloom_hash(edge_key(lo, hi));
  const uint64_t bits = bloom_bits(h);
  if ((ld_g(g.bloom + (h & g.bloom_mask)) & bits) != bits) return false;
  return hash_find(g, lo, hi) >= 0;
}

// v in the sorted row (binary search); G: the row is in device memory
template <bool G = false>
__device__ __forceinline__ bool row_contains(const int32_t* row, int64_t n, int32_t v) {
  auto at = [&](int64_t i) { return G ? ld_g(row + i) : row[i]; };
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (at(mid) < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && at(lo) == v;
}

// De-dup table: entry = (value, smallest attempt index that drew it).
__device__ __forceinline__ int table_insert(int2* tab, uint32_t cap_mask, int32_t v) {
  uint32_t s = (uint32_t)mix64((uint64_t)(uint32_t)v) & cap_mask;
  while (true) {
    const int prev = atomicCAS(&tab[s].x, -1, v);
    if (prev == -1 || prev == v) return (int)s;
    s = (s + 1) & cap_mask;
  }
}

// The same table in device memory, as ints (entry i at 2i, the value, and
// 2i + 1, the smallest attempt index that drew it; global atomics): insert
// v, lower its attempt to ai, return the slot.
__device__ __forceinline__ int table_insert_min_g(gptr<int> tab, uint32_t cap_mask, int32_t v,
                                                  int ai) {
  uint32_t s = (uint32_t)mix64((uint64_t)(uint32_t)v) & cap_mask;
  while (true) {
    int prev = -1;
    __hip_atomic_compare_exchange_strong(tab + 2 * s, &prev, v, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == -1 || prev == v) break;
    s = (s + 1) & cap_mask;
  }
  __hip_atomic_fetch_min(tab + 2 * s + 1, ai, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return (int)s;
}
// The attempt index recorded in slot s of a device-memory table.
__device__ __forceinline__ int table_attempt_g(gptr<int> tab, int s) {
  return __hip_atomic_load(tab + 2 * s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Hash set of node ids (-1 = empty): true when v was not in the set.
__device__ __forceinline__ bool set_insert(int* set, uint32_t cap_mask, int32_t v) {
  uint32_t s = (uint32_t)mix64((uint64_t)(uint32_t)v) & cap_mask;
  while (true) {
    const int prev = atomicCAS(&set[s], -1, v);
    if (prev == -1) return true;
    if (prev == v) return false;
    s = (s + 1) & cap_mask;
  }
}

// De-dup table of 64-bit pair keys (random-edge minibatches).
__device__ __forceinline__ int table_insert64(unsigned long long* keys, uint32_t cap_mask,
                                              unsigned long long k) {
  uint32_t s = (uint32_t)mix64(k) & cap_mask;
  while (true) {
    const unsigned long long prev = atomicCAS(&keys[s], ~0ull, k);
    if (prev == ~0ull || prev == k) return (int)s;
    s = (s + 1) & cap_mask;
  }
}

__device__ __forceinline__ int load_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Ordering of a de-dup table (or minibatch buffer) private to one workgroup,
// in LDS or in global memory: a workgroup-scope fence.  The waves of a
// workgroup share one CU and its vector L1, so the memory model needs no
// cache maintenance at this scope; an agent-scope fence would also write
// back the whole L2 of the XCD (buffer_wbl2) on every call.
__device__ __forceinline__ void wg_fence() { __threadfence_block(); }

// Exclusive block scan of a 0/1 predicate (blockDim multiple of 64).
__device__ __forceinline__ int block_scan(int pred, int* s_wave, int* total) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nw = blockDim.x >> 6;
  const uint64_t m = __ballot(pred);
  const int within = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) s_wave[w] = __popcll(m);
  __syncthreads();
  int off = 0, tot = 0;
  for (int i = 0; i < nw; ++i) {
    const int c = s_wave[i];
    off += (i < w) ? c : 0;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return off + within;
}

constexpr int kPhiNbLds = 2048;  // neighbour ids staged in LDS by k_update_phi
constexpr int kMbThreads = 1024;
constexpr int kMbWaves = kMbThreads / 64;
constexpr int kMaxSources = 1024;
constexpr int kRowLds = 4096;
constexpr int kSrcTab = 2 * kMaxSources;  // LDS set of link sources (power of two)

// First source index of node v in the source set, INT_MAX if v is no source.
__device__ __forceinline__ int src_first(const int2* tab, int32_t v) {
  uint32_t s = (uint32_t)mix64((uint64_t)(uint32_t)v) & (kSrcTab - 1);
  while (true) {
    const int2 e = tab[s];
    if (e.x == v) return e.y;
    if (e.x == -1) return INT_MAX;
    s = (s + 1) & (kSrcTab - 1);
  }
}

// ---------------------------------------------------------------------------
// Minibatch sampler: one workgroup.  Sequential definition (the host replica
// implements exactly this):
//   flag = u64(FLAG; 0,0,t) % 2                          [stratified only]
//   non-links: node = u64(SOURCE; 0,0,t) % N; attempt j = 0,1,...:
//       v = u64(NONLINK; j,0,t) % N, accepted iff v != node, (node,v) not a
//       link, not held-out/test, and v not accepted before; until m
//       accepted.  scale = N*N/m * E/(N(N-1)/2 - E)       (network.cc:341-342)
//   links: round r = 0,1,...: for s < max_source: node = u64(SOURCE; r,s,t)%N;
//       stop if s > 0 and |S| + deg(node) > m; all training edges of node in
//       CSR order not already in S; repeat the round while S is empty.
//       scale = N if max_source == 1 else E/|S|           (network.cc:385-390)
// Non-link attempts are processed 1024 at a time; within a batch a value is
// won by its smallest attempt index (atomicMin in the de-dup table, kept in
// LDS when it fits), and accepted values are appended in attempt order
// through a block scan, which yields the sequential result.
// Link sources are processed 16 at a time, one wavefront each: an edge of
// source s is new iff its other end is not a source of index < s (and s is
// the first occurrence of its node); a single thread then advances |S| over
// the 16 counts and applies the stop rule, and the kept edges are written at
// their sequential offsets.
// Iteration t into ring slot `slot`; set_base: this slot is slot 0 of a
// fill starting at t (DevState::ring_base = t).
__device__ __forceinline__ void minibatch_body(const SamplerArgs& a, const int slot,
                                               const int64_t t, const bool set_base) {
  extern __shared__ int2 s_dyn[];  // non-link de-dup table when table_in_lds
  __shared__ int s_wave[kMbWaves];
  __shared__ int s_row[kRowLds];    // non-links: row of the node; links: degrees
  __shared__ int s_src[kMaxSources];
  __shared__ int s_kept[kMaxSources];
  __shared__ int s_cb[kMaxSources];
  __shared__ int2 s_stab[kSrcTab];
  __shared__ int s_nodes, s_total, s_used, s_err;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  DevState* st = a.st;
  const uint32_t t32 = (uint32_t)t;
  const Ring& R = a.ring;
  // the ring's arrays, the graph's and the HBM tables are device memory:
  // global accesses (a batch reads these pointers from an argument array)
  const gptr<int32_t> mb_edges = (gptr<int32_t>)(R.edges + (int64_t)slot * R.edge_cap * 2);
  const gptr<uint8_t> mb_y = (gptr<uint8_t>)(R.y + (int64_t)slot * R.edge_cap);
  const gptr<int32_t> out_nodes = (gptr<int32_t>)(R.nodes + (int64_t)slot * R.node_cap);
  const gptr<uint32_t> stamp = (gptr<uint32_t>)(R.stamp + (int64_t)slot * R.N);
  const int64_t edge_cap = R.edge_cap, node_cap = R.node_cap;
  const uint32_t N = (uint32_t)a.g.N;
  const int64_t m = a.m;
  int err = 0;
  int flag;
  if (a.strategy == 0)
    flag = (int)(draw_u64(a.key, 0, 0, t32, TAG_FLAG) % 2ull);
  else if (a.strategy == 1)
    flag = 1;
  else if (a.strategy == 2)
    flag = 0;
  else if (a.strategy == 3)
    flag = 2;
  else
    flag = -1;
  int64_t c = 0;
  double scale = 0.0;
  if (tid == 0) {
    s_nodes = 0;
    s_err = 0;
  }
  if (flag == 0) {
    const int32_t node =
        (int32_t)randint_fast(draw_u64(a.key, 0, 0, t32, TAG_SOURCE), N, a.g.n_recip);
    const int64_t rs = ld_g(a.g.row_ptr + node);
    const int64_t deg = ld_g(a.g.row_ptr + node + 1) - rs;
    const bool in_lds = deg <= kRowLds;
    if (in_lds)
      for (int64_t i = tid; i < deg; i += kMbThreads) s_row[i] = ld_g(a.g.cols + rs + i);
    const uint32_t cmask = (uint32_t)(a.table_cap - 1);
    const int64_t max_batches = 4096 + 64 * ((m + kMbThreads - 1) / kMbThreads);
    // the de-dup table in LDS (LDS atomics) or in HBM (global atomics;
    // entry i at ints 2i, the value, and 2i + 1, the attempt)
    auto nonlinks = [&](auto in_lds_table) {
      constexpr bool L = decltype(in_lds_table)::value;
      typedef int gi2 __attribute__((ext_vector_type(2)));
      int2* const ltab = s_dyn;
      const gptr<int> gtab = (gptr<int>)(a.table + (int64_t)slot * a.table_cap);
      for (int64_t i = tid; i < a.table_cap; i += kMbThreads) {
        if constexpr (L)
          ltab[i] = make_int2(-1, INT_MAX);
        else
          ((gptr<gi2>)gtab)[i] = gi2{-1, INT_MAX};
      }
      wg_fence();
      __syncthreads();
      for (int64_t batch = 0;; ++batch) {
        if (batch >= max_batches) {
          err |= ERR_NONLINK_ATTEMPTS;
          break;
        }
        const uint32_t ai = (uint32_t)(batch * kMbThreads + tid);
        const int32_t v =
            (int32_t)randint_fast(draw_u64(a.key, ai, 0, t32, TAG_NONLINK), N, a.g.n_recip);
        bool valid = v != node;
        const uint32_t lo = (uint32_t)min(node, v), hi = (uint32_t)max(node, v);
        if (valid) valid = !(in_lds ? row_contains(s_row, deg, v) : row_contains<true>(a.g.cols + rs, deg, v));
        if (valid) valid = !pair_held_out(a.g, lo, hi);
        int ts = -1;
        if (valid) {
          if constexpr (L) {
            ts = table_insert(ltab, cmask, v);
            atomicMin(&ltab[ts].y, (int)ai);
          } else {
            ts = table_insert_min_g(gtab, cmask, v, (int)ai);
          }
        }
        wg_fence();
        __syncthreads();
        int win = 0;
        if (valid) {
          if constexpr (L)
            win = ltab[ts].y == (int)ai;
          else
            win = table_attempt_g(gtab, ts) == (int)ai;
        }
        int total;
        const int pos = block_scan(win, s_wave, &total);
        if (win && c + pos < m) {
          mb_edges[2 * (c + pos)] = (int32_t)lo;
          mb_edges[2 * (c + pos) + 1] = (int32_t)hi;
          mb_y[c + pos] = 0;
          // node set: the source, then the (distinct) partners in order
          if (1 + c + pos < node_cap) out_nodes[1 + c + pos] = v;
        }
        c = min(m, c + (int64_t)total);
        if (c >= m) break;
      }
    };
    if (a.table_in_lds)
      nonlinks(std::integral_constant<bool, true>());
    else
      nonlinks(std::integral_constant<bool, false>());
    if (tid == 0) {
      if (c > 0) out_nodes[0] = node;
      s_nodes = c > 0 ? (int)c + 1 : 0;
    }
    const double dN = (double)a.g.N, dE = (double)a.g.E;
    scale = dN * dN / (double)m;
    scale *= dE / (dN * (dN - 1.0) / 2.0 - dE);
  } else if (flag == 2) {
    // random edge (network.cc:237-268).  Full training set when m >=
    // E - |held-out| - |test| (the reference's entry counts): the training
    // links a<b in CSR order.  Else attempt j: q = u64(RANDOM_EDGE; j,0,t) %
    // 2E, the CSR entry q (row esrc[q], column cols[q]) as (min, max),
    // accepted iff not held-out/test and not accepted before, until m.
    const int64_t E2 = 2 * a.g.E;
    if (m >= a.g.E - a.held_test) {
      for (int64_t q0 = 0; q0 < E2; q0 += kMbThreads) {
        const int64_t q = q0 + tid;
        bool keep = false;
        int32_t u = 0, v = 0;
        if (q < E2) {
          u = ld_g(a.g.esrc + q);
          v = ld_g(a.g.cols + q);
          keep = u < v && !pair_held_out(a.g, (uint32_t)u, (uint32_t)v);
        }
        int total;
        const int pos = block_scan(keep, s_wave, &total);
        if (keep && c + pos < edge_cap) {
          mb_edges[2 * (c + pos)] = u;
          mb_edges[2 * (c + pos) + 1] = v;
          mb_y[c + pos] = 1;
        }
        if (c + total > edge_cap) err |= ERR_MB_OVERFLOW;
        c = min(edge_cap, c + (int64_t)total);
      }
    } else {
      const gptr<unsigned long long> keys =
          (gptr<unsigned long long>)(a.re_keys + (int64_t)slot * a.re_cap);
      const gptr<int32_t> att = (gptr<int32_t>)(a.re_att + (int64_t)slot * a.re_cap);
      for (int64_t i = tid; i < a.re_cap; i += kMbThreads) {
        keys[i] = ~0ull;
        att[i] = INT_MAX;
      }
      wg_fence();
      __syncthreads();
      const uint32_t cmask = (uint32_t)(a.re_cap - 1);
      const int64_t max_batches = 4096 + 64 * ((m + kMbThreads - 1) / kMbThreads);
      for (int64_t batch = 0;; ++batch) {
        if (batch >= max_batches) {
          err |= ERR_NONLINK_ATTEMPTS;
          break;
        }
        const uint32_t ai = (uint32_t)(batch * kMbThreads + tid);
        const int64_t q = (int64_t)randint_fast(draw_u64(a.key, ai, 0, t32, TAG_RANDOM_EDGE),
                                                (uint32_t)E2, a.g.e2_recip);
        const int32_t u = ld_g(a.g.esrc + q), v = ld_g(a.g.cols + q);
        const uint32_t lo = (uint32_t)min(u, v), hi = (uint32_t)max(u, v);
        const bool valid = !pair_held_out(a.g, lo, hi);
        int ts = -1;
        if (valid) {
          const unsigned long long key = (unsigned long long)edge_key(lo, hi);
          uint32_t h = (uint32_t)mix64(key) & cmask;
          while (true) {
            unsigned long long prev = ~0ull;
            __hip_atomic_compare_exchange_strong(keys + h, &prev, key, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (prev == ~0ull || prev == key) break;
            h = (h + 1) & cmask;
          }
          ts = (int)h;
          __hip_atomic_fetch_min(att + ts, (int)ai, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        wg_fence();
        __syncthreads();
        const int win = valid && __hip_atomic_load(att + ts, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT) == (int)ai;
        int total;
        const int pos = block_scan(win, s_wave, &total);
        if (win && c + pos < m) {
          mb_edges[2 * (c + pos)] = (int32_t)lo;
          mb_edges[2 * (c + pos) + 1] = (int32_t)hi;
          mb_y[c + pos] = 1;
        }
        c = min(m, c + (int64_t)total);
        if (c >= m) break;
      }
    }
    const double dN = (double)a.g.N;
    scale = (dN - 1.0) * dN / 2.0 / (double)(c > 0 ? c : 1);  // network.cc:262
  } else if (flag == 1) {
    int* s_deg = s_row;
    const int nsrc = (int)min<int64_t>(a.max_source, kMaxSources);
    int64_t r = 0;
    const int64_t max_rounds = 1 << 20;
    for (; r < max_rounds; ++r) {
      for (int i = tid; i < kSrcTab; i += kMbThreads) s_stab[i] = make_int2(-1, INT_MAX);
      __syncthreads();
      for (int s = tid; s < nsrc; s += kMbThreads) {
        const int32_t node =
            (int32_t)randint_fast(draw_u64(a.key, (uint32_t)r, (uint32_t)s, t32, TAG_SOURCE), N,
                                  a.g.n_recip);
        s_src[s] = node;
        s_deg[s] = (int)(ld_g(a.g.row_ptr + node + 1) - ld_g(a.g.row_ptr + node));
        const int slot = table_insert(s_stab, kSrcTab - 1, node);
        atomicMin(&s_stab[slot].y, s);
      }
      if (tid == 0) {
        s_total = 0;
        s_used = -1;
      }
      __syncthreads();
      int used = -1;
      c = 0;
      for (int base = 0; base < nsrc && used < 0; base += kMbWaves) {
        const int s = base + w;
        if (s < nsrc) {
          const int32_t node = s_src[s];
          int kept = 0;
          if (src_first(s_stab, node) == s) {
            const int64_t rs = ld_g(a.g.row_ptr + node);
            const int deg = s_deg[s];
            for (int j0 = 0; j0 < deg; j0 += 64) {
              const int j = j0 + lane;
              bool keep = false;
              if (j < deg) {
                const int32_t v = ld_g(a.g.cols + rs + j);
                keep = src_first(s_stab, v) >= s &&
                       !pair_held_out(a.g, (uint32_t)min(node, v), (uint32_t)max(node, v));
              }
              kept += __popcll(__ballot(keep));
            }
          }
          if (lane == 0) s_kept[s] = kept;
        }
        __syncthreads();
        if (tid == 0) {
          int64_t cc = s_total;
          int u = -1;
          const int hi = min(base + kMbWaves, nsrc);
          for (int q = base; q < hi; ++q) {
            if (q > 0 && cc + s_deg[q] > m) {
              u = q;
              break;
            }
            if (cc + s_kept[q] > edge_cap) {
              s_err |= ERR_MB_OVERFLOW;
              u = q;
              break;
            }
            s_cb[q] = (int)cc;
            cc += s_kept[q];
          }
          s_total = (int)cc;
          s_used = u;
        }
        __syncthreads();
        c = s_total;
        used = s_used;
      }
      if (used < 0) used = nsrc;
      if (c > 0) {
        for (int s = w; s < used; s += kMbWaves) {
          const int32_t node = s_src[s];
          if (src_first(s_stab, node) != s) continue;
          const int64_t rs = ld_g(a.g.row_ptr + node);
          const int deg = s_deg[s];
          int off = s_cb[s];
          for (int j0 = 0; j0 < deg; j0 += 64) {
            const int j = j0 + lane;
            bool keep = false;
            int32_t v = 0;
            if (j < deg) {
              v = ld_g(a.g.cols + rs + j);
              keep = src_first(s_stab, v) >= s &&
                     !pair_held_out(a.g, (uint32_t)min(node, v), (uint32_t)max(node, v));
            }
            const uint64_t mk = __ballot(keep);
            if (keep) {
              const int pos = off + __popcll(mk & ((1ull << lane) - 1ull));
              mb_edges[2 * pos] = min(node, v);
              mb_edges[2 * pos + 1] = max(node, v);
              mb_y[pos] = 1;
            }
            off += __popcll(mk);
          }
        }
        break;
      }
      __syncthreads();
    }
    if (r >= max_rounds) err |= ERR_LINK_ROUNDS;
    scale = a.max_source == 1 ? (double)a.g.N : (double)a.g.E / (double)(c > 0 ? c : 1);
  } else {
    err |= ERR_BAD_STRATEGY;
  }
  wg_fence();
  __syncthreads();
  // nodes_in_batch: unique endpoints (mcmc_sampler_stochastic.cc:315-327).
  // Order is not part of the contract: every per-node draw is keyed by the
  // node id, not by its position.  Non-link minibatches wrote theirs while
  // sampling; for the others a node is new when it enters an LDS hash set
  // (the non-link table's space, at load <= 3/4), else when it swaps the
  // slot's stamp generation into its global stamp.
  const uint32_t epoch = R.meta[slot].epoch + 1u;
  const int64_t set_ints = 2 * a.table_cap;
  const bool lds_set = a.table_in_lds && 3 * set_ints >= 8 * c;
  int* const nset = reinterpret_cast<int*>(s_dyn);
  if (flag != 0 && lds_set) {
    for (int64_t i = tid; i < set_ints; i += kMbThreads) nset[i] = -1;
    __syncthreads();
  }
  for (int64_t e0 = 0; flag != 0 && e0 < c; e0 += kMbThreads) {
    const int64_t e = e0 + tid;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      int32_t v = 0;
      bool isnew = false;
      if (e < c) {
        v = mb_edges[2 * e + q];
        isnew = lds_set ? set_insert(nset, (uint32_t)(set_ints - 1), v)
                        : __hip_atomic_exchange(stamp + v, epoch, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT) != epoch;
      }
      // one LDS atomic per wavefront, positions by ballot prefix
      const uint64_t mk = __ballot(isnew);
      int base = 0;
      if (lane == 0 && mk) base = atomicAdd(&s_nodes, __popcll(mk));
      base = __shfl(base, 0, 64);
      const int pos = base + __popcll(mk & ((1ull << lane) - 1ull));
      if (isnew && pos < node_cap) out_nodes[pos] = v;
    }
  }
  __syncthreads();
  if (tid == 0) {
    err |= s_err;
    if (s_nodes > node_cap) err |= ERR_MB_OVERFLOW;
    IterMeta mt;
    mt.step = t;
    mt.flag = flag;
    mt.mb_count = (int32_t)c;
    mt.node_count = (int32_t)min<int64_t>(s_nodes, node_cap);
    mt.epoch = epoch;
    mt.scale = scale;
    R.meta[slot] = mt;
    if (err) atomicOr(&st->error, err);
    if (set_base) st->ring_base = t;
  }
}

__global__ __launch_bounds__(kMbThreads) void k_minibatch(SamplerArgs a) {
  const int64_t t0 = a.t0 >= 0 ? a.t0 : a.st->step;
  minibatch_body(a, blockIdx.x, t0 + blockIdx.x, blockIdx.x == 0);
}

// ---------------------------------------------------------------------------
// Neighbour sampler: one wavefront per minibatch node.  Sequential
// definition: attempt j = 0,1,...: v = u64(NEIGHBOR; j,node,t) % N, accepted
// iff v != node, (node,v) not held-out/test, v not accepted before; until
// n+1 accepted (mcmc_sampler_stochastic.h:93-107: n+1 draws, links allowed).
// Output order = acceptance order; bit 31 flags a link (the y_ab of
// update_phi, mcmc_sampler_stochastic.cc:279-283).
// LDS: the de-dup table (and the node's row) in LDS, else the table in
// device memory; either way private to the wavefront and ordered by a
// workgroup-scope fence.  The accesses name their memory (LDS or global)
// so that none is a flat access (a pointer read from a batch's argument
// array is generic to the compiler).
template <int WAVES, bool LDS>
__device__ __forceinline__ void neighbors_body(const NeighborArgs& a, int32_t slot_offset,
                                               int32_t slot_end, int32_t rslot0, int bx, int by) {
  extern __shared__ int2 s_tab[];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int rslot = rslot0 + by;  // ring slot (iteration)
  const int slot = slot_offset + bx * WAVES + w;  // node slot
  const IterMeta* mt = a.ring.meta + rslot;
  // the node id is read together with the slot's count (its index clamped
  // into the slot), not after the count test: one dependent load less
  const int32_t node_any =
      ld_g(a.ring.nodes + (int64_t)rslot * a.ring.node_cap + min(slot, (int32_t)a.ring.node_cap - 1));
  const int count = ld_g(&mt->node_count);
  if (slot >= count || slot >= slot_end) return;
  const uint32_t t32 = (uint32_t)ld_g(&mt->step);
  const uint32_t N = (uint32_t)a.g.N;
  const int32_t node = node_any;
  // LDS per wavefront: the de-dup table, then room for the node's row
  int2* const ltab = s_tab + (size_t)w * (a.tcap + kNbrRowLds / 2);
  // in device memory: entry i at ints 2i (value) and 2i + 1 (attempt)
  typedef int gi2 __attribute__((ext_vector_type(2)));
  const gptr<int> gtab = (gptr<int>)(a.gtable + (size_t)(slot - slot_offset) * a.tcap);
  for (int i = lane; i < a.tcap; i += 64) {
    if constexpr (LDS)
      ltab[i] = make_int2(-1, INT_MAX);
    else
      ((gptr<gi2>)gtab)[i] = gi2{-1, INT_MAX};
  }
  wg_fence();
  __builtin_amdgcn_wave_barrier();
  const uint32_t cmask = (uint32_t)(a.tcap - 1);
  const int64_t rs = ld_g(a.g.row_ptr + node);
  const int64_t deg = ld_g(a.g.row_ptr + node + 1) - rs;
  const int nnb = a.ring.nnb;
  const gptr<int32_t> out =
      (gptr<int32_t>)(a.ring.nbrs + ((int64_t)rslot * a.ring.node_cap + slot) * nnb);
  // The node's adjacency row in LDS (up to kNbrRowLds entries) for the
  // link test of the accepted draws; longer rows are searched in HBM.
  const bool row_in_lds = LDS && deg <= kNbrRowLds;
  int32_t* const srow = reinterpret_cast<int32_t*>(ltab + a.tcap);
  if (row_in_lds) {
    for (int64_t q = lane; q < deg; q += 64) srow[q] = ld_g(a.g.cols + rs + q);
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
  }
  int c = 0;
  const int max_batches = 4096 + 64 * ((nnb + 63) / 64);
  for (int batch = 0; c < nnb; ++batch) {
    if (batch >= max_batches) {
      if (lane == 0) atomicOr(&a.st->error, ERR_NEIGHBOR_ATTEMPTS);
      break;
    }
    const uint32_t ai = (uint32_t)(batch * 64 + lane);
    const int32_t v = (int32_t)randint_fast(draw_u64(a.key, ai, (uint32_t)node, t32, TAG_NEIGHBOR), N,
                                         a.g.n_recip);
    bool valid = v != node;
    if (valid) valid = !pair_held_out(a.g, (uint32_t)min(node, v), (uint32_t)max(node, v));
    int s = -1;
    if (valid) {
      if constexpr (LDS) {
        s = table_insert(ltab, cmask, v);
        atomicMin(&ltab[s].y, (int)ai);
      } else {
        s = table_insert_min_g(gtab, cmask, v, (int)ai);
      }
    }
    wg_fence();
    __builtin_amdgcn_wave_barrier();
    bool win = false;
    if (valid) {
      if constexpr (LDS)
        win = ltab[s].y == (int)ai;
      else
        win = table_attempt_g(gtab, s) == (int)ai;
    }
    const uint64_t mk = __ballot(win);
    const int pos = c + __popcll(mk & ((1ull << lane) - 1ull));
    if (win && pos < nnb) {
      const bool lk = row_in_lds ? row_contains(srow, deg, v) : row_contains<true>(a.g.cols + rs, deg, v);
      out[pos] = v | (lk ? (int32_t)0x80000000 : 0);
    }
    c += __popcll(mk);
  }
}

template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_neighbors(NeighborArgs a, int32_t slot_offset,
                                                          int32_t slot_end, int32_t rslot0) {
  if (a.use_lds)
    neighbors_body<WAVES, true>(a, slot_offset, slot_end, rslot0, blockIdx.x, blockIdx.y);
  else
    neighbors_body<WAVES, false>(a, slot_offset, slot_end, rslot0, blockIdx.x, blockIdx.y);
}

// ---------------------------------------------------------------------------
// update_phi + fused update_pi (one workgroup per minibatch node).
// For node i with neighbours b (label y_b):
//   e = eps or 1-eps, f_k = beta_k - eps or eps - beta_k
//   x_bk = pi_bk f_k + e,  Z_b = sum_k pi_ik x_bk
//   grads_k = sum_b (x_bk / Z_b - 1) / phi_sum
//     (= ((probs_k/Z_b)/pi_ik - 1)/phi_sum with probs_k = pi_ik x_bk)
//   phi_k = pi_ik phi_sum
//   phi'_k = |phi_k + eps_t/2 (alpha - phi_k + N/n grads_k)
//            + sqrt(eps_t phi_k) xi_k|,  floored at 1e-24
//   phi_sum' = sum_k phi'_k (double), pi'_k = phi'_k / phi_sum'
// Waves split the neighbours; each row group streams whole neighbour rows,
// reduces Z_b across its lanes and accumulates x_bk/Z_b - 1 per lane; the
// waves' partial gradients meet in LDS.
// phi step of one node (mcmc_sampler_stochastic_distr.cc:1207-1226) from the
// row group's accumulated gradient sums acc (x_bk/Z_b - 1 summed over the
// neighbours), written with its renormalised pi and phi_sum to stg (K+1).
// LPR = 0: the row group has lpr_rt lanes (a uniform runtime value).  t is
// the step (the noise counter) and eps_t_d its step size (phi_step_size),
// read by the caller before its row loads so that their latency is hidden.
template <typename F, int LPR, int CH>
__device__ __forceinline__ void phi_epilogue(const PhiArgs& a, int slot, int32_t node,
                                             const F phi_sum, const F* pii, const F* acc, int lg,
                                             F* stg, int64_t t, double eps_t_d, int lpr_rt = 0) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CH;
  const int L = LPR > 0 ? LPR : lpr_rt;
  const int K = a.K;
  const F eps_t = (F)eps_t_d;
  const F half_eps = eps_t / (F)2;
  const F alpha = (F)a.hp.alpha;
  const F Nn = (F)a.hp.Nn;
  const F inv_phi_sum = recip(phi_sum);
  const F* noise = (const F*)a.noise;
  F phin[EPL];
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
#pragma unroll
    for (int v = 0; v < V; v += 2) {
      const int k = (c * L + lg) * V + v;
      F xi0 = (F)0, xi1 = (F)0;
      if (k < K) {
        if (noise) {
          xi0 = noise[(int64_t)slot * K + k];
          if (k + 1 < K) xi1 = noise[(int64_t)slot * K + k + 1];
        } else {
          normal_pair(draw4(a.key, (uint32_t)(k >> 1), (uint32_t)node, (uint32_t)t, TAG_PHI_NOISE),
                      &xi0, &xi1);
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = c * V + v + h;
        const F phik = pii[j] * phi_sum;
        const F grad = acc[j] * inv_phi_sum;
        F val = fabs(phik + half_eps * (alpha - phik + Nn * grad) + sqrt(eps_t * phik) * (h ? xi1 : xi0));
        if (val < (F)1e-24) val = (F)1e-24;
        if (k + h >= K) val = (F)0;
        phin[j] = val;
        s += (double)val;
      }
    }
  }
  if constexpr (LPR > 0)
    s = group_sum<LPR>(s);
  else
    s = group_sum_rt(s, lpr_rt);
  const double inv_s = 1.0 / s;
  // one non-temporal 16-byte store per vector (the row is 16-byte aligned;
  // streaming stores cost the concurrent row gathers of other workgroups
  // less: N 4M K 1024 FP32 update_phi 383 -> 373 us); a vector that reaches past the
  // row is stored element by element, so that phi_sum at K is written once
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k0 = (c * L + lg) * V;
    F out[V];
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = (F)((double)phin[c * V + v] * inv_s);
    if (k0 + V <= K) {
      store_vec_nt<F>(stg + k0, out);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (k0 + v < K) stg[k0 + v] = out[v];
    }
  }
  if (lg == 0) stg[K] = (F)s;
}

// The iteration's step size: the injected one, else the one computed once
// per step on the device.
__device__ __forceinline__ double phi_step_size(const PhiArgs& a) {
  return a.eps_override >= 0.0 ? a.eps_override : a.st->eps_t;
}

// The phi step is applied here and the new row staged; k_scatter publishes
// it once every gather of the iteration is done.
template <typename F, int LPR, int CH, int WAVES, int U>
__global__ __launch_bounds__(WAVES * 64) void k_update_phi(PhiArgs a) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CH;
  constexpr int RPW = 64 / LPR;
  constexpr int STRIDE = WAVES * RPW;
  // The neighbour ids staged in LDS (more than 64*WAVES neighbours) during
  // the row loop, then the wavefronts' partial sums (EPL per lane of a row
  // group and the scalar sum C): one buffer, a barrier between the uses.
  constexpr int kAccWords = WAVES * (EPL + 1) * LPR * (int)sizeof(F) / 4;
  __shared__ int32_t s_buf[kAccWords > kPhiNbLds ? kAccWords : kPhiNbLds];
  int32_t* const s_nb = s_buf;
  F* const s_acc = reinterpret_cast<F*>(s_buf);
  __shared__ double s_sum[WAVES];
  __shared__ F s_pii[EPL * LPR];  // the node's pi row, for the phi step
  __shared__ F s_f[EPL * LPR];    // f_k = beta_k - eps (0 past the row)
  const int slot = blockIdx.x;
  int count;
  const int32_t* nodes;
  const int32_t* nbrs;
  if (a.use_ring) {
    const int rs = ring_slot_of(a.ring_slot_host, a.st);
    count = a.ring.meta[rs].node_count;
    nodes = a.ring.nodes + (int64_t)rs * a.ring.node_cap;
    nbrs = a.ring.nbrs + (int64_t)rs * a.ring.node_cap * a.nnb;
  } else {
    count = a.count_host;
    nodes = a.nodes;
    nbrs = a.nbrs;
  }
  if (slot >= count) return;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int g = lane / LPR, lg = lane % LPR;
  const F* pi = (const F*)a.pi;
  const int K = a.K;
  const int32_t node = nodes[slot];
  const F* prow = pi + (int64_t)node * a.ld;
  const F eps = (F)a.hp.epsilon;
  const int32_t* nb_src = nbrs + (int64_t)slot * a.nnb;
  // Up to 64*WAVES neighbours (uniform test): each wavefront holds its row
  // ids in one register, lane j the id of row w*RPW + j%RPW + STRIDE*(j/RPW),
  // read together with the node's own row and phi_sum (no LDS staging and no
  // barrier before the first neighbour row is requested).
  const bool reg_ids = a.nnb <= 64 * WAVES;
  int32_t ids = node;
  if (reg_ids) {
    const int r = w * RPW + lane % RPW + STRIDE * (lane / RPW);
    if (r < a.nnb) ids = nb_src[r];
  }
  const F phi_sum = prow[K];
  const int64_t t = a.st->step;
  const double eps_t = phi_step_size(a);

  // The gradient is regrouped so that the row loop holds three arrays:
  //   x_bk = s_b f_k pi_bk + e_b  (link: s = 1, e = eps; non-link: s = -1,
  //   e = 1 - eps; f_k = beta_k - eps),
  //   Z_b = sum_k pi_ak x_bk = s_b P_b + e_b S,  P_b = sum_k q_k pi_bk,
  //   q_k = pi_ak f_k,  S = sum_k pi_ak,
  //   grads_k = sum_b (x_bk / Z_b - 1) = f_k acc_k + C,
  //   acc_k = sum_b s_b pi_bk / Z_b,  C = sum_b (e_b / Z_b - 1),
  // the same sums as update_phi_node's (mcmc_sampler_stochastic_distr.cc:
  // 1184-1205) in another order.
  F q[EPL], acc[EPL];
  F Sp = (F)0;
  {
    F pii[EPL];
    load_row<F, LPR, CH>(prow, K, lg, pii);
    // f_k into LDS, coalesced, element k at (k / V) % LPR * V + ... as the
    // lanes read it: vector c*LPR+lg holds k = (c*LPR+lg)*V .. +V-1
    for (int k = tid; k < EPL * LPR; k += WAVES * 64)
      s_f[k] = beta_at<F>(a.beta, k, K) - (k < K ? eps : (F)0);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      F f[V];
      load_vec<F>(s_f + (c * LPR + lg) * V, f);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int j = c * V + v;
        q[j] = pii[j] * f[v];
        Sp += pii[j];
        acc[j] = (F)0;
      }
    }
    if (w == 0 && g == 0) {
#pragma unroll
      for (int c = 0; c < CH; ++c) store_vec<F>(s_pii + (c * LPR + lg) * V, pii + c * V);
    }
  }
  const F S = group_sum<LPR>(Sp);
  const F e1 = eps, e0 = (F)1 - eps;
  const F zs1 = e1 * S, zs0 = e0 * S;
  F C = (F)0;
  // Otherwise the node's neighbour ids (bit 31 = link) are staged in LDS,
  // chunk by chunk.  The U ids of an iteration are read before any of its
  // rows is requested, so that the U rows' loads are all in flight together.
  // The trip count is uniform across the wavefront (row group 0's), so every
  // lane takes part in the cross-lane id reads.
  for (int cb = 0; cb < a.nnb; cb += kPhiNbLds) {
    const int cn = min(kPhiNbLds, a.nnb - cb);
    if (!reg_ids) {
      __syncthreads();
      for (int i = tid; i < cn; i += WAVES * 64) s_nb[i] = nb_src[cb + i];
      __syncthreads();
    }
    for (int it = 0; w * RPW + it * STRIDE < cn; it += U) {
      const int b0 = w * RPW + g + it * STRIDE;
      F row[U][EPL];
      int32_t nbv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int b = b0 + u * STRIDE;
        int32_t v;
        if (reg_ids) {
          if constexpr (RPW == 1)
            v = __builtin_amdgcn_readlane(ids, min(it + u, 63));
          else
            v = __shfl(ids, ((it + u) * RPW + g) & 63, 64);
        } else {
          v = b < cn ? s_nb[b] : node;
        }
        nbv[u] = b < cn ? v : node;  // past the end: own row, not counted
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (LPR == 64) nbv[u] = __builtin_amdgcn_readfirstlane(nbv[u]);  // one row per wavefront
        issue_row<F, LPR, CH, true>(pi + (int64_t)(nbv[u] & 0x7fffffff) * a.ld, K, lg, row[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        // q is zero past the row, so the loads past it need no mask here;
        // acc past the row is never used
        const bool ok = (b0 + u * STRIDE < cn) && ((nbv[u] & 0x7fffffff) != node);
        F p = (F)0;
#pragma unroll
        for (int j = 0; j < EPL; ++j) p = fma(q[j], row[u][j], p);
        p = group_sum<LPR>(p);
        const bool link = nbv[u] < 0;
        const F wz = recip(link ? p + zs1 : zs0 - p);
        if (ok) {
          const F sw = link ? wz : -wz;
#pragma unroll
          for (int j = 0; j < EPL; ++j) acc[j] = fma(row[u][j], sw, acc[j]);
          C += fma(link ? e1 : e0, wz, (F)-1);
        }
      }
    }
  }
  if (RPW > 1) {
#pragma unroll
    for (int j = 0; j < EPL; ++j) acc[j] = cross_group_sum<LPR>(acc[j]);
    C = cross_group_sum<LPR>(C);
  }
  if (!reg_ids) __syncthreads();  // s_nb's last reads before s_acc is written
  F* stg = (F*)a.staging + (int64_t)slot * a.ld;
  if constexpr (LPR == 64 && WAVES > 1) {
    // The phi step shared by the wavefronts: wavefront w takes the row's
    // vector chunks c = w, w + WAVES, ... (its lanes' registers then hold a
    // fraction of the row, not all of it).  The sums over wavefronts are
    // added in wavefront order, the renormalisation sum per wavefront and
    // then over wavefronts in order.
#pragma unroll
    for (int j = 0; j < EPL; ++j) s_acc[(w * (EPL + 1) + j) * 64 + lane] = acc[j];
    s_acc[(w * (EPL + 1) + EPL) * 64 + lane] = C;
    __syncthreads();
    F Ct = (F)0;
#pragma unroll
    for (int q2 = 0; q2 < WAVES; ++q2) Ct += s_acc[(q2 * (EPL + 1) + EPL) * 64 + lane];
    constexpr int CPW = (CH + WAVES - 1) / WAVES;  // chunks per wavefront
    const F eps_tf = (F)eps_t;
    const F half_eps = eps_tf / (F)2;
    const F alpha = (F)a.hp.alpha;
    const F Nn = (F)a.hp.Nn;
    const F inv_phi_sum = recip(phi_sum);
    const F* noise = (const F*)a.noise;
    F phin[CPW * V];
    double ssum = 0.0;
#pragma unroll
    for (int cc = 0; cc < CPW; ++cc) {
      const int c = w + cc * WAVES;
#pragma unroll
      for (int v = 0; v < V; ++v) phin[cc * V + v] = (F)0;
      if (c >= CH) continue;  // uniform over the wavefront
      const int k0 = (c * 64 + lane) * V;
      F f[V], pv[V];
      load_vec<F>(s_f + k0, f);
      load_vec<F>(s_pii + k0, pv);
#pragma unroll
      for (int v = 0; v < V; v += 2) {
        const int k = k0 + v;
        F xi0 = (F)0, xi1 = (F)0;
        if (k < K) {
          if (noise) {
            xi0 = noise[(int64_t)slot * K + k];
            if (k + 1 < K) xi1 = noise[(int64_t)slot * K + k + 1];
          } else {
            normal_pair(draw4(a.key, (uint32_t)(k >> 1), (uint32_t)node, (uint32_t)t, TAG_PHI_NOISE),
                        &xi0, &xi1);
          }
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          F accv = (F)0;
#pragma unroll
          for (int q2 = 0; q2 < WAVES; ++q2)
            accv += s_acc[(q2 * (EPL + 1) + c * V + v + h) * 64 + lane];
          const F grad = fma(f[v + h], accv, Ct) * inv_phi_sum;
          const F phik = pv[v + h] * phi_sum;
          F val = fabs(phik + half_eps * (alpha - phik + Nn * grad) +
                       sqrt(eps_tf * phik) * (h ? xi1 : xi0));
          if (val < (F)1e-24) val = (F)1e-24;
          if (k + h >= K) val = (F)0;
          phin[cc * V + v + h] = val;
          ssum += (double)val;
        }
      }
    }
    ssum = group_sum<64>(ssum);
    if (lane == 0) s_sum[w] = ssum;
    __syncthreads();
    double stot = 0.0;
#pragma unroll
    for (int q2 = 0; q2 < WAVES; ++q2) stot += s_sum[q2];
    const double inv_s = 1.0 / stot;
#pragma unroll
    for (int cc = 0; cc < CPW; ++cc) {
      const int c = w + cc * WAVES;
      if (c >= CH) continue;
      const int k0 = (c * 64 + lane) * V;
      F out[V];
#pragma unroll
      for (int v = 0; v < V; ++v) out[v] = (F)((double)phin[cc * V + v] * inv_s);
      if (k0 + V <= K) {
        store_vec_nt<F>(stg + k0, out);
      } else {
#pragma unroll
        for (int v = 0; v < V; ++v)
          if (k0 + v < K) stg[k0 + v] = out[v];
      }
    }
    if (w == 0 && lane == 0) stg[K] = (F)stot;
  } else {
    if (WAVES > 1) {
      if (w > 0 && g == 0) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) s_acc[((w - 1) * (EPL + 1) + j) * LPR + lg] = acc[j];
        s_acc[((w - 1) * (EPL + 1) + EPL) * LPR + lg] = C;
      }
      __syncthreads();
      if (w != 0) return;
#pragma unroll
      for (int q2 = 1; q2 < WAVES; ++q2) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) acc[j] += s_acc[((q2 - 1) * (EPL + 1) + j) * LPR + lg];
        C += s_acc[((q2 - 1) * (EPL + 1) + EPL) * LPR + lg];
      }
    }
    if (g != 0) return;
    // grads_k = f_k acc_k + C, and the node's pi row, for the phi step
    F pii[EPL];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      F f[V];
      load_vec<F>(s_f + (c * LPR + lg) * V, f);
      load_vec<F>(s_pii + (c * LPR + lg) * V, pii + c * V);
#pragma unroll
      for (int v = 0; v < V; ++v) acc[c * V + v] = fma(f[v], acc[c * V + v], C);
    }
    phi_epilogue<F, LPR, CH>(a, slot, node, phi_sum, pii, acc, lg, stg, t, eps_t);
  }
}

// Publish staged rows: pi[nodes[slot]] = staging[slot] (K+1 elements).
// One wavefront per staged row, four rows per workgroup.
template <typename F>
__global__ __launch_bounds__(256) void k_scatter(ScatterArgs a) {
  int count;
  const int32_t* nodes;
  if (a.use_ring) {
    const int rs = ring_slot_of(a.ring_slot_host, a.st);
    count = a.ring.meta[rs].node_count;
    nodes = a.ring.nodes + (int64_t)rs * a.ring.node_cap;
  } else {
    count = a.count_host;
    nodes = a.nodes;
  }
  const int slot = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (slot >= count) return;
  constexpr int V = VT<F>::V;
  const int nv = (a.K + 1 + V - 1) / V;  // row and phi_sum; ld is a multiple of 128 B
  typedef typename NTV<F>::type T;
  const T* src = reinterpret_cast<const T*>((const F*)a.staging + (int64_t)slot * a.ld);
  T* dst = reinterpret_cast<T*>((F*)a.pi + (int64_t)nodes[slot] * a.ld);
  // non-temporal stores: the rows are read again only by the next iteration
  // (N 4M, K 1024, FP32: the next update_phi 373 -> 368 us against plain stores)
  for (int i = lane; i < nv; i += 64) __builtin_nontemporal_store(src[i], dst + i);
}

// ---------------------------------------------------------------------------
// update_beta gradient (mcmc_sampler_stochastic.cc:212-240): per edge
//   f_k = pi_ak pi_bk, probs_k = beta_k f_k (y=1) or (1-beta_k) f_k (y=0)
//   prob_0 = (eps or 1-eps) (1 - sum_k f_k), Z = sum_k probs_k + prob_0
//   G_k0 += probs_k/Z ((1-y)/theta_k0 - 1/S_k)
//   G_k1 += probs_k/Z (y/theta_k1 - 1/S_k),   S_k = theta_k0 + theta_k1
// accumulated in double; one partial [2][K] per workgroup.
template <typename F, int LPR, int CH, int WAVES>
__device__ __forceinline__ void beta_body(const BetaArgs& a, const int bx) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CH;
  constexpr int RPW = 64 / LPR;
  __shared__ double s_g[(WAVES > 1 ? WAVES - 1 : 1) * EPL * LPR];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int g = lane / LPR, lg = lane % LPR;
  const int K = a.K;
  const F* pi = (const F*)a.pi;
  const F eps = (F)a.epsilon;
  int count;
  const int32_t* edges;
  const uint8_t* ys;
  if (a.use_ring) {
    const int rs = ring_slot_of(a.ring_slot_host, a.st);
    count = a.ring.meta[rs].mb_count;
    edges = a.ring.edges + (int64_t)rs * a.ring.edge_cap * 2;
    ys = a.ring.y + (int64_t)rs * a.ring.edge_cap;
  } else {
    count = a.count_host;
    edges = a.edges;
    ys = a.y;
  }
  // g1[j] / g0[j]: sum of probs_k/Z over the edges with y = 1 / y = 0; the
  // theta factors ((1-y)/theta_k0 - 1/S_k, y/theta_k1 - 1/S_k) depend on k
  // and y only, and are applied once when the partial row is written.
  F bt[EPL];
  double g0[EPL], g1[EPL];
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = (c * LPR + lg) * V + v;
      const int j = c * V + v;
      bt[j] = beta_at<F>(a.beta, k, K);
      g0[j] = 0.0;
      g1[j] = 0.0;
    }
  // Several edges per row group in flight: the edges' ids are read, then
  // all their rows are requested, then the edges are reduced.
  // edges in flight per row group: fewer for big rows (registers; the
  // launch has about one wavefront per SIMD, so occupancy is not the limit)
  constexpr int ROWREG = EPL * (int)sizeof(F) / 4;  // VGPRs of one row per lane
  constexpr int UE = ROWREG <= 16 ? 4 : (ROWREG < 32 ? 2 : 1);
  const int gstride = a.blocks * WAVES * RPW;
  for (int e0 = (bx * WAVES + w) * RPW + g; e0 < count; e0 += UE * gstride) {
    int32_t ea[UE], eb[UE];
    int yv[UE];
    bool val[UE];
#pragma unroll
    for (int u = 0; u < UE; ++u) {
      const int e = e0 + u * gstride;
      val[u] = e < count;
      const int ee = val[u] ? e : e0;
      ea[u] = ld_g(edges + 2 * ee);
      eb[u] = ld_g(edges + 2 * ee + 1);
      yv[u] = ld_g(ys + ee) ? 1 : 0;
    }
    F ra[UE][EPL], rb[UE][EPL];
#pragma unroll
    for (int u = 0; u < UE; ++u) {
      if (LPR == 64) {
        ea[u] = __builtin_amdgcn_readfirstlane(ea[u]);
        eb[u] = __builtin_amdgcn_readfirstlane(eb[u]);
        yv[u] = __builtin_amdgcn_readfirstlane(yv[u]);
      }
      issue_row<F, LPR, CH>(pi + (int64_t)ea[u] * a.ld, K, lg, ra[u]);
      issue_row<F, LPR, CH>(pi + (int64_t)eb[u] * a.ld, K, lg, rb[u]);
    }
#pragma unroll
    for (int u = 0; u < UE; ++u) {
      if (!val[u]) continue;  // uniform within the row group
      mask_row<F, LPR, CH>(K, lg, ra[u]);
      mask_row<F, LPR, CH>(K, lg, rb[u]);
      const int y = yv[u];
      F pisum = (F)0, psum = (F)0;
#pragma unroll
      for (int j = 0; j < EPL; ++j) {
        const F f = ra[u][j] * rb[u][j];
        pisum += f;
        ra[u][j] = y ? bt[j] * f : ((F)1 - bt[j]) * f;  // probs_k
        psum += ra[u][j];
      }
      pisum = group_sum<LPR>(pisum);
      psum = group_sum<LPR>(psum);
      const F prob0 = (y ? eps : (F)1 - eps) * ((F)1 - pisum);
      const F rz = recip(psum + prob0);
      if (y) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) g1[j] += (double)(ra[u][j] * rz);
      } else {
#pragma unroll
        for (int j = 0; j < EPL; ++j) g0[j] += (double)(ra[u][j] * rz);
      }
    }
  }
  if (RPW > 1) {
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
      g0[j] = cross_group_sum<LPR>(g0[j]);
      g1[j] = cross_group_sum<LPR>(g1[j]);
    }
  }
  // wavefront partials meet in LDS (g0 then g1 through one buffer): one
  // partial row [2K] per workgroup
  if (WAVES > 1) {
    if (w > 0 && g == 0) {
#pragma unroll
      for (int j = 0; j < EPL; ++j) s_g[((w - 1) * EPL + j) * LPR + lg] = g0[j];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int q = 1; q < WAVES; ++q)
#pragma unroll
        for (int j = 0; j < EPL; ++j) g0[j] += s_g[((q - 1) * EPL + j) * LPR + lg];
    }
    __syncthreads();
    if (w > 0 && g == 0) {
#pragma unroll
      for (int j = 0; j < EPL; ++j) s_g[((w - 1) * EPL + j) * LPR + lg] = g1[j];
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int q = 1; q < WAVES; ++q)
#pragma unroll
        for (int j = 0; j < EPL; ++j) g1[j] += s_g[((q - 1) * EPL + j) * LPR + lg];
    }
  }
  if (w != 0 || g != 0) return;
  double* out = a.partials + (int64_t)bx * 2 * K;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = (c * LPR + lg) * V + v;
      if (k < K) {
        const int j = c * V + v;
        const double th0 = a.theta[2 * k], th1 = a.theta[2 * k + 1];
        const double inv_s = 1.0 / (th0 + th1);
        // G_k0 = sum_{y=1} f (0/th0 - 1/S) + sum_{y=0} f (1/th0 - 1/S)
        // G_k1 = sum_{y=1} f (1/th1 - 1/S) + sum_{y=0} f (0/th1 - 1/S)
        out[k] = g1[j] * (0.0 / th0 - inv_s) + g0[j] * (1.0 / th0 - inv_s);
        out[K + k] = g1[j] * (1.0 / th1 - inv_s) + g0[j] * (0.0 / th1 - inv_s);
      }
    }
}

template <typename F, int LPR, int CH, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_beta(BetaArgs a) {
  beta_body<F, LPR, CH, WAVES>(a, blockIdx.x);
}

// theta step (mcmc_sampler_stochastic.cc:245-266).  Workgroup b owns
// k in [kThetaK*b, kThetaK*(b+1)); its 1024/kThetaK lane rows sum strided
// partial rows, then the row sums meet in LDS by a pairwise tree in a fixed
// order (deterministic).  The last workgroup to finish advances step_count,
// so every workgroup reads the same step.
constexpr int kThetaThreads = 256;  // 1024 measured slower: fewer resident workgroups
constexpr int kThetaK = 16;                        // k per workgroup
constexpr int kThetaRows = kThetaThreads / kThetaK;  // partial-row strides (power of two)
__device__ __forceinline__ void theta_body(const ThetaArgs& a, const int bx, const int gx) {
  __shared__ double s_g[kThetaRows][2][kThetaK];
  DevState* st = a.st;
  const int64_t t = st->step;
  const int tid = threadIdx.x, kl = tid % kThetaK, rg = tid / kThetaK;
  const int K = a.K;
  const int k = bx * kThetaK + kl;
  double G0 = 0.0, G1 = 0.0;
  if (k < K) {
#pragma unroll 8
    for (int b = rg; b < a.blocks; b += kThetaRows) {
      G0 += ld_g(a.partials + (int64_t)b * 2 * K + k);
      G1 += ld_g(a.partials + (int64_t)b * 2 * K + K + k);
    }
  }
  s_g[rg][0][kl] = G0;
  s_g[rg][1][kl] = G1;
  __syncthreads();
#pragma unroll
  for (int h = kThetaRows / 2; h >= 1; h >>= 1) {
    if (rg < h) {
      s_g[rg][0][kl] += s_g[rg + h][0][kl];
      s_g[rg][1][kl] += s_g[rg + h][1][kl];
    }
    __syncthreads();
  }
  if (rg == 0 && k < K) {
    G0 = s_g[0][0][kl];
    G1 = s_g[0][1][kl];
    const double eps_t = a.eps_override >= 0.0 ? a.eps_override : st->eps_t;
    const double scale =
        a.scale_override >= 0.0 ? a.scale_override : a.meta[ring_slot(st)].scale;
    double xi0, xi1;
    if (a.noise) {
      xi0 = a.noise[2 * k];
      xi1 = a.noise[2 * k + 1];
    } else {
      normal_pair(draw4(a.key, (uint32_t)k, 0, (uint32_t)t, TAG_THETA_NOISE), &xi0, &xi1);
    }
    const double G[2] = {G0, G1};
    const double xi[2] = {xi0, xi1};
    const double eta[2] = {a.hp.eta0, a.hp.eta1};
    double th[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double th_old = a.theta[2 * k + i];
      const double f = sqrt(eps_t * th_old);
      double v = fabs(th_old + eps_t / 2.0 * (eta[i] - th_old + scale * G[i]) + f * xi[i]);
      if (v < 1e-24) v = 1e-24;
      th[i] = v;
      a.theta[2 * k + i] = v;
    }
    a.beta[k] = th[1] / (th[0] + th[1]);
  }
  // The last workgroup to count itself advances the step.  The counter is a
  // device-scope atomic; theta, beta and the state are read by later kernels
  // only, so no fence is needed (an agent-scope one writes back the XCD's L2
  // in every workgroup: 36 -> 13 us for the theta step of 64 chains).
  if (a.bump_step) {
    __syncthreads();
    if (tid == 0) {
      if (atomicAdd(&st->done_count, 1) == gx - 1) {
        const IterMeta& mt = a.meta[ring_slot(st)];
        st->nodes_total += mt.node_count;
        st->edges_total += mt.mb_count;
        st->done_count = 0;
        st->eps_t = eps_of_step(a.hp, t + 1);
        st->step = t + 1;
      }
    }
  }
}

__global__ __launch_bounds__(kThetaThreads) void k_theta(ThetaArgs a) {
  theta_body(a, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Held-out perplexity partials (learner.cc:144-173): per edge the
// likelihood of cal_edge_likelihood (learner.h:100-139) floored at 1e-30,
// running mean R_e = (R_e (c-1) + L)/c, log R_e summed per label in double.
template <typename F, int LPR, int CH, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_ppx(PpxArgs a) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CH;
  constexpr int RPW = 64 / LPR;
  __shared__ double s_p[WAVES][4];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int g = lane / LPR, lg = lane % LPR;
  const int K = a.K;
  const F* pi = (const F*)a.pi;
  const F eps = (F)a.epsilon;
  F bt[EPL];
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = (c * LPR + lg) * V + v;
      bt[c * V + v] = beta_at<F>(a.beta, k, K);
    }
  double ll1 = 0.0, ll0 = 0.0, n1 = 0.0, n0 = 0.0;
  const double cnt = a.avg_count;
  // The held-out pairs are one node sequence, position t the pair
  // (seq[t], seq[t + 1]) (a bridge position between two trails is no
  // pair).  Each workgroup takes a contiguous range of positions and each
  // of its wavefronts a contiguous part of that range.  The ids, labels and
  // running averages of a wavefront's next 64 positions are read at once
  // (lane j: position w_lo + i0 RPW + j), so requesting a pair's rows never
  // waits for a load, and the new running averages are kept in the same
  // layout and written together after those positions (a store in the loop
  // would be waited for with the next load).
  //
  // One row group per wavefront: the wavefront walks its part of the
  // sequence UE positions at a time, requesting the rows seq[t + 1 .. t +
  // UE] into slots 1..UE, while slot 0 holds seq[t] (the last row of the
  // previous batch): every row of the sequence is requested once, and pair
  // u is slots (u, u + 1).  UE = 8 up to 4 vectors per lane, else 2 (N 4M,
  // K 1024, FP64, 8 vectors: 12.3 ms per evaluation, 12.2 with UE = 4 at one
  // wavefront per SIMD, 13.9 requesting both rows of each pair; above 8
  // vectors the slots cost a wavefront per SIMD and still gain: N 400k,
  // FP32 K 2560 1.55 against 1.86 ms, FP64 K 1200 1.49 against 1.68 ms).
  // Several row groups per wavefront (short rows): row group g takes
  // positions w_lo + g + i RPW, both rows of each requested (the row shared
  // with the previous position is found in L2), two pairs in flight.
  constexpr bool kSeq = LPR == 64;
  constexpr int UE = kSeq ? (CH <= 4 ? 8 : 2) : 2;
  constexpr int64_t kAlign = (int64_t)WAVES * RPW;
  constexpr int IPL = 64 / RPW;  // positions per row group and preload
  const int64_t chunk = ((a.H + a.blocks - 1) / a.blocks + kAlign - 1) / kAlign * kAlign;
  const int64_t wchunk = chunk / WAVES;  // a multiple of RPW
  const int64_t e_lo = (int64_t)blockIdx.x * chunk;
  const int64_t e_hi = min(a.H, e_lo + chunk);
  const int64_t w_lo = e_lo + w * wchunk;
  const int64_t w_hi = min(e_hi, w_lo + wchunk);
  auto take_i = [&](int32_t v, int src) -> int32_t {
    if constexpr (RPW == 1)
      return __builtin_amdgcn_readlane(v, src);
    else
      return __shfl(v, src, 64);
  };
  double rnew = 0.0;  // lane j: the new average of its preloaded position
  // the likelihood of one pair from its two (masked) rows, added to the
  // per-label sums; returns the new running average (held by every lane of
  // the row group)
  auto pair_ll = [&](const F (&ra)[EPL], const F (&rb)[EPL], const bool y,
                     const double rold) -> double {
    F s = (F)0, sum = (F)0;
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
      const F f = ra[j] * rb[j];
      s += y ? f * bt[j] : f * ((F)1 - bt[j]);
      sum += f;
    }
    s = group_sum<LPR>(s);
    sum = group_sum<LPR>(sum);
    if (!y) s += ((F)1 - sum) * ((F)1 - eps);
    if (s < (F)1e-30) s = (F)1e-30;
    // every lane of the row group holds s and rold, hence r
    const double r = (rold * (cnt - 1.0) + (double)s) / cnt;
    if (lg == 0) {
      const double lr = log(r);
      if (y) {
        ll1 += lr;
        n1 += 1.0;
      } else {
        ll0 += lr;
        n0 += 1.0;
      }
    }
    return r;
  };
  F rs[kSeq ? UE + 1 : 1][EPL];  // row slots of the sequence walk
  for (int64_t i0 = 0; w_lo + i0 * RPW < w_hi; i0 += IPL) {
    int32_t pa, pb, py, prl, prh;
    {
      const int64_t e = w_lo + i0 * RPW + lane;
      const int64_t ee = e < w_hi ? e : w_lo;  // (a valid index; not used past w_hi)
      pa = a.seq[ee];
      pb = a.seq[ee + 1];
      py = a.labels[ee];
      const double r = a.R[ee];
      prl = (int32_t)__double2loint(r);
      prh = (int32_t)__double2hiint(r);
    }
    rnew = 0.0;
    for (int i = 0; i < IPL && w_lo + (i0 + i) * RPW < w_hi; i += UE) {
      // (every lane takes part in the exchanges: a shuffle reading an
      // inactive lane returns nothing useful)
      int32_t ea[UE], eb[UE];
      int yv[UE];
      bool val[UE];
      double rold[UE];
#pragma unroll
      for (int u = 0; u < UE; ++u) {
        const int src = (i + u) * RPW + g;
        val[u] = w_lo + g + (int64_t)(i0 + i + u) * RPW < w_hi;
        ea[u] = take_i(pa, src);
        eb[u] = take_i(pb, src);
        yv[u] = take_i(py, src);
        rold[u] = __hiloint2double(take_i(prh, src), take_i(prl, src));
        if (!val[u]) ea[u] = eb[u] = 0;  // (row 0: in bounds, unused)
      }
      if constexpr (kSeq) {
        const bool first = i0 == 0 && i == 0;
        if (first) {
          const int32_t s0 = __builtin_amdgcn_readfirstlane(ea[0]);
          issue_row<F, LPR, CH>(pi + (int64_t)s0 * a.ld, K, lg, rs[0]);
        }
#pragma unroll
        for (int u = 0; u < UE; ++u) {
          eb[u] = __builtin_amdgcn_readfirstlane(eb[u]);
          yv[u] = __builtin_amdgcn_readfirstlane(yv[u]);
          issue_row<F, LPR, CH>(pi + (int64_t)eb[u] * a.ld, K, lg, rs[u + 1]);
        }
        if (first) mask_row<F, LPR, CH>(K, lg, rs[0]);
#pragma unroll
        for (int u = 0; u < UE; ++u) {
          if (!val[u]) continue;  // uniform
          mask_row<F, LPR, CH>(K, lg, rs[u + 1]);
          if (yv[u] == kPpxBridge) continue;
          const double r = pair_ll(rs[u], rs[u + 1], yv[u] != 0, rold[u]);
          if (lane == i + u) rnew = r;  // (one row group: uniform control flow)
        }
        // (a batch with val[UE - 1] false is the wavefront's last)
#pragma unroll
        for (int j = 0; j < EPL; ++j) rs[0][j] = rs[UE][j];
      } else {
        F ra[UE][EPL], rb[UE][EPL];
#pragma unroll
        for (int u = 0; u < UE; ++u) {
          // (a bridge position reads row 0 twice, from L2, and is not counted)
          if (yv[u] == kPpxBridge) ea[u] = eb[u] = 0;
          if (LPR == 64) {
            ea[u] = __builtin_amdgcn_readfirstlane(ea[u]);
            eb[u] = __builtin_amdgcn_readfirstlane(eb[u]);
            yv[u] = __builtin_amdgcn_readfirstlane(yv[u]);
          }
          issue_row<F, LPR, CH>(pi + (int64_t)ea[u] * a.ld, K, lg, ra[u]);
          issue_row<F, LPR, CH>(pi + (int64_t)eb[u] * a.ld, K, lg, rb[u]);
        }
#pragma unroll
        for (int u = 0; u < UE; ++u) {
          double r = 0.0;
          if (val[u] && yv[u] != kPpxBridge) {  // uniform within the row group
            mask_row<F, LPR, CH>(K, lg, ra[u]);
            mask_row<F, LPR, CH>(K, lg, rb[u]);
            r = pair_ll(ra[u], rb[u], yv[u] != 0, rold[u]);
          }
          // lane j keeps the average of position i0 + j / RPW, row group
          // j % RPW; every lane takes part (the row groups diverge above)
          if constexpr (RPW == 1) {
            if (lane == i + u) rnew = r;
          } else {
            const double rv = __shfl(r, (lane % RPW) * LPR, 64);
            if (lane / RPW == i + u) rnew = rv;
          }
        }
      }
    }
    {
      const int64_t e = w_lo + i0 * RPW + lane;
      if (e < w_hi) a.R[e] = rnew;
    }
  }
  // lanes with lg != 0 hold zeros
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    ll1 += __shfl_xor(ll1, off, 64);
    ll0 += __shfl_xor(ll0, off, 64);
    n1 += __shfl_xor(n1, off, 64);
    n0 += __shfl_xor(n0, off, 64);
  }
  if (lane == 0) {
    s_p[w][0] = ll1;
    s_p[w][1] = ll0;
    s_p[w][2] = n1;
    s_p[w][3] = n0;
  }
  __syncthreads();
  if (tid < 4) {
    double v = 0.0;
    for (int q = 0; q < WAVES; ++q) v += s_p[q][tid];
    a.partials[(int64_t)blockIdx.x * 4 + tid] = v;
  }
}

__global__ __launch_bounds__(256) void k_ppx_final(const double* partials, int blocks,
                                                   DevState* st) {
  __shared__ double s[4][256];
  double v[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < blocks; b += 256)
    for (int q = 0; q < 4; ++q) v[q] += partials[(int64_t)b * 4 + q];
  for (int q = 0; q < 4; ++q) s[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int q = 0; q < 4; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = s[2][0] + s[3][0];
    const double avg = n != 0.0 ? (s[0][0] + s[1][0]) / n : 0.0;
    st->ppx = -avg;
  }
}

// phi ~ Gamma(1,1) = -log(U) per (node, k), pi = phi / sum (init,
// mcmc_sampler_stochastic.cc:66-76).  One wavefront per node.
template <typename F>
__global__ __launch_bounds__(256) void k_init_phi(F* pi, int64_t ld, int K, int N, Key key) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t node = wave; node < N; node += nwaves) {
    double s = 0.0;
    for (int k = 2 * lane; k < K; k += 128) {
      const u32x4 r = draw4(key, (uint32_t)(k >> 1), (uint32_t)node, 0, TAG_INIT_PHI);
      s += -log(u01_from_u64((uint64_t)r.x | ((uint64_t)r.y << 32)));
      if (k + 1 < K) s += -log(u01_from_u64((uint64_t)r.z | ((uint64_t)r.w << 32)));
    }
    s = group_sum<64>(s);
    F* row = pi + node * ld;
    for (int k = 2 * lane; k < K; k += 128) {
      const u32x4 r = draw4(key, (uint32_t)(k >> 1), (uint32_t)node, 0, TAG_INIT_PHI);
      row[k] = (F)(-log(u01_from_u64((uint64_t)r.x | ((uint64_t)r.y << 32))) / s);
      if (k + 1 < K) row[k + 1] = (F)(-log(u01_from_u64((uint64_t)r.z | ((uint64_t)r.w << 32))) / s);
    }
    if (lane == 0) row[K] = (F)s;
  }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_f64(double* p, double v, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_set_ring(DevState* st, int64_t base, int32_t slots) {
  if (threadIdx.x == 0) {
    st->ring_base = base;
    st->ring_slots = slots;
  }
}

__global__ void k_set_step(DevState* st, int64_t step, DevHyper hp) {
  if (threadIdx.x == 0) {
    st->step = step;
    st->eps_t = eps_of_step(hp, step);
    st->error = 0;
  }
}

constexpr int kPhiWaves = 4;
constexpr int kPhiUnroll = 1;

constexpr int kPpxWaves = 4;
constexpr int kNbrWaves = 4;

// (LPR, CH) dispatch: LPR in {4,8,16,32} with CH=1, or LPR=64 with
// CH in {1,2,3,4,6,8,12,16}.
#define AMMSB_ROW_CONFIGS(X) \
  X(4, 1)                    \
  X(8, 1)                    \
  X(16, 1)                   \
  X(32, 1)                   \
  X(64, 1)                   \
  X(64, 2)                   \
  X(64, 3)                   \
  X(64, 4)                   \
  X(64, 6)                   \
  X(64, 8)                   \
  X(64, 12)                  \
  X(64, 16)

}  // namespace

bool row_config(int32_t K, int precision, int* lpr, int* ch) {
  const int V = precision == 0 ? 2 : 4;
  const int vecs = (K + V - 1) / V;
  if (K < 1) return false;
  if (vecs <= 32) {
    int l = 4;
    while (l < vecs) l <<= 1;
    *lpr = l;
    *ch = 1;
    return true;
  }
  const int need = (vecs + 63) / 64;
  static const int opts[] = {1, 2, 3, 4, 6, 8, 12, 16};
  for (int o : opts)
    if (o >= need) {
      *lpr = 64;
      *ch = o;
      return true;
    }
  return false;
}

hipError_t launch_minibatch(const SamplerArgs& a, int32_t fills, hipStream_t s) {
  if (fills < 1 || fills > a.ring.slots) return hipErrorInvalidValue;
  const size_t lds = a.table_in_lds ? (size_t)a.table_cap * sizeof(int2) : 0;
  hipLaunchKernelGGL(k_minibatch, dim3(fills), dim3(kMbThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_neighbors(const NeighborArgs& a, int32_t fills, hipStream_t s) {
  const int32_t grid_nodes = (int32_t)a.ring.node_cap;
  if (a.use_lds) {
    const size_t lds = neighbor_lds_bytes(a.tcap);
    const int blocks = (grid_nodes + kNbrWaves - 1) / kNbrWaves;
    hipLaunchKernelGGL(k_neighbors<kNbrWaves>, dim3(blocks, fills), dim3(kNbrWaves * 64), lds, s,
                       a, 0, grid_nodes, 0);
  } else {
    // global de-dup tables exist for gtable_slots nodes at a time
    for (int32_t rs = 0; rs < fills; ++rs)
      for (int32_t off = 0; off < grid_nodes; off += a.gtable_slots) {
        const int n = std::min<int32_t>(a.gtable_slots, grid_nodes - off);
        hipLaunchKernelGGL(k_neighbors<1>, dim3(n, 1), dim3(64), 0, s, a, off, off + n, rs);
      }
  }
  return hipGetLastError();
}

namespace {
// Launch shapes of k_update_phi: WAVES wavefronts per node, U neighbour rows
// in flight per row group.  Full-wave rows with CH in {1,2,4,8} (the common
// K of FP64 K<=1024 and FP32 K<=2048) get every shape; other row configs
// run the default (4, 1).
bool phi_shape_ok(int lpr, int ch, int waves, int unroll) {
  if (waves == kPhiWaves && unroll == kPhiUnroll) return true;
  if (!(lpr == 64 && (ch == 1 || ch == 2 || ch == 4 || ch == 8))) return false;
  return (waves == 2 && (unroll == 2 || unroll == 4)) ||
         (waves == 4 && (unroll == 2 || unroll == 4)) ||
         (waves == 8 && (unroll == 1 || unroll == 2 || unroll == 4));
}

template <typename F, int L, int C>
hipError_t phi_launch(const PhiArgs& a, int grid, hipStream_t s, int waves, int unroll) {
#define V(W, U)                                                                        \
  if (waves == W && unroll == U) {                                                     \
    hipLaunchKernelGGL((k_update_phi<F, L, C, W, U>), dim3(grid), dim3(W * 64), 0, s, a); \
    return hipGetLastError();                                                          \
  }
  V(kPhiWaves, kPhiUnroll)
  if constexpr (L == 64 && (C == 1 || C == 2 || C == 4 || C == 8)) {
    V(2, 2) V(2, 4) V(4, 2) V(4, 4) V(8, 1) V(8, 2) V(8, 4)
  }
#undef V
  return hipErrorInvalidValue;
}
}  // namespace

bool update_phi_shape_supported(int32_t K, int precision, int waves, int unroll) {
  int lpr, ch;
  if (!row_config(K, precision, &lpr, &ch)) return false;
  return phi_shape_ok(lpr, ch, waves, unroll);
}

hipError_t launch_update_phi(const PhiArgs& a, int precision, int32_t grid_nodes,
                             hipStream_t s, int waves, int unroll) {
  int lpr, ch;
  if (!row_config(a.K, precision, &lpr, &ch)) return hipErrorInvalidValue;
  if (!phi_shape_ok(lpr, ch, waves, unroll)) return hipErrorInvalidValue;
  if (grid_nodes <= 0) return hipSuccess;
#define X(L, C)                                                               \
  if (lpr == L && ch == C) {                                                  \
    if (precision == 0)                                                       \
      return phi_launch<double, L, C>(a, grid_nodes, s, waves, unroll); \
    return phi_launch<float, L, C>(a, grid_nodes, s, waves, unroll);    \
  }
  AMMSB_ROW_CONFIGS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_scatter(const ScatterArgs& a, int precision, int32_t grid_nodes,
                          hipStream_t s) {
  if (grid_nodes <= 0) return hipSuccess;
  const dim3 grid((grid_nodes + 3) / 4);
  if (precision == 0)
    hipLaunchKernelGGL(k_scatter<double>, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_scatter<float>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_beta(const BetaArgs& a, int precision, hipStream_t s) {
  int lpr, ch;
  if (!row_config(a.K, precision, &lpr, &ch)) return hipErrorInvalidValue;
#define X(L, C)                                                                          \
  if (lpr == L && ch == C) {                                                             \
    if (precision == 0)                                                                  \
      hipLaunchKernelGGL((k_beta<double, L, C, kBetaWaves>), dim3(a.blocks),             \
                         dim3(kBetaWaves * 64), 0, s, a);                                \
    else                                                                                 \
      hipLaunchKernelGGL((k_beta<float, L, C, kBetaWaves>), dim3(a.blocks),              \
                         dim3(kBetaWaves * 64), 0, s, a);                                \
    return hipGetLastError();                                                            \
  }
  AMMSB_ROW_CONFIGS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_theta(const ThetaArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_theta, dim3((a.K + kThetaK - 1) / kThetaK), dim3(kThetaThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_perplexity(const PpxArgs& a, int precision, hipStream_t s) {
  int lpr, ch;
  if (!row_config(a.K, precision, &lpr, &ch)) return hipErrorInvalidValue;
#define X(L, C)                                                                          \
  if (lpr == L && ch == C) {                                                             \
    if (precision == 0)                                                                  \
      hipLaunchKernelGGL((k_ppx<double, L, C, kPpxWaves>), dim3(a.blocks),               \
                         dim3(kPpxWaves * 64), 0, s, a);                                 \
    else                                                                                 \
      hipLaunchKernelGGL((k_ppx<float, L, C, kPpxWaves>), dim3(a.blocks),                \
                         dim3(kPpxWaves * 64), 0, s, a);                                 \
    hipError_t e = hipGetLastError();                                                    \
    if (e != hipSuccess) return e;                                                       \
    hipLaunchKernelGGL(k_ppx_final, dim3(1), dim3(256), 0, s, a.partials, a.blocks,      \
                       a.st);                                                            \
    return hipGetLastError();                                                            \
  }
  AMMSB_ROW_CONFIGS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_init_phi(void* pi, int64_t ld, int32_t K, int32_t N, Key key, int precision,
                           hipStream_t s) {
  const int blocks = (int)std::min<int64_t>(65536, ((int64_t)N + 3) / 4);
  if (precision == 0)
    hipLaunchKernelGGL(k_init_phi<double>, dim3(blocks), dim3(256), 0, s, (double*)pi, ld, K, N,
                       key);
  else
    hipLaunchKernelGGL(k_init_phi<float>, dim3(blocks), dim3(256), 0, s, (float*)pi, ld, K, N,
                       key);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s) {
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256));
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_fill_f64(double* p, double v, int64_t n, hipStream_t s) {
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256));
  hipLaunchKernelGGL(k_fill_f64, dim3(blocks), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t launch_set_ring(DevState* st, int64_t base, int32_t slots, hipStream_t s) {
  hipLaunchKernelGGL(k_set_ring, dim3(1), dim3(64), 0, s, st, base, slots);
  return hipGetLastError();
}

hipError_t launch_set_state_step(DevState* st, int64_t step, const DevHyper& hp, hipStream_t s) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, st, step, hp);
  return hipGetLastError();
}

}  // namespace ammsb

namespace ammsb {
namespace {
__global__ void k_label_nbrs(DevGraph g, const int32_t* nodes, int32_t n_nodes, int32_t* nbrs,
                             int32_t nnb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_nodes * nnb) return;
  const int32_t node = nodes[i / nnb];
  const int32_t v = nbrs[i] & 0x7fffffff;
  const int64_t rs = g.row_ptr[node];
  const bool lk = row_contains(g.cols + rs, g.row_ptr[node + 1] - rs, v);
  nbrs[i] = v | (lk ? (int32_t)0x80000000 : 0);
}
}  // namespace

hipError_t launch_label_nbrs(DevGraph g, const int32_t* nodes, int32_t n_nodes, int32_t* nbrs,
                             int32_t nnb, hipStream_t s) {
  const int64_t n = (int64_t)n_nodes * nnb;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_label_nbrs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, nodes,
                     n_nodes, nbrs, nnb);
  return hipGetLastError();
}
}  // namespace ammsb

namespace ammsb {
namespace {
// Contiguous read of `n` 16-byte vectors (bandwidth probe and counter
// calibration: a read of known size).  The sum is stored only if it equals
// an impossible value, so the loads cannot be removed.
__global__ __launch_bounds__(256) void k_read_probe(const double2* __restrict__ p, int64_t n,
                                                    double* sink) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = p[i];
    s += v.x + v.y;
  }
  if (s == -1.2345678e300) sink[0] = s;
}
}  // namespace

hipError_t launch_read_probe(const void* p, int64_t bytes, double* sink, hipStream_t s) {
  const int64_t n = bytes / 16;
  hipLaunchKernelGGL(k_read_probe, dim3(256 * 64), dim3(256), 0, s, (const double2*)p, n, sink);
  return hipGetLastError();
}

namespace {
// Bandwidth probe of k_update_phi's row gather: the same workgroup per
// minibatch node, wavefront split, row groups, rows in flight, neighbour-id
// registers and streaming row loads as k_update_phi (the context's launch
// shape), without the arithmetic and without the staged row write.  The rows
// are summed into a value stored only if it equals an impossible value, so
// no load can be removed.
template <typename F, int LPR, int CH, int WAVES, int U>
__global__ __launch_bounds__(WAVES * 64) void k_gather_probe(PhiArgs a, double* sink) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CH;
  constexpr int RPW = 64 / LPR;
  constexpr int STRIDE = WAVES * RPW;
  const int slot = blockIdx.x;
  const int rs = ring_slot(a.st);
  const int count = a.ring.meta[rs].node_count;
  if (slot >= count) return;
  const int32_t* nodes = a.ring.nodes + (int64_t)rs * a.ring.node_cap;
  const int32_t* nb_src = a.ring.nbrs + ((int64_t)rs * a.ring.node_cap + slot) * a.nnb;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int g = lane / LPR, lg = lane % LPR;
  const F* pi = (const F*)a.pi;
  const int K = a.K;
  const int32_t node = nodes[slot];
  const bool reg_ids = a.nnb <= 64 * WAVES;
  int32_t ids = node;
  if (reg_ids) {
    const int r = w * RPW + lane % RPW + STRIDE * (lane / RPW);
    if (r < a.nnb) ids = nb_src[r];
  }
  F row[EPL];
  load_row<F, LPR, CH>(pi + (int64_t)node * a.ld, K, lg, row);
  F acc = (F)0;
#pragma unroll
  for (int j = 0; j < EPL; ++j) acc += row[j];
  for (int it = 0; w * RPW + it * STRIDE < a.nnb; it += U) {
    F rows[U][EPL];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = w * RPW + g + (it + u) * STRIDE;
      int32_t v;
      if (reg_ids) {
        if constexpr (RPW == 1)
          v = __builtin_amdgcn_readlane(ids, min(it + u, 63));
        else
          v = __shfl(ids, ((it + u) * RPW + g) & 63, 64);
      } else {
        v = b < a.nnb ? nb_src[b] : node;
      }
      v = b < a.nnb ? v : node;
      if (LPR == 64) v = __builtin_amdgcn_readfirstlane(v);
      issue_row<F, LPR, CH, true>(pi + (int64_t)(v & 0x7fffffff) * a.ld, K, lg, rows[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int j = 0; j < EPL; ++j) acc += rows[u][j];
  }
  if (acc == (F)-1.2345e30) sink[0] = (double)acc;
}
}  // namespace

namespace {
template <typename F, int L, int C>
hipError_t probe_launch(const PhiArgs& a, int grid, double* sink, hipStream_t s, int waves,
                        int unroll) {
  if (waves == 4 && unroll == 1) {
    hipLaunchKernelGGL((k_gather_probe<F, L, C, 4, 1>), dim3(grid), dim3(256), 0, s, a, sink);
    return hipGetLastError();
  }
  if constexpr (L == 64 && (C == 1 || C == 2 || C == 4 || C == 8)) {
    if (waves == 8 && unroll == 2) {
      hipLaunchKernelGGL((k_gather_probe<F, L, C, 8, 2>), dim3(grid), dim3(512), 0, s, a, sink);
      return hipGetLastError();
    }
  }
  return hipErrorInvalidValue;
}
}  // namespace

hipError_t launch_gather_probe(const PhiArgs& a, int precision, int32_t grid_nodes, double* sink,
                               hipStream_t s, int waves, int unroll) {
  int lpr, ch;
  if (!row_config(a.K, precision, &lpr, &ch)) return hipErrorInvalidValue;
  if (grid_nodes <= 0) return hipSuccess;
#define X(L, C)                                                                       \
  if (lpr == L && ch == C) {                                                          \
    if (precision == 0) return probe_launch<double, L, C>(a, grid_nodes, sink, s, waves, unroll); \
    return probe_launch<float, L, C>(a, grid_nodes, sink, s, waves, unroll);           \
  }
  AMMSB_ROW_CONFIGS(X)
#undef X
  return hipErrorInvalidValue;
}
}  // namespace ammsb

// ===========================================================================
// Differentiation: the minibatch log-likelihood
//   L = sum_s sum_{p in pairs(s)} log p(y_p | pi_s, pi_{b_p}, beta(theta))
//   p(y) = sum_k pi_sk (pi_bk f_k + e),  f_k = beta_k - eps (y=1) / eps - beta_k
//          (y=0), e = eps / 1 - eps  -- the likelihood of update_phi
//          (mcmc_sampler_stochastic.cc:285-292) and of update_beta (:220-233)
//   pi_s = phi_s / sum(phi_s) from the caller's rows, pi_b the resident rows
//   (constants, as in update_phi), beta_k = theta_k1 / (theta_k0 + theta_k1).
// Adjoint (recomputes every x_bk and Z instead of storing them):
//   dL/dphi_sk = sum_p (x_pk / Z_p - 1) / sum(phi_s)   (update_phi's grads)
//   dL/dbeta_k = sum_p (2y_p - 1) pi_sk pi_bk / Z_p
//   dL/dtheta_k0 = -theta_k1/S^2 dL/dbeta_k, dL/dtheta_k1 = theta_k0/S^2 dL/dbeta_k
//   (equal to update_beta's G_k0, G_k1, mcmc_sampler_stochastic.cc:234-239)
namespace ammsb {
namespace {

template <typename F, int LPR, int CH, int WAVES, bool BWD>
__global__ __launch_bounds__(WAVES * 64) void k_loglik(LogLikArgs a) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CH;
  constexpr int RPW = 64 / LPR;
  constexpr int STRIDE = WAVES * RPW;
  __shared__ double s_red[(WAVES > 1 ? WAVES - 1 : 1) * EPL * LPR];
  __shared__ double s_ll[WAVES];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int g = lane / LPR, lg = lane % LPR;
  const int K = a.K;
  const F* pi = (const F*)a.pi;
  const F eps = (F)a.epsilon;
  F f1[EPL];
  double dbeta[EPL];
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = (c * LPR + lg) * V + v;
      const int j = c * V + v;
      double bk = 0.0;
      if (k < K) bk = a.theta[2 * k + 1] / (a.theta[2 * k] + a.theta[2 * k + 1]);
      f1[j] = (F)bk - eps;
      dbeta[j] = 0.0;
    }
  const F e1 = eps, e0 = (F)1 - eps;
  for (int64_t s = blockIdx.x; s < a.M; s += gridDim.x) {
    const int32_t node = a.nodes[s];
    const F* phis = (const F*)a.phi + s * a.phi_ld;
    F pis[EPL];
    double sp = 0.0;
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int k = (c * LPR + lg) * V + v;
        pis[c * V + v] = k < K ? phis[k] : (F)0;
        sp += (double)pis[c * V + v];
      }
    sp = group_sum<LPR>(sp);  // sum(phi_s) in double
    if (RPW > 1) sp = read_lane(sp, 0);
#pragma unroll
    for (int j = 0; j < EPL; ++j) pis[j] = (F)((double)pis[j] / sp);
    F acc[EPL];
#pragma unroll
    for (int j = 0; j < EPL; ++j) acc[j] = (F)0;
    double ll = 0.0;
    // pair offsets clamped to [0, P]; neighbour ids outside [0, N) are
    // skipped and reported (caller-provided device arrays)
    const int64_t p0 = min(max(a.pair_ptr[s], (int64_t)0), a.P);
    const int64_t p1 = min(max(a.pair_ptr[s + 1], p0), a.P);
    for (int64_t p = p0 + w * RPW + g; p < p1; p += STRIDE) {
      int32_t b = a.pair_nbr[p];
      int y = a.pair_y[p] ? 1 : 0;
      if (LPR == 64) {
        b = __builtin_amdgcn_readfirstlane(b);
        y = __builtin_amdgcn_readfirstlane(y);
      }
      if (b < 0 || b >= a.N) {
        if (lg == 0) atomicOr(a.error, ERR_BAD_INPUT);
        continue;
      }
      if (b == node) continue;  // mcmc_sampler_stochastic.cc:275-277
      F rb[EPL];
      issue_row<F, LPR, CH, true>(pi + (int64_t)b * a.ld, K, lg, rb);
      mask_row<F, LPR, CH>(K, lg, rb);
      F z = (F)0;
#pragma unroll
      for (int j = 0; j < EPL; ++j)
        z = fma(pis[j], y ? fma(rb[j], f1[j], e1) : fma(-rb[j], f1[j], e0), z);
      z = group_sum<LPR>(z);
      if (!BWD) {
        if (lg == 0) ll += log((double)z);
      } else {
        const F rz = recip(z);
        const double sg = y ? 1.0 : -1.0;
#pragma unroll
        for (int j = 0; j < EPL; ++j) {
          const F x = y ? fma(rb[j], f1[j], e1) : fma(-rb[j], f1[j], e0);  // recomputed
          acc[j] += fma(x, rz, (F)-1);
          dbeta[j] += sg * (double)(pis[j] * rb[j] * rz);
        }
      }
    }
    if (!BWD) {
      // lanes with lg != 0 hold zeros
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) ll += __shfl_xor(ll, off, 64);
      if (lane == 0) s_ll[w] = ll;
      __syncthreads();
      if (tid == 0) {
        double t = 0.0;
        for (int q = 0; q < WAVES; ++q) t += s_ll[q];
        a.per_source[s] = t;
      }
      __syncthreads();
    } else {
      if (RPW > 1) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) acc[j] = cross_group_sum<LPR>(acc[j]);
      }
      if (WAVES > 1) {
        if (w > 0 && g == 0) {
#pragma unroll
          for (int j = 0; j < EPL; ++j) s_red[((w - 1) * EPL + j) * LPR + lg] = (double)acc[j];
        }
        __syncthreads();
        if (w == 0) {
#pragma unroll
          for (int q = 1; q < WAVES; ++q)
#pragma unroll
            for (int j = 0; j < EPL; ++j) acc[j] += (F)s_red[((q - 1) * EPL + j) * LPR + lg];
        }
        __syncthreads();
      }
      if (w == 0 && g == 0) {
        const double go = a.gout[0];
        F* gp = (F*)a.grad_phi + s * a.phi_ld;
#pragma unroll
        for (int c = 0; c < CH; ++c)
#pragma unroll
          for (int v = 0; v < V; ++v) {
            const int k = (c * LPR + lg) * V + v;
            if (k < K) gp[k] = (F)(go * (double)acc[c * V + v] / sp);
          }
      }
    }
  }
  if (BWD) {
    if (RPW > 1) {
#pragma unroll
      for (int j = 0; j < EPL; ++j) dbeta[j] = cross_group_sum<LPR>(dbeta[j]);
    }
    if (WAVES > 1) {
      if (w > 0 && g == 0) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) s_red[((w - 1) * EPL + j) * LPR + lg] = dbeta[j];
      }
      __syncthreads();
      if (w == 0) {
#pragma unroll
        for (int q = 1; q < WAVES; ++q)
#pragma unroll
          for (int j = 0; j < EPL; ++j) dbeta[j] += s_red[((q - 1) * EPL + j) * LPR + lg];
      }
    }
    if (w == 0 && g == 0) {
      double* out = a.dbeta_partials + (int64_t)blockIdx.x * K;
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int k = (c * LPR + lg) * V + v;
          if (k < K) out[k] = dbeta[c * V + v];
        }
    }
  }
}

// K-chunked form for rows too wide for one wavefront's registers (K*F above
// 2 KB): a block of W = blockDim/64 wavefronts works on one source at a
// time, wavefront w owning the vectors [w*CHC*64, (w+1)*CHC*64) of every row
// (CHC coalesced 1 KB slices).  Pairs go P at a time: each wavefront loads
// its slice of the P neighbour rows, reduces its partial Z of each pair, and
// the W partials meet in LDS (double-buffered, one barrier per P pairs).
// Every wavefront then holds the full Z of the P pairs and finishes its own
// slice of dL/dphi and dL/dbeta with no further exchange; grad_phi and the
// dbeta partials are written slice by slice.
constexpr int kLlMaxWaves = 8;

// P pairs' slices of neighbour rows, their labels and validity.
template <typename F, int EPL, int P>
struct LlGroup {
  F rb[P][EPL];
  int y[P];
  bool ok[P];
};

// Issue the loads of pairs [p, p+P) (block-uniform ids).  A pair that is
// absent (past p1), out of range (reported) or the source itself reads row 0
// and is marked not ok.
template <typename F, int CHC, int P, int EPL>
__device__ __forceinline__ void ll_load(const LogLikArgs& a, const F* pi, int32_t node, int64_t p,
                                        int64_t p1, int w, int lane, LlGroup<F, EPL, P>& gr) {
  constexpr int V = VT<F>::V;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const bool in = p + i < p1;
    int32_t b = in ? a.pair_nbr[p + i] : node;
    int y = in ? (a.pair_y[p + i] ? 1 : 0) : 0;
    b = __builtin_amdgcn_readfirstlane(b);
    gr.y[i] = __builtin_amdgcn_readfirstlane(y);
    const bool bad = b < 0 || b >= a.N;
    if (bad && w == 0 && lane == 0) atomicOr(a.error, ERR_BAD_INPUT);
    gr.ok[i] = !bad && b != node;  // mcmc_sampler_stochastic.cc:275-277
    const F* row = pi + (int64_t)(gr.ok[i] ? b : 0) * a.ld;
#pragma unroll
    for (int c = 0; c < CHC; ++c) {
      const int k0 = ((w * CHC + c) * 64 + lane) * V;
      load_vec_nt<F>(row + (k0 < a.K ? k0 : 0), gr.rb[i] + c * V);
    }
  }
}

// Partial Z of each pair over this wavefront's slice, exchange through LDS
// (one barrier), then the forward log-sum or this slice's adjoint terms.
template <typename F, int CHC, int P, int EPL, bool BWD>
__device__ __forceinline__ void ll_process(LlGroup<F, EPL, P>& gr, const F* pis, const F* f1, F e1,
                                           F e0, int K, int w, int lane, int W,
                                           F (*s_z)[P][kLlMaxWaves], int& buf, double& ll,
                                           F* acc, double* dbeta) {
  constexpr int V = VT<F>::V;
#pragma unroll
  for (int i = 0; i < P; ++i) {
#pragma unroll
    for (int c = 0; c < CHC; ++c)
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int k = ((w * CHC + c) * 64 + lane) * V + v;
        gr.rb[i][c * V + v] = k < K ? gr.rb[i][c * V + v] : (F)0;
      }
    F z = (F)0;
#pragma unroll
    for (int j = 0; j < EPL; ++j)
      z = fma(pis[j], gr.y[i] ? fma(gr.rb[i][j], f1[j], e1) : fma(-gr.rb[i][j], f1[j], e0), z);
    z = group_sum<64>(z);
    if (lane == 0) s_z[buf][i][w] = z;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < P; ++i) {
    if (!gr.ok[i]) continue;
    F Z = (F)0;
    for (int q = 0; q < W; ++q) Z += s_z[buf][i][q];
    if (!BWD) {
      if (w == 0) ll += log((double)Z);
    } else {
      const F rz = recip(Z);
      const double sg = gr.y[i] ? 1.0 : -1.0;
#pragma unroll
      for (int j = 0; j < EPL; ++j) {
        const F x = gr.y[i] ? fma(gr.rb[i][j], f1[j], e1) : fma(-gr.rb[i][j], f1[j], e0);
        acc[j] += fma(x, rz, (F)-1);
        dbeta[j] += sg * (double)(pis[j] * gr.rb[i][j] * rz);
      }
    }
  }
  buf ^= 1;
}

// PF: the next P pairs' loads are issued before this group's barrier, so
// they are in flight while the block waits and finishes the current group.
template <typename F, int CHC, int P, bool PF, bool BWD>
__global__ __launch_bounds__(kLlMaxWaves * 64) void k_loglik_kc(LogLikArgs a) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CHC;
  __shared__ F s_z[2][P][kLlMaxWaves];
  __shared__ double s_sp[kLlMaxWaves];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int W = blockDim.x >> 6;
  const int K = a.K;
  const F* pi = (const F*)a.pi;
  const F eps = (F)a.epsilon;
  const F e1 = eps, e0 = (F)1 - eps;
  F f1[EPL];
  double dbeta[EPL];
#pragma unroll
  for (int c = 0; c < CHC; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = ((w * CHC + c) * 64 + lane) * V + v;
      double bk = 0.0;
      if (k < K) bk = a.theta[2 * k + 1] / (a.theta[2 * k] + a.theta[2 * k + 1]);
      f1[c * V + v] = (F)bk - eps;
      dbeta[c * V + v] = 0.0;
    }
  int buf = 0;
  for (int64_t s = blockIdx.x; s < a.M; s += gridDim.x) {
    const int32_t node = a.nodes[s];
    const F* phis = (const F*)a.phi + s * a.phi_ld;
    F pis[EPL];
    double sp = 0.0;
#pragma unroll
    for (int c = 0; c < CHC; ++c)
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int k = ((w * CHC + c) * 64 + lane) * V + v;
        pis[c * V + v] = k < K ? phis[k] : (F)0;
        sp += (double)pis[c * V + v];
      }
    sp = group_sum<64>(sp);
    if (lane == 0) s_sp[w] = sp;
    __syncthreads();
    sp = 0.0;
    for (int q = 0; q < W; ++q) sp += s_sp[q];  // sum(phi_s), fixed order
#pragma unroll
    for (int j = 0; j < EPL; ++j) pis[j] = (F)((double)pis[j] / sp);
    F acc[EPL];
#pragma unroll
    for (int j = 0; j < EPL; ++j) acc[j] = (F)0;
    double ll = 0.0;
    const int64_t p0 = min(max(a.pair_ptr[s], (int64_t)0), a.P);
    const int64_t p1 = min(max(a.pair_ptr[s + 1], p0), a.P);
    if (PF) {
      LlGroup<F, EPL, P> A, B;
      int64_t p = p0;
      if (p < p1) ll_load<F, CHC, P, EPL>(a, pi, node, p, p1, w, lane, A);
      while (p < p1) {
        const bool more = p + P < p1;
        if (more) ll_load<F, CHC, P, EPL>(a, pi, node, p + P, p1, w, lane, B);
        ll_process<F, CHC, P, EPL, BWD>(A, pis, f1, e1, e0, K, w, lane, W, s_z, buf, ll, acc,
                                        dbeta);
        p += P;
        if (!more) break;
        const bool more2 = p + P < p1;
        if (more2) ll_load<F, CHC, P, EPL>(a, pi, node, p + P, p1, w, lane, A);
        ll_process<F, CHC, P, EPL, BWD>(B, pis, f1, e1, e0, K, w, lane, W, s_z, buf, ll, acc,
                                        dbeta);
        p += P;
        if (!more2) break;
      }
    } else {
      for (int64_t p = p0; p < p1; p += P) {
        LlGroup<F, EPL, P> A;
        ll_load<F, CHC, P, EPL>(a, pi, node, p, p1, w, lane, A);
        ll_process<F, CHC, P, EPL, BWD>(A, pis, f1, e1, e0, K, w, lane, W, s_z, buf, ll, acc,
                                        dbeta);
      }
    }
    if (!BWD) {
      if (w == 0 && lane == 0) a.per_source[s] = ll;
    } else {
      const double go = a.gout[0];
      F* gp = (F*)a.grad_phi + s * a.phi_ld;
#pragma unroll
      for (int c = 0; c < CHC; ++c)
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int k = ((w * CHC + c) * 64 + lane) * V + v;
          if (k < K) gp[k] = (F)(go * (double)acc[c * V + v] / sp);
        }
    }
    __syncthreads();  // s_sp and s_z are rewritten by the next source
  }
  if (BWD) {
    double* out = a.dbeta_partials + (int64_t)blockIdx.x * K;
#pragma unroll
    for (int c = 0; c < CHC; ++c)
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int k = ((w * CHC + c) * 64 + lane) * V + v;
        if (k < K) out[k] = dbeta[c * V + v];
      }
  }
}

// Log-likelihood and its adjoint for rows wider than one wavefront's 2 KB,
// whole rows per wavefront.  Pair p of a source goes to wavefront
// (p - p0) mod 4, which holds the ids and labels of its next 64 pairs in
// two registers (lane j: its j-th pair), so that requesting a row never
// waits for a load; up to 8 KB rows it keeps two neighbour rows in flight
// (the next pair's row is requested before the current one is reduced).
// The source row is held in registers in the regrouped form of update_phi's
// gradient:
//   q_k = pi_sk f1_k,  S = sum_k pi_sk,  P_p = sum_k q_k pi_bk,
//   Z_p = P_p + eps S (link) or (1 - eps) S - P_p (non-link),
// one multiply-add per element and pair.  The adjoint accumulates per
// wavefront and source
//   A_k = sum_p s_p pi_bk / Z_p  (s_p = +1 link, -1 non-link),
//   C = sum_p (e_p / Z_p - 1),
// so that dL/dphi_sk = (f1_k A_k + C) / sum(phi_s) and the source's share
// of dL/dbeta_k is pi_sk A_k (the sums of the x/Z - 1 and
// (2y-1) pi_s pi_b / Z terms above, in another order).  At the end of a
// source the wavefronts' A and C meet through two LDS buffers as
// (A0 + A1) + (A2 + A3); wavefront w then writes dL/dphi of the row's
// vector chunks w, w + 4, ... and adds their dL/dbeta terms into the
// workgroup's partial row (each element always by the same lane).  The forward adds the wavefronts' log-sums in wavefront order.
// Dynamic LDS: f1, q (the source row is staged through it, read
// coalesced) and, for the adjoint, pi_s and one more combine buffer, each
// NCH*64*V Floats, and 2 C.
template <typename F, int NCH, bool BWD>
__global__ __launch_bounds__(kLlRowWaves * 64) void k_loglik_rows(LogLikArgs a) {
  constexpr int V = VT<F>::V;
  constexpr int W = kLlRowWaves;
  constexpr int E = NCH * V;  // elements per lane
  constexpr int KP = E * 64;
  // two rows in flight up to 8 KB rows; wider rows leave the registers to
  // q and A (with q read from LDS for each pair instead, the rows in flight
  // per SIMD do not grow: two rows but one wavefront)
  constexpr bool DB = NCH <= 8;
  static_assert(W == 4, "the adjoint's combine tree");
  extern __shared__ __align__(16) unsigned char s_dynb[];
  F* const s_f1 = reinterpret_cast<F*>(s_dynb);  // f1_k = beta_k - eps, 0 past K
  F* const s_q = s_f1 + KP;                       // q_k; adjoint: then a combine buffer
  F* const s_pk = s_q + KP;                       // adjoint: pi_sk
  F* const s_b1 = s_pk + KP;                      // adjoint: combine buffer
  F* const s_c = s_b1 + KP;                       // adjoint: [2]
  __shared__ double s_red[W];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int K = a.K;
  const F* pi = (const F*)a.pi;
  const F eps = (F)a.epsilon;
  const F e1 = eps, e0 = (F)1 - eps;
  for (int k = tid; k < KP; k += W * 64) {
    double bk = 0.0;
    if (k < K) bk = a.theta[2 * k + 1] / (a.theta[2 * k] + a.theta[2 * k + 1]);
    s_f1[k] = k < K ? (F)bk - eps : (F)0;
  }
  // the adjoint's partial dL/dbeta row: vector chunk c of every source is
  // added by wavefront c mod 4, each element by the same lane
  double* const dbr = BWD ? a.dbeta_partials + (int64_t)blockIdx.x * K : nullptr;
  if constexpr (BWD) {
    for (int c = w; c < NCH; c += W)
#pragma unroll
      for (int v = 0; v < V; ++v) {
        const int k = (c * 64 + lane) * V + v;
        if (k < K) dbr[k] = 0.0;
      }
  }
  const double go = BWD ? a.gout[0] : 0.0;
  // each lane's byte offsets in a neighbour row (a vector past the row
  // re-reads vector 0; q is 0 there)
  uint32_t voff[NCH];
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    const int k0 = (c * 64 + lane) * V;
    voff[c] = (uint32_t)((k0 < K ? k0 : 0) * (int)sizeof(F));
  }
  const uint32_t row_bytes = (uint32_t)(a.ld * (int64_t)sizeof(F));
  for (int64_t s = blockIdx.x; s < a.M; s += gridDim.x) {
    const int32_t node = a.nodes[s];
    const F* phis = (const F*)a.phi + s * a.phi_ld;
    // the source row, read coalesced: sum(phi_s) (per thread, per
    // wavefront, then in wavefront order), pi_s = phi_s / sum(phi_s), q and
    // S = sum(pi_s) the same way
    constexpr int KT = (KP + W * 64 - 1) / (W * 64);  // elements per thread
    F ph[KT];
    double part = 0.0;
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      const int k = tid + i * W * 64;
      ph[i] = k < K ? phis[k < K ? k : 0] : (F)0;
      part += (double)ph[i];
    }
    part = group_sum<64>(part);
    __syncthreads();  // the previous source's reads of s_red, s_q and s_c are done
    if (lane == 0) s_red[w] = part;
    __syncthreads();
    double sp = 0.0;
    for (int q2 = 0; q2 < W; ++q2) sp += s_red[q2];  // sum(phi_s)
    double spi = 0.0;
#pragma unroll
    for (int i = 0; i < KT; ++i) {
      const int k = tid + i * W * 64;
      if (k < KP) {
        const F pk = k < K ? (F)((double)ph[i] / sp) : (F)0;
        spi += (double)pk;
        s_q[k] = pk * s_f1[k];
        if constexpr (BWD) s_pk[k] = pk;
      }
    }
    spi = group_sum<64>(spi);
    __syncthreads();  // every wavefront has read s_red
    if (lane == 0) s_red[w] = spi;
    __syncthreads();
    double st = 0.0;
    for (int q2 = 0; q2 < W; ++q2) st += s_red[q2];
    const F S = (F)st;
    const F zs1 = e1 * S, zs0 = e0 * S;
    F q[E];
#pragma unroll
    for (int c = 0; c < NCH; ++c) load_vec<F>(s_q + (c * 64 + lane) * V, q + c * V);
    const int64_t p0 = min(max(a.pair_ptr[s], (int64_t)0), a.P);
    const int64_t p1 = min(max(a.pair_ptr[s + 1], p0), a.P);
    double ll = 0.0;
    F A[BWD ? E : 1];
#pragma unroll
    for (int j = 0; j < (BWD ? E : 1); ++j) A[j] = (F)0;
    F C = (F)0;
    // this wavefront's pairs: p0 + w + W j, in chunks of 64
    for (int64_t cb = p0 + w; cb < p1; cb += 64 * W) {
      const int n = (int)min((int64_t)64, (p1 - cb + W - 1) / W);
      int32_t idv = 0, yv = 0;
      {
        const int64_t pl = cb + (int64_t)lane * W;
        const int64_t pc = pl < p1 ? pl : cb;
        idv = a.pair_nbr[pc];
        yv = a.pair_y[pc] ? 1 : 0;
      }
      // request the j-th pair's row
      auto fetch = [&](int j, F* r, int& y, bool& ok) {
        const int32_t b = __builtin_amdgcn_readlane(idv, j);
        y = __builtin_amdgcn_readlane(yv, j);
        const bool bad = b < 0 || b >= a.N;
        if (bad && lane == 0) atomicOr(a.error, ERR_BAD_INPUT);
        ok = !bad && b != node;  // mcmc_sampler_stochastic.cc:275-277
        const __amdgpu_buffer_rsrc_t rr = row_rsrc(pi + (int64_t)(ok ? b : 0) * a.ld, row_bytes);
#pragma unroll
        for (int c = 0; c < NCH; ++c) load_vec_buf<F, true>(rr, voff[c], r + c * V);
      };
      // reduce a requested row (q is 0 past K, so no mask is needed)
      auto reduce = [&](const F* r, int y, bool ok) {
        F P = (F)0;
#pragma unroll
        for (int j = 0; j < E; ++j) P = fma(q[j], r[j], P);
        P = group_sum<64>(P);
        const F z = y ? P + zs1 : zs0 - P;
        if (!ok) return;  // uniform
        if constexpr (!BWD) {
          ll += log((double)z);
        } else {
          const F rz = recip(z);
          const F sw = y ? rz : -rz;
#pragma unroll
          for (int j = 0; j < E; ++j) A[j] = fma(r[j], sw, A[j]);
          C += fma(y ? e1 : e0, rz, (F)-1);
        }
      };
      if constexpr (DB) {
        F ra[E], rb[E];
        int ya = 0, yb = 0;
        bool oka = false, okb = false;
        fetch(0, ra, ya, oka);
        for (int j = 0; j < n; j += 2) {
          if (j + 1 < n) fetch(j + 1, rb, yb, okb);
          reduce(ra, ya, oka);
          if (j + 1 >= n) break;
          if (j + 2 < n) fetch(j + 2, ra, ya, oka);
          reduce(rb, yb, okb);
        }
      } else {
        for (int j = 0; j < n; ++j) {
          F r[E];
          int y;
          bool ok;
          fetch(j, r, y, ok);
          reduce(r, y, ok);
        }
      }
    }
    if constexpr (!BWD) {
      __syncthreads();  // every wavefront has read s_red
      if (lane == 0) s_red[w] = ll;
      __syncthreads();
      if (tid == 0) {
        double t = 0.0;
        for (int q2 = 0; q2 < W; ++q2) t += s_red[q2];
        a.per_source[s] = t;
      }
    } else {
      // (A0 + A1) + (A2 + A3) in wavefront 0; s_q serves as a buffer
      F* const s_b0 = s_q;
      auto put = [&](F* buf, int slot) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) store_vec<F>(buf + (c * 64 + lane) * V, A + c * V);
        if (lane == 0) s_c[slot] = C;
      };
      auto add = [&](const F* buf, int slot) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
          F t[V];
          load_vec<F>(buf + (c * 64 + lane) * V, t);
#pragma unroll
          for (int v = 0; v < V; ++v) A[c * V + v] += t[v];
        }
        C += s_c[slot];
      };
      __syncthreads();  // every wavefront has read s_q
      if (w == 1) put(s_b0, 0);
      if (w == 3) put(s_b1, 1);
      __syncthreads();
      if (w == 0) add(s_b0, 0);
      if (w == 2) add(s_b1, 1);
      __syncthreads();
      if (w == 2) put(s_b0, 0);
      __syncthreads();
      if (w == 0) {
        add(s_b0, 0);
        put(s_b1, 1);
      }
      __syncthreads();
      // the sums in s_b1 and s_c[1]; wavefront w finishes chunks w, w+4, ...
      const F Ct = s_c[1];
      F* gp = (F*)a.grad_phi + s * a.phi_ld;
      for (int c = w; c < NCH; c += W) {
        const int k0 = (c * 64 + lane) * V;
        F f[V], pk[V], At[V];
        load_vec<F>(s_f1 + k0, f);
        load_vec<F>(s_pk + k0, pk);
        load_vec<F>(s_b1 + k0, At);
#pragma unroll
        for (int v = 0; v < V; ++v) {
          const int k = k0 + v;
          if (k < K) {
            gp[k] = (F)(go * (double)fma(f[v], At[v], Ct) / sp);
            dbr[k] += (double)(pk[v] * At[v]);
          }
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_sum(const double* x, int64_t n, double* out) {
  __shared__ double s[256];
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) v += x[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// dL/dtheta from the [blocks][K] dbeta partials: 64 k per block, 16 lane
// rows splitting the partials (b = r mod 16), fixed-order LDS combine.
__global__ __launch_bounds__(1024) void k_theta_grad(const double* part, int blocks, int K,
                                                     const double* theta, const double* gout,
                                                     double* grad) {
  __shared__ double s[16][64];
  const int kl = threadIdx.x & 63, r = threadIdx.x >> 6;
  const int k = blockIdx.x * 64 + kl;
  double d = 0.0;
  if (k < K)
    for (int b = r; b < blocks; b += 16) d += part[(int64_t)b * K + k];
  s[r][kl] = d;
  __syncthreads();
  if (r != 0 || k >= K) return;
  d = 0.0;
  for (int q = 0; q < 16; ++q) d += s[q][kl];
  const double t0 = theta[2 * k], t1 = theta[2 * k + 1];
  const double S = t0 + t1;
  const double go = gout[0];
  grad[2 * k] = go * d * (-t1 / (S * S));
  grad[2 * k + 1] = go * d * (t0 / (S * S));
}

}  // namespace

int loglik_kc_waves(int32_t K, int precision) {
  const int V = precision == 0 ? 2 : 4;
  const int vecs = (K + V - 1) / V;
  if (vecs <= 2 * 64) return 0;  // one wavefront holds the row
  const int w = (vecs + 2 * 64 - 1) / (2 * 64);
  return w <= kLlMaxWaves ? w : 0;
}

size_t loglik_rows_lds(int32_t K, int precision, bool backward) {
  // f1 and q, and for the adjoint pi_s and a combine buffer: NCH 16-byte
  // vectors per lane each, and 2 C
  const size_t row = (size_t)loglik_rows_chunks(K, precision) * 64 * 16;
  return (backward ? 4 : 2) * row + 2 * (precision == 0 ? 8 : 4);
}

int loglik_rows_blocks_per_cu(int32_t K, int precision) {
  const int nch = loglik_rows_chunks(K, precision);
  const size_t lds = loglik_rows_lds(K, precision, true);
  const void* fn = nullptr;
#define RB(NC)                                                                       \
  if (nch == NC)                                                                     \
    fn = precision == 0 ? (const void*)k_loglik_rows<double, NC, true>               \
                        : (const void*)k_loglik_rows<float, NC, true>;
  RB(4)
  RB(8)
  RB(16)
#undef RB
  if (!fn) return 1;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return 1;
  int nb = 0;
  const hipError_t e =
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, kLlRowWaves * 64, lds);
  return e == hipSuccess && nb > 0 ? nb : 1;
}

int loglik_rows_chunks(int32_t K, int precision) {
  const int V = precision == 0 ? 2 : 4;
  const int vecs = (K + V - 1) / V;
  if (vecs <= 2 * 64) return 0;  // one wavefront holds the row (k_loglik)
  for (int n : {4, 8, 16})
    if (vecs <= n * 64) return n;
  return 0;
}

bool loglik_kc_shape_supported(int pairs, int prefetch) {
  return (pairs == 2 || pairs == 4 || pairs == 8) && (prefetch == 0 || prefetch == 1);
}

hipError_t launch_loglik(const LogLikArgs& a, int precision, bool backward, hipStream_t s,
                         int pairs, int prefetch, bool rows) {
  int lpr, ch;
  if (!row_config(a.K, precision, &lpr, &ch)) return hipErrorInvalidValue;
  if (a.M <= 0) return hipSuccess;
  const int nch = loglik_rows_chunks(a.K, precision);
  if (rows && nch > 0) {
    // backward: a.blocks workgroups (a.blocks beta partial rows); forward:
    // up to 2048, over the sources
    const size_t lds = loglik_rows_lds(a.K, precision, backward);
    const dim3 grid(backward ? (unsigned)a.blocks : (unsigned)std::min<int64_t>(a.M, 2048)),
        block(kLlRowWaves * 64);
#define RB(NC, B)                                                                             \
  if (nch == NC && backward == B) {                                                          \
    const void* fn = precision == 0 ? (const void*)k_loglik_rows<double, NC, B>              \
                                    : (const void*)k_loglik_rows<float, NC, B>;              \
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,       \
                                       (int)lds);                                            \
    if (e != hipSuccess) return e;                                                           \
    if (precision == 0)                                                                      \
      hipLaunchKernelGGL((k_loglik_rows<double, NC, B>), grid, block, lds, s, a);            \
    else                                                                                     \
      hipLaunchKernelGGL((k_loglik_rows<float, NC, B>), grid, block, lds, s, a);             \
    return hipGetLastError();                                                                \
  }
    RB(4, false)
    RB(8, false)
    RB(16, false)
    RB(4, true)
    RB(8, true)
    RB(16, true)
#undef RB
    return hipErrorInvalidValue;
  }
  const int waves = loglik_kc_waves(a.K, precision);
  if (waves > 0) {
    const dim3 grid(a.blocks), block(waves * 64);
#define KC(P, PF)                                                                               \
  if (pairs == P && prefetch == PF) {                                                           \
    if (precision == 0) {                                                                       \
      if (backward)                                                                             \
        hipLaunchKernelGGL((k_loglik_kc<double, 2, P, PF, true>), grid, block, 0, s, a);        \
      else                                                                                      \
        hipLaunchKernelGGL((k_loglik_kc<double, 2, P, PF, false>), grid, block, 0, s, a);       \
    } else {                                                                                    \
      if (backward)                                                                             \
        hipLaunchKernelGGL((k_loglik_kc<float, 2, P, PF, true>), grid, block, 0, s, a);         \
      else                                                                                      \
        hipLaunchKernelGGL((k_loglik_kc<float, 2, P, PF, false>), grid, block, 0, s, a);        \
    }                                                                                           \
    return hipGetLastError();                                                                   \
  }
    KC(2, false)
    KC(2, true)
    KC(4, false)
    KC(4, true)
    KC(8, false)
    KC(8, true)
#undef KC
    return hipErrorInvalidValue;
  }
  const dim3 grid(a.blocks), block(4 * 64);
#define X(L, C)                                                                              \
  if (lpr == L && ch == C) {                                                                    \
    if (precision == 0) {                                                                       \
      if (backward)                                                                             \
        hipLaunchKernelGGL((k_loglik<double, L, C, 4, true>), grid, block, 0, s, a);            \
      else                                                                                      \
        hipLaunchKernelGGL((k_loglik<double, L, C, 4, false>), grid, block, 0, s, a);           \
    } else {                                                                                    \
      if (backward)                                                                             \
        hipLaunchKernelGGL((k_loglik<float, L, C, 4, true>), grid, block, 0, s, a);             \
      else                                                                                      \
        hipLaunchKernelGGL((k_loglik<float, L, C, 4, false>), grid, block, 0, s, a);            \
    }                                                                                           \
    return hipGetLastError();                                                                   \
  }
  AMMSB_ROW_CONFIGS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_sum(const double* x, int64_t n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, x, n, out);
  return hipGetLastError();
}

hipError_t launch_theta_grad(const double* part, int32_t blocks, int32_t K, const double* theta,
                             const double* gout, double* grad, hipStream_t s) {
  hipLaunchKernelGGL(k_theta_grad, dim3((K + 63) / 64), dim3(1024), 0, s, part, blocks, K, theta,
                     gout, grad);
  return hipGetLastError();
}

}  // namespace ammsb

// ===========================================================================
// Batched chains.  Every phase of an iteration is one launch over all chains
// of a batch (a chain = one independent sampler with its own graph, state,
// K and seed); kernels read their chain's arguments from a device array, so
// the launch sequence of an iteration is fixed and is captured once as a
// hipGraph.  update_phi is a persistent kernel fed by a work queue, because
// the chains' minibatches differ in size and cost.
namespace ammsb {
namespace {

bool shape_rows_fit(int lpr, int ch, int32_t K, int precision) {
  const int V = precision == 0 ? 2 : 4;
  const int vecs = (K + V - 1) / V;
  // every vector of the row is covered, and no configuration with half the
  // lanes would cover it too (no all-idle lanes)
  return lpr * ch >= vecs && (lpr == 4 || (lpr / 2) * ch < vecs);
}

// Chain c of an item: the largest c with off[c] <= item.
__device__ __forceinline__ int chain_of(const int32_t* off, int n, int item) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// update_phi of npb minibatch nodes of one chain: nodes [item*npb,
// item*npb + npb).  Rows are read by row groups of lpr lanes (a uniform
// runtime value), CH 16-byte vectors per lane.  The workgroup's
// G = kBatchWaves*64/lpr row groups are split evenly, gpn = G/npb groups
// per node: gpn < 64/lpr packs several nodes into one wavefront, gpn >
// 64/lpr spreads a node over several wavefronts.  Each group streams whole
// neighbour rows (U in flight), reduces P_b over its lanes and accumulates
// the regrouped gradient sums of k_update_phi (acc_k = sum_b s_b pi_bk /
// Z_b and C = sum_b (e_b / Z_b - 1), so grads_k = f_k acc_k + C); the
// node's groups are combined by a butterfly inside the wavefront, then
// across its wavefronts through LDS.  The node's leader group applies the
// phi step (noise, renormalisation) and stages the new row with its
// phi_sum; k_scatter_batch publishes it once every gather of the iteration
// is done.
template <typename F, int CH, int U>
__device__ __forceinline__ void phi_batch_body(const PhiArgs& a, int item, int npb, int lpr,
                                               int rs, int count, F* s_red) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CH;
  const int rpw = 64 / lpr;
  const int G = kBatchWaves * rpw;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int g = lane / lpr, lg = lane % lpr;
  const int gpn = G / npb;
  const int gid = w * rpw + g;
  const int j = gid / gpn, r = gid % gpn;
  const int first = item * npb;  // < count: the item exists
  const int slot = first + j;
  const bool active = slot < count;
  const int32_t* nodes = a.ring.nodes + (int64_t)rs * a.ring.node_cap;
  const int32_t node = ld_g(nodes + (active ? slot : first));
  const int32_t* nb = a.ring.nbrs + ((int64_t)rs * a.ring.node_cap + (active ? slot : first)) * a.nnb;
  const F* pi = (const F*)a.pi;
  const int K = a.K;
  const F eps = (F)a.hp.epsilon;
  F q[EPL], acc[EPL];
  int koff[CH];  // first element of each of the lane's vectors (K: past the row)
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k0 = (c * lpr + lg) * V;
    koff[c] = k0 < K ? k0 : K;
  }
  // The node's own row: q_k = pi_ak f_k and S here; the leader reads it
  // again for the phi step (fewer registers in the row loop).
  const F* prow = pi + (int64_t)node * a.ld;
#pragma unroll
  for (int c = 0; c < CH; ++c) load_vec_g<F>(prow + (koff[c] < K ? koff[c] : 0), q + c * V);
  const F phi_sum = ld_g(prow + K);
  const int64_t t = ld_g(&a.st->step);
  const double eps_t = phi_step_size(a);
  F Sp = (F)0;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = koff[c] + v;
      const bool in = koff[c] < K && k < K;
      const int jj = c * V + v;
      const F pk = in ? q[jj] : (F)0;
      q[jj] = pk * (beta_at<F>(a.beta, k, K) - eps);
      Sp += pk;
      acc[jj] = (F)0;
    }
  const F S = group_sum_rt(Sp, lpr);
  const F e1 = eps, e0 = (F)1 - eps;
  const F zs1 = e1 * S, zs0 = e0 * S;
  F C = (F)0;
  const int nnb = a.nnb;
  // neighbour ids one iteration ahead, so that an iteration's row loads do
  // not wait for its id loads
  int32_t nxt[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int b = r + u * gpn;
    nxt[u] = b < nnb ? ld_g(nb + b) : node;  // past the end: own row, not counted
  }
  for (int b0 = r; b0 < nnb; b0 += U * gpn) {
    int32_t nbv[U];
    F row[U][EPL];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      nbv[u] = nxt[u];
      const F* rr = pi + (int64_t)(nbv[u] & 0x7fffffff) * a.ld;
#pragma unroll
      for (int c = 0; c < CH; ++c) load_vec_g<F>(rr + (koff[c] < K ? koff[c] : 0), row[u] + c * V);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + (U + u) * gpn;
      nxt[u] = b < nnb ? ld_g(nb + b) : node;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // q is zero past the row, so p needs no mask; acc past the row is
      // never used
      const bool ok = (b0 + u * gpn < nnb) && ((nbv[u] & 0x7fffffff) != node);
      F p = (F)0;
#pragma unroll
      for (int jj = 0; jj < EPL; ++jj) p = fma(q[jj], row[u][jj], p);
      p = group_sum_rt(p, lpr);
      const bool link = nbv[u] < 0;
      const F wz = recip(link ? p + zs1 : zs0 - p);
      if (ok) {
        const F sw = link ? wz : -wz;
#pragma unroll
        for (int jj = 0; jj < EPL; ++jj) acc[jj] = fma(row[u][jj], sw, acc[jj]);
        C += fma(link ? e1 : e0, wz, (F)-1);
      }
    }
  }
  // combine the node's groups inside the wavefront: lanes of a node are
  // contiguous, min(gpn, rpw)*lpr of them
  const int span = (gpn < rpw ? gpn : rpw) * lpr;
  for (int off = lpr; off < span; off <<= 1) {
#pragma unroll
    for (int jj = 0; jj < EPL; ++jj) acc[jj] += __shfl_xor(acc[jj], off, 64);
    C += __shfl_xor(C, off, 64);
  }
  bool leader = r == 0;
  if (gpn > rpw) {  // block-uniform: the node spans wpn wavefronts
    const int wpn = gpn / rpw;
    if ((w % wpn) != 0 && g == 0) {
#pragma unroll
      for (int jj = 0; jj < EPL; ++jj) s_red[(w * (EPL + 1) + jj) * 64 + lg] = acc[jj];
      s_red[(w * (EPL + 1) + EPL) * 64 + lg] = C;
    }
    __syncthreads();
    if ((w % wpn) == 0 && g == 0) {
      for (int q2 = 1; q2 < wpn; ++q2) {
#pragma unroll
        for (int jj = 0; jj < EPL; ++jj) acc[jj] += s_red[((w + q2) * (EPL + 1) + jj) * 64 + lg];
        C += s_red[((w + q2) * (EPL + 1) + EPL) * 64 + lg];
      }
    }
    leader = (w % wpn) == 0 && g == 0;
  }
  if (!leader || !active) return;
  // grads_k = f_k acc_k + C, and the node's pi row (past the row: unused)
  F pii[EPL];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    load_vec_g<F>(prow + (koff[c] < K ? koff[c] : 0), pii + c * V);
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int jj = c * V + v;
      acc[jj] = fma(beta_at<F>(a.beta, koff[c] + v, K) - eps, acc[jj], C);
    }
  }
  phi_epilogue<F, 0, CH>(a, slot, node, phi_sum, pii, acc, lg,
                         (F*)a.staging + (int64_t)slot * a.ld, t, eps_t, lpr);
}

// The work layout of the n positions staged in dynamic LDS: info [n] (int4)
// then offsets [n+1] (int32).
__host__ __device__ constexpr size_t batch_layout_lds(int n) {
  return (size_t)n * sizeof(int4) + (size_t)(n + 1) * sizeof(int32_t);
}

// Resident wavefronts per SIMD: five in FP32 (registers bounded to 96, one
// register spilled outside the row loop; six would spill in it), three in
// FP64 (168 registers).
template <typename F, int U>
__global__ __launch_bounds__(kBatchWaves * 64)
__attribute__((amdgpu_waves_per_eu(sizeof(F) == 4 ? 5 : 3))) void k_update_phi_batch(
    const PhiArgs* __restrict__ args, const BatchPhiShape* __restrict__ shapes, int n,
    BatchIndex ix) {
  __shared__ F s_red[kBatchWaves * (8 + 1) * 64];
  __shared__ int s_item;
  extern __shared__ int4 s_info[];  // batch_layout_lds(n) bytes
  int32_t* s_off = reinterpret_cast<int32_t*>(s_info + n);
  for (int i = threadIdx.x; i <= n; i += blockDim.x) s_off[i] = ix.items[i];
  for (int i = threadIdx.x; i < n; i += blockDim.x) s_info[i] = ix.info[i];
  int next = 0;
  if (threadIdx.x == 0) next = atomicAdd(ix.counter, 1);
  __syncthreads();
  const int total = s_off[n];
  while (true) {
    if (threadIdx.x == 0) s_item = next;
    __syncthreads();
    // uniform values in scalar registers: the chain's arguments are then
    // scalar loads and the shape switch a scalar branch
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (it >= total) break;
    // claim the next item now: its index arrives while this one runs
    if (threadIdx.x == 0) next = atomicAdd(ix.counter, 1);
    const int p = __builtin_amdgcn_readfirstlane(chain_of(s_off, n, it));
    const int c = __builtin_amdgcn_readfirstlane(s_info[p].x);
    const int rs = __builtin_amdgcn_readfirstlane(s_info[p].y);
    const int count = __builtin_amdgcn_readfirstlane(s_info[p].z);
    const PhiArgs& a = args[c];
    const int cfg = __builtin_amdgcn_readfirstlane(shapes[c].cfg);
    const int npb = __builtin_amdgcn_readfirstlane(shapes[c].npb);
    const int local = it - __builtin_amdgcn_readfirstlane(s_off[p]);
    const int lpr = 4 << (cfg % 5);  // the kBatchCfgLpr / kBatchCfgCh tables
    // neighbour rows in flight per row group: 4 of one 16-byte vector per
    // lane, 2 of two (4 KB per wavefront either way)
    if (cfg < 5)
      phi_batch_body<F, 1, 2 * U>(a, local, npb, lpr, rs, count, s_red);
    else
      phi_batch_body<F, 2, U>(a, local, npb, lpr, rs, count, s_red);
  }
}

// Work layout of the current iteration (one workgroup): per position of
// the work order, the chain's ring slot and node count, its first work item
// and scatter task, and the work queue reset.  Runs after the samplers,
// before update_phi.
__global__ __launch_bounds__(1024) void k_batch_index(const PhiArgs* __restrict__ args,
                                                      const BatchPhiShape* __restrict__ shapes,
                                                      const int32_t* __restrict__ cfg, int n,
                                                      BatchIndex ix) {
  __shared__ int32_t s_items[kMaxBatchChains], s_pub[kMaxBatchChains];
  const int p = threadIdx.x;
  if (p < n) {
    const int c = ix.order[p];
    const PhiArgs& a = args[c];
    const int rs = ring_slot(a.st);
    const int cnt = a.ring.meta[rs].node_count;
    const int rpw = 64 / (4 << (cfg[c] % 5));  // scatter: row groups per wavefront
    s_items[p] = (cnt + shapes[c].npb - 1) / shapes[c].npb;
    s_pub[p] = (cnt + rpw - 1) / rpw;
    ix.info[p] = make_int4(c, rs, cnt, 0);
  }
  __syncthreads();
  if (p == 0) {
    int32_t ti = 0, tp = 0;
    for (int i = 0; i < n; ++i) {
      ix.items[i] = ti;
      ix.pub[i] = tp;
      ti += s_items[i];
      tp += s_pub[i];
    }
    ix.items[n] = ti;
    ix.pub[n] = tp;
    *ix.counter = 0;
  }
}

// update_pi of every chain's minibatch nodes: pi[node] = staged row (K
// elements and phi_sum, in 16-byte vectors; ld is a multiple of 128 B).  A
// wavefront takes 64/lpr consecutive nodes of one chain, one per row group
// of lpr lanes (the chain's update_beta row shape).  Runs after every
// neighbour-row read of the iteration (previous launch), as the reference
// refreshes pi only after the node loop (mcmc_sampler_stochastic.cc:162-166).
template <typename F>
__global__ __launch_bounds__(256) void k_scatter_batch(const PhiArgs* __restrict__ args,
                                                       const int32_t* __restrict__ cfg, int n,
                                                       BatchIndex ix) {
  constexpr int V = VT<F>::V;
  typedef typename NTV<F>::type T;
  extern __shared__ int4 s_info[];  // batch_layout_lds(n) bytes
  int32_t* s_off = reinterpret_cast<int32_t*>(s_info + n);
  for (int i = threadIdx.x; i <= n; i += blockDim.x) s_off[i] = ix.pub[i];
  for (int i = threadIdx.x; i < n; i += blockDim.x) s_info[i] = ix.info[i];
  __syncthreads();
  const int total = s_off[n];
  const int lane = threadIdx.x & 63;
  const int gw = blockIdx.x * 4 + (threadIdx.x >> 6), nw = gridDim.x * 4;
  for (int it = gw; it < total; it += nw) {
    const int p = __builtin_amdgcn_readfirstlane(chain_of(s_off, n, it));
    const int c = __builtin_amdgcn_readfirstlane(s_info[p].x);
    const int rs = __builtin_amdgcn_readfirstlane(s_info[p].y);
    const int count = __builtin_amdgcn_readfirstlane(s_info[p].z);
    const int lpr = 4 << (__builtin_amdgcn_readfirstlane(cfg[c]) % 5);  // kBatchCfgLpr
    const PhiArgs& a = args[c];
    const int slot = (it - __builtin_amdgcn_readfirstlane(s_off[p])) * (64 / lpr) + lane / lpr;
    if (slot >= count) continue;
    const int32_t node = ld_g(a.ring.nodes + (int64_t)rs * a.ring.node_cap + slot);
    const int nv = (a.K + 1 + V - 1) / V;
    gptr<const T> src = (gptr<const T>)((const F*)a.staging + (int64_t)slot * a.ld);
    gptr<T> dst = (gptr<T>)((F*)a.pi + (int64_t)node * a.ld);
    for (int i = lane % lpr; i < nv; i += lpr) __builtin_nontemporal_store(src[i], dst + i);
  }
}

// Samplers of iterations t0 + j (j < fills) of every chain into ring slots
// (slot0 + j) mod slots: workgroup (chain, j).
__global__ __launch_bounds__(kMbThreads) void k_minibatch_batch(const SamplerArgs* __restrict__ args,
                                                               int64_t t0, int slot0, int slots) {
  const int j = blockIdx.y;
  minibatch_body(args[blockIdx.x], (slot0 + j) % slots, t0 + j, false);
}

__global__ __launch_bounds__(kNbrWaves * 64) void k_neighbors_batch(
    const NeighborArgs* __restrict__ args, int slot0, int fills, int slots) {
  const int c = blockIdx.y / fills, j = blockIdx.y % fills;
  if (args[c].use_lds)
    neighbors_body<kNbrWaves, true>(args[c], 0, INT_MAX, 0, blockIdx.x, (slot0 + j) % slots);
  else
    neighbors_body<kNbrWaves, false>(args[c], 0, INT_MAX, 0, blockIdx.x, (slot0 + j) % slots);
}

template <typename F>
__global__ __launch_bounds__(kBetaWaves * 64) void k_beta_batch(const BetaArgs* __restrict__ args,
                                                                const int32_t* __restrict__ cfg) {
  const BetaArgs& a = args[blockIdx.y];
  if ((int)blockIdx.x >= a.blocks) return;
  switch (cfg[blockIdx.y]) {
#define B(I)                                                                  \
  case I:                                                                     \
    beta_body<F, kBatchCfgLpr[I], kBatchCfgCh[I], kBetaWaves>(a, blockIdx.x); \
    break;
    B(0) B(1) B(2) B(3) B(4) B(5) B(6) B(7) B(8) B(9)
#undef B
    default:
      break;
  }
}

__global__ __launch_bounds__(kThetaThreads) void k_theta_batch(const ThetaArgs* __restrict__ args) {
  const ThetaArgs& a = args[blockIdx.y];
  const int gx = (a.K + kThetaK - 1) / kThetaK;
  if ((int)blockIdx.x >= gx) return;
  theta_body(a, blockIdx.x, gx);
}

}  // namespace

bool batch_cfg_valid(int cfg, int32_t K, int precision) {
  if (cfg < 0 || cfg >= kBatchCfgs) return false;
  return shape_rows_fit(kBatchCfgLpr[cfg], kBatchCfgCh[cfg], K, precision);
}

int batch_cfg_default(int32_t K, int precision) {
  int lpr, ch;
  if (!row_config(K, precision, &lpr, &ch) || ch > 2) return -1;
  for (int i = 0; i < kBatchCfgs; ++i)
    if (kBatchCfgLpr[i] == lpr && kBatchCfgCh[i] == ch) return i;
  return -1;
}

hipError_t launch_minibatch_batch(const SamplerArgs* args, int n, int64_t t0, int slot0, int fills,
                                  int slots, size_t lds, hipStream_t s) {
  if (n < 1 || fills < 1 || slots < fills || slot0 < 0 || t0 < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_minibatch_batch, dim3(n, fills), dim3(kMbThreads), lds, s, args, t0, slot0,
                     slots);
  return hipGetLastError();
}

hipError_t launch_neighbors_batch(const NeighborArgs* args, int n, int slot0, int fills, int slots,
                                  int max_node_cap, size_t lds, hipStream_t s) {
  if (n < 1 || fills < 1 || slots < fills || slot0 < 0) return hipErrorInvalidValue;
  const int blocks = (max_node_cap + kNbrWaves - 1) / kNbrWaves;
  hipLaunchKernelGGL(k_neighbors_batch, dim3(blocks, n * fills), dim3(kNbrWaves * 64), lds, s,
                     args, slot0, fills, slots);
  return hipGetLastError();
}

hipError_t launch_batch_index(const PhiArgs* args, const BatchPhiShape* shapes,
                              const int32_t* cfg, int n, BatchIndex ix, hipStream_t s) {
  if (n < 1 || n > kMaxBatchChains) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_batch_index, dim3(1), dim3(1024), 0, s, args, shapes, cfg, n, ix);
  return hipGetLastError();
}

hipError_t launch_update_phi_batch(const PhiArgs* args, const BatchPhiShape* shapes, int n,
                                   BatchIndex ix, int precision, int grid, hipStream_t s) {
  if (n < 1 || n > kMaxBatchChains || grid < 1) return hipErrorInvalidValue;
  const size_t lds = batch_layout_lds(n);
  if (precision == 0)
    hipLaunchKernelGGL((k_update_phi_batch<double, 2>), dim3(grid), dim3(kBatchWaves * 64), lds,
                       s, args, shapes, n, ix);
  else
    hipLaunchKernelGGL((k_update_phi_batch<float, 2>), dim3(grid), dim3(kBatchWaves * 64), lds, s,
                       args, shapes, n, ix);
  return hipGetLastError();
}

int update_phi_batch_blocks_per_cu(int n, int precision) {
  int nb = 0;
  const size_t lds = batch_layout_lds(n);
  const hipError_t e =
      precision == 0
          ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, (const void*)k_update_phi_batch<double, 2>, kBatchWaves * 64, lds)
          : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, (const void*)k_update_phi_batch<float, 2>, kBatchWaves * 64, lds);
  return e == hipSuccess && nb > 0 ? nb : 1;
}

hipError_t launch_scatter_batch(const PhiArgs* args, const int32_t* cfg, int n, BatchIndex ix,
                                int precision, int grid, hipStream_t s) {
  if (n < 1 || n > kMaxBatchChains || grid < 1) return hipErrorInvalidValue;
  const size_t lds = batch_layout_lds(n);
  if (precision == 0)
    hipLaunchKernelGGL(k_scatter_batch<double>, dim3(grid), dim3(256), lds, s, args, cfg, n, ix);
  else
    hipLaunchKernelGGL(k_scatter_batch<float>, dim3(grid), dim3(256), lds, s, args, cfg, n, ix);
  return hipGetLastError();
}

int beta_batch_blocks_per_cu(int precision) {
  int nb = 0;
  const hipError_t e =
      precision == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                           &nb, (const void*)k_beta_batch<double>, kBetaWaves * 64, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                           &nb, (const void*)k_beta_batch<float>, kBetaWaves * 64, 0);
  return e == hipSuccess && nb > 0 ? nb : 1;
}

hipError_t launch_beta_batch(const BetaArgs* args, const int32_t* cfg, int n, int max_blocks,
                             int precision, hipStream_t s) {
  if (n < 1 || max_blocks < 1) return hipErrorInvalidValue;
  if (precision == 0)
    hipLaunchKernelGGL(k_beta_batch<double>, dim3(max_blocks, n), dim3(kBetaWaves * 64), 0, s,
                       args, cfg);
  else
    hipLaunchKernelGGL(k_beta_batch<float>, dim3(max_blocks, n), dim3(kBetaWaves * 64), 0, s,
                       args, cfg);
  return hipGetLastError();
}

hipError_t launch_theta_batch(const ThetaArgs* args, int n, int max_K, hipStream_t s) {
  if (n < 1 || max_K < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_theta_batch, dim3((max_K + kThetaK - 1) / kThetaK, n), dim3(kThetaThreads), 0,
                     s, args);
  return hipGetLastError();
}

}  // namespace ammsb

namespace ammsb {
hipError_t prepare_minibatch_lds(size_t bytes) {
  static size_t granted = 64 * 1024 - kMinibatchStaticLds;
  if (bytes <= granted) return hipSuccess;
  hipError_t e = hipFuncSetAttribute((const void*)k_minibatch,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_minibatch_batch,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) granted = bytes;
  return e;
}
}  // namespace ammsb
