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
#include <limits.h>
#include <math.h>

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

// Row of K elements into registers: vector vi = c*LPR + lg, masked to k < K.
template <typename F, int LPR, int CH>
__device__ __forceinline__ void load_row(const F* row, int K, int lg, F* out) {
  constexpr int V = VT<F>::V;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int k0 = (c * LPR + lg) * V;
    if (k0 < K) {
      load_vec<F>(row + k0, out + c * V);
#pragma unroll
      for (int v = 0; v < V; ++v)
        if (k0 + v >= K) out[c * V + v] = (F)0;
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) out[c * V + v] = (F)0;
    }
  }
}

template <int LPR, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = LPR / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
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
    const uint64_t k = g.hkeys[s];
    if (k == key) return g.hvals[s];
    if (k == kHashEmpty) return -1;
    s = (s + 1) & g.hmask;
  }
}

__device__ __forceinline__ bool row_contains(const int32_t* row, int64_t n, int32_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (row[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && row[lo] == v;
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

__device__ __forceinline__ int load_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

constexpr int kMbThreads = 1024;
constexpr int kMaxSources = 1024;
constexpr int kRowLds = 4096;

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
// Attempts are processed 1024 at a time; within a batch, a value is won by
// its smallest attempt index (atomicMin in the de-dup table), and accepted
// values are appended in attempt order through a block scan, which yields
// the sequential result.
__global__ __launch_bounds__(kMbThreads) void k_minibatch(SamplerArgs a) {
  __shared__ int s_wave[kMbThreads / 64];
  __shared__ int s_row[kRowLds];
  __shared__ int s_srcs[kMaxSources];
  __shared__ int s_nodes;
  const int tid = threadIdx.x;
  DevState* st = a.st;
  const int64_t t = a.step_override >= 0 ? a.step_override : st->step;
  const uint32_t t32 = (uint32_t)t;
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
  else
    flag = -1;
  int64_t c = 0;
  double scale = 0.0;
  if (tid == 0) s_nodes = 0;
  if (flag == 0) {
    const int32_t node = (int32_t)randint_mod(draw_u64(a.key, 0, 0, t32, TAG_SOURCE), N);
    const int64_t rs = a.g.row_ptr[node];
    const int64_t deg = a.g.row_ptr[node + 1] - rs;
    const bool in_lds = deg <= kRowLds;
    if (in_lds)
      for (int64_t i = tid; i < deg; i += kMbThreads) s_row[i] = a.g.cols[rs + i];
    for (int64_t i = tid; i < a.table_cap; i += kMbThreads) a.table[i] = make_int2(-1, INT_MAX);
    __threadfence();
    __syncthreads();
    const uint32_t cmask = (uint32_t)(a.table_cap - 1);
    const int64_t max_batches = 4096 + 64 * ((m + kMbThreads - 1) / kMbThreads);
    for (int64_t batch = 0;; ++batch) {
      if (batch >= max_batches) {
        err |= ERR_NONLINK_ATTEMPTS;
        break;
      }
      const uint32_t ai = (uint32_t)(batch * kMbThreads + tid);
      const int32_t v = (int32_t)randint_mod(draw_u64(a.key, ai, 0, t32, TAG_NONLINK), N);
      bool valid = v != node;
      const uint32_t lo = (uint32_t)min(node, v), hi = (uint32_t)max(node, v);
      if (valid) valid = !(in_lds ? row_contains(s_row, deg, v) : row_contains(a.g.cols + rs, deg, v));
      if (valid) valid = hash_find(a.g, lo, hi) < 0;
      int slot = -1;
      if (valid) {
        slot = table_insert(a.table, cmask, v);
        atomicMin(&a.table[slot].y, (int)ai);
      }
      __threadfence();
      __syncthreads();
      const int win = valid && load_agent(&a.table[slot].y) == (int)ai;
      int total;
      const int pos = block_scan(win, s_wave, &total);
      if (win && c + pos < m) {
        a.mb_edges[2 * (c + pos)] = (int32_t)lo;
        a.mb_edges[2 * (c + pos) + 1] = (int32_t)hi;
        a.mb_y[c + pos] = 0;
      }
      c = min(m, c + (int64_t)total);
      if (c >= m) break;
    }
    const double dN = (double)a.g.N, dE = (double)a.g.E;
    scale = dN * dN / (double)m;
    scale *= dE / (dN * (dN - 1.0) / 2.0 - dE);
  } else if (flag == 1) {
    int64_t r = 0;
    const int64_t max_rounds = 1 << 20;
    for (; r < max_rounds; ++r) {
      c = 0;
      int nsrc = 0;
      for (int64_t s = 0; s < a.max_source; ++s) {
        const int32_t node =
            (int32_t)randint_mod(draw_u64(a.key, (uint32_t)r, (uint32_t)s, t32, TAG_SOURCE), N);
        const int64_t rs = a.g.row_ptr[node];
        const int64_t deg = a.g.row_ptr[node + 1] - rs;
        if (s > 0 && c + deg > m) break;
        bool dup = false;
        for (int i = 0; i < nsrc; ++i) dup |= s_srcs[i] == node;
        __syncthreads();
        if (tid == 0) s_srcs[nsrc] = node;
        __syncthreads();
        ++nsrc;
        if (dup) continue;
        if (c + deg > a.edge_cap) {
          err |= ERR_MB_OVERFLOW;
          break;
        }
        for (int64_t j0 = 0; j0 < deg; j0 += kMbThreads) {
          const int64_t j = j0 + tid;
          int keep = 0;
          uint32_t lo = 0, hi = 0;
          if (j < deg) {
            const int32_t v = a.g.cols[rs + j];
            lo = (uint32_t)min(node, v);
            hi = (uint32_t)max(node, v);
            keep = hash_find(a.g, lo, hi) < 0;
            for (int i = 0; i + 1 < nsrc && keep; ++i) keep = s_srcs[i] != v;
          }
          int total;
          const int pos = block_scan(keep, s_wave, &total);
          if (keep) {
            a.mb_edges[2 * (c + pos)] = (int32_t)lo;
            a.mb_edges[2 * (c + pos) + 1] = (int32_t)hi;
            a.mb_y[c + pos] = 1;
          }
          c += total;
        }
      }
      if (c > 0 || err) break;
    }
    if (r >= max_rounds) err |= ERR_LINK_ROUNDS;
    scale = a.max_source == 1 ? (double)a.g.N : (double)a.g.E / (double)(c > 0 ? c : 1);
  } else {
    err |= ERR_BAD_STRATEGY;
  }
  __threadfence();
  __syncthreads();
  // nodes_in_batch: unique endpoints (mcmc_sampler_stochastic.cc:315-327).
  // Order is not part of the contract: every per-node draw is keyed by the
  // node id, not by its position.
  const uint32_t epoch = st->epoch + 1u;
  for (int64_t e = tid; e < c; e += kMbThreads) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int32_t v = a.mb_edges[2 * e + q];
      if (atomicExch(&a.stamp[v], epoch) != epoch) {
        const int pos = atomicAdd(&s_nodes, 1);
        if (pos < a.node_cap) a.nodes[pos] = v;
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    if (s_nodes > a.node_cap) err |= ERR_MB_OVERFLOW;
    st->epoch = epoch;
    st->flag = flag;
    st->mb_count = (int32_t)c;
    st->node_count = min(s_nodes, (int)a.node_cap);
    st->scale = scale;
    st->error |= err;
    if (a.step_override < 0) {
      st->nodes_total += min(s_nodes, (int)a.node_cap);
      st->edges_total += c;
    }
  }
}

// ---------------------------------------------------------------------------
// Neighbour sampler: one wavefront per minibatch node.  Sequential
// definition: attempt j = 0,1,...: v = u64(NEIGHBOR; j,node,t) % N, accepted
// iff v != node, (node,v) not held-out/test, v not accepted before; until
// n+1 accepted (mcmc_sampler_stochastic.h:93-107: n+1 draws, links allowed).
// Output order = acceptance order; bit 31 flags a link (the y_ab of
// update_phi, mcmc_sampler_stochastic.cc:279-283).
template <int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_neighbors(NeighborArgs a, int32_t slot_offset,
                                                          int32_t slot_end) {
  extern __shared__ int2 s_tab[];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int slot = slot_offset + blockIdx.x * WAVES + w;
  const int count = a.st->node_count;
  if (slot >= count || slot >= slot_end) return;
  const int64_t t = a.step_override >= 0 ? a.step_override : a.st->step;
  const uint32_t t32 = (uint32_t)t;
  const uint32_t N = (uint32_t)a.g.N;
  const int32_t node = a.nodes[slot];
  int2* tab = a.use_lds ? s_tab + (size_t)w * a.tcap
                        : a.gtable + (size_t)(slot - slot_offset) * a.tcap;
  for (int i = lane; i < a.tcap; i += 64) tab[i] = make_int2(-1, INT_MAX);
  __threadfence_block();
  __builtin_amdgcn_wave_barrier();
  const uint32_t cmask = (uint32_t)(a.tcap - 1);
  const int64_t rs = a.g.row_ptr[node];
  const int64_t deg = a.g.row_ptr[node + 1] - rs;
  int32_t* out = a.nbrs + (int64_t)slot * a.nnb;
  int c = 0;
  const int max_batches = 4096 + 64 * ((a.nnb + 63) / 64);
  for (int batch = 0; c < a.nnb; ++batch) {
    if (batch >= max_batches) {
      if (lane == 0) atomicOr(&a.st->error, ERR_NEIGHBOR_ATTEMPTS);
      break;
    }
    const uint32_t ai = (uint32_t)(batch * 64 + lane);
    const int32_t v = (int32_t)randint_mod(draw_u64(a.key, ai, (uint32_t)node, t32, TAG_NEIGHBOR), N);
    bool valid = v != node;
    if (valid) valid = hash_find(a.g, (uint32_t)min(node, v), (uint32_t)max(node, v)) < 0;
    int s = -1;
    if (valid) {
      s = table_insert(tab, cmask, v);
      atomicMin(&tab[s].y, (int)ai);
    }
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
    const bool win = valid && (a.use_lds ? tab[s].y : load_agent(&tab[s].y)) == (int)ai;
    const uint64_t mk = __ballot(win);
    const int pos = c + __popcll(mk & ((1ull << lane) - 1ull));
    if (win && pos < a.nnb) {
      const bool lk = row_contains(a.g.cols + rs, deg, v);
      out[pos] = v | (lk ? (int32_t)0x80000000 : 0);
    }
    c += __popcll(mk);
  }
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
template <typename F, int LPR, int CH, int WAVES, int U>
__global__ __launch_bounds__(WAVES * 64) void k_update_phi(PhiArgs a) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CH;
  constexpr int RPW = 64 / LPR;
  constexpr int STRIDE = WAVES * RPW;
  __shared__ F s_acc[(WAVES > 1 ? WAVES - 1 : 1) * EPL * LPR];
  const int slot = blockIdx.x;
  const int count = a.count_dev ? *a.count_dev : a.count_host;
  if (slot >= count) return;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int g = lane / LPR, lg = lane % LPR;
  const F* pi = (const F*)a.pi;
  const int K = a.K;
  const int32_t node = a.nodes[slot];
  const F* prow = pi + (int64_t)node * a.ld;
  const F eps = (F)a.hp.epsilon;

  F pii[EPL], f1[EPL], f0[EPL], acc[EPL];
  load_row<F, LPR, CH>(prow, K, lg, pii);
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = (c * LPR + lg) * V + v;
      const F bk = k < K ? (F)a.beta[k] : (F)0;
      f1[c * V + v] = bk - eps;
      f0[c * V + v] = eps - bk;
      acc[c * V + v] = (F)0;
    }
  const F e1 = eps, e0 = (F)1 - eps;
  const int32_t* nb_base = a.nbrs + (int64_t)slot * a.nnb;

  for (int b0 = w * RPW + g; b0 < a.nnb; b0 += U * STRIDE) {
    F row[U][EPL];
    int lab[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = b0 + u * STRIDE;
      const int32_t nb = b < a.nnb ? nb_base[b] : node;
      const int32_t v = nb & 0x7fffffff;
      ok[u] = (b < a.nnb) && (v != node);
      lab[u] = (int)((uint32_t)nb >> 31);
      load_row<F, LPR, CH>(pi + (int64_t)v * a.ld, K, lg, row[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const F e = lab[u] ? e1 : e0;
      F p = (F)0;
#pragma unroll
      for (int j = 0; j < EPL; ++j) {
        const F f = lab[u] ? f1[j] : f0[j];
        row[u][j] = row[u][j] * f + e;  // x_bk
        p += pii[j] * row[u][j];
      }
      p = group_sum<LPR>(p);
      const F wz = (F)1 / p;
      if (ok[u]) {
#pragma unroll
        for (int j = 0; j < EPL; ++j) acc[j] += row[u][j] * wz - (F)1;
      }
    }
  }
  if (RPW > 1) {
#pragma unroll
    for (int j = 0; j < EPL; ++j) acc[j] = cross_group_sum<LPR>(acc[j]);
  }
  if (WAVES > 1) {
    if (w > 0 && g == 0) {
#pragma unroll
      for (int j = 0; j < EPL; ++j) s_acc[((w - 1) * EPL + j) * LPR + lg] = acc[j];
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int q = 1; q < WAVES; ++q)
#pragma unroll
      for (int j = 0; j < EPL; ++j) acc[j] += s_acc[((q - 1) * EPL + j) * LPR + lg];
  }
  if (g != 0) return;

  // epilogue: phi step (mcmc_sampler_stochastic_distr.cc:1207-1226)
  const int64_t t = a.st->step;
  const double eps_t_d = a.eps_override >= 0.0 ? a.eps_override : eps_of_step(a.hp, t);
  const F eps_t = (F)eps_t_d;
  const F half_eps = eps_t / (F)2;
  const F alpha = (F)a.hp.alpha;
  const F Nn = (F)a.hp.Nn;
  const F phi_sum = prow[K];
  const F* noise = (const F*)a.noise;
  F phin[EPL];
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
#pragma unroll
    for (int v = 0; v < V; v += 2) {
      const int k = (c * LPR + lg) * V + v;
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
        const F grad = acc[j] / phi_sum;
        F val = fabs(phik + half_eps * (alpha - phik + Nn * grad) + sqrt(eps_t * phik) * (h ? xi1 : xi0));
        if (val < (F)1e-24) val = (F)1e-24;
        if (k + h >= K) val = (F)0;
        phin[j] = val;
        s += (double)val;
      }
    }
  }
  s = group_sum<LPR>(s);
  F* stg = (F*)a.staging + (int64_t)slot * a.ld;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = (c * LPR + lg) * V + v;
      if (k < K) stg[k] = (F)((double)phin[c * V + v] / s);
    }
  if (lg == 0) stg[K] = (F)s;
}

// Publish staged rows: pi[nodes[slot]] = staging[slot] (K+1 elements).
template <typename F>
__global__ __launch_bounds__(256) void k_scatter(ScatterArgs a) {
  const int count = a.count_dev ? *a.count_dev : a.count_host;
  const int slot = blockIdx.x;
  if (slot >= count) return;
  constexpr int V = VT<F>::V;
  const int nv = (a.K + 1 + V - 1) / V;
  const F* src = (const F*)a.staging + (int64_t)slot * a.ld;
  F* dst = (F*)a.pi + (int64_t)a.nodes[slot] * a.ld;
  typedef typename VT<F>::type T;
  for (int i = threadIdx.x; i < nv; i += blockDim.x)
    reinterpret_cast<T*>(dst)[i] = reinterpret_cast<const T*>(src)[i];
}

// ---------------------------------------------------------------------------
// update_beta gradient (mcmc_sampler_stochastic.cc:212-240): per edge
//   f_k = pi_ak pi_bk, probs_k = beta_k f_k (y=1) or (1-beta_k) f_k (y=0)
//   prob_0 = (eps or 1-eps) (1 - sum_k f_k), Z = sum_k probs_k + prob_0
//   G_k0 += probs_k/Z ((1-y)/theta_k0 - 1/S_k)
//   G_k1 += probs_k/Z (y/theta_k1 - 1/S_k),   S_k = theta_k0 + theta_k1
// accumulated in double; one partial [2][K] per wavefront.
template <typename F, int LPR, int CH, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void k_beta(BetaArgs a) {
  constexpr int V = VT<F>::V;
  constexpr int EPL = V * CH;
  constexpr int RPW = 64 / LPR;
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int g = lane / LPR, lg = lane % LPR;
  const int K = a.K;
  const F* pi = (const F*)a.pi;
  const F eps = (F)a.epsilon;
  const int count = a.count_dev ? *a.count_dev : a.count_host;
  F bt[EPL];
  double c0y[EPL], c0n[EPL], c1y[EPL], c1n[EPL], g0[EPL], g1[EPL];
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = (c * LPR + lg) * V + v;
      const int j = c * V + v;
      double th0 = 1.0, th1 = 1.0;
      if (k < K) {
        th0 = a.theta[2 * k];
        th1 = a.theta[2 * k + 1];
      }
      const double inv_s = 1.0 / (th0 + th1);
      bt[j] = k < K ? (F)a.beta[k] : (F)0;
      c0y[j] = 0.0 / th0 - inv_s;
      c0n[j] = 1.0 / th0 - inv_s;
      c1y[j] = 1.0 / th1 - inv_s;
      c1n[j] = 0.0 / th1 - inv_s;
      g0[j] = 0.0;
      g1[j] = 0.0;
    }
  const int gstride = a.blocks * WAVES * RPW;
  for (int e = (blockIdx.x * WAVES + w) * RPW + g; e < count; e += gstride) {
    const int32_t ea = a.edges[2 * e], eb = a.edges[2 * e + 1];
    const int y = a.y[e] ? 1 : 0;
    F ra[EPL], rb[EPL];
    load_row<F, LPR, CH>(pi + (int64_t)ea * a.ld, K, lg, ra);
    load_row<F, LPR, CH>(pi + (int64_t)eb * a.ld, K, lg, rb);
    F pisum = (F)0, psum = (F)0;
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
      const F f = ra[j] * rb[j];
      pisum += f;
      ra[j] = y ? bt[j] * f : ((F)1 - bt[j]) * f;  // probs_k
      psum += ra[j];
    }
    pisum = group_sum<LPR>(pisum);
    psum = group_sum<LPR>(psum);
    const F prob0 = (y ? eps : (F)1 - eps) * ((F)1 - pisum);
    const F Z = psum + prob0;
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
      const double f = (double)(ra[j] / Z);
      g0[j] += f * (y ? c0y[j] : c0n[j]);
      g1[j] += f * (y ? c1y[j] : c1n[j]);
    }
  }
  if (RPW > 1) {
#pragma unroll
    for (int j = 0; j < EPL; ++j) {
      g0[j] = cross_group_sum<LPR>(g0[j]);
      g1[j] = cross_group_sum<LPR>(g1[j]);
    }
  }
  if (g != 0) return;
  // one partial row per wavefront: [blocks * WAVES][2K]
  double* out = a.partials + ((int64_t)blockIdx.x * WAVES + w) * 2 * K;
#pragma unroll
  for (int c = 0; c < CH; ++c)
#pragma unroll
    for (int v = 0; v < V; ++v) {
      const int k = (c * LPR + lg) * V + v;
      if (k < K) {
        out[k] = g0[c * V + v];
        out[K + k] = g1[c * V + v];
      }
    }
}

// theta step (mcmc_sampler_stochastic.cc:245-266), one workgroup.
__global__ __launch_bounds__(256) void k_theta(ThetaArgs a) {
  DevState* st = a.st;
  const int64_t t = st->step;
  const double eps_t = a.eps_override >= 0.0 ? a.eps_override : eps_of_step(a.hp, t);
  const double scale = a.scale_override >= 0.0 ? a.scale_override : st->scale;
  const int K = a.K;
  for (int k = threadIdx.x; k < K; k += blockDim.x) {
    double G0 = 0.0, G1 = 0.0;
    for (int b = 0; b < a.blocks; ++b) {
      G0 += a.partials[(int64_t)b * 2 * K + k];
      G1 += a.partials[(int64_t)b * 2 * K + K + k];
    }
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
  if (a.bump_step) {
    __syncthreads();
    if (threadIdx.x == 0) st->step = t + 1;
  }
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
      bt[c * V + v] = k < K ? (F)a.beta[k] : (F)0;
    }
  double ll1 = 0.0, ll0 = 0.0, n1 = 0.0, n0 = 0.0;
  const double cnt = a.avg_count;
  const int64_t gstride = (int64_t)a.blocks * WAVES * RPW;
  for (int64_t e = ((int64_t)blockIdx.x * WAVES + w) * RPW + g; e < a.H; e += gstride) {
    const int32_t ea = a.edges[2 * e], eb = a.edges[2 * e + 1];
    const bool y = a.labels[e] != 0;
    F ra[EPL], rb[EPL];
    load_row<F, LPR, CH>(pi + (int64_t)ea * a.ld, K, lg, ra);
    load_row<F, LPR, CH>(pi + (int64_t)eb * a.ld, K, lg, rb);
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
    if (lg == 0) {
      const double r = (a.R[e] * (cnt - 1.0) + (double)s) / cnt;
      a.R[e] = r;
      const double lr = log(r);
      if (y) {
        ll1 += lr;
        n1 += 1.0;
      } else {
        ll0 += lr;
        n0 += 1.0;
      }
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
__global__ void k_set_step(DevState* st, int64_t step) {
  if (threadIdx.x == 0) {
    st->step = step;
    st->error = 0;
  }
}

constexpr int kPhiWaves = 4;
constexpr int kPhiUnroll = 2;

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

hipError_t launch_minibatch(const SamplerArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_minibatch, dim3(1), dim3(kMbThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_neighbors(const NeighborArgs& a, int32_t grid_nodes, hipStream_t s) {
  if (a.use_lds) {
    const size_t lds = (size_t)kNbrWaves * a.tcap * sizeof(int2);
    const int blocks = (grid_nodes + kNbrWaves - 1) / kNbrWaves;
    hipLaunchKernelGGL(k_neighbors<kNbrWaves>, dim3(blocks), dim3(kNbrWaves * 64), lds, s, a, 0,
                       grid_nodes);
  } else {
    // global de-dup tables exist for gtable_slots nodes at a time
    for (int32_t off = 0; off < grid_nodes; off += a.gtable_slots) {
      const int n = std::min<int32_t>(a.gtable_slots, grid_nodes - off);
      hipLaunchKernelGGL(k_neighbors<1>, dim3(n), dim3(64), 0, s, a, off, off + n);
    }
  }
  return hipGetLastError();
}

hipError_t launch_update_phi(const PhiArgs& a, int precision, int32_t grid_nodes,
                             hipStream_t s) {
  int lpr, ch;
  if (!row_config(a.K, precision, &lpr, &ch)) return hipErrorInvalidValue;
  if (grid_nodes <= 0) return hipSuccess;
#define X(L, C)                                                                           \
  if (lpr == L && ch == C) {                                                              \
    if (precision == 0)                                                                   \
      hipLaunchKernelGGL((k_update_phi<double, L, C, kPhiWaves, kPhiUnroll>),             \
                         dim3(grid_nodes), dim3(kPhiWaves * 64), 0, s, a);                \
    else                                                                                  \
      hipLaunchKernelGGL((k_update_phi<float, L, C, kPhiWaves, kPhiUnroll>),              \
                         dim3(grid_nodes), dim3(kPhiWaves * 64), 0, s, a);                \
    return hipGetLastError();                                                             \
  }
  AMMSB_ROW_CONFIGS(X)
#undef X
  return hipErrorInvalidValue;
}

hipError_t launch_scatter(const ScatterArgs& a, int precision, int32_t grid_nodes,
                          hipStream_t s) {
  if (grid_nodes <= 0) return hipSuccess;
  if (precision == 0)
    hipLaunchKernelGGL(k_scatter<double>, dim3(grid_nodes), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_scatter<float>, dim3(grid_nodes), dim3(256), 0, s, a);
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
  hipLaunchKernelGGL(k_theta, dim3(1), dim3(256), 0, s, a);
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

hipError_t launch_set_state_step(DevState* st, int64_t step, hipStream_t s) {
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, st, step);
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
