// ===========================================================================
// ORACLE -- test infrastructure only.
//
// A CPU restatement of the reference's SGRLD a-MMSB hot path, used by
// tests/ (as the checker), by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg.  Nothing in the product (mcmc_for_ammsb_amd/) includes,
// links or calls this file.
//
// What it restates, function by function (all paths relative to the
// reference repository root):
//   orc_update_phi   MCMCSamplerStochasticDistributed::update_phi_node
//                    c++/src/mcmc/learning/mcmc_sampler_stochastic_distr.cc:1158-1227
//                    (== MCMCSamplerStochastic::update_phi,
//                     mcmc_sampler_stochastic.cc:269-313, in the pi ++ phi_sum
//                     record layout), followed by pi_from_phi (:743-751) for
//                    every node (update_pi, mcmc_sampler_stochastic.cc:162-166)
//   orc_update_beta  MCMCSamplerStochastic::update_beta
//                    mcmc_sampler_stochastic.cc:200-267
//   orc_perplexity   Learner::cal_perplexity, learner.cc:138-183, with
//                    Learner::cal_edge_likelihood, learner.h:100-139
//   orc_eps_t        get_eps_t, mcmc_sampler_stochastic.h:112-115
//   orc_run          MCMCSamplerStochastic::run, mcmc_sampler_stochastic.cc:93-183
//   orc_phi_grads / orc_beta_grads   the closed-form gradients of update_phi
//                    (distr.cc:1165-1197) and update_beta (:202-240)
//   orc_sample_minibatch   Network::stratified_random_node_sampling{,_links,
//                    _nonlinks}, network.cc:274-421, and random_edge_sampling
//                    with sample_random_edges / sample_full_training_set,
//                    network.cc:237-268, 448-500, with the counter-based draw
//                    addressing of this project (see below)
//   orc_sample_neighbors   sample_neighbor_nodes, mcmc_sampler_stochastic.h:87-108
//   orc_init_theta / orc_init_phi_row   init()'s state draw,
//                    mcmc_sampler_stochastic.cc:54-76 (theta = gamma(eta0, eta1,
//                    K, 2), GSL parameter order of Random::gamma,
//                    random.cc:472-511; phi = gamma(1, 1, N, K); pi = phi
//                    row-normalised)
// With threads > 1, orc_run parallelises the loops the distributed reference
// runs under OpenMP (DrawNeighbors, update_phi_node, beta_calc_grads,
// cal_perplexity: mcmc_sampler_stochastic_distr.cc:1068,1142,1378,1543) with
// per-thread accumulators summed in thread order.
// Float is a template parameter (config.h.in:14-20); the code keeps the
// reference's operation order and accumulation types (np::sum is a
// sequential std::accumulate in Float, np.h:34-37; pi_from_phi accumulates
// in double, distr.cc:745).
//
// Random draws.  The reference draws from one sequential xorshift128+ stream
// (random.cc:136-176).  This project replaces that stream by counter-based
// Philox4x32-10 (Salmon et al., SC'11): every draw is addressed by
// (idx, id, iter, tag) -- see the sequential definitions in orc_sample_*.
// This file holds its own Philox implementation (independent of the
// product's) so that the device samplers are checked against a separate
// restatement; the Philox function itself is pinned by the Random123
// known-answer vectors in tests/.
// ===========================================================================
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

// ---- Philox4x32-10 ---------------------------------------------------------
struct Blk {
  uint32_t v[4];
};

Blk philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  uint32_t c[4] = {c0, c1, c2, c3};
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  Blk b;
  memcpy(b.v, c, sizeof c);
  return b;
}

enum : uint32_t {
  T_FLAG = 1,
  T_SOURCE = 2,
  T_NONLINK = 3,
  T_NEIGHBOR = 4,
  T_PHI_NOISE = 5,
  T_THETA_NOISE = 6,
  T_INIT_PHI = 7,
  T_INIT_THETA = 8,
  T_RANDOM_EDGE = 11
};

struct Stream {
  uint32_t k0, k1;
  explicit Stream(uint64_t seed) : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)) {}
  Blk blk(uint32_t idx, uint32_t id, uint32_t iter, uint32_t tag) const {
    return philox(idx, id, iter, tag, k0, k1);
  }
  uint64_t u64(uint32_t idx, uint32_t id, uint32_t iter, uint32_t tag) const {
    const Blk b = blk(idx, id, iter, tag);
    return (uint64_t)b.v[0] | ((uint64_t)b.v[1] << 32);
  }
};

double unit53(uint64_t u) { return ((double)(u >> 11) + 0.5) / 9007199254740992.0; }
float unit24(uint32_t u) { return ((float)(u >> 8) + 0.5f) / 16777216.0f; }

// Box-Muller (Box & Muller 1958) on one block, in the precision of Float.
void gauss2(const Blk& b, double* z0, double* z1) {
  const double u1 = unit53((uint64_t)b.v[0] | ((uint64_t)b.v[1] << 32));
  const double u2 = unit53((uint64_t)b.v[2] | ((uint64_t)b.v[3] << 32));
  const double r = std::sqrt(-2.0 * std::log(u1));
  const double a = 2.0 * 3.14159265358979323846264338327950288 * u2;
  *z0 = r * std::cos(a);
  *z1 = r * std::sin(a);
}
void gauss2(const Blk& b, float* z0, float* z1) {
  const float u1 = unit24(b.v[1]);
  const float u2 = unit24(b.v[3]);
  const float r = std::sqrt(-2.0f * std::log(u1));
  const float a = 6.28318530717958647692f * u2;
  *z0 = r * std::cos(a);
  *z1 = r * std::sin(a);
}

// ---- network view --------------------------------------------------------
struct Net {
  int32_t N = 0;
  int64_t E = 0;
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> cols;
  std::unordered_map<uint64_t, uint8_t> heldtest;  // bit0 held-out, bit1 test
  std::vector<int32_t> ho_edges;                   // held-out in given order
  std::vector<uint8_t> ho_labels;
  int64_t n_test = 0;                              // test entries (links + non-links)

  static uint64_t key(int32_t a, int32_t b) {
    if (a > b) std::swap(a, b);
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
  }
  bool linked(int32_t a, int32_t b) const {
    return std::binary_search(cols.begin() + row_ptr[a], cols.begin() + row_ptr[a + 1], b);
  }
  bool in_heldtest(int32_t a, int32_t b) const { return heldtest.count(key(a, b)) != 0; }
  int64_t degree(int32_t a) const { return row_ptr[a + 1] - row_ptr[a]; }
  // row holding directed entry q of cols (np::find_le on the cumulative
  // degrees, network.cc:459-468)
  int32_t row_of(int64_t q) const {
    return (int32_t)(std::upper_bound(row_ptr.begin(), row_ptr.end(), q) - row_ptr.begin() - 1);
  }
};

// ---- samplers (sequential definitions) -------------------------------------
// Stratified random node (network.cc:410-421), draws by counter:
//   flag      = u64(0, 0, t, FLAG) % 2
//   non-links = node u64(0,0,t,SOURCE) % N; attempt j: v = u64(j,0,t,NONLINK) % N,
//               accepted iff v != node, not linked, not held-out/test, not
//               accepted before; until m accepted (network.cc:286-337);
//               scale = N*N/m * E/(N(N-1)/2 - E) (network.cc:341-342)
//   links     = round r: for s < max_source: node = u64(r,s,t,SOURCE) % N,
//               stop if s>0 and |S|+deg(node) > m; add every edge of node
//               that is not held-out/test and not already in S, in CSR
//               order; repeat rounds while S is empty (network.cc:361-383);
//               scale = N if max_source == 1 else E/|S| (network.cc:385-390)
// Random edge (network.cc:237-268): if m >= E - |held-out| - |test| (entry
//   counts, links and non-links, as the reference compares), every training
//   link (a<b, not held-out/test) in CSR order (sample_full_training_set);
//   else attempt j: q = u64(j,0,t,RANDOM_EDGE) % 2E, the directed entry q of
//   the CSR (row a, column b; sample_random_edges draws over both
//   directions, network.cc:448-500) gives (min, max), accepted iff not
//   held-out/test and not accepted before; until m accepted.  All y = 1;
//   scale = N(N-1)/2 / |S| (network.cc:262).
int sample_minibatch(const Net& g, const Stream& rs, int strategy, int64_t m, int64_t max_source,
                     int64_t t, int* flag_out, double* scale_out, std::vector<int32_t>* edges,
                     std::vector<uint8_t>* ys) {
  const uint32_t t32 = (uint32_t)t;
  const uint64_t N = (uint64_t)g.N;
  int flag;
  if (strategy == 0)
    flag = (int)(rs.u64(0, 0, t32, T_FLAG) % 2);
  else if (strategy == 1)
    flag = 1;
  else if (strategy == 2)
    flag = 0;
  else if (strategy == 3)
    flag = 2;
  else
    return -1;
  edges->clear();
  ys->clear();
  if (flag == 2) {
    const int64_t held_test = (int64_t)g.ho_labels.size() + g.n_test;
    if (m >= g.E - held_test) {
      for (int32_t a = 0; a < g.N; ++a)
        for (int64_t q = g.row_ptr[a]; q < g.row_ptr[a + 1]; ++q) {
          const int32_t b = g.cols[q];
          if (a < b && !g.in_heldtest(a, b)) {
            edges->push_back(a);
            edges->push_back(b);
            ys->push_back(1);
          }
        }
    } else {
      std::unordered_set<uint64_t> taken;
      const uint64_t E2 = 2 * (uint64_t)g.E;
      for (uint64_t j = 0; (int64_t)taken.size() < m; ++j) {
        if (j > (uint64_t)(1u << 31)) return -2;
        const int64_t q = (int64_t)(rs.u64((uint32_t)j, 0, t32, T_RANDOM_EDGE) % E2);
        const int32_t a = g.row_of(q), b = g.cols[q];
        if (g.in_heldtest(a, b)) continue;
        if (!taken.insert(Net::key(a, b)).second) continue;
        edges->push_back(std::min(a, b));
        edges->push_back(std::max(a, b));
        ys->push_back(1);
      }
    }
    const double dN = (double)g.N;
    *scale_out = (dN - 1.0) * dN / 2.0 / (double)ys->size();
  } else if (flag == 0) {
    const int32_t node = (int32_t)(rs.u64(0, 0, t32, T_SOURCE) % N);
    std::unordered_set<int32_t> taken;
    for (uint64_t j = 0; (int64_t)taken.size() < m; ++j) {
      if (j > (uint64_t)(1u << 31)) return -2;
      const int32_t v = (int32_t)(rs.u64((uint32_t)j, 0, t32, T_NONLINK) % N);
      if (v == node) continue;
      if (g.linked(node, v) || g.in_heldtest(node, v)) continue;
      if (!taken.insert(v).second) continue;
      edges->push_back(std::min(node, v));
      edges->push_back(std::max(node, v));
      ys->push_back(0);
    }
    const double dN = (double)g.N, dE = (double)g.E;
    double scale = dN * dN / (double)m;
    scale *= dE / (dN * (dN - 1.0) / 2.0 - dE);
    *scale_out = scale;
  } else {
    int64_t c = 0;
    for (uint64_t r = 0;; ++r) {
      if (r > (1u << 20)) return -3;
      std::vector<int32_t> srcs;
      std::unordered_set<uint64_t> in_s;
      edges->clear();
      ys->clear();
      c = 0;
      for (int64_t s = 0; s < max_source; ++s) {
        const int32_t node = (int32_t)(rs.u64((uint32_t)r, (uint32_t)s, t32, T_SOURCE) % N);
        if (s > 0 && c + g.degree(node) > m) break;
        srcs.push_back(node);
        for (int64_t q = g.row_ptr[node]; q < g.row_ptr[node + 1]; ++q) {
          const int32_t v = g.cols[q];
          if (g.in_heldtest(node, v)) continue;
          if (!in_s.insert(Net::key(node, v)).second) continue;
          edges->push_back(std::min(node, v));
          edges->push_back(std::max(node, v));
          ys->push_back(1);
          ++c;
        }
      }
      if (c > 0) break;
    }
    *scale_out = max_source == 1 ? (double)g.N : (double)g.E / (double)c;
  }
  *flag_out = flag;
  return 0;
}

// n+1 distinct neighbours != node, not held-out/test; links allowed
// (mcmc_sampler_stochastic.h:93-107).  attempt j: v = u64(j,node,t,NEIGHBOR) % N.
int sample_neighbors(const Net& g, const Stream& rs, int32_t node, int64_t nnb, int64_t t,
                     int32_t* out) {
  std::unordered_set<int32_t> taken;
  int64_t c = 0;
  for (uint64_t j = 0; c < nnb; ++j) {
    if (j > (uint64_t)(1u << 31)) return -1;
    const int32_t v = (int32_t)(rs.u64((uint32_t)j, (uint32_t)node, (uint32_t)t, T_NEIGHBOR) %
                                (uint64_t)g.N);
    if (v == node || g.in_heldtest(node, v)) continue;
    if (!taken.insert(v).second) continue;
    out[c++] = v | (g.linked(node, v) ? (int32_t)0x80000000 : 0);
  }
  return 0;
}

// ---- threads ---------------------------------------------------------------
// [0, n) in `threads` contiguous chunks (the static schedule of the
// distributed reference's `#pragma omp parallel for` loops); fn(lo, hi, tid).
template <typename Fn>
void parallel_chunks(int64_t n, int threads, Fn fn) {
  if (threads <= 1 || n < 2) {
    fn((int64_t)0, n, 0);
    return;
  }
  std::vector<std::thread> th;
  for (int tt = 0; tt < threads; ++tt)
    th.emplace_back(fn, n * tt / threads, n * (tt + 1) / threads, tt);
  for (auto& x : th) x.join();
}

// ---- state ---------------------------------------------------------------
struct Hyper {
  double alpha, eta0, eta1, epsilon, a, b, c;
  int64_t num_node_sample;
};

template <typename Float>
struct State {
  const Net* g = nullptr;
  int64_t K = 0;
  Hyper hp{};
  std::vector<Float> pi;                   // [N][K+1], slot K = phi_sum
  std::vector<std::vector<Float>> theta;  // [K][2]
  std::vector<Float> beta;                 // [K]
  std::vector<Float> ppx_R;                // ppx_per_heldout_edge_
  size_t average_count = 1;
  size_t step_count = 1;
  Float* row(int64_t i) { return &pi[(size_t)i * (K + 1)]; }
};

template <typename Float>
Float eps_t_of(const Hyper& hp, size_t step_count) {
  // get_eps_t (mcmc_sampler_stochastic.h:112-115)
  const Float a = (Float)hp.a, b = (Float)hp.b, c = (Float)hp.c;
  return a * std::pow(1 + step_count / b, -c);
}

// update_phi_node (distr.cc:1158-1227) for node i, reading the old pi rows.
template <typename Float>
void update_phi_node(State<Float>& S, int32_t i, const int32_t* nbrs, int64_t nnb,
                     const Float* noise, Float eps_t, std::vector<Float>* phi_node) {
  const int64_t K = S.K;
  const Float* pi_node = S.row(i);
  const Float epsilon = (Float)S.hp.epsilon;
  const Float alpha = (Float)S.hp.alpha;
  Float phi_i_sum = pi_node[K];
  std::vector<Float> grads(K, 0.0);
  std::vector<Float> probs(K);
  for (int64_t ix = 0; ix < nnb; ++ix) {
    const int32_t neighbor = nbrs[ix] & 0x7fffffff;
    if (i == neighbor) continue;
    const int y_ab = S.g->linked(std::min(i, neighbor), std::max(i, neighbor)) ? 1 : 0;
    const Float* pi_b = S.row(neighbor);
    Float e = (y_ab == 1) ? epsilon : 1.0 - epsilon;
    for (int64_t k = 0; k < K; ++k) {
      Float f = (y_ab == 1) ? (S.beta[k] - epsilon) : (epsilon - S.beta[k]);
      probs[k] = pi_node[k] * (pi_b[k] * f + e);
    }
    Float prob_sum = std::accumulate(probs.begin(), probs.end(), static_cast<Float>(0));
    for (int64_t k = 0; k < K; ++k) {
      grads[k] += ((probs[k] / prob_sum) / pi_node[k] - 1.0) / phi_i_sum;
    }
  }
  Float Nn = (1.0 * S.g->N) / S.hp.num_node_sample;
  phi_node->resize(K);
  for (int64_t k = 0; k < K; ++k) {
    Float phi_node_k = pi_node[k] * phi_i_sum;
    phi_node_k = std::abs(phi_node_k + eps_t / 2 * (alpha - phi_node_k + Nn * grads[k]) +
                          std::sqrt(eps_t * phi_node_k) * noise[k]);
    (*phi_node)[k] = phi_node_k < (Float)1.0e-24 ? (Float)1.0e-24 : phi_node_k;
  }
}

// pi_from_phi (distr.cc:743-751)
template <typename Float>
void pi_from_phi(Float* pi, const std::vector<Float>& phi, int64_t K) {
  Float phi_sum = std::accumulate(phi.begin(), phi.begin() + K, 0.0);
  for (int64_t k = 0; k < K; ++k) pi[k] = phi[k] / phi_sum;
  pi[K] = phi_sum;
}

template <typename Float>
void noise_phi(const Stream& rs, int32_t node, int64_t t, int64_t K, Float* out) {
  for (int64_t k = 0; k < K; k += 2) {
    Float z0, z1;
    gauss2(rs.blk((uint32_t)(k >> 1), (uint32_t)node, (uint32_t)t, T_PHI_NOISE), &z0, &z1);
    out[k] = z0;
    if (k + 1 < K) out[k + 1] = z1;
  }
}

// update_phi for a node set (all reads see the old pi) + update_pi.
template <typename Float>
void update_phi_set(State<Float>& S, const int32_t* nodes, int64_t n, const int32_t* nbrs,
                    int64_t nnb, const Float* noise, Float eps_t, const Stream* rs, int64_t t,
                    int threads) {
  const int64_t K = S.K;
  std::vector<std::vector<Float>> out(n);
  // distr.cc:1142-1150 (omp parallel over the nodes)
  parallel_chunks(n, threads, [&](int64_t lo, int64_t hi, int) {
    std::vector<Float> nz(K);
    for (int64_t q = lo; q < hi; ++q) {
      const Float* nzp;
      if (noise) {
        nzp = noise + q * K;
      } else {
        noise_phi<Float>(*rs, nodes[q], t, K, nz.data());
        nzp = nz.data();
      }
      update_phi_node<Float>(S, nodes[q], nbrs + q * nnb, nnb, nzp, eps_t, &out[q]);
    }
  });
  for (int64_t q = 0; q < n; ++q) pi_from_phi<Float>(S.row(nodes[q]), out[q], K);
}

// update_beta (mcmc_sampler_stochastic.cc:200-267) with injected y and noise.
// threads > 1: per-thread gradient rows over contiguous edge ranges, summed in
// thread order (beta_calc_grads / beta_sum_grads, distr.cc:1377-1426).
template <typename Float>
void update_beta(State<Float>& S, const int32_t* edges, const uint8_t* ys, int64_t E,
                 Float scale, const Float* noise /* [K][2] */, Float eps_t, int threads = 1) {
  const int64_t K = S.K;
  const Float epsilon = (Float)S.hp.epsilon;
  std::vector<Float> theta_sum(K);
  for (int64_t k = 0; k < K; ++k)
    theta_sum[k] = std::accumulate(S.theta[k].begin(), S.theta[k].end(), static_cast<Float>(0));
  const int nt = std::max(1, std::min<int>(threads, (int)std::max<int64_t>(E, 1)));
  std::vector<std::vector<std::vector<Float>>> tg(
      nt, std::vector<std::vector<Float>>(K, std::vector<Float>(2, (Float)0.0)));
  parallel_chunks(E, nt, [&](int64_t lo, int64_t hi, int tid) {
    std::vector<std::vector<Float>>& grads = tg[tid];
    std::vector<Float> probs(K);
    for (int64_t e = lo; e < hi; ++e) {
      const int y = ys[e] ? 1 : 0;
      const Float* pi_i = S.row(edges[2 * e]);
      const Float* pi_j = S.row(edges[2 * e + 1]);
      Float pi_sum = (Float)0.0;
      for (int64_t k = 0; k < K; ++k) {
        pi_sum += pi_i[k] * pi_j[k];
        Float f = pi_i[k] * pi_j[k];
        if (y == 1)
          probs[k] = S.beta[k] * f;
        else
          probs[k] = ((Float)1.0 - S.beta[k]) * f;
      }
      Float prob_0 = ((y == 1) ? epsilon : ((Float)1.0 - epsilon)) * ((Float)1.0 - pi_sum);
      Float prob_sum = std::accumulate(probs.begin(), probs.end(), static_cast<Float>(0)) + prob_0;
      for (int64_t k = 0; k < K; ++k) {
        Float f = probs[k] / prob_sum;
        Float one_over_theta_sum = (Float)1.0 / theta_sum[k];
        grads[k][0] += f * (((Float)1.0 - y) / S.theta[k][0] - one_over_theta_sum);
        grads[k][1] += f * (y / S.theta[k][1] - one_over_theta_sum);
      }
    }
  });
  std::vector<std::vector<Float>>& grads = tg[0];
  for (int tt = 1; tt < nt; ++tt)
    for (int64_t k = 0; k < K; ++k) {
      grads[k][0] += tg[tt][k][0];
      grads[k][1] += tg[tt][k][1];
    }
  const Float eta[2] = {(Float)S.hp.eta0, (Float)S.hp.eta1};
  for (int64_t k = 0; k < K; ++k) {
    for (int i = 0; i < 2; ++i) {
      Float f = std::sqrt(eps_t * S.theta[k][i]);
      S.theta[k][i] = std::abs(S.theta[k][i] +
                               eps_t / (Float)2.0 * (eta[i] - S.theta[k][i] + scale * grads[k][i]) +
                               f * noise[2 * k + i]);
      if (S.theta[k][i] < (Float)1.0e-24) S.theta[k][i] = (Float)1.0e-24;
    }
  }
  for (int64_t k = 0; k < K; ++k) {
    Float s = std::accumulate(S.theta[k].begin(), S.theta[k].end(), static_cast<Float>(0));
    S.beta[k] = S.theta[k][1] / s;
  }
}

// cal_edge_likelihood (learner.h:100-139), DONT_FOLD_Y branch
template <typename Float>
Float edge_likelihood(const Float* pi_a, const Float* pi_b, bool y, const std::vector<Float>& beta,
                      int64_t K, Float epsilon) {
  Float s = (Float)0.0;
  if (y) {
    for (int64_t k = 0; k < K; k++) s += pi_a[k] * pi_b[k] * beta[k];
  } else {
    Float sum = (Float)0.0;
    for (int64_t k = 0; k < K; k++) {
      Float f = pi_a[k] * pi_b[k];
      s += f * ((Float)1.0 - beta[k]);
      sum += f;
    }
    s += ((Float)1.0 - sum) * ((Float)1.0 - epsilon);
  }
  if (s < (Float)1.0e-30) s = (Float)1.0e-30;
  return s;
}

// cal_perplexity (learner.cc:138-183) over the held-out list in its order
// threads > 1: per-thread link / non-link accumulators over contiguous
// ranges, summed in thread order (distr.cc:1543-1600).
template <typename Float>
Float perplexity(State<Float>& S, int threads = 1) {
  const Net& g = *S.g;
  const int64_t H = (int64_t)g.ho_labels.size();
  const int nt = std::max(1, std::min<int>(threads, (int)std::max<int64_t>(H, 1)));
  std::vector<Float> ll(nt, 0), nl(nt, 0);
  std::vector<size_t> lc(nt, 0), nc(nt, 0);
  parallel_chunks(H, nt, [&](int64_t lo, int64_t hi, int tid) {
    for (int64_t i = lo; i < hi; ++i) {
      const bool y = g.ho_labels[i] != 0;
      Float L = edge_likelihood<Float>(S.row(g.ho_edges[2 * i]), S.row(g.ho_edges[2 * i + 1]), y,
                                       S.beta, S.K, (Float)S.hp.epsilon);
      S.ppx_R[i] = (S.ppx_R[i] * (S.average_count - 1) + L) / (S.average_count);
      if (y) {
        lc[tid]++;
        ll[tid] += std::log(S.ppx_R[i]);
      } else {
        nc[tid]++;
        nl[tid] += std::log(S.ppx_R[i]);
      }
    }
  });
  Float link_likelihood = 0, non_link_likelihood = 0;
  size_t link_count = 0, non_link_count = 0;
  for (int tt = 0; tt < nt; ++tt) {
    link_likelihood += ll[tt];
    non_link_likelihood += nl[tt];
    link_count += lc[tt];
    non_link_count += nc[tt];
  }
  Float avg = 0;
  if (link_count + non_link_count != 0)
    avg = (link_likelihood + non_link_likelihood) / (link_count + non_link_count);
  S.average_count = S.average_count + 1;
  return -avg;
}

// Marsaglia & Tsang (ACM TOMS 26(3), 2000) on the counter stream; log value.
double log_gamma(const Stream& rs, double shape, uint32_t id, uint32_t iter, uint32_t tag) {
  double boost = 0.0;
  double a = shape;
  if (a < 1.0) {
    const Blk b = rs.blk(0xFFFFFFFFu, id, iter, tag);
    boost = std::log(unit53((uint64_t)b.v[0] | ((uint64_t)b.v[1] << 32))) / a;
    a += 1.0;
  }
  const double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
  for (uint32_t j = 0;; ++j) {
    double x, unused;
    gauss2(rs.blk(2 * j, id, iter, tag), &x, &unused);
    double v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    const Blk b = rs.blk(2 * j + 1, id, iter, tag);
    const double u = unit53((uint64_t)b.v[0] | ((uint64_t)b.v[1] << 32));
    if (u < 1.0 - 0.0331 * x * x * x * x || std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v)))
      return std::log(d * v) + boost;
  }
}

struct Handle {
  Net net;
  int prec = 0;
  State<double> s64;
  State<float> s32;
};

Hyper hyper_from(const double* h) {
  Hyper hp;
  hp.alpha = h[0];
  hp.eta0 = h[1];
  hp.eta1 = h[2];
  hp.epsilon = h[3];
  hp.a = h[4];
  hp.b = h[5];
  hp.c = h[6];
  hp.num_node_sample = (int64_t)h[7];
  return hp;
}

template <typename Float>
State<Float>& st(Handle* h);
template <>
State<double>& st<double>(Handle* h) {
  return h->s64;
}
template <>
State<float>& st<float>(Handle* h) {
  return h->s32;
}

template <typename Float>
int run_impl(Handle* h, int64_t iters, int strategy, int64_t m, int64_t max_source,
             int64_t interval, uint64_t seed, int threads, double* ppx_out, int64_t* ppx_steps,
             int64_t cap, int64_t* n_ppx, double* phase_secs) {
  State<Float>& S = st<Float>(h);
  const Net& g = h->net;
  const Stream rs(seed);
  const int64_t nnb = S.hp.num_node_sample + 1;
  int64_t np = 0;
  std::vector<int32_t> edges;
  std::vector<uint8_t> ys;
  using clk = std::chrono::steady_clock;
  for (int64_t it = 0; it < iters; ++it) {
    const int64_t t = (int64_t)S.step_count;
    auto t0 = clk::now();
    if ((t - 1) % interval == 0) {
      const double p = (double)perplexity<Float>(S, threads);
      if (ppx_out && np < cap) {
        ppx_out[np] = p;
        ppx_steps[np] = t;
      }
      ++np;
    }
    auto t1 = clk::now();
    int flag;
    double scale;
    if (sample_minibatch(g, rs, strategy, m, max_source, t, &flag, &scale, &edges, &ys) != 0)
      return -1;
    std::vector<int32_t> nodes;
    {
      std::unordered_set<int32_t> ns(edges.begin(), edges.end());
      nodes.assign(ns.begin(), ns.end());
      std::sort(nodes.begin(), nodes.end());
    }
    auto t2 = clk::now();
    const Float eps_t = eps_t_of<Float>(S.hp, S.step_count);
    // DrawNeighbors (distr.cc:1061-1100, omp parallel over the nodes); every
    // node's draws are keyed by its id, so the split changes nothing
    std::vector<int32_t> nb((size_t)nodes.size() * nnb);
    std::vector<int> nb_rc(std::max(threads, 1), 0);
    parallel_chunks((int64_t)nodes.size(), threads, [&](int64_t lo, int64_t hi, int tid) {
      for (int64_t q = lo; q < hi; ++q)
        if (sample_neighbors(g, rs, nodes[q], nnb, t, &nb[q * nnb]) != 0) nb_rc[tid] = -2;
    });
    for (int rc : nb_rc)
      if (rc != 0) return rc;
    auto t3 = clk::now();
    update_phi_set<Float>(S, nodes.data(), (int64_t)nodes.size(), nb.data(), nnb, nullptr, eps_t,
                          &rs, t, threads);
    auto t4 = clk::now();
    std::vector<Float> tn(2 * S.K);
    for (int64_t k = 0; k < S.K; ++k) {
      Float z0, z1;
      gauss2(rs.blk((uint32_t)k, 0, (uint32_t)t, T_THETA_NOISE), &z0, &z1);
      tn[2 * k] = z0;
      tn[2 * k + 1] = z1;
    }
    update_beta<Float>(S, edges.data(), ys.data(), (int64_t)ys.size(), (Float)scale, tn.data(),
                       eps_t, threads);
    auto t5 = clk::now();
    S.step_count++;
    if (phase_secs) {
      phase_secs[0] += std::chrono::duration<double>(t1 - t0).count();  // perplexity
      phase_secs[1] += std::chrono::duration<double>(t2 - t1).count();  // minibatch+nodes
      phase_secs[2] += std::chrono::duration<double>(t3 - t2).count();  // neighbours
      phase_secs[3] += std::chrono::duration<double>(t4 - t3).count();  // phi+pi
      phase_secs[4] += std::chrono::duration<double>(t5 - t4).count();  // beta
      phase_secs[5] += (double)nodes.size();                            // nodes updated
    }
  }
  if (n_ppx) *n_ppx = np;
  return 0;
}

}  // namespace

extern "C" {

void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  const Blk b = philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
  memcpy(out, b.v, 16);
}

void* orc_create(int32_t N, int64_t E, const int64_t* row_ptr, const int32_t* cols,
                 const int32_t* ho_edges, const uint8_t* ho_labels, int64_t H,
                 const int32_t* te_edges, const uint8_t* te_labels, int64_t T, int64_t K,
                 int32_t precision, const double* hyper) {
  Handle* h = new Handle();
  Net& g = h->net;
  g.N = N;
  g.E = E;
  g.row_ptr.assign(row_ptr, row_ptr + N + 1);
  g.cols.assign(cols, cols + row_ptr[N]);
  g.heldtest.reserve((size_t)(2 * (H + T)));
  for (int64_t i = 0; i < H; ++i) {
    g.heldtest[Net::key(ho_edges[2 * i], ho_edges[2 * i + 1])] |= 1;
    g.ho_edges.push_back(std::min(ho_edges[2 * i], ho_edges[2 * i + 1]));
    g.ho_edges.push_back(std::max(ho_edges[2 * i], ho_edges[2 * i + 1]));
    g.ho_labels.push_back(ho_labels[i]);
  }
  for (int64_t i = 0; i < T; ++i) g.heldtest[Net::key(te_edges[2 * i], te_edges[2 * i + 1])] |= 2;
  g.n_test = T;
  (void)te_labels;
  h->prec = precision;
  const Hyper hp = hyper_from(hyper);
  auto init = [&](auto& S) {
    S.g = &h->net;
    S.K = K;
    S.hp = hp;
    S.pi.assign((size_t)N * (K + 1), 0);
    S.theta.assign(K, {0, 0});
    S.beta.assign(K, 0);
    S.ppx_R.assign(H, 0);
  };
  if (precision == 0)
    init(h->s64);
  else
    init(h->s32);
  return h;
}

void orc_destroy(void* hv) { delete (Handle*)hv; }

// pi: [N][K+1] in the handle's precision; theta [K][2] double
void orc_set_state(void* hv, const void* pi, const double* theta) {
  Handle* h = (Handle*)hv;
  auto set = [&](auto& S) {
    using F = typename std::decay_t<decltype(S.pi)>::value_type;
    memcpy(S.pi.data(), pi, S.pi.size() * sizeof(F));
    for (int64_t k = 0; k < S.K; ++k) {
      S.theta[k][0] = (F)theta[2 * k];
      S.theta[k][1] = (F)theta[2 * k + 1];
      F s = std::accumulate(S.theta[k].begin(), S.theta[k].end(), static_cast<F>(0));
      S.beta[k] = S.theta[k][1] / s;
    }
  };
  if (h->prec == 0)
    set(h->s64);
  else
    set(h->s32);
}

void orc_get_state(void* hv, void* pi, double* theta, double* beta) {
  Handle* h = (Handle*)hv;
  auto get = [&](auto& S) {
    using F = typename std::decay_t<decltype(S.pi)>::value_type;
    if (pi) memcpy(pi, S.pi.data(), S.pi.size() * sizeof(F));
    for (int64_t k = 0; k < S.K; ++k) {
      if (theta) {
        theta[2 * k] = S.theta[k][0];
        theta[2 * k + 1] = S.theta[k][1];
      }
      if (beta) beta[k] = S.beta[k];
    }
  };
  if (h->prec == 0)
    get(h->s64);
  else
    get(h->s32);
}

void orc_set_counters(void* hv, int64_t step_count, int64_t average_count) {
  Handle* h = (Handle*)hv;
  h->s64.step_count = h->s32.step_count = (size_t)step_count;
  h->s64.average_count = h->s32.average_count = (size_t)average_count;
}

// update_phi + update_pi with injected neighbours and noise ([n][K], handle
// precision; NULL: Philox noise of iteration `step`)
int orc_update_phi(void* hv, const int32_t* nodes, int64_t n, const int32_t* nbrs, int64_t nnb,
                   const void* noise, double eps_t, uint64_t seed, int64_t step, int threads) {
  Handle* h = (Handle*)hv;
  const Stream rs(seed);
  if (h->prec == 0)
    update_phi_set<double>(h->s64, nodes, n, nbrs, nnb, (const double*)noise, (double)eps_t, &rs,
                           step, threads);
  else
    update_phi_set<float>(h->s32, nodes, n, nbrs, nnb, (const float*)noise, (float)eps_t, &rs,
                          step, threads);
  return 0;
}

int orc_update_beta(void* hv, const int32_t* edges, const uint8_t* y, int64_t E, double scale,
                    const double* noise, double eps_t) {
  Handle* h = (Handle*)hv;
  if (h->prec == 0) {
    update_beta<double>(h->s64, edges, y, E, scale, noise, eps_t);
  } else {
    std::vector<float> nz(2 * h->s32.K);
    for (size_t i = 0; i < nz.size(); ++i) nz[i] = (float)noise[i];
    update_beta<float>(h->s32, edges, y, E, (float)scale, nz.data(), (float)eps_t);
  }
  return 0;
}

double orc_perplexity(void* hv) {
  Handle* h = (Handle*)hv;
  return h->prec == 0 ? (double)perplexity<double>(h->s64) : (double)perplexity<float>(h->s32);
}

double orc_eps_t(const double* hyper, int64_t step, int32_t precision) {
  const Hyper hp = hyper_from(hyper);
  return precision == 0 ? (double)eps_t_of<double>(hp, (size_t)step)
                        : (double)eps_t_of<float>(hp, (size_t)step);
}

// The reference's closed-form gradients, before any step is taken:
// update_phi_node's grads (distr.cc:1165-1197) for one node and its
// neighbours, and update_beta's grads (mcmc_sampler_stochastic.cc:202-240)
// for a list of edges.  Both read the handle's state.
void orc_phi_grads(void* hv, int32_t node, const int32_t* nbrs, int64_t nnb, double* out) {
  Handle* h = (Handle*)hv;
  auto run = [&](auto& S) {
    using Float = typename std::decay_t<decltype(S.pi)>::value_type;
    const int64_t K = S.K;
    const Float* pi_node = S.row(node);
    const Float epsilon = (Float)S.hp.epsilon;
    Float phi_i_sum = pi_node[K];
    std::vector<Float> grads(K, 0.0), probs(K);
    for (int64_t ix = 0; ix < nnb; ++ix) {
      const int32_t neighbor = nbrs[ix] & 0x7fffffff;
      if (node == neighbor) continue;
      const int y_ab = S.g->linked(std::min(node, neighbor), std::max(node, neighbor)) ? 1 : 0;
      const Float* pi_b = S.row(neighbor);
      Float e = (y_ab == 1) ? epsilon : 1.0 - epsilon;
      for (int64_t k = 0; k < K; ++k) {
        Float f = (y_ab == 1) ? (S.beta[k] - epsilon) : (epsilon - S.beta[k]);
        probs[k] = pi_node[k] * (pi_b[k] * f + e);
      }
      Float prob_sum = std::accumulate(probs.begin(), probs.end(), static_cast<Float>(0));
      for (int64_t k = 0; k < K; ++k) grads[k] += ((probs[k] / prob_sum) / pi_node[k] - 1.0) / phi_i_sum;
    }
    for (int64_t k = 0; k < K; ++k) out[k] = grads[k];
  };
  if (h->prec == 0)
    run(h->s64);
  else
    run(h->s32);
}

void orc_beta_grads(void* hv, const int32_t* edges, const uint8_t* ys, int64_t E, double* out) {
  Handle* h = (Handle*)hv;
  auto run = [&](auto& S) {
    using Float = typename std::decay_t<decltype(S.pi)>::value_type;
    const int64_t K = S.K;
    const Float epsilon = (Float)S.hp.epsilon;
    std::vector<std::vector<Float>> grads(K, std::vector<Float>(2, (Float)0.0));
    std::vector<Float> probs(K), theta_sum(K);
    for (int64_t k = 0; k < K; ++k)
      theta_sum[k] = std::accumulate(S.theta[k].begin(), S.theta[k].end(), static_cast<Float>(0));
    for (int64_t e = 0; e < E; ++e) {
      const int y = ys[e] ? 1 : 0;
      const Float* pi_i = S.row(edges[2 * e]);
      const Float* pi_j = S.row(edges[2 * e + 1]);
      Float pi_sum = (Float)0.0;
      for (int64_t k = 0; k < K; ++k) {
        pi_sum += pi_i[k] * pi_j[k];
        Float f = pi_i[k] * pi_j[k];
        probs[k] = (y == 1) ? S.beta[k] * f : ((Float)1.0 - S.beta[k]) * f;
      }
      Float prob_0 = ((y == 1) ? epsilon : ((Float)1.0 - epsilon)) * ((Float)1.0 - pi_sum);
      Float prob_sum = std::accumulate(probs.begin(), probs.end(), static_cast<Float>(0)) + prob_0;
      for (int64_t k = 0; k < K; ++k) {
        Float f = probs[k] / prob_sum;
        Float one_over_theta_sum = (Float)1.0 / theta_sum[k];
        grads[k][0] += f * (((Float)1.0 - y) / S.theta[k][0] - one_over_theta_sum);
        grads[k][1] += f * (y / S.theta[k][1] - one_over_theta_sum);
      }
    }
    for (int64_t k = 0; k < K; ++k) {
      out[2 * k] = grads[k][0];
      out[2 * k + 1] = grads[k][1];
    }
  };
  if (h->prec == 0)
    run(h->s64);
  else
    run(h->s32);
}

// Sampler replicas.  edges_out needs room for 2*cap ints.
int orc_sample_minibatch(void* hv, int32_t strategy, int64_t m, int64_t max_source, uint64_t seed,
                         int64_t step, int32_t* flag, double* scale, int32_t* edges_out,
                         uint8_t* y_out, int64_t cap, int64_t* count) {
  Handle* h = (Handle*)hv;
  std::vector<int32_t> e;
  std::vector<uint8_t> y;
  int f = 0;
  double sc = 0;
  int rc = sample_minibatch(h->net, Stream(seed), strategy, m, max_source, step, &f, &sc, &e, &y);
  if (rc != 0) return rc;
  *flag = f;
  *scale = sc;
  *count = (int64_t)y.size();
  if ((int64_t)y.size() > cap) return -10;
  memcpy(edges_out, e.data(), e.size() * 4);
  if (y_out) memcpy(y_out, y.data(), y.size());
  return 0;
}

int orc_sample_neighbors(void* hv, int32_t node, int64_t nnb, uint64_t seed, int64_t step,
                         int32_t* out) {
  Handle* h = (Handle*)hv;
  return sample_neighbors(h->net, Stream(seed), node, nnb, step, out);
}

// Philox noise of update_phi for one node (handle precision)
void orc_phi_noise(int32_t precision, uint64_t seed, int32_t node, int64_t step, int64_t K,
                   void* out) {
  const Stream rs(seed);
  if (precision == 0)
    noise_phi<double>(rs, node, step, K, (double*)out);
  else
    noise_phi<float>(rs, node, step, K, (float*)out);
}

void orc_theta_noise(uint64_t seed, int64_t step, int64_t K, double* out) {
  const Stream rs(seed);
  for (int64_t k = 0; k < K; ++k)
    gauss2(rs.blk((uint32_t)k, 0, (uint32_t)step, T_THETA_NOISE), &out[2 * k], &out[2 * k + 1]);
}

double orc_log_gamma(uint64_t seed, double shape, uint32_t id, uint32_t iter, uint32_t tag) {
  return log_gamma(Stream(seed), shape, id, iter, tag);
}

// init() state draw (mcmc_sampler_stochastic.cc:54-76) on the counter
// stream: theta[k][i] = Gamma(shape eta0, scale eta1) = exp(log G) * eta1
// with G ~ Gamma(eta0, 1) from stream (2k+i, 0, INIT_THETA) -- the GSL order
// of Random::gamma(a, b) is shape a, scale b (random.cc:472-511).
void orc_init_theta(uint64_t seed, double eta0, double eta1, int64_t K, double* theta) {
  const Stream rs(seed);
  for (int64_t k = 0; k < K; ++k)
    for (int i = 0; i < 2; ++i)
      theta[2 * k + i] =
          std::exp(log_gamma(rs, eta0, (uint32_t)(2 * k + i), 0, T_INIT_THETA)) * eta1;
}

// phi[node][k] ~ Gamma(1, 1) = -log U, U from block (k>>1, node, 0, INIT_PHI):
// words (0,1) for even k, (2,3) for odd k; row = phi / sum(phi) ++ sum
// (np::row_normalize, mcmc_sampler_stochastic.cc:66-76).  row: K+1 doubles.
void orc_init_phi_row(uint64_t seed, int32_t node, int64_t K, double* row) {
  const Stream rs(seed);
  std::vector<double> phi(K);
  for (int64_t k = 0; k < K; ++k) {
    const Blk b = rs.blk((uint32_t)(k >> 1), (uint32_t)node, 0, T_INIT_PHI);
    const uint64_t u = (k & 1) ? ((uint64_t)b.v[2] | ((uint64_t)b.v[3] << 32))
                               : ((uint64_t)b.v[0] | ((uint64_t)b.v[1] << 32));
    phi[k] = -std::log(unit53(u));
  }
  const double s = std::accumulate(phi.begin(), phi.end(), 0.0);
  for (int64_t k = 0; k < K; ++k) row[k] = phi[k] / s;
  row[K] = s;
}

// Full loop (the host replica of the device iteration); perplexity at
// every `interval` steps.  phase_secs[6]: perplexity, minibatch+nodes,
// neighbours, update_phi+pi, update_beta, nodes updated (count).
int orc_run(void* hv, int64_t iters, int32_t strategy, int64_t m, int64_t max_source,
            int64_t interval, uint64_t seed, int32_t threads, double* ppx_out, int64_t* ppx_steps,
            int64_t cap, int64_t* n_ppx, double* phase_secs) {
  Handle* h = (Handle*)hv;
  if (h->prec == 0)
    return run_impl<double>(h, iters, strategy, m, max_source, interval, seed, threads, ppx_out,
                            ppx_steps, cap, n_ppx, phase_secs);
  return run_impl<float>(h, iters, strategy, m, max_source, interval, seed, threads, ppx_out,
                         ppx_steps, cap, n_ppx, phase_secs);
}

}  // extern "C"
