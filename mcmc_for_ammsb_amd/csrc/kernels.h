// Launcher interface between the host driver (capi.hip) and the gfx950
// kernels (kernels.hip).  Everything here is plain C++ (no templates cross
// the boundary); the precision is a runtime switch that picks a template
// instantiation inside kernels.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "philox.h"

namespace ammsb {

// Device-resident per-context iteration state.  Every kernel of an
// iteration reads `step` from here, so an iteration has no host-side
// arguments that change between iterations (it can be replayed as a
// captured graph).
struct DevState {
  int64_t step;        // Learner::step_count (1-based)
  int32_t error;       // sampler error bits (ERR_*)
  int32_t pad;
  double ppx;          // last held-out perplexity
  int64_t nodes_total; // sum of node_count over all minibatch draws
  int64_t edges_total; // sum of mb_count over all minibatch draws
  int32_t done_count;  // k_theta workgroups finished (last one bumps step)
  int32_t ring_slots;  // slots of the sample-ahead ring
  int64_t ring_base;   // step held by slot 0 of the sample-ahead ring
  double eps_t;        // get_eps_t of `step`, set with it (k_set_step, k_theta)
};

enum : int32_t {
  ERR_NONLINK_ATTEMPTS = 1,   // non-link sampler exceeded its attempt budget
  ERR_LINK_ROUNDS = 2,        // link sampler found only empty minibatches
  ERR_NEIGHBOR_ATTEMPTS = 4,  // neighbour sampler exceeded its budget
  ERR_MB_OVERFLOW = 8,        // minibatch larger than its buffer
  ERR_BAD_STRATEGY = 16,
  ERR_BAD_INPUT = 32          // caller-provided device ids out of range
};

// Read-only graph view on the device.
struct DevGraph {
  int32_t N;
  int64_t E;                 // undirected edges
  const int64_t* row_ptr;    // N+1
  const int32_t* cols;       // 2E, rows ascending
  const uint64_t* hkeys;     // held-out ∪ test hash (host_util.h layout)
  const uint8_t* hvals;
  uint64_t hmask;
  const uint64_t* bloom;     // blocked Bloom filter of the same keys
  uint64_t bloom_mask;
  const int32_t* esrc;       // [2E] row of each CSR entry (random-edge strategy only)
  uint64_t n_recip;          // fastmod_recip(N): randint(0, N-1) without a division
  uint64_t e2_recip;         // fastmod_recip(2E)
};

// Hyper-parameters as the kernels see them (all double; each kernel casts
// to its Float).
struct DevHyper {
  double alpha, epsilon, a, b, c, eta0, eta1;
  double Nn;  // N / num_node_sample (mcmc_sampler_stochastic.cc:300)
};

// One sampled iteration (a slot of the sample-ahead ring).
struct IterMeta {
  int64_t step;        // iteration whose draws the slot holds
  int32_t flag;        // 0 = non-link minibatch, 1 = link minibatch
  int32_t mb_count;    // minibatch edges
  int32_t node_count;  // unique minibatch nodes
  uint32_t epoch;      // node de-dup stamp generation of this slot
  double scale;        // unbiasing scale (network.cc:341-342,385-390)
};

// Sample-ahead ring: the draws of an iteration depend only on (graph, seed,
// step), never on the state, so `slots` future iterations are drawn by one
// launch and consumed one per iteration.  Iteration t is held by slot
// (t - DevState::ring_base) mod slots (batches refill one half of the ring
// while the other half is consumed).
struct Ring {
  int32_t slots;
  int64_t edge_cap;   // edges per slot
  int64_t node_cap;   // nodes per slot
  int32_t nnb;        // neighbours per node (n + 1)
  int32_t N;
  int32_t* edges;     // [slots][edge_cap][2]
  uint8_t* y;         // [slots][edge_cap]
  int32_t* nodes;     // [slots][node_cap]
  int32_t* nbrs;      // [slots][node_cap][nnb], bit 31 = link
  IterMeta* meta;     // [slots]
  uint32_t* stamp;    // [slots][N] node de-dup stamps
};

__device__ __forceinline__ int ring_slot(const DevState* st) {
  return (int)((st->step - st->ring_base) % st->ring_slots);
}
// The ring slot of the current iteration: given by the host (which mirrors
// the step) so that the slot's loads need not wait for the state, else read
// from the state.
__device__ __forceinline__ int ring_slot_of(int32_t host, const DevState* st) {
  return host >= 0 ? host : ring_slot(st);
}

struct SamplerArgs {
  DevGraph g;
  DevState* st;
  Key key;
  int32_t strategy;
  int64_t m;             // mini_batch_size
  int64_t max_source;
  Ring ring;
  int64_t t0;            // first step to draw (<0: st->step); block i draws t0+i into slot i
  int2* table;           // non-link de-dup tables in HBM [slots][table_cap] (if not in LDS)
  int64_t table_cap;     // power of two, > 1.25 x (m + 1024) entries
  int32_t table_in_lds;  // 1: the table lives in dynamic LDS instead
  // random-edge strategy: held-out + test entry count and the edge de-dup
  // tables in HBM [slots][re_cap] (keys a<<32|b, smallest attempt)
  int64_t held_test;
  unsigned long long* re_keys;
  int32_t* re_att;
  int64_t re_cap;
};

struct NeighborArgs {
  DevGraph g;
  DevState* st;
  Key key;
  Ring ring;
  int2* gtable;           // global de-dup tables [gtable_slots][tcap] or null
  int32_t tcap;           // per-node table capacity (power of two)
  int32_t use_lds;
  int32_t gtable_slots;   // number of global tables (when !use_lds)
};

struct PhiArgs {
  const void* pi;         // [N][ld]
  int64_t ld;
  int32_t K;
  int32_t N;
  int32_t use_ring;          // nodes / nbrs / count of the current ring slot
  Ring ring;
  const int32_t* nodes;      // injected path
  int32_t count_host;        // injected path
  const int32_t* nbrs;       // injected path [slot][nnb]
  int32_t nnb;
  const double* beta;        // [K]
  const DevState* st;
  DevHyper hp;
  double eps_override;       // <0: derive eps_t from st->step
  const void* noise;         // injected [slot][K] or null
  Key key;
  void* staging;             // [slot][ld]
  int32_t ring_slot_host;    // the current ring slot when the host knows it, else -1
};

struct ScatterArgs {
  void* pi;
  int64_t ld;
  int32_t K;
  int32_t use_ring;
  Ring ring;
  const DevState* st;
  const int32_t* nodes;
  int32_t count_host;
  const void* staging;
  int32_t ring_slot_host;    // as PhiArgs
};

struct BetaArgs {
  const void* pi;
  int64_t ld;
  int32_t K;
  int32_t use_ring;         // edges / y / count of the current ring slot
  Ring ring;
  const DevState* st;
  const int32_t* edges;     // injected path [E][2]
  const uint8_t* y;         // injected path [E]
  int32_t count_host;
  const double* beta;
  const double* theta;      // [K][2]
  double epsilon;
  double* partials;         // [blocks][2K]
  int32_t ring_slot_host;   // as PhiArgs
  int32_t blocks;
};

constexpr int kBetaWaves = 4;   // wavefronts per k_beta workgroup

struct ThetaArgs {
  double* theta;
  double* beta;
  int32_t K;
  const double* partials;  // [rows][2K]
  int32_t blocks;          // rows of partials
  DevState* st;
  DevHyper hp;
  double eps_override;
  double scale_override;   // <0: scale of the current ring slot
  const IterMeta* meta;    // ring metadata (scale, totals) when scale_override < 0
  const double* noise;     // injected [K][2] or null
  Key key;
  int32_t bump_step;       // 1: st->step += 1 at the end (end of an iteration)
};

// label of a held-out position that joins two trails (no pair)
constexpr uint8_t kPpxBridge = 2;

struct PpxArgs {
  const void* pi;
  int64_t ld;
  int32_t K;
  const int32_t* seq;     // [H + 1]: position t is the pair (seq[t], seq[t + 1])
  const uint8_t* labels;  // [H]: 0 or 1, kPpxBridge for a position that is no pair
  int64_t H;              // positions
  double* R;              // running average per position
  double avg_count;       // Learner::average_count
  const double* beta;
  double epsilon;
  double* partials;       // [blocks][4]
  int32_t blocks;
  DevState* st;
};

// launchers (return hipError_t of the launch)
// Draw `fills` consecutive iterations into ring slots 0..fills-1.
hipError_t launch_minibatch(const SamplerArgs& a, int32_t fills, hipStream_t s);
// Lets the minibatch kernels use `bytes` of dynamic LDS (above the default
// 64 KB per workgroup where the device allows it).
hipError_t prepare_minibatch_lds(size_t bytes);
hipError_t launch_neighbors(const NeighborArgs& a, int32_t fills, hipStream_t s);
// The neighbour sampler's dynamic LDS per workgroup (4 wavefronts): each
// wavefront's de-dup table of tcap entries and a node row of kNbrRowLds.
constexpr int kNbrRowLds = 256;
inline size_t neighbor_lds_bytes(int32_t tcap) {
  return (size_t)4 * ((size_t)tcap * 8 + (size_t)kNbrRowLds * 4);
}
// update_phi + the phi step of grid_nodes minibatch nodes, new rows staged
// (launch_scatter publishes them)
hipError_t launch_update_phi(const PhiArgs& a, int precision, int32_t grid_nodes,
                             hipStream_t s, int waves, int unroll);
// Launch shape (wavefronts per node, neighbour rows in flight) available for K.
bool update_phi_shape_supported(int32_t K, int precision, int waves, int unroll);
constexpr int kPhiWavesDefault = 4;
constexpr int kPhiUnrollDefault = 1;
hipError_t launch_scatter(const ScatterArgs& a, int precision, int32_t grid_nodes,
                          hipStream_t s);
// Static LDS of k_minibatch (bytes), for sizing its dynamic de-dup table.
constexpr int kMinibatchStaticLds = 46 * 1024;
hipError_t launch_beta(const BetaArgs& a, int precision, hipStream_t s);
hipError_t launch_theta(const ThetaArgs& a, hipStream_t s);
hipError_t launch_perplexity(const PpxArgs& a, int precision, hipStream_t s);
hipError_t launch_init_phi(void* pi, int64_t ld, int32_t K, int32_t N, Key key,
                           int precision, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, int64_t n, hipStream_t s);
hipError_t launch_fill_f64(double* p, double v, int64_t n, hipStream_t s);
// DevState::step = step (with its eps_t) and the error word cleared, in stream order
hipError_t launch_set_state_step(DevState* st, int64_t step, const DevHyper& hp, hipStream_t s);
// DevState::ring_base and ring_slots, in stream order
hipError_t launch_set_ring(DevState* st, int64_t base, int32_t slots, hipStream_t s);
// Which (lanes-per-row, chunks) configuration serves K; false if K too big.
bool row_config(int32_t K, int precision, int* lpr, int* ch);

}  // namespace ammsb

namespace ammsb {
// Sets bit 31 of nbrs[i][j] when (nodes[i], nbr) is a link of the CSR
// (injected-draw path of update_phi).
hipError_t launch_label_nbrs(DevGraph g, const int32_t* nodes, int32_t n_nodes, int32_t* nbrs,
                             int32_t nnb, hipStream_t s);
}  // namespace ammsb

namespace ammsb {
// Contiguous read of `bytes` (a multiple of 16) starting at p.
hipError_t launch_read_probe(const void* p, int64_t bytes, double* sink, hipStream_t s);
// update_phi's row gather of the current ring iteration without the
// arithmetic or the row write, in launch shape (waves, unroll): (4, 1), or
// (8, 2) where update_phi has it.
hipError_t launch_gather_probe(const PhiArgs& a, int precision, int32_t grid_nodes, double* sink,
                               hipStream_t s, int waves, int unroll);
}  // namespace ammsb

namespace ammsb {
// Minibatch log-likelihood and its adjoint (differentiation).
struct LogLikArgs {
  const void* pi;            // resident state [N][ld] (neighbour rows)
  int64_t ld;
  int32_t K;
  int64_t M;                 // source nodes
  const int32_t* nodes;      // [M] source node ids (a pair with nbr == node is skipped)
  const void* phi;           // [M][phi_ld] caller's phi rows of the sources
  int64_t phi_ld;
  const double* theta;       // [K][2]
  const int64_t* pair_ptr;   // [M+1]
  const int32_t* pair_nbr;   // [P]
  const uint8_t* pair_y;     // [P]
  double epsilon;
  double* per_source;        // fwd: [M] log-likelihood of each source's pairs
  const double* gout;        // bwd: dOut/dL (one double on the device)
  void* grad_phi;            // bwd: [M][phi_ld]
  double* dbeta_partials;    // bwd: [blocks][K]
  int32_t blocks;
  int32_t N;                 // node ids must lie in [0, N)
  int64_t P;                 // pairs (pair_ptr is clamped to [0, P])
  int32_t* error;            // ERR_BAD_INPUT is or-ed in for out-of-range ids
};
// pairs/prefetch: launch shape of the K-chunked kernel (pairs per LDS
// exchange, next pairs' loads issued before the exchange); ignored when one
// wavefront holds a row
// rows: rows wider than 2 KB run whole rows per wavefront against the
// source row held in registers (k_loglik_rows) instead of the K-chunked
// kernel
hipError_t launch_loglik(const LogLikArgs& a, int precision, bool backward, hipStream_t s,
                         int pairs, int prefetch, bool rows);
constexpr int kLlRowWaves = 4;  // wavefronts per workgroup of k_loglik_rows(_bwd)
constexpr int64_t kLdsPerCu = 160 * 1024;  // LDS bytes per compute unit (gfx950)
// 16-byte vectors per lane of a whole row (4, 8 or 16) for the whole-row
// kernels; 0 where one wavefront's 2 KB holds the row or the row is too wide.
// With rows, the adjoint writes one beta partial row per workgroup.
int loglik_rows_chunks(int32_t K, int precision);
// Dynamic LDS bytes of one k_loglik_rows workgroup (0: not used for K)
size_t loglik_rows_lds(int32_t K, int precision, bool backward);
// Resident workgroups per compute unit of the whole-row adjoint
int loglik_rows_blocks_per_cu(int32_t K, int precision);
bool loglik_kc_shape_supported(int pairs, int prefetch);
// Wavefronts per source of the K-chunked log-likelihood kernel used for rows
// wider than 2 KB, 0 where one wavefront per row is used.
int loglik_kc_waves(int32_t K, int precision);
// total = sum_s per_source[s] (fixed order)
hipError_t launch_sum(const double* x, int64_t n, double* out, hipStream_t s);
// grad_theta[k][0..1] = gout * dbeta_k * (-theta_k1, theta_k0) / (theta_k0+theta_k1)^2,
// dbeta_k = sum over blocks of dbeta_partials
hipError_t launch_theta_grad(const double* dbeta_partials, int32_t blocks, int32_t K,
                             const double* theta, const double* gout, double* grad_theta,
                             hipStream_t s);
// ---------------------------------------------------------------------------
// Batched chains: one launch per iteration phase covers every chain of a
// batch; each kernel reads its chain's arguments from a device array.
constexpr int kBatchWaves = 4;       // wavefronts per workgroup of the batched update_phi
constexpr int kMaxBatchChains = 1024;
// Row shapes of the batched update_phi and update_beta: (LPR, CH) =
// lanes per row and 16-byte vectors per lane.
constexpr int kBatchCfgs = 10;
constexpr int kBatchCfgLpr[kBatchCfgs] = {4, 8, 16, 32, 64, 4, 8, 16, 32, 64};
constexpr int kBatchCfgCh[kBatchCfgs] = {1, 1, 1, 1, 1, 2, 2, 2, 2, 2};
// Per-chain launch shape of the batched update_phi: cfg indexes the tables
// above, npb = minibatch nodes per workgroup (power of two, at most the
// workgroup's kBatchWaves * 64/LPR row groups).
struct BatchPhiShape {
  int32_t cfg;
  int32_t npb;
};
// Per-iteration work layout of a batch.  Chains are laid out in a work
// order set by the host (order[p] = the chain at position p, the most
// expensive work items first, so that the last items to finish are short):
// items[p] = first work item of position p (items[n] = total), pub[p]
// likewise for the batched scatter's wavefront tasks, info[p] = (chain,
// ring slot, minibatch node count, 0) of the current iteration; counter =
// work queue head.
struct BatchIndex {
  const int32_t* order;
  int32_t* items;
  int32_t* pub;
  int4* info;
  int32_t* counter;
};
// Samplers of `fills` iterations ahead into ring slots 0..fills-1.
// Iterations t0 .. t0+fills-1 of every chain into ring slots
// (slot0 + j) mod slots; DevState::ring_base is left as it is.
hipError_t launch_minibatch_batch(const SamplerArgs* args, int n, int64_t t0, int slot0, int fills,
                                  int slots, size_t lds, hipStream_t s);
hipError_t launch_neighbors_batch(const NeighborArgs* args, int n, int slot0, int fills, int slots,
                                  int max_node_cap, size_t lds, hipStream_t s);
// cfg: each chain's update_beta row shape, also the batched scatter's
hipError_t launch_batch_index(const PhiArgs* args, const BatchPhiShape* shapes,
                              const int32_t* cfg, int n, BatchIndex ix, hipStream_t s);
hipError_t launch_update_phi_batch(const PhiArgs* args, const BatchPhiShape* shapes, int n,
                                   BatchIndex ix, int precision, int grid, hipStream_t s);
// Resident workgroups per CU of the batched update_phi for n chains (its
// persistent grid is this times the CU count).
int update_phi_batch_blocks_per_cu(int n, int precision);
// update_pi of the batch: the persistent gather stages each node's new row
// (phi step applied, phi_sum included), the scatter writes it into pi
hipError_t launch_scatter_batch(const PhiArgs* args, const int32_t* cfg, int n, BatchIndex ix,
                                int precision, int grid, hipStream_t s);
hipError_t launch_beta_batch(const BetaArgs* args, const int32_t* cfg, int n, int max_blocks,
                             int precision, hipStream_t s);
hipError_t launch_theta_batch(const ThetaArgs* args, int n, int max_K, hipStream_t s);
// Resident workgroups per CU of the batched update_beta.
int beta_batch_blocks_per_cu(int precision);
// Batched update_phi shape validity for K / precision.
bool batch_cfg_valid(int cfg, int32_t K, int precision);
int batch_cfg_default(int32_t K, int precision);  // -1: K too large for batching

}  // namespace ammsb
