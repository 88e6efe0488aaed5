// Host driver and C-ABI of libammsb.so.
//
// The context is the MI355X-native counterpart of MCMCSamplerStochastic +
// Learner (c++/src/mcmc/learning/mcmc_sampler_stochastic.{h,cc},
// learner.{h,cc}): options and derived defaults, HBM-resident state
// (pi ++ phi_sum rows, theta, beta), graph (CSR + held-out∪test hash),
// and the run() loop with its perplexity schedule, convergence test,
// stdout log line and per-phase timers.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <initializer_list>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ammsb.h"
#include "host_util.h"
#include "kernels.h"

// Device scratch of the no-allocation adjoint (ammsb_loglik_bwd_noalloc):
// a file-scope array, one copy per GPU, held until the process ends.  Its
// first 256 bytes hold the error word, the rest the adjoint's beta partial
// rows ([rows][K] doubles).  Host code never passes `scratch` itself to a
// kernel (on the host it names a stand-in, not device memory): it uses the
// per-device address from hipGetSymbolAddress (ammsb::noalloc_scratch).
static constexpr size_t kNoAllocHeader = 256;
static constexpr size_t kNoAllocScratchBytes = ((size_t)32 << 20) + 4096;  // 32 MiB + 4 KiB
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace ammsb {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* get_error() { return g_err.c_str(); }

namespace {

enum Phase {
  PH_OUTER = 0,
  PH_PERPLEXITY,
  PH_MINIBATCH,
  PH_NODES,
  PH_NEIGHBORS,
  PH_UPDATE_PHI,
  PH_UPDATE_PI,
  PH_UPDATE_BETA,
  PH_COUNT
};
const char* kPhaseNames[PH_COUNT] = {"  outer",           "  perplexity",
                                     "  sample_mini_batch", "  nodes_in_mini_batch",
                                     "  sample_neighbor_nodes", "  update_phi",
                                     "  update_pi",         "  update_beta"};

struct PendingEvent {
  int phase;
  hipEvent_t a, b;
};

}  // namespace
}  // namespace ammsb

using namespace ammsb;

struct ammsb_ctx {
  ammsb_options opt;
  int32_t N = 0;
  int64_t E = 0;
  int32_t K = 0;
  int64_t ld = 0;
  int prec = 0;
  size_t fsize = 8;
  int64_t m = 0, n = 0, nnb = 0;
  int64_t interval = 50, stats_interval = 65536;
  int64_t edge_cap = 0, node_cap = 0;
  int64_t max_degree = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;           // false: the stream of the batch that owns the chain
  ammsb_batch* owner = nullptr;     // batch that owns this chain, if any
  int32_t max_ring_slots = 16;
  int64_t ring_gen = 0;             // bumped by every single-context ring refill
  Key key{};
  DevHyper hp{};
  // device graph
  int64_t* d_row_ptr = nullptr;
  int32_t* d_cols = nullptr;
  uint64_t* d_hkeys = nullptr;
  uint8_t* d_hvals = nullptr;
  uint64_t hmask = 0;
  uint64_t* d_bloom = nullptr;
  uint64_t bloom_mask = 0;
  int32_t* d_esrc = nullptr;                 // random-edge: row of each CSR entry
  unsigned long long* d_re_keys = nullptr;   // random-edge de-dup tables
  int32_t* d_re_att = nullptr;
  int64_t re_cap = 0;
  // held-out set (perplexity)
  int64_t H = 0, T = 0;
  // The held-out pairs as one node sequence: position t is the pair
  // (seq[t], seq[t + 1]) with label 0 or 1, or kPpxBridge where the sequence
  // passes from one trail to the next (no pair).  ppx_len positions.
  int64_t ppx_len = 0;
  int32_t* d_ho_seq = nullptr;
  uint8_t* d_ho_labels = nullptr;
  double* d_R = nullptr;  // running average per position
  // state
  char* d_pi = nullptr;       // [N][ld] of the precision type
  double* d_theta = nullptr;
  double* d_beta = nullptr;
  DevState* d_st = nullptr;
  // iteration scratch
  Ring ring{};              // sample-ahead ring (device pointers)
  int64_t ring_base = 1;    // host mirror of DevState::ring_base
  bool ring_valid = false;
  char* d_staging = nullptr;  // [node_cap][ld]
  int2* d_mbtable = nullptr;
  int64_t mbtable_cap = 0;
  int32_t mbtable_in_lds = 0;
  int32_t phi_waves = kPhiWavesDefault, phi_unroll = kPhiUnrollDefault;
  // K-chunked log-likelihood launch shapes [0] forward, [1] adjoint
  int32_t ll_pairs[2] = {0, 0}, ll_prefetch[2] = {0, 0};
  int32_t ll_fwd_rows = 1;  // forward of wide rows: whole rows per wavefront (else K-chunked)
  int32_t ll_bwd_rows = 1;  // adjoint of wide rows: the same
  int32_t cus = 256;        // compute units of the device
  double* d_ll_scratch = nullptr;   // differentiation: per-source values / dbeta partials
  size_t ll_scratch_bytes = 0;
  int2* d_nbr_gtable = nullptr;
  int32_t nbr_tcap = 0, nbr_use_lds = 1, nbr_gslots = 0;
  double* d_beta_partials = nullptr;
  int32_t beta_blocks = 1;
  double* d_ppx_partials = nullptr;
  int32_t ppx_blocks = 1;
  // device workspace of the injected-draw operators (ammsb_update_phi /
  // _beta): grown on demand, kept with the context
  char* d_opws = nullptr;
  size_t opws_bytes = 0;
  // host mirrors of Learner counters
  int64_t step_count = 1;
  int64_t average_count = 1;
  std::deque<double> ppx_cb;  // boost::circular_buffer<Float>(10), learner.cc:8
  bool state_ready = false;
  // timers
  bool timers_on = false;
  uint32_t timer_mask = 0;  // phases timed (bit = Phase)
  double t_secs[PH_COUNT] = {0};
  int64_t t_ticks[PH_COUNT] = {0};
  std::vector<PendingEvent> pending;
  std::vector<hipEvent_t> event_pool;
};

namespace {

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                \
      return AMMSB_ERR_HIP;                                                        \
    }                                                                              \
  } while (0)

template <typename T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
              " bytes): " + hipGetErrorString(e));
    *p = nullptr;
    return AMMSB_ERR_OOM;
  }
  return AMMSB_OK;
}

// Device address of `scratch` on `device` (the current device), looked up
// once per device and kept: every GPU holds its own copy of the array.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("no-allocation scratch: device ordinal out of range");
    return AMMSB_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return AMMSB_ERR_HIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return AMMSB_OK;
}

#define ALLOC_TRY(p, count)           \
  do {                                \
    int _rc = dalloc(&(p), (count));  \
    if (_rc != AMMSB_OK) return _rc;  \
  } while (0)

// Carves the operators' device buffers out of the context's workspace,
// 256-byte aligned; grows it (after the stream has drained) when the call
// needs more than it holds.
int op_workspace(ammsb_ctx* c, std::initializer_list<std::pair<void**, size_t>> parts) {
  size_t need = 0;
  for (const auto& p : parts) need += (p.second + 255) / 256 * 256;
  if (need > c->opws_bytes) {
    if (c->d_opws) {
      HIP_TRY(hipStreamSynchronize(c->stream));
      HIP_TRY(hipFree(c->d_opws));
    }
    c->d_opws = nullptr;
    c->opws_bytes = 0;
    const int rc = dalloc(&c->d_opws, need);
    if (rc != AMMSB_OK) return rc;
    c->opws_bytes = need;
  }
  char* q = c->d_opws;
  for (const auto& p : parts) {
    *p.first = p.second ? q : nullptr;
    q += (p.second + 255) / 256 * 256;
  }
  return AMMSB_OK;
}

uint64_t pow2_at_least(uint64_t x) {
  uint64_t c = 1;
  while (c < x) c <<= 1;
  return c;
}

DevGraph dev_graph(const ammsb_ctx* c) {
  DevGraph g{};
  g.N = c->N;
  g.E = c->E;
  g.row_ptr = c->d_row_ptr;
  g.cols = c->d_cols;
  g.hkeys = c->d_hkeys;
  g.hvals = c->d_hvals;
  g.hmask = c->hmask;
  g.bloom = c->d_bloom;
  g.bloom_mask = c->bloom_mask;
  g.esrc = c->d_esrc;
  g.n_recip = fastmod_recip((uint32_t)c->N);
  g.e2_recip = c->E > 0 ? fastmod_recip((uint32_t)(2 * c->E)) : 0;
  return g;
}

hipEvent_t take_event(ammsb_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int flush_timers(ammsb_ctx* c) {
  if (c->pending.empty()) return AMMSB_OK;
  HIP_TRY(hipEventSynchronize(c->pending.back().b));
  for (auto& p : c->pending) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
    c->t_secs[p.phase] += ms * 1e-3;
    c->t_ticks[p.phase] += 1;
    c->event_pool.push_back(p.a);
    c->event_pool.push_back(p.b);
  }
  c->pending.clear();
  return AMMSB_OK;
}

struct PhaseScope {
  ammsb_ctx* c;
  int phase;
  hipEvent_t a = nullptr;
  PhaseScope(ammsb_ctx* c_, int ph) : c(c_), phase(ph) {
    if (c->timers_on && (c->timer_mask >> ph & 1u)) {
      a = take_event(c);
      if (a) (void)hipEventRecord(a, c->stream);
    }
  }
  ~PhaseScope() {
    if (c->timers_on && a) {
      hipEvent_t b = take_event(c);
      if (b) {
        (void)hipEventRecord(b, c->stream);
        c->pending.push_back({phase, a, b});
        if (c->pending.size() > 4096) flush_timers(c);
      }
    }
  }
};

// ---- kernel arguments of one iteration, built from the context
SamplerArgs sampler_args(const ammsb_ctx* c, int64_t t0) {
  SamplerArgs sa{};
  sa.g = dev_graph(c);
  sa.st = c->d_st;
  sa.key = c->key;
  sa.strategy = c->opt.strategy;
  sa.m = c->m;
  sa.max_source = c->opt.sampler_max_source;
  sa.ring = c->ring;
  sa.t0 = t0;
  sa.table = c->d_mbtable;
  sa.table_cap = c->mbtable_cap;
  sa.table_in_lds = c->mbtable_in_lds;
  sa.held_test = c->H + c->T;
  sa.re_keys = c->d_re_keys;
  sa.re_att = c->d_re_att;
  sa.re_cap = c->re_cap;
  return sa;
}

NeighborArgs neighbor_args(const ammsb_ctx* c) {
  NeighborArgs na{};
  na.g = dev_graph(c);
  na.st = c->d_st;
  na.key = c->key;
  na.ring = c->ring;
  na.gtable = c->d_nbr_gtable;
  na.tcap = c->nbr_tcap;
  na.use_lds = c->nbr_use_lds;
  na.gtable_slots = c->nbr_gslots;
  return na;
}

PhiArgs phi_args(const ammsb_ctx* c) {
  PhiArgs pa{};
  pa.pi = c->d_pi;
  pa.ld = c->ld;
  pa.K = c->K;
  pa.N = c->N;
  pa.use_ring = 1;
  pa.ring = c->ring;
  pa.nodes = nullptr;
  pa.count_host = 0;
  pa.nbrs = nullptr;
  pa.nnb = (int32_t)c->nnb;
  pa.beta = c->d_beta;
  pa.st = c->d_st;
  pa.hp = c->hp;
  pa.eps_override = -1.0;
  pa.noise = nullptr;
  pa.key = c->key;
  pa.staging = c->d_staging;
  pa.ring_slot_host = -1;
  return pa;
}

ScatterArgs scatter_args(const ammsb_ctx* c) {
  ScatterArgs sc{};
  sc.pi = c->d_pi;
  sc.ld = c->ld;
  sc.K = c->K;
  sc.use_ring = 1;
  sc.ring = c->ring;
  sc.st = c->d_st;
  sc.nodes = nullptr;
  sc.count_host = 0;
  sc.staging = c->d_staging;
  sc.ring_slot_host = -1;
  return sc;
}

BetaArgs beta_args(const ammsb_ctx* c) {
  BetaArgs ba{};
  ba.pi = c->d_pi;
  ba.ld = c->ld;
  ba.K = c->K;
  ba.use_ring = 1;
  ba.ring = c->ring;
  ba.st = c->d_st;
  ba.edges = nullptr;
  ba.y = nullptr;
  ba.count_host = 0;
  ba.beta = c->d_beta;
  ba.theta = c->d_theta;
  ba.epsilon = c->hp.epsilon;
  ba.partials = c->d_beta_partials;
  ba.blocks = c->beta_blocks;
  ba.ring_slot_host = -1;
  return ba;
}

ThetaArgs theta_args(const ammsb_ctx* c) {
  ThetaArgs ta{};
  ta.theta = c->d_theta;
  ta.beta = c->d_beta;
  ta.K = c->K;
  ta.partials = c->d_beta_partials;
  ta.blocks = c->beta_blocks;
  ta.st = c->d_st;
  ta.hp = c->hp;
  ta.eps_override = -1.0;
  ta.scale_override = -1.0;
  ta.meta = c->ring.meta;
  ta.noise = nullptr;
  ta.key = c->key;
  ta.bump_step = 1;
  return ta;
}

// ---- sample-ahead: draw `fills` iterations starting at step t0 (t0 < 0: the
// device's step_count) into ring slots 0..fills-1
int fill_ring(ammsb_ctx* c, int64_t t0, int32_t fills) {
  c->ring_gen++;
  {
    PhaseScope ps(c, PH_MINIBATCH);
    HIP_TRY(launch_minibatch(sampler_args(c, t0), fills, c->stream));
  }
  {
    PhaseScope ps(c, PH_NEIGHBORS);
    HIP_TRY(launch_neighbors(neighbor_args(c), fills, c->stream));
  }
  return AMMSB_OK;
}

// ---- one SGRLD iteration on the device (mcmc_sampler_stochastic.cc:131-174)
// The minibatch and neighbour draws come from the ring, refilled on this
// stream every ring.slots iterations.
int enqueue_iteration(ammsb_ctx* c) {
  if (!c->ring_valid || c->step_count < c->ring_base ||
      c->step_count >= c->ring_base + c->ring.slots) {
    int rc = fill_ring(c, c->step_count, c->ring.slots);
    if (rc != AMMSB_OK) return rc;
    c->ring_base = c->step_count;
    c->ring_valid = true;
  }
  // the ring holds [ring_base, ring_base + slots) here (the sampler set the
  // device's ring_base to the same step), so the slot is known on the host
  const int32_t rs = (int32_t)(c->step_count - c->ring_base);
  PhiArgs pa = phi_args(c);
  pa.ring_slot_host = rs;
  ScatterArgs sa = scatter_args(c);
  sa.ring_slot_host = rs;
  BetaArgs ba = beta_args(c);
  ba.ring_slot_host = rs;
  {
    PhaseScope ps(c, PH_UPDATE_PHI);
    HIP_TRY(launch_update_phi(pa, c->prec, (int32_t)c->node_cap, c->stream, c->phi_waves,
                              c->phi_unroll));
  }
  {
    PhaseScope ps(c, PH_UPDATE_PI);
    HIP_TRY(launch_scatter(sa, c->prec, (int32_t)c->node_cap, c->stream));
  }
  {
    PhaseScope ps(c, PH_UPDATE_BETA);
    HIP_TRY(launch_beta(ba, c->prec, c->stream));
    HIP_TRY(launch_theta(theta_args(c), c->stream));
  }
  return AMMSB_OK;
}

int read_devstate(ammsb_ctx* c, DevState* h) {
  HIP_TRY(hipMemcpyAsync(h, c->d_st, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMMSB_OK;
}

int check_devstate(ammsb_ctx* c) {
  DevState h;
  int rc = read_devstate(c, &h);
  if (rc != AMMSB_OK) return rc;
  if (h.error) {
    std::string msg = "device sampler error:";
    if (h.error & ERR_NONLINK_ATTEMPTS) msg += " non-link attempt budget exceeded;";
    if (h.error & ERR_LINK_ROUNDS) msg += " link sampler found no training edges;";
    if (h.error & ERR_NEIGHBOR_ATTEMPTS) msg += " neighbour attempt budget exceeded;";
    if (h.error & ERR_MB_OVERFLOW) msg += " #edge in minibatch exceeds buffer size upper bound;";
    if (h.error & ERR_BAD_STRATEGY) msg += " invalid sampling strategy;";
    set_error(msg);
    return AMMSB_ERR_INVARIANT;
  }
  if (h.step != c->step_count) {
    set_error("device step_count diverged from host mirror");
    return AMMSB_ERR_STATE;
  }
  return AMMSB_OK;
}

int do_perplexity(ammsb_ctx* c, double* out) {
  PpxArgs p{};
  p.pi = c->d_pi;
  p.ld = c->ld;
  p.K = c->K;
  p.seq = c->d_ho_seq;
  p.labels = c->d_ho_labels;
  p.H = c->ppx_len;
  p.R = c->d_R;
  p.avg_count = (double)c->average_count;
  p.beta = c->d_beta;
  p.epsilon = c->hp.epsilon;
  p.partials = c->d_ppx_partials;
  p.blocks = c->ppx_blocks;
  p.st = c->d_st;
  {
    PhaseScope ps(c, PH_PERPLEXITY);
    HIP_TRY(launch_perplexity(p, c->prec, c->stream));
  }
  DevState h;
  int rc = read_devstate(c, &h);
  if (rc != AMMSB_OK) return rc;
  c->average_count += 1;  // learner.cc:180
  *out = h.ppx;
  return AMMSB_OK;
}

void print_stats(ammsb_ctx* c, FILE* f) {
  flush_timers(c);
  // timer.cc:18-56 tabular layout
  fprintf(f, "%-37s%13s%9s%15s\n", "timer", "total (s)", "ticks", "per tick (us)");
  for (int p = 0; p < PH_COUNT; ++p) {
    if (c->t_ticks[p] == 0) {
      fprintf(f, "%-36s<unused>\n", kPhaseNames[p]);
    } else {
      fprintf(f, "%-36s %12.3f %8lld %14.3f\n", kPhaseNames[p], c->t_secs[p],
              (long long)c->t_ticks[p], c->t_secs[p] * 1e6 / (double)c->t_ticks[p]);
    }
  }
  fflush(f);
}

bool is_converged(const ammsb_ctx* c) {
  // learner.cc:130-136
  const size_t n = c->ppx_cb.size();
  if (n < 2) return false;
  return std::abs(c->ppx_cb[n - 1] - c->ppx_cb[n - 2]) / c->ppx_cb[n - 2] <=
         c->opt.convergence_threshold;
}

void free_ctx(ammsb_ctx* c) {
  if (!c) return;
  void* ptrs[] = {c->d_row_ptr, c->d_cols,    c->d_hkeys,    c->d_hvals,     c->d_ho_seq,
                  c->d_ho_labels, c->d_R,     c->d_pi,       c->d_theta,     c->d_beta,
                  c->d_st,      c->ring.edges, c->ring.y,    c->ring.nodes,  c->ring.nbrs,
                  c->d_staging, c->d_mbtable, c->ring.stamp, c->d_nbr_gtable, c->d_beta_partials,
                  c->ring.meta, c->d_ll_scratch, c->d_bloom, c->d_esrc, c->d_re_keys, c->d_re_att,
                  c->d_ppx_partials, c->d_opws};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int create_impl(const ammsb_options* opts, const ammsb_graph* g, const ammsb_split* split,
                ammsb_ctx* c) {
  c->opt = *opts;
  ammsb_options& o = c->opt;
  if (!g || !g->row_ptr || !g->cols || g->N < 2) {
    set_error("ammsb_create: bad graph");
    return AMMSB_ERR_ARG;
  }
  if (o.K < 1) {
    set_error("ammsb_create: K must be >= 1");
    return AMMSB_ERR_ARG;
  }
  if (o.precision != AMMSB_FP64 && o.precision != AMMSB_FP32) {
    set_error("ammsb_create: unknown precision");
    return AMMSB_ERR_ARG;
  }
  if (o.strategy < 0 || o.strategy > 3) {
    set_error("ammsb_create: Invalid sampling strategy");
    return AMMSB_ERR_ARG;
  }
  if (o.strategy == AMMSB_RANDOM_EDGE && 2 * g->num_edges > (int64_t)INT32_MAX) {
    set_error("ammsb_create: random-edge minibatches need 2*num_edges < 2^31");
    return AMMSB_ERR_ARG;
  }
  c->N = g->N;
  c->E = g->num_edges;
  c->K = (int32_t)o.K;
  c->prec = o.precision;
  c->fsize = c->prec == AMMSB_FP64 ? 8 : 4;
  // K-chunked log-likelihood shapes measured at N=1M, K=2048 (2.5M pairs):
  // FP64 (4 pairs, no prefetch) both ways; FP32 forward (8, no prefetch),
  // adjoint (2, prefetch), whose registers leave 4 wavefronts per SIMD
  c->ll_pairs[0] = c->prec == AMMSB_FP64 ? 4 : 8;
  c->ll_prefetch[0] = 0;
  c->ll_pairs[1] = c->prec == AMMSB_FP64 ? 4 : 2;
  c->ll_prefetch[1] = c->prec == AMMSB_FP64 ? 0 : 1;
  // whole-row adjoint in both precisions (N=1M, K=2048, 2.5M pairs, against
  // the K-chunked one: FP64 6.9 against 8.6 ms, FP32 3.5 against 4.2 ms)
  c->ll_bwd_rows = 1;
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, o.device) ==
            hipSuccess &&
        cus > 0)
      c->cus = cus;
  }
  int lpr, ch;
  if (!row_config(c->K, c->prec, &lpr, &ch)) {
    set_error("ammsb_create: K too large for the device kernels (max 2048 FP64 / 4096 FP32)");
    return AMMSB_ERR_ARG;
  }
  const int rpw = 64 / lpr;
  // row stride: K+1 elements rounded up to 128 bytes
  const int64_t per128 = 128 / (int64_t)c->fsize;
  c->ld = ((int64_t)c->K + 1 + per128 - 1) / per128 * per128;
  // update_phi: 8 wavefronts per node with 2 rows in flight where the row
  // shape offers it (whole-wavefront rows), measured 1-5% faster than 4 x 1
  // for K 256..2048 in both precisions on one MI355X; else 4 x 1
  if (update_phi_shape_supported(c->K, c->prec, 8, 2)) {
    c->phi_waves = 8;
    c->phi_unroll = 2;
  }
  // derived defaults (options.h:257-262, mcmc_sampler_stochastic.cc:17-44,
  // learner.cc:87-91)
  if (o.a == 0.0) o.a = std::pow(o.b, -o.c);
  if (o.alpha == 0.0) o.alpha = 1.0 / (double)o.K;
  c->interval = o.interval == 0 ? 50 : o.interval;
  c->stats_interval = o.stats_print_interval == 0 ? 64 * 1024 : o.stats_print_interval;
  c->n = o.num_node_sample == 0 ? c->N / 50 : o.num_node_sample;
  c->m = o.mini_batch_size == 0 ? c->N / 10 : o.mini_batch_size;
  if (o.held_out_ratio == 0.0) o.held_out_ratio = 0.1;
  if (c->n < 1 || c->n + 1 > c->N - 1) {
    set_error("ammsb_create: num_node_sample must be in [1, N-2]");
    return AMMSB_ERR_ARG;
  }
  if (c->m < 1 || c->m > (int64_t)c->N - 1) {
    set_error("ammsb_create: mini_batch_size must be in [1, N-1]");
    return AMMSB_ERR_ARG;
  }
  if (o.sampler_max_source < 1 || o.sampler_max_source > 1024) {
    set_error("ammsb_create: sampler.max-source must be in [1, 1024]");
    return AMMSB_ERR_ARG;
  }
  c->nnb = c->n + 1;
  c->key = make_key((uint64_t)(uint32_t)o.random_seed);
  c->hp.alpha = o.alpha;
  c->hp.epsilon = o.epsilon;
  c->hp.a = o.a;
  c->hp.b = o.b;
  c->hp.c = o.c;
  c->hp.eta0 = o.eta0;
  c->hp.eta1 = o.eta1;
  c->hp.Nn = (1.0 * (double)c->N) / (double)c->n;
  if (o.precision == AMMSB_FP32) c->hp.Nn = (double)(float)((1.0f * (float)c->N) / (float)c->n);

  HIP_TRY(hipSetDevice(o.device));
  if (!c->stream) HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  {
    // the no-allocation adjoint's scratch address on this device, once
    void* sp = nullptr;
    const int rc = noalloc_scratch(o.device, &sp);
    if (rc != AMMSB_OK) return rc;
  }

  // ---- graph
  for (int32_t i = 0; i < c->N; ++i)
    c->max_degree = std::max<int64_t>(c->max_degree, g->row_ptr[i + 1] - g->row_ptr[i]);
  if (g->row_ptr[c->N] != 2 * c->E) {
    set_error("ammsb_create: row_ptr[N] != 2*num_edges");
    return AMMSB_ERR_ARG;
  }
  ALLOC_TRY(c->d_row_ptr, (size_t)c->N + 1);
  ALLOC_TRY(c->d_cols, (size_t)(2 * c->E));
  HIP_TRY(hipMemcpy(c->d_row_ptr, g->row_ptr, sizeof(int64_t) * (c->N + 1), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_cols, g->cols, sizeof(int32_t) * 2 * c->E, hipMemcpyHostToDevice));

  // ---- split
  int32_t *ho = nullptr, *te = nullptr;
  uint8_t *hol = nullptr, *tel = nullptr;
  int64_t hc = 0, tc = 0;
  bool own_split = false;
  if (split) {
    ho = const_cast<int32_t*>(split->held_out_edges);
    hol = const_cast<uint8_t*>(split->held_out_labels);
    hc = split->held_out_count;
    te = const_cast<int32_t*>(split->test_edges);
    tel = const_cast<uint8_t*>(split->test_labels);
    tc = split->test_count;
    if (hc < 0 || tc < 0 || (hc && (!ho || !hol)) || (tc && (!te || !tel))) {
      set_error("ammsb_create: split arrays missing or negative counts");
      return AMMSB_ERR_ARG;
    }
    for (int64_t i = 0; i < 2 * hc + 2 * tc; ++i) {
      const int32_t v = i < 2 * hc ? ho[i] : te[i - 2 * hc];
      if (v < 0 || v >= c->N) {
        set_error("ammsb_create: split pair with a node id outside [0, N)");
        return AMMSB_ERR_ARG;
      }
    }
  } else {
    int rc = ammsb_split_generate(g, o.held_out_ratio, (uint64_t)(uint32_t)o.random_seed, &ho,
                                  &hol, &hc, &te, &tel, &tc);
    if (rc != AMMSB_OK) return rc;
    own_split = true;
  }
  c->H = hc;
  c->T = tc;
  EdgeHashHost hh;
  hh.init(hc + tc);
  for (int64_t i = 0; i < hc; ++i) {
    const int32_t a = std::min(ho[2 * i], ho[2 * i + 1]), b = std::max(ho[2 * i], ho[2 * i + 1]);
    hh.insert(edge_key((uint32_t)a, (uint32_t)b), (uint8_t)(1 | (hol[i] ? 4 : 0)));
  }
  for (int64_t i = 0; i < tc; ++i) {
    const int32_t a = std::min(te[2 * i], te[2 * i + 1]), b = std::max(te[2 * i], te[2 * i + 1]);
    hh.insert(edge_key((uint32_t)a, (uint32_t)b), (uint8_t)(2 | (tel[i] ? 4 : 0)));
  }
  c->hmask = hh.mask;
  {
    BloomHost bl;
    bl.init(hc + tc);
    for (uint64_t k : hh.keys)
      if (k != kHashEmpty) bl.insert(k);
    c->bloom_mask = bl.mask;
    ALLOC_TRY(c->d_bloom, bl.words.size());
    HIP_TRY(hipMemcpy(c->d_bloom, bl.words.data(), 8 * bl.words.size(), hipMemcpyHostToDevice));
  }
  ALLOC_TRY(c->d_hkeys, hh.keys.size());
  ALLOC_TRY(c->d_hvals, hh.vals.size());
  HIP_TRY(hipMemcpy(c->d_hkeys, hh.keys.data(), 8 * hh.keys.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_hvals, hh.vals.data(), hh.vals.size(), hipMemcpyHostToDevice));
  bool ho_order_ok = true;
  {
    // Held-out pairs as trails: pair t + 1 starts at the node pair t ends
    // at, wherever the held-out graph allows, so the perplexity pass reads
    // one row per pair plus one per trail instead of two per pair.  The
    // edge likelihood is symmetric in its two rows (the same products in
    // the same k order), and the perplexity is a sum over pairs, so neither
    // the orientation nor the order of the pairs changes anything but the
    // summation order.
    //
    // Walks over unused pairs (Euler trails): first from each node with an
    // odd number of unused pairs, in id order (such a walk ends at another
    // such node), then from each node with unused pairs left (closed walks).
    // The number of trails is about half the number of odd-degree nodes: on
    // 8.1M uniform pairs over 4M nodes, 1.0M trails, 9.1M rows per pass
    // against 10.5M when each pair's smaller endpoint row is shared by the
    // pairs sorted next to it.  The trails are joined into one sequence with
    // a bridge position between consecutive trails.
    const int64_t N = c->N;
    std::vector<int64_t> st(N + 1, 0);
    for (int64_t i = 0; i < 2 * hc; ++i) ++st[ho[i] + 1];
    for (int64_t v = 0; v < N; ++v) st[v + 1] += st[v];
    std::vector<int64_t> inc(2 * hc);  // pair ids incident to each node, in pair order
    std::vector<int64_t> next(st.begin(), st.end() - 1);
    for (int64_t i = 0; i < 2 * hc; ++i) inc[next[ho[i]]++] = i / 2;
    for (int64_t v = 0; v < N; ++v) next[v] = st[v];
    std::vector<int64_t> rem(N);
    for (int64_t v = 0; v < N; ++v) rem[v] = st[v + 1] - st[v];
    std::vector<uint8_t> used(hc, 0);
    std::vector<int32_t> seq;
    std::vector<uint8_t> lab;
    seq.reserve(hc + hc / 4 + 1);
    lab.reserve(hc + hc / 4);
    int64_t pairs = 0;
    auto walk = [&](int32_t v) {
      bool started = false;
      while (true) {
        int64_t e = -1;
        while (next[v] < st[v + 1]) {
          const int64_t q = inc[next[v]++];
          if (!used[q]) {
            e = q;
            break;
          }
        }
        if (e < 0) return;
        if (!started) {
          if (!seq.empty()) lab.push_back(kPpxBridge);
          seq.push_back(v);
          started = true;
        }
        used[e] = 1;
        const int32_t w = ho[2 * e] == v ? ho[2 * e + 1] : ho[2 * e];
        --rem[v];
        --rem[w];
        lab.push_back(hol[e] != 0 ? 1 : 0);
        seq.push_back(w);
        ++pairs;
        v = w;
      }
    };
    for (int32_t v = 0; v < (int32_t)N; ++v)
      if (rem[v] & 1) walk(v);
    for (int32_t v = 0; v < (int32_t)N; ++v)
      while (rem[v] > 0) walk(v);
    ho_order_ok = pairs == hc && (int64_t)seq.size() == (int64_t)lab.size() + (hc ? 1 : 0);
    c->ppx_len = (int64_t)lab.size();
    ALLOC_TRY(c->d_ho_seq, seq.size());
    ALLOC_TRY(c->d_ho_labels, lab.size());
    ALLOC_TRY(c->d_R, lab.size());
    if (!seq.empty())
      HIP_TRY(hipMemcpy(c->d_ho_seq, seq.data(), 4 * seq.size(), hipMemcpyHostToDevice));
    if (!lab.empty())
      HIP_TRY(hipMemcpy(c->d_ho_labels, lab.data(), lab.size(), hipMemcpyHostToDevice));
  }
  if (own_split) {
    free(ho);
    free(hol);
    free(te);
    free(tel);
  }
  if (!ho_order_ok) {
    set_error("ammsb_create: internal error ordering the held-out pairs");
    return AMMSB_ERR_INVARIANT;
  }
  HIP_TRY(launch_fill_f64(c->d_R, 0.0, std::max<int64_t>(c->ppx_len, 1), c->stream));

  // ---- state
  ALLOC_TRY(c->d_pi, (size_t)c->N * c->ld * c->fsize);
  ALLOC_TRY(c->d_theta, (size_t)2 * c->K);
  ALLOC_TRY(c->d_beta, (size_t)c->K);
  ALLOC_TRY(c->d_st, 1);
  DevState h0{};
  h0.step = 1;
  h0.ring_slots = 1;
  HIP_TRY(hipMemcpy(c->d_st, &h0, sizeof(DevState), hipMemcpyHostToDevice));
  HIP_TRY(launch_set_state_step(c->d_st, 1, c->hp, c->stream));  // step 1 and its eps_t

  // ---- iteration scratch (sizes: network.cc:172-201)
  c->edge_cap = c->m + c->max_degree + 1;
  // random-edge with m >= E - |held-out| - |test| takes the whole training
  // set, which can hold more than m links (network.cc:238-240)
  if (o.strategy == AMMSB_RANDOM_EDGE && c->m >= c->E - (c->H + c->T))
    c->edge_cap = std::max<int64_t>(c->edge_cap, c->E + 1);
  c->node_cap = std::min<int64_t>(c->N, c->edge_cap + std::min<int64_t>(o.sampler_max_source, c->N) + 1);
  // random edges: up to two new endpoints per edge
  if (o.strategy == AMMSB_RANDOM_EDGE) c->node_cap = std::min<int64_t>(c->N, 2 * c->edge_cap);
  {
    // sample-ahead ring: up to 16 iterations, at most ~1 GiB
    const int64_t slot_bytes = c->edge_cap * 9 + c->node_cap * 4 + c->node_cap * c->nnb * 4 +
                               (int64_t)c->N * 4 + 64;
    // single contexts: two halves of up to 16 (the samplers draw one half
    // while the other is used)
    const int64_t max_slots = c->owner ? c->max_ring_slots : 2 * c->max_ring_slots;
    const int64_t slots = std::max<int64_t>(
        1, std::min<int64_t>(max_slots, ((int64_t)1 << 30) / slot_bytes));
    Ring& R = c->ring;
    R.slots = (int32_t)slots;
    HIP_TRY(hipMemcpy(&c->d_st->ring_slots, &R.slots, sizeof(int32_t), hipMemcpyHostToDevice));
    R.edge_cap = c->edge_cap;
    R.node_cap = c->node_cap;
    R.nnb = (int32_t)c->nnb;
    R.N = c->N;
    ALLOC_TRY(R.edges, (size_t)(slots * 2 * c->edge_cap));
    ALLOC_TRY(R.y, (size_t)(slots * c->edge_cap));
    ALLOC_TRY(R.nodes, (size_t)(slots * c->node_cap));
    ALLOC_TRY(R.nbrs, (size_t)(slots * c->node_cap * c->nnb));
    ALLOC_TRY(R.meta, (size_t)slots);
    ALLOC_TRY(R.stamp, (size_t)(slots * c->N));
    HIP_TRY(launch_fill_u32(R.stamp, 0u, slots * c->N, c->stream));
    HIP_TRY(hipMemsetAsync(R.meta, 0, sizeof(IterMeta) * slots, c->stream));
  }
  ALLOC_TRY(c->d_staging, (size_t)(c->node_cap * c->ld * c->fsize));
  {
    // non-link de-dup table: in LDS (load <= 0.8) when it fits next to the
    // sampler's ~45 KB of static LDS, else in HBM (load <= 0.5)
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeSharedMemPerBlockOptin, o.device) !=
            hipSuccess ||
        lds_max <= 0)
      HIP_TRY(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, o.device));
    const int64_t entries = c->m + 1024;
    const int64_t lds_cap = (int64_t)pow2_at_least((uint64_t)(entries + entries / 4 + 1));
    const int64_t static_lds = kMinibatchStaticLds;
    if (lds_cap * (int64_t)sizeof(int2) + static_lds <= std::min<int64_t>(lds_max, 160 * 1024) &&
        prepare_minibatch_lds((size_t)lds_cap * sizeof(int2)) == hipSuccess) {
      c->mbtable_cap = lds_cap;
      c->mbtable_in_lds = 1;
    } else {
      c->mbtable_cap = (int64_t)pow2_at_least((uint64_t)(2 * entries));
      c->mbtable_in_lds = 0;
    }
  }
  ALLOC_TRY(c->d_mbtable, c->mbtable_in_lds ? 1 : (size_t)(c->ring.slots * c->mbtable_cap));
  if (o.strategy == AMMSB_RANDOM_EDGE) {
    // row of every CSR entry, and the per-slot edge de-dup tables
    std::vector<int32_t> esrc((size_t)(2 * c->E));
    for (int32_t a = 0; a < c->N; ++a)
      for (int64_t q = g->row_ptr[a]; q < g->row_ptr[a + 1]; ++q) esrc[(size_t)q] = a;
    ALLOC_TRY(c->d_esrc, esrc.size());
    HIP_TRY(hipMemcpy(c->d_esrc, esrc.data(), 4 * esrc.size(), hipMemcpyHostToDevice));
    c->re_cap = (int64_t)pow2_at_least((uint64_t)(2 * (c->m + 1024)));
    ALLOC_TRY(c->d_re_keys, (size_t)(c->ring.slots * c->re_cap));
    ALLOC_TRY(c->d_re_att, (size_t)(c->ring.slots * c->re_cap));
  }
  c->nbr_tcap = (int32_t)pow2_at_least((uint64_t)(2 * (c->nnb + 64)));
  c->nbr_use_lds = neighbor_lds_bytes(c->nbr_tcap) <= 64 * 1024 ? 1 : 0;
  if (!c->nbr_use_lds) {
    const int64_t budget = (int64_t)1 << 30;
    c->nbr_gslots = (int32_t)std::max<int64_t>(
        1, std::min<int64_t>(c->node_cap, budget / ((int64_t)c->nbr_tcap * 8)));
    ALLOC_TRY(c->d_nbr_gtable, (size_t)c->nbr_gslots * c->nbr_tcap);
  }
  c->beta_blocks = (int32_t)std::max<int64_t>(
      1, std::min<int64_t>(256, (c->edge_cap + 4 * rpw * 4 - 1) / (4 * rpw * 4)));
  ALLOC_TRY(c->d_beta_partials, (size_t)c->beta_blocks * 2 * c->K);
  c->ppx_blocks = (int32_t)std::max<int64_t>(
      1, std::min<int64_t>(2048, (c->ppx_len + 4 * rpw * 8 - 1) / (4 * rpw * 8)));
  ALLOC_TRY(c->d_ppx_partials, (size_t)c->ppx_blocks * 4);
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (o.verbose) {
    // learner.cc:105-120, mcmc_sampler_stochastic.cc:84-91
    const double link_ratio = (double)c->E / (((double)c->N * (double)(c->N - 1)) / 2.0);
    fprintf(stderr, "Floating point precision: %dbit\n", (int)(c->fsize * 8));
    fprintf(stderr, "N %d E %lld link ratio %g K %d\n", c->N, (long long)c->E, link_ratio, c->K);
    fprintf(stderr, "minibatch size %lld epsilon %g alpha %g iterations %lld convergence %g\n",
            (long long)c->m, o.epsilon, o.alpha, (long long)o.max_iteration,
            o.convergence_threshold);
    fprintf(stderr, "num_node_sample %lld\n", (long long)c->n);
    fprintf(stderr, "a %g b %g c %g eta (%g,%g)\n", o.a, o.b, o.c, o.eta0, o.eta1);
  }
  return AMMSB_OK;
}

// pi ++ phi_sum rows between the device layout [N][ld] and the host layout
// [N][K+1]: contiguous pieces of whole device rows through a pinned staging
// buffer, repacked on the host (a 2D copy of millions of short rows into
// pageable memory is far slower).  Device padding past K+1 is written as 0.
int transfer_state(ammsb_ctx* c, void* host, bool to_host) {
  const size_t drow = (size_t)c->ld * c->fsize, hrow = (size_t)(c->K + 1) * c->fsize;
  const int64_t rows_per = std::max<int64_t>(1, ((int64_t)64 << 20) / (int64_t)drow);
  const int64_t nr_max = std::min<int64_t>(rows_per, c->N);
  char* stage = nullptr;
  HIP_TRY(hipHostMalloc((void**)&stage, (size_t)nr_max * drow, hipHostMallocDefault));
  memset(stage, 0, (size_t)nr_max * drow);
  hipError_t e = hipSuccess;
  for (int64_t r0 = 0; r0 < c->N && e == hipSuccess; r0 += rows_per) {
    const int64_t nr = std::min<int64_t>(rows_per, c->N - r0);
    char* dev = c->d_pi + (size_t)r0 * drow;
    char* h = (char*)host + (size_t)r0 * hrow;
    if (to_host) {
      e = hipMemcpyAsync(stage, dev, (size_t)nr * drow, hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e == hipSuccess)
        parallel_for(nr, [&](int64_t lo, int64_t hi, int) {
          for (int64_t r = lo; r < hi; ++r) memcpy(h + r * hrow, stage + r * drow, hrow);
        }, 4096);
    } else {
      parallel_for(nr, [&](int64_t lo, int64_t hi, int) {
        for (int64_t r = lo; r < hi; ++r) memcpy(stage + r * drow, h + r * hrow, hrow);
      }, 4096);
      e = hipMemcpyAsync(dev, stage, (size_t)nr * drow, hipMemcpyHostToDevice, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
  }
  const hipError_t ef = hipHostFree(stage);
  HIP_TRY(e);
  HIP_TRY(ef);
  return AMMSB_OK;
}

int set_beta_from_theta(ammsb_ctx* c, const double* theta) {
  std::vector<double> beta(c->K);
  for (int k = 0; k < c->K; ++k) {
    // np::row_normalize + SelectColumn(1) (mcmc_sampler_stochastic.cc:59-63)
    const double s = 0.0 + theta[2 * k] + theta[2 * k + 1];
    beta[k] = theta[2 * k + 1] / s;
  }
  HIP_TRY(hipMemcpy(c->d_theta, theta, sizeof(double) * 2 * c->K, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_beta, beta.data(), sizeof(double) * c->K, hipMemcpyHostToDevice));
  return AMMSB_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* ammsb_last_error(void) { return get_error(); }
const char* ammsb_version(void) { return "ammsb-mi355x 0.1 (gfx950)"; }

int ammsb_options_default(ammsb_options* o) {
  if (!o) return AMMSB_ERR_ARG;
  memset(o, 0, sizeof(*o));
  o->alpha = 0.0;
  o->eta0 = 1.0;
  o->eta1 = 1.0;
  o->epsilon = 0.0000001;
  o->a = 0.0;
  o->b = 1024;
  o->c = 0.5;
  o->K = 300;
  o->mini_batch_size = 0;
  o->num_node_sample = 0;
  o->strategy = AMMSB_STRATIFIED_RANDOM_NODE;
  o->sampler_max_source = 1;
  o->max_iteration = 10000000;
  o->interval = 0;
  o->stats_print_interval = 0;
  o->num_updates = 1000;
  o->held_out_ratio = 0.0;
  o->random_seed = 42;
  o->convergence_threshold = 0.000000000001;
  o->precision = AMMSB_FP64;
  o->device = 0;
  o->verbose = 0;
  return AMMSB_OK;
}

int ammsb_create(const ammsb_options* opts, const ammsb_graph* g, const ammsb_split* split,
                 ammsb_ctx** out) {
  if (!opts || !out) {
    set_error("ammsb_create: null argument");
    return AMMSB_ERR_ARG;
  }
  *out = nullptr;
  ammsb_ctx* c = nullptr;
  try {
    c = new ammsb_ctx();
  } catch (...) {
    set_error("ammsb_create: out of host memory");
    return AMMSB_ERR_OOM;
  }
  int rc;
  try {
    rc = create_impl(opts, g, split, c);
  } catch (const std::bad_alloc&) {
    set_error("ammsb_create: out of host memory");
    rc = AMMSB_ERR_OOM;
  }
  if (rc != AMMSB_OK) {
    free_ctx(c);
    return rc;
  }
  *out = c;
  return AMMSB_OK;
}

int ammsb_destroy(ammsb_ctx* c) {
  if (!c) return AMMSB_OK;
  if (c->owner) {
    set_error("ammsb_destroy: the context is a chain of a batch (ammsb_batch_destroy frees it)");
    return AMMSB_ERR_ARG;
  }
  (void)hipSetDevice(c->opt.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_ctx(c);
  return AMMSB_OK;
}

int ammsb_get_info(const ammsb_ctx* c, ammsb_info* info) {
  if (!c || !info) return AMMSB_ERR_ARG;
  memset(info, 0, sizeof(*info));
  info->N = c->N;
  info->K = c->K;
  info->num_edges = c->E;
  info->ld = c->ld;
  info->mini_batch_size = c->m;
  info->num_node_sample = c->n;
  info->interval = c->interval;
  info->stats_print_interval = c->stats_interval;
  info->max_minibatch_edges = c->edge_cap;
  info->max_minibatch_nodes = c->node_cap;
  info->held_out_count = c->H;
  info->test_count = c->T;
  info->step_count = c->step_count;
  info->average_count = c->average_count;
  info->alpha = c->opt.alpha;
  info->a = c->opt.a;
  info->b = c->opt.b;
  info->c = c->opt.c;
  info->epsilon = c->opt.epsilon;
  info->eta0 = c->opt.eta0;
  info->eta1 = c->opt.eta1;
  info->precision = c->prec;
  info->strategy = c->opt.strategy;
  info->sampler_max_source = c->opt.sampler_max_source;
  info->max_iteration = c->opt.max_iteration;
  return AMMSB_OK;
}

int ammsb_init_state(ammsb_ctx* c, uint64_t seed) {
  if (!c) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const Key key = make_key(seed);
  // theta = gamma(eta0, eta1, K, 2): shape eta0, scale eta1 for every entry
  // (mcmc_sampler_stochastic.cc:54; random.cc:472-511 parameter order)
  std::vector<double> theta(2 * c->K);
  for (int k = 0; k < c->K; ++k)
    for (int i = 0; i < 2; ++i)
      theta[2 * k + i] =
          std::exp(log_gamma_draw(key, c->opt.eta0, (uint32_t)(2 * k + i), 0, TAG_INIT_THETA)) *
          c->opt.eta1;
  int rc = set_beta_from_theta(c, theta.data());
  if (rc != AMMSB_OK) return rc;
  HIP_TRY(launch_init_phi(c->d_pi, c->ld, c->K, c->N, key, c->prec, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->state_ready = true;
  return AMMSB_OK;
}

int ammsb_set_state(ammsb_ctx* c, const void* pi, const double* theta) {
  if (!c || !pi || !theta) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  int rc = transfer_state(c, const_cast<void*>(pi), false);
  if (rc != AMMSB_OK) return rc;
  rc = set_beta_from_theta(c, theta);
  if (rc != AMMSB_OK) return rc;
  c->state_ready = true;
  return AMMSB_OK;
}

int ammsb_get_state(ammsb_ctx* c, void* pi, double* theta, double* beta) {
  if (!c) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (pi) {
    const int rc = transfer_state(c, pi, true);
    if (rc != AMMSB_OK) return rc;
  }
  if (theta) HIP_TRY(hipMemcpy(theta, c->d_theta, sizeof(double) * 2 * c->K, hipMemcpyDeviceToHost));
  if (beta) HIP_TRY(hipMemcpy(beta, c->d_beta, sizeof(double) * c->K, hipMemcpyDeviceToHost));
  return AMMSB_OK;
}

int ammsb_get_rows(ammsb_ctx* c, const int32_t* nodes, int64_t n, void* out) {
  if (!c || (n > 0 && (!nodes || !out))) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const size_t row = (c->K + 1) * c->fsize;
  for (int64_t i = 0; i < n; ++i) {
    if (nodes[i] < 0 || nodes[i] >= c->N) {
      set_error("ammsb_get_rows: node out of range");
      return AMMSB_ERR_ARG;
    }
    HIP_TRY(hipMemcpy((char*)out + i * row, (char*)c->d_pi + (size_t)nodes[i] * c->ld * c->fsize,
                      row, hipMemcpyDeviceToHost));
  }
  return AMMSB_OK;
}

int ammsb_set_counters(ammsb_ctx* c, int64_t step_count, int64_t average_count) {
  if (!c || step_count < 1 || average_count < 1) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  c->step_count = step_count;
  c->average_count = average_count;
  c->ring_valid = false;
  HIP_TRY(launch_set_state_step(c->d_st, step_count, c->hp, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMMSB_OK;
}

int ammsb_step(ammsb_ctx* c, int64_t iters, int32_t sync) {
  if (!c || iters < 0) return AMMSB_ERR_ARG;
  if (!c->state_ready) {
    set_error("ammsb_step: state not initialised (call ammsb_init_state or ammsb_set_state)");
    return AMMSB_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(c->opt.device));
  for (int64_t i = 0; i < iters; ++i) {
    PhaseScope ps(c, PH_OUTER);
    int rc = enqueue_iteration(c);
    if (rc != AMMSB_OK) return rc;
    c->step_count++;
  }
  if (sync) return check_devstate(c);
  return AMMSB_OK;
}

int ammsb_perplexity(ammsb_ctx* c, double* ppx) {
  if (!c || !ppx) return AMMSB_ERR_ARG;
  if (!c->state_ready) {
    set_error("ammsb_perplexity: state not initialised");
    return AMMSB_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(c->opt.device));
  return do_perplexity(c, ppx);
}

int ammsb_run(ammsb_ctx* c, int64_t max_steps, int32_t print, double* ppx_out,
              int64_t* ppx_steps, int64_t cap, int64_t* n_ppx) {
  if (!c) return AMMSB_ERR_ARG;
  if (!c->state_ready) {
    set_error("ammsb_run: state not initialised");
    return AMMSB_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(c->opt.device));
  const auto t_start = std::chrono::system_clock::now();
  int64_t done = 0, np = 0;
  int rc = AMMSB_OK;
  // mcmc_sampler_stochastic.cc:110-180
  while (c->step_count < c->opt.max_iteration && !is_converged(c)) {
    if (max_steps > 0 && done >= max_steps) break;
    if ((c->step_count - 1) % c->interval == 0) {
      double ppx = 0.0;
      rc = do_perplexity(c, &ppx);
      if (rc != AMMSB_OK) return rc;
      rc = check_devstate(c);
      if (rc != AMMSB_OK) return rc;
      if (print) {
        const auto t_ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                              std::chrono::system_clock::now() - t_start)
                              .count();
        printf("average_count is: %lld step count: %lld time: %.3f perplexity for hold out set: "
               "%.12f\n",
               (long long)c->average_count, (long long)c->step_count, t_ms / 1000.0, ppx);
        fflush(stdout);
      }
      c->ppx_cb.push_back(ppx);
      if (c->ppx_cb.size() > 10) c->ppx_cb.pop_front();
      if (ppx_out && ppx_steps && np < cap) {
        ppx_out[np] = ppx;
        ppx_steps[np] = c->step_count;
      }
      ++np;
    }
    {
      PhaseScope ps(c, PH_OUTER);
      rc = enqueue_iteration(c);
      if (rc != AMMSB_OK) return rc;
    }
    c->step_count++;
    ++done;
    if (print && c->step_count % c->stats_interval == 0) print_stats(c, stdout);
  }
  rc = check_devstate(c);
  if (rc != AMMSB_OK) return rc;
  if (print) print_stats(c, stdout);
  if (n_ppx) *n_ppx = np;
  return AMMSB_OK;
}

int ammsb_enable_timers(ammsb_ctx* c, int32_t on) {
  if (!c) return AMMSB_ERR_ARG;
  if (!on) flush_timers(c);
  c->timers_on = on != 0;
  // 1: every phase; 2: update_phi only (two events per iteration)
  c->timer_mask = on == 2 ? (1u << PH_UPDATE_PHI) : 0xFFFFFFFFu;
  return AMMSB_OK;
}

int ammsb_timers(ammsb_ctx* c, double* secs, int64_t* ticks) {
  if (!c) return AMMSB_ERR_ARG;
  int rc = flush_timers(c);
  if (rc != AMMSB_OK) return rc;
  for (int p = 0; p < PH_COUNT; ++p) {
    if (secs) secs[p] = c->t_secs[p];
    if (ticks) ticks[p] = c->t_ticks[p];
  }
  return AMMSB_OK;
}

int ammsb_update_phi(ammsb_ctx* c, const int32_t* nodes, int64_t n_nodes, const int32_t* nbrs,
                     int64_t n_nb, const void* noise, double eps_t) {
  if (!c || n_nodes < 0 || (n_nodes > 0 && (!nodes || !nbrs)) || n_nb < 1 || eps_t < 0.0)
    return AMMSB_ERR_ARG;
  if (n_nodes == 0) return AMMSB_OK;
  if (n_nodes > INT32_MAX / 2 || n_nb > INT32_MAX / 2) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  {
    std::vector<int32_t> srt(nodes, nodes + n_nodes);
    std::sort(srt.begin(), srt.end());
    for (int64_t i = 0; i < n_nodes; ++i)
      if (srt[i] < 0 || srt[i] >= c->N || (i && srt[i] == srt[i - 1])) {
        set_error("ammsb_update_phi: nodes must be distinct and in [0, N)");
        return AMMSB_ERR_ARG;
      }
    for (int64_t i = 0; i < n_nodes * n_nb; ++i)
      if (nbrs[i] < 0 || nbrs[i] >= c->N) {
        set_error("ammsb_update_phi: neighbour out of range");
        return AMMSB_ERR_ARG;
      }
  }
  int32_t *d_nodes = nullptr, *d_nbrs = nullptr;
  void* d_noise = nullptr;
  void* d_stage = nullptr;
  const int rc = op_workspace(
      c, {{(void**)&d_nodes, (size_t)(4 * n_nodes)},
          {(void**)&d_nbrs, (size_t)(4 * n_nodes * n_nb)},
          {&d_stage, (size_t)(n_nodes * c->ld * c->fsize)},
          {&d_noise, noise ? (size_t)(n_nodes * c->K * c->fsize) : (size_t)0}});
  if (rc != AMMSB_OK) return rc;
  hipError_t e = hipMemcpy(d_nodes, nodes, 4 * n_nodes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_nbrs, nbrs, 4 * n_nodes * n_nb, hipMemcpyHostToDevice);
  if (e == hipSuccess && noise)
    e = hipMemcpy(d_noise, noise, n_nodes * c->K * c->fsize, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_label_nbrs(dev_graph(c), d_nodes, (int32_t)n_nodes, d_nbrs, (int32_t)n_nb, c->stream);
  if (e == hipSuccess) {
    PhiArgs pa{};
    pa.pi = c->d_pi;
    pa.ld = c->ld;
    pa.K = c->K;
    pa.N = c->N;
    pa.use_ring = 0;
    pa.ring = c->ring;
    pa.nodes = d_nodes;
    pa.count_host = (int32_t)n_nodes;
    pa.nbrs = d_nbrs;
    pa.nnb = (int32_t)n_nb;
    pa.beta = c->d_beta;
    pa.st = c->d_st;
    pa.hp = c->hp;
    pa.eps_override = eps_t;
    pa.noise = d_noise;
    pa.key = c->key;
    pa.staging = d_stage;
    e = launch_update_phi(pa, c->prec, (int32_t)n_nodes, c->stream, c->phi_waves, c->phi_unroll);
  }
  if (e == hipSuccess) {
    ScatterArgs sc{};
    sc.pi = c->d_pi;
    sc.ld = c->ld;
    sc.K = c->K;
    sc.use_ring = 0;
    sc.ring = c->ring;
    sc.st = c->d_st;
    sc.nodes = d_nodes;
    sc.count_host = (int32_t)n_nodes;
    sc.staging = d_stage;
    e = launch_scatter(sc, c->prec, (int32_t)n_nodes, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    set_error(std::string("ammsb_update_phi: ") + hipGetErrorString(e));
    return AMMSB_ERR_HIP;
  }
  return AMMSB_OK;
}

int ammsb_update_beta(ammsb_ctx* c, const int32_t* edges, const uint8_t* y, int64_t E,
                      double scale, const double* noise, double eps_t) {
  if (!c || E < 0 || (E > 0 && (!edges || !y)) || eps_t < 0.0) return AMMSB_ERR_ARG;
  if (E > INT32_MAX / 2) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < 2 * E; ++i)
    if (edges[i] < 0 || edges[i] >= c->N) {
      set_error("ammsb_update_beta: edge endpoint out of range");
      return AMMSB_ERR_ARG;
    }
  int32_t* d_edges = nullptr;
  uint8_t* d_y = nullptr;
  double* d_noise = nullptr;
  const int rc = op_workspace(c, {{(void**)&d_edges, (size_t)(8 * E)},
                                  {(void**)&d_y, (size_t)E},
                                  {(void**)&d_noise, noise ? (size_t)(16 * c->K) : (size_t)0}});
  hipError_t e = hipSuccess;
  if (rc == AMMSB_OK) {
    if (E) e = hipMemcpy(d_edges, edges, 8 * E, hipMemcpyHostToDevice);
    if (e == hipSuccess && E) e = hipMemcpy(d_y, y, E, hipMemcpyHostToDevice);
    if (e == hipSuccess && noise)
      e = hipMemcpy(d_noise, noise, 16 * c->K, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      BetaArgs ba{};
      ba.pi = c->d_pi;
      ba.ld = c->ld;
      ba.K = c->K;
      ba.use_ring = 0;
      ba.ring = c->ring;
      ba.st = c->d_st;
      ba.edges = d_edges;
      ba.y = d_y;
      ba.count_host = (int32_t)E;
      ba.beta = c->d_beta;
      ba.theta = c->d_theta;
      ba.epsilon = c->hp.epsilon;
      ba.partials = c->d_beta_partials;
      ba.blocks = c->beta_blocks;
      e = launch_beta(ba, c->prec, c->stream);
    }
    if (e == hipSuccess) {
      ThetaArgs ta{};
      ta.theta = c->d_theta;
      ta.beta = c->d_beta;
      ta.K = c->K;
      ta.partials = c->d_beta_partials;
      ta.blocks = c->beta_blocks;
      ta.st = c->d_st;
      ta.hp = c->hp;
      ta.eps_override = eps_t;
      ta.scale_override = scale;
      ta.meta = c->ring.meta;
      ta.noise = d_noise;
      ta.key = c->key;
      ta.bump_step = 0;
      e = launch_theta(ta, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  if (rc != AMMSB_OK) return rc;
  if (e != hipSuccess) {
    set_error(std::string("ammsb_update_beta: ") + hipGetErrorString(e));
    return AMMSB_ERR_HIP;
  }
  return AMMSB_OK;
}

int ammsb_debug_sample(ammsb_ctx* c, int64_t step, int32_t* flag, double* scale, int32_t* edges,
                       int64_t* n_edges, int32_t* nodes, int64_t* n_nodes, int32_t* nbrs) {
  if (!c || step < 1) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  // draw iteration `step` into ring slot 0; the ring is refilled before the
  // next iteration runs
  c->ring_valid = false;
  int rc = fill_ring(c, step, 1);
  if (rc != AMMSB_OK) return rc;
  DevState h;
  rc = read_devstate(c, &h);
  if (rc != AMMSB_OK) return rc;
  if (h.error) {
    set_error("ammsb_debug_sample: device sampler error");
    // clear the flag so later calls start clean
    (void)launch_set_state_step(c->d_st, c->step_count, c->hp, c->stream);
    (void)hipStreamSynchronize(c->stream);
    return AMMSB_ERR_INVARIANT;
  }
  IterMeta mt;
  HIP_TRY(hipMemcpy(&mt, c->ring.meta, sizeof(IterMeta), hipMemcpyDeviceToHost));
  if (flag) *flag = mt.flag;
  if (scale) *scale = mt.scale;
  if (n_edges) *n_edges = mt.mb_count;
  if (n_nodes) *n_nodes = mt.node_count;
  if (edges && mt.mb_count)
    HIP_TRY(hipMemcpy(edges, c->ring.edges, 8 * (size_t)mt.mb_count, hipMemcpyDeviceToHost));
  if (nodes && mt.node_count)
    HIP_TRY(hipMemcpy(nodes, c->ring.nodes, 4 * (size_t)mt.node_count, hipMemcpyDeviceToHost));
  if (nbrs && mt.node_count)
    HIP_TRY(hipMemcpy(nbrs, c->ring.nbrs, 4 * (size_t)mt.node_count * c->nnb,
                      hipMemcpyDeviceToHost));
  return AMMSB_OK;
}

int ammsb_get_totals(ammsb_ctx* c, int64_t* nodes_updated, int64_t* minibatch_edges) {
  if (!c) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  DevState h;
  int rc = read_devstate(c, &h);
  if (rc != AMMSB_OK) return rc;
  if (nodes_updated) *nodes_updated = h.nodes_total;
  if (minibatch_edges) *minibatch_edges = h.edges_total;
  return AMMSB_OK;
}

int ammsb_set_tuning(ammsb_ctx* c, const char* name, int64_t value) {
  if (!c || !name) return AMMSB_ERR_ARG;
  const std::string n(name);
  if (n == "ll_fwd_rows" || n == "ll_bwd_rows") {
    if (value != 0 && value != 1) return AMMSB_ERR_ARG;
    (n == "ll_fwd_rows" ? c->ll_fwd_rows : c->ll_bwd_rows) = (int32_t)value;
    return AMMSB_OK;
  }
  if (n == "ll_fwd_pairs" || n == "ll_fwd_prefetch" || n == "ll_bwd_pairs" ||
      n == "ll_bwd_prefetch") {
    const int d = n.compare(0, 6, "ll_bwd") == 0 ? 1 : 0;
    const bool is_pairs = n.size() > 6 && n.compare(n.size() - 5, 5, "pairs") == 0;
    const int pairs = is_pairs ? (int)value : c->ll_pairs[d];
    const int pf = is_pairs ? c->ll_prefetch[d] : (int)value;
    if (!loglik_kc_shape_supported(pairs, pf)) {
      set_error("ammsb_set_tuning: log-likelihood shape (pairs, prefetch) not available");
      return AMMSB_ERR_ARG;
    }
    c->ll_pairs[d] = pairs;
    c->ll_prefetch[d] = pf;
    return AMMSB_OK;
  }
  int waves = c->phi_waves, unroll = c->phi_unroll;
  if (n == "phi_waves")
    waves = (int)value;
  else if (n == "phi_unroll")
    unroll = (int)value;
  else {
    set_error("ammsb_set_tuning: unknown parameter " + n);
    return AMMSB_ERR_ARG;
  }
  // a shape that K's row configuration does not offer is rejected here, the
  // pair is validated when both are set (phi_waves first, then phi_unroll)
  if (!update_phi_shape_supported(c->K, c->prec, waves, unroll) && n == "phi_unroll") {
    set_error("ammsb_set_tuning: update_phi shape (waves, unroll) not available for this K");
    return AMMSB_ERR_ARG;
  }
  c->phi_waves = waves;
  c->phi_unroll = unroll;
  return AMMSB_OK;
}

int ammsb_get_tuning(ammsb_ctx* c, const char* name, int64_t* value) {
  if (!c || !name || !value) return AMMSB_ERR_ARG;
  const std::string n(name);
  if (n == "phi_waves")
    *value = c->phi_waves;
  else if (n == "phi_unroll")
    *value = c->phi_unroll;
  else if (n == "ll_fwd_pairs")
    *value = c->ll_pairs[0];
  else if (n == "ll_fwd_prefetch")
    *value = c->ll_prefetch[0];
  else if (n == "ll_bwd_pairs")
    *value = c->ll_pairs[1];
  else if (n == "ll_bwd_prefetch")
    *value = c->ll_prefetch[1];
  else if (n == "ll_fwd_rows")
    *value = c->ll_fwd_rows;
  else if (n == "ll_bwd_rows")
    *value = c->ll_bwd_rows;
  else if (n == "sampler_lds")  // read-only: bit 0 minibatch table, bit 1 neighbour tables in LDS
    *value = (c->mbtable_in_lds ? 1 : 0) | (c->nbr_use_lds ? 2 : 0);
  else if (n == "ring_slots")  // read-only: iterations the sample-ahead ring holds
    *value = c->ring.slots;
  else {
    set_error("ammsb_get_tuning: unknown parameter " + n);
    return AMMSB_ERR_ARG;
  }
  return AMMSB_OK;
}

int ammsb_probe_read(ammsb_ctx* c, int32_t reps, double* seconds, int64_t* bytes) {
  if (!c || reps < 1 || !seconds || !bytes) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  const int64_t n = ((int64_t)c->N * c->ld * (int64_t)c->fsize) / 16 * 16;
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) {
    set_error("ammsb_probe_read: event creation failed");
    return AMMSB_ERR_HIP;
  }
  HIP_TRY(launch_read_probe(c->d_pi, n, c->d_ppx_partials, c->stream));  // warm
  HIP_TRY(hipEventRecord(a, c->stream));
  for (int32_t r = 0; r < reps; ++r) HIP_TRY(launch_read_probe(c->d_pi, n, c->d_ppx_partials, c->stream));
  HIP_TRY(hipEventRecord(b, c->stream));
  HIP_TRY(hipEventSynchronize(b));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, a, b));
  c->event_pool.push_back(a);
  c->event_pool.push_back(b);
  *seconds = ms * 1e-3 / reps;
  *bytes = n;
  return AMMSB_OK;
}

int ammsb_probe_gather(ammsb_ctx* c, int32_t reps, double* seconds, int64_t* bytes) {
  if (!c || reps < 1 || !seconds || !bytes) return AMMSB_ERR_ARG;
  if (!c->state_ready) {
    set_error("ammsb_probe_gather: state not initialised");
    return AMMSB_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(c->opt.device));
  // the draws of iteration step_count into ring slot 0; the ring is refilled
  // before the next iteration runs
  c->ring_valid = false;
  int rc = fill_ring(c, c->step_count, 1);
  if (rc != AMMSB_OK) return rc;
  IterMeta mt;
  HIP_TRY(hipMemcpyAsync(&mt, c->ring.meta, sizeof(IterMeta), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) {
    set_error("ammsb_probe_gather: event creation failed");
    return AMMSB_ERR_HIP;
  }
  const PhiArgs pa = phi_args(c);
  // update_phi's own launch shape when the probe has it, else 4 x 1
  const bool s82 = c->phi_waves == 8 && c->phi_unroll == 2;
  const int pw = s82 ? 8 : 4, pu = s82 ? 2 : 1;
  HIP_TRY(launch_gather_probe(pa, c->prec, (int32_t)c->node_cap, c->d_ppx_partials, c->stream, pw,
                              pu));
  HIP_TRY(hipEventRecord(a, c->stream));
  for (int32_t r = 0; r < reps; ++r)
    HIP_TRY(launch_gather_probe(pa, c->prec, (int32_t)c->node_cap, c->d_ppx_partials, c->stream,
                                pw, pu));
  HIP_TRY(hipEventRecord(b, c->stream));
  HIP_TRY(hipEventSynchronize(b));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, a, b));
  c->event_pool.push_back(a);
  c->event_pool.push_back(b);
  *seconds = ms * 1e-3 / reps;
  // own row + n+1 neighbour rows of K elements, and the n+1 neighbour ids
  *bytes = (int64_t)mt.node_count * ((c->nnb + 1) * (int64_t)c->K * (int64_t)c->fsize + 4 * c->nnb);
  return AMMSB_OK;
}

namespace {
int loglik_common(ammsb_ctx* c, int64_t M, const int32_t* nodes, const void* phi, int64_t phi_ld,
                  const double* theta, const int64_t* pair_ptr, const int32_t* pair_nbr,
                  const uint8_t* pair_y, int64_t P, bool bwd, double* per_source, double* total,
                  const double* gout, void* grad_phi, double* grad_theta, void* stream,
                  bool noalloc = false) {
  if (!c || M < 0 || P < 0 || (M > 0 && (!nodes || !phi || !pair_ptr)) || (P > 0 && (!pair_nbr || !pair_y)) ||
      phi_ld < c->K) {
    set_error("ammsb_loglik: bad arguments");
    return AMMSB_ERR_ARG;
  }
  if (!bwd && (!total && !per_source)) return AMMSB_ERR_ARG;
  if (bwd && (!gout || (!grad_phi && !grad_theta))) return AMMSB_ERR_ARG;
  if (M > INT32_MAX) return AMMSB_ERR_ARG;
  HIP_TRY(hipSetDevice(c->opt.device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  if (stream) HIP_TRY(hipStreamSynchronize(c->stream));  // resident state settled
  // the whole-row adjoint writes one beta partial row per workgroup: as
  // many workgroups as are resident at once (FP64 K=2048: one per CU)
  const bool rows_bwd = bwd && c->ll_bwd_rows && loglik_rows_chunks(c->K, c->prec) > 0;
  int64_t max_blocks = 1024;
  if (rows_bwd) max_blocks = (int64_t)loglik_rows_blocks_per_cu(c->K, c->prec) * c->cus;
  const int32_t blocks = (int32_t)std::max<int64_t>(1, std::min<int64_t>(M, max_blocks));
  const int64_t prows = blocks;
  // scratch: per-source values (fwd) or [prows][K] dbeta partials (bwd), plus
  // one double for an absent total and the error word
  const int64_t body = bwd ? prows * c->K : M;
  double* scr = nullptr;
  int32_t* d_err = nullptr;
  if (noalloc) {
    // the file-scope array: error word in the header, then the body
    const size_t need = kNoAllocHeader + sizeof(double) * (size_t)body;
    if (need > kNoAllocScratchBytes) {
      set_error("ammsb_loglik_bwd_noalloc: the adjoint needs " + std::to_string(need) +
                " bytes of scratch (" + std::to_string(prows) + " partial rows of K=" +
                std::to_string(c->K) + "), more than the " +
                std::to_string(kNoAllocScratchBytes) + " the no-allocation array holds");
      return AMMSB_ERR_ARG;
    }
    void* base = nullptr;
    const int rc = noalloc_scratch(c->opt.device, &base);
    if (rc != AMMSB_OK) return rc;
    d_err = (int32_t*)base;
    scr = (double*)((char*)base + kNoAllocHeader);
  } else {
    const size_t need = sizeof(double) * (size_t)(body + 2);
    if (need > c->ll_scratch_bytes) {
      if (c->d_ll_scratch) HIP_TRY(hipFree(c->d_ll_scratch));
      c->d_ll_scratch = nullptr;
      c->ll_scratch_bytes = 0;
      HIP_TRY(hipMalloc((void**)&c->d_ll_scratch, need));
      c->ll_scratch_bytes = need;
    }
    scr = c->d_ll_scratch;
    d_err = (int32_t*)(c->d_ll_scratch + body + 1);
  }
  HIP_TRY(launch_fill_u32((uint32_t*)d_err, 0u, 1, s));  // a kernel, not a memset node
  LogLikArgs a{};
  a.pi = c->d_pi;
  a.ld = c->ld;
  a.K = c->K;
  a.M = M;
  a.nodes = nodes;
  a.phi = phi;
  a.phi_ld = phi_ld;
  a.theta = theta ? theta : c->d_theta;
  a.pair_ptr = pair_ptr;
  a.pair_nbr = pair_nbr;
  a.pair_y = pair_y;
  a.epsilon = c->hp.epsilon;
  a.per_source = per_source ? per_source : scr;
  a.gout = gout;
  a.grad_phi = grad_phi;
  a.dbeta_partials = scr;
  a.blocks = blocks;
  a.N = c->N;
  a.P = P;
  a.error = d_err;
  if (!bwd) {
    HIP_TRY(launch_loglik(a, c->prec, false, s, c->ll_pairs[0], c->ll_prefetch[0],
                          c->ll_fwd_rows != 0));
    double* out = total ? total : scr + body;
    HIP_TRY(launch_sum(a.per_source, M, out, s));
  } else {
    if (!grad_phi) {
      set_error("ammsb_loglik_bwd: grad_phi is required");
      return AMMSB_ERR_ARG;
    }
    HIP_TRY(launch_loglik(a, c->prec, true, s, c->ll_pairs[1], c->ll_prefetch[1], rows_bwd));
    if (grad_theta)
      HIP_TRY(launch_theta_grad(scr, (int32_t)prows, c->K, a.theta, gout, grad_theta, s));
  }
  int32_t herr = 0;
  HIP_TRY(hipMemcpyAsync(&herr, d_err, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (herr) {
    set_error("ammsb_loglik: neighbour id outside [0, N)");
    return AMMSB_ERR_ARG;
  }
  return AMMSB_OK;
}
}  // namespace

int ammsb_loglik_fwd(ammsb_ctx* c, int64_t M, const int32_t* nodes, const void* phi,
                     int64_t phi_ld, const double* theta, const int64_t* pair_ptr,
                     const int32_t* pair_nbr, const uint8_t* pair_y, int64_t P, double* per_source,
                     double* total, void* stream) {
  return loglik_common(c, M, nodes, phi, phi_ld, theta, pair_ptr, pair_nbr, pair_y, P, false,
                       per_source, total, nullptr, nullptr, nullptr, stream);
}

int ammsb_loglik_bwd(ammsb_ctx* c, int64_t M, const int32_t* nodes, const void* phi,
                     int64_t phi_ld, const double* theta, const int64_t* pair_ptr,
                     const int32_t* pair_nbr, const uint8_t* pair_y, int64_t P, const double* gout,
                     void* grad_phi, double* grad_theta, void* stream) {
  return loglik_common(c, M, nodes, phi, phi_ld, theta, pair_ptr, pair_nbr, pair_y, P, true,
                       nullptr, nullptr, gout, grad_phi, grad_theta, stream);
}

int ammsb_loglik_bwd_noalloc(ammsb_ctx* c, int64_t M, const int32_t* nodes, const void* phi,
                             int64_t phi_ld, const double* theta, const int64_t* pair_ptr,
                             const int32_t* pair_nbr, const uint8_t* pair_y, int64_t P,
                             const double* gout, void* grad_phi, double* grad_theta,
                             void* stream) {
  return loglik_common(c, M, nodes, phi, phi_ld, theta, pair_ptr, pair_nbr, pair_y, P, true,
                       nullptr, nullptr, gout, grad_phi, grad_theta, stream, true);
}

int ammsb_device_ptrs(ammsb_ctx* c, void** pi, void** theta, void** beta) {
  if (!c) return AMMSB_ERR_ARG;
  if (pi) *pi = c->d_pi;
  if (theta) *theta = c->d_theta;
  if (beta) *beta = c->d_beta;
  return AMMSB_OK;
}

}  // extern "C"

// ===========================================================================
// Batches of independent chains.  The chains are full contexts (graph,
// split, state, ring of one slot) on the batch's stream; the batch keeps
// device arrays of each phase's per-chain kernel arguments, so one launch
// per phase serves every chain and the iteration's launch sequence is fixed:
// it is captured once as a hipGraph and replayed.
struct ammsb_batch {
  int32_t n = 0;
  int prec = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<ammsb_ctx*> chains;
  SamplerArgs* d_sargs = nullptr;
  NeighborArgs* d_nargs = nullptr;
  PhiArgs* d_pargs = nullptr;
  BetaArgs* d_bargs = nullptr;
  ThetaArgs* d_targs = nullptr;
  BatchPhiShape* d_shapes = nullptr;
  int32_t* d_beta_cfg = nullptr;
  int32_t* d_index = nullptr;       // work layout (batch_index_ints(n) ints)
  int32_t* d_tune_index = nullptr;  // the same for one chain (autotuner)
  // Sample-ahead ring of every chain: 2R slots in two halves.  Block k =
  // iterations [ring_B + kR, ring_B + (k+1)R) lives in half k mod 2; the
  // samplers of block k+1 run on fill_stream while block k computes.
  int32_t R = 1;
  hipStream_t fill_stream = nullptr;
  hipEvent_t ev_fill[2] = {nullptr, nullptr};  // samplers of the half done
  hipEvent_t ev_done[2] = {nullptr, nullptr};  // iterations of the half done
  bool ring_ok = false;
  int64_t ring_B = 0, ring_hi = 0;  // ring holds iterations [ring_hi - 2R, ring_hi)
  std::vector<int64_t> ring_gen;    // chains' ring_gen when the ring was set up
  std::vector<BatchPhiShape> shapes;
  std::vector<double> tuned_us;
  size_t mb_lds = 0, nb_lds = 0;
  int32_t max_node_cap = 0, max_beta_blocks = 0, max_K = 0;
  int32_t phi_grid = 1, scatter_grid = 1;
  bool use_graph = true;
  bool refill_in_order = false;  // sampler refills on the main stream (else a second stream)
  hipGraph_t graph = nullptr;  // one iteration (samplers excluded)
  hipGraphExec_t exec = nullptr;
};

namespace {

// Work layout in one device buffer of int32: order [n], items [n+1],
// pub [n+1], counter, then info [n] (int4, 16-byte aligned).
size_t batch_index_info_offset(int32_t n) { return ((size_t)3 * n + 3 + 3) / 4 * 4; }
size_t batch_index_ints(int32_t n) { return batch_index_info_offset(n) + (size_t)4 * n; }

BatchIndex batch_index(int32_t* base, int32_t n) {
  BatchIndex ix;
  ix.order = base;
  ix.items = base + n;
  ix.pub = base + 2 * n + 1;
  ix.counter = base + 3 * n + 2;
  ix.info = reinterpret_cast<int4*>(base + batch_index_info_offset(n));
  return ix;
}

void batch_drop_graph(ammsb_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

void free_batch(ammsb_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->fill_stream) (void)hipStreamSynchronize(b->fill_stream);
  batch_drop_graph(b);
  for (auto* c : b->chains) free_ctx(c);
  for (int h = 0; h < 2; ++h) {
    if (b->ev_fill[h]) (void)hipEventDestroy(b->ev_fill[h]);
    if (b->ev_done[h]) (void)hipEventDestroy(b->ev_done[h]);
  }
  if (b->fill_stream) (void)hipStreamDestroy(b->fill_stream);
  void* ptrs[] = {b->d_sargs,  b->d_nargs,    b->d_pargs, b->d_bargs,      b->d_targs,
                  b->d_shapes, b->d_beta_cfg, b->d_index, b->d_tune_index};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// Work order of the batched update_phi: chains whose work items move the
// most row bytes first (npb nodes x nnb rows of K x sizeof(Float), in
// 64-byte lines), so that the last items handed out are the short ones and
// the persistent workgroups finish together.  Ties keep the chain order.
std::vector<int32_t> batch_work_order(const ammsb_batch* b) {
  std::vector<int64_t> cost(b->n);
  for (int32_t i = 0; i < b->n; ++i) {
    const ammsb_ctx* c = b->chains[i];
    const int64_t row = ((int64_t)c->K * (c->prec == 0 ? 8 : 4) + 63) / 64;
    cost[i] = (int64_t)b->shapes[i].npb * c->nnb * row;
  }
  std::vector<int32_t> order(b->n);
  for (int32_t i = 0; i < b->n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
  return order;
}

int batch_upload_shapes(ammsb_batch* b) {
  const std::vector<int32_t> order = batch_work_order(b);
  HIP_TRY(hipMemcpyAsync(b->d_shapes, b->shapes.data(), sizeof(BatchPhiShape) * b->n,
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d_index, order.data(), sizeof(int32_t) * b->n,
                         hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return AMMSB_OK;
}

// The samplers of block k (iterations ring_B + kR ..) of every chain into
// half k mod 2 of its ring, on stream s; no state changes (draws depend on
// the step only).
int batch_enqueue_fill(ammsb_batch* b, int64_t k, hipStream_t s) {
  const int64_t t0 = b->ring_B + k * b->R;
  const int slot0 = (int)(k % 2) * b->R;
  HIP_TRY(launch_minibatch_batch(b->d_sargs, b->n, t0, slot0, b->R, 2 * b->R, b->mb_lds, s));
  HIP_TRY(launch_neighbors_batch(b->d_nargs, b->n, slot0, b->R, 2 * b->R, b->max_node_cap,
                                 b->nb_lds, s));
  return AMMSB_OK;
}

// Ring set-up at step S (all chains at S): ring_base = S on every chain's
// device state, block 0 drawn on the main stream.
int batch_ring_reset(ammsb_batch* b, int64_t S) {
  HIP_TRY(hipStreamSynchronize(b->fill_stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  const int32_t slots = 2 * b->R;
  for (auto* c : b->chains) {
    HIP_TRY(hipMemcpy(&c->d_st->ring_base, &S, sizeof(int64_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(&c->d_st->ring_slots, &slots, sizeof(int32_t), hipMemcpyHostToDevice));
  }
  b->ring_B = S;
  const int rc = batch_enqueue_fill(b, 0, b->stream);
  if (rc != AMMSB_OK) return rc;
  HIP_TRY(hipEventRecord(b->ev_fill[0], b->stream));
  b->ring_hi = S + b->R;
  b->ring_ok = true;
  for (int32_t i = 0; i < b->n; ++i) b->ring_gen[i] = b->chains[i]->ring_gen;
  return AMMSB_OK;
}

// Every chain at the same step: the batch's common step, else -1.
int64_t batch_common_step(const ammsb_batch* b) {
  const int64_t S = b->chains[0]->step_count;
  for (const auto* c : b->chains)
    if (c->step_count != S) return -1;
  return S;
}

bool batch_ring_valid(const ammsb_batch* b, int64_t S) {
  if (!b->ring_ok || S < b->ring_hi - 2 * b->R || S >= b->ring_hi) return false;
  for (int32_t i = 0; i < b->n; ++i)
    if (b->chains[i]->ring_gen != b->ring_gen[i]) return false;
  return true;
}

int batch_enqueue_iteration(ammsb_batch* b) {
  const BatchIndex ix = batch_index(b->d_index, b->n);
  HIP_TRY(launch_batch_index(b->d_pargs, b->d_shapes, b->d_beta_cfg, b->n, ix, b->stream));
  HIP_TRY(launch_update_phi_batch(b->d_pargs, b->d_shapes, b->n, ix, b->prec, b->phi_grid,
                                  b->stream));
  HIP_TRY(launch_scatter_batch(b->d_pargs, b->d_beta_cfg, b->n, ix, b->prec, b->scatter_grid,
                               b->stream));
  HIP_TRY(launch_beta_batch(b->d_bargs, b->d_beta_cfg, b->n, b->max_beta_blocks, b->prec,
                            b->stream));
  HIP_TRY(launch_theta_batch(b->d_targs, b->n, b->max_K, b->stream));
  return AMMSB_OK;
}

int batch_capture(ammsb_batch* b) {
  batch_drop_graph(b);
  HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = batch_enqueue_iteration(b);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != AMMSB_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(e);
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return AMMSB_OK;
}

int batch_check_states(ammsb_batch* b) {
  for (auto* c : b->chains) {
    const int rc = check_devstate(c);
    if (rc != AMMSB_OK) return rc;
  }
  return AMMSB_OK;
}

int shape_cfg(const ammsb_batch_shape& s) {
  for (int i = 0; i < kBatchCfgs; ++i)
    if (kBatchCfgLpr[i] == s.lanes_per_row && kBatchCfgCh[i] == s.vectors_per_lane) return i;
  return -1;
}

bool shape_valid(const ammsb_ctx* c, int cfg, int npb) {
  if (!batch_cfg_valid(cfg, c->K, c->prec)) return false;
  const int groups = kBatchWaves * (64 / kBatchCfgLpr[cfg]);
  return npb >= 1 && npb <= groups && (npb & (npb - 1)) == 0;
}

// Default shape of the batched update_phi (what the autotuner picks most
// often on chains with K in {16, 64, 256, 512}): two vectors per lane where
// that leaves no lane idle, and 64 lanes per node for rows of up to 32
// vectors, else 128 (two wavefronts).
int default_cfg(const ammsb_ctx* c) {
  for (int cfg = 0; cfg < kBatchCfgs; ++cfg)
    if (kBatchCfgCh[cfg] == 2 && batch_cfg_valid(cfg, c->K, c->prec)) return cfg;
  return batch_cfg_default(c->K, c->prec);
}

int default_npb(const ammsb_ctx* c, int cfg) {
  const int lpr = kBatchCfgLpr[cfg];
  const int groups = kBatchWaves * (64 / lpr);
  const int V = c->prec == AMMSB_FP64 ? 2 : 4;
  const int vecs = (c->K + V - 1) / V;
  int gpn = (vecs <= 32 ? 64 : 128) / lpr;
  gpn = std::max(1, std::min(groups, gpn));
  return groups / gpn;
}

}  // namespace

extern "C" {

int ammsb_batch_create(int32_t n, const ammsb_options* opts, const ammsb_graph* graphs,
                       const ammsb_split* splits, ammsb_batch** out) {
  if (!out || !opts || !graphs || n < 1 || n > kMaxBatchChains) {
    set_error("ammsb_batch_create: need 1..1024 chains with options and graphs");
    return AMMSB_ERR_ARG;
  }
  *out = nullptr;
  for (int32_t i = 0; i < n; ++i) {
    if (opts[i].precision != opts[0].precision || opts[i].device != opts[0].device) {
      set_error("ammsb_batch_create: all chains must share precision and device");
      return AMMSB_ERR_ARG;
    }
    if (batch_cfg_default((int32_t)opts[i].K, opts[i].precision) < 0) {
      set_error("ammsb_batch_create: K of chain " + std::to_string(i) +
                " exceeds the batched kernels (256 FP64 / 512 FP32)");
      return AMMSB_ERR_ARG;
    }
  }
  ammsb_batch* b = nullptr;
  try {
    b = new ammsb_batch();
  } catch (...) {
    set_error("ammsb_batch_create: out of host memory");
    return AMMSB_ERR_OOM;
  }
  b->n = n;
  b->prec = opts[0].precision;
  b->device = opts[0].device;
  auto fail = [&](int rc) {
    free_batch(b);
    return rc;
  };
  if (hipSetDevice(b->device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("ammsb_batch_create: cannot create a stream on the device");
    return fail(AMMSB_ERR_HIP);
  }
  std::vector<SamplerArgs> sargs(n);
  std::vector<NeighborArgs> nargs(n);
  std::vector<PhiArgs> pargs(n);
  std::vector<BetaArgs> bargs(n);
  std::vector<ThetaArgs> targs(n);
  std::vector<int32_t> beta_cfg(n);
  b->shapes.resize(n);
  b->tuned_us.assign(n, 0.0);
  for (int32_t i = 0; i < n; ++i) {
    ammsb_ctx* c = nullptr;
    try {
      c = new ammsb_ctx();
    } catch (...) {
      set_error("ammsb_batch_create: out of host memory");
      return fail(AMMSB_ERR_OOM);
    }
    c->stream = b->stream;
    c->own_stream = false;
    c->owner = b;
    c->max_ring_slots = 32;  // two halves of up to 16 iterations
    b->chains.push_back(c);
    const int rc = create_impl(&opts[i], &graphs[i], splits ? &splits[i] : nullptr, c);
    if (rc != AMMSB_OK) {
      set_error("ammsb_batch_create: chain " + std::to_string(i) + ": " + get_error());
      return fail(rc);
    }
  }
  // one refill cadence for the batch: every chain's ring is used as two
  // halves of R slots, R from the smallest chain's ring
  int32_t slots = b->chains[0]->ring.slots;
  for (auto* c : b->chains) slots = std::min(slots, c->ring.slots);
  b->R = std::max(1, slots / 2);
  if (2 * b->R > slots) {
    set_error("ammsb_batch_create: a chain's sample-ahead ring is too large for HBM budget");
    return fail(AMMSB_ERR_OOM);
  }
  b->ring_gen.assign(n, -1);
  if (hipStreamCreateWithFlags(&b->fill_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fill[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fill[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_done[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_done[1], hipEventDisableTiming) != hipSuccess) {
    set_error("ammsb_batch_create: cannot create the sampler stream and its events");
    return fail(AMMSB_ERR_HIP);
  }
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess)
    cus = 256;
  // update_beta workgroups per chain: the chains share one launch, so about
  // as many workgroups in all as stay resident (each wavefront then takes
  // several groups of edges, and fewer partial rows reach the theta step);
  // a chain's own bound (its edge capacity) still applies.  64 chains: 41 us
  // with each chain's own bound, 28 us at the resident count, 38 us at twice
  // and 41 us at half of it.
  const int32_t beta_per_chain =
      std::max(1, (cus * beta_batch_blocks_per_cu(b->prec) + n - 1) / n);
  for (int32_t i = 0; i < n; ++i) {
    ammsb_ctx* c = b->chains[i];
    c->ring.slots = 2 * b->R;
    if (hipMemcpy(&c->d_st->ring_slots, &c->ring.slots, sizeof(int32_t),
                  hipMemcpyHostToDevice) != hipSuccess) {
      set_error("ammsb_batch_create: device state update failed");
      return fail(AMMSB_ERR_HIP);
    }
    if (!c->nbr_use_lds) {
      set_error("ammsb_batch_create: chain " + std::to_string(i) +
                ": num_node_sample too large for the batched neighbour sampler");
      return fail(AMMSB_ERR_ARG);
    }
    sargs[i] = sampler_args(c, -1);
    nargs[i] = neighbor_args(c);
    pargs[i] = phi_args(c);
    bargs[i] = beta_args(c);
    targs[i] = theta_args(c);
    bargs[i].blocks = targs[i].blocks = std::min(c->beta_blocks, beta_per_chain);
    beta_cfg[i] = batch_cfg_default(c->K, c->prec);
    b->shapes[i].cfg = default_cfg(c);
    b->shapes[i].npb = default_npb(c, b->shapes[i].cfg);
    if (c->mbtable_in_lds) b->mb_lds = std::max(b->mb_lds, (size_t)c->mbtable_cap * sizeof(int2));
    b->nb_lds = std::max(b->nb_lds, neighbor_lds_bytes(c->nbr_tcap));
    b->max_node_cap = std::max<int32_t>(b->max_node_cap, (int32_t)c->node_cap);
    b->max_beta_blocks = std::max(b->max_beta_blocks, bargs[i].blocks);
    b->max_K = std::max(b->max_K, c->K);
  }
  int rc;
  if ((rc = dalloc(&b->d_sargs, n)) || (rc = dalloc(&b->d_nargs, n)) ||
      (rc = dalloc(&b->d_pargs, n)) || (rc = dalloc(&b->d_bargs, n)) ||
      (rc = dalloc(&b->d_targs, n)) || (rc = dalloc(&b->d_shapes, n)) ||
      (rc = dalloc(&b->d_beta_cfg, n)) || (rc = dalloc(&b->d_index, batch_index_ints(n))) ||
      (rc = dalloc(&b->d_tune_index, batch_index_ints(1))))
    return fail(rc);
  static const int32_t kZero = 0;
  const struct {
    void* d;
    const void* h;
    size_t bytes;
  } up[] = {{b->d_sargs, sargs.data(), sizeof(SamplerArgs) * n},
            {b->d_nargs, nargs.data(), sizeof(NeighborArgs) * n},
            {b->d_pargs, pargs.data(), sizeof(PhiArgs) * n},
            {b->d_bargs, bargs.data(), sizeof(BetaArgs) * n},
            {b->d_targs, targs.data(), sizeof(ThetaArgs) * n},
            {b->d_beta_cfg, beta_cfg.data(), sizeof(int32_t) * n},
            {b->d_tune_index, &kZero, sizeof(int32_t)}};  // the tuner's order: chain 0
  for (const auto& u : up)
    if (hipMemcpy(u.d, u.h, u.bytes, hipMemcpyHostToDevice) != hipSuccess) {
      set_error("ammsb_batch_create: upload of kernel arguments failed");
      return fail(AMMSB_ERR_HIP);
    }
  if ((rc = batch_upload_shapes(b)) != AMMSB_OK) return fail(rc);
  // persistent: as many workgroups as stay resident
  b->phi_grid = std::max(1, cus * update_phi_batch_blocks_per_cu(n, b->prec));
  // eight resident wavefronts per SIMD for the copy
  b->scatter_grid = std::max(1, cus * 8);
  *out = b;
  return AMMSB_OK;
}

int ammsb_batch_destroy(ammsb_batch* b) {
  free_batch(b);
  return AMMSB_OK;
}

int ammsb_batch_chain(ammsb_batch* b, int32_t i, ammsb_ctx** chain) {
  if (!b || !chain || i < 0 || i >= b->n) return AMMSB_ERR_ARG;
  *chain = b->chains[i];
  return AMMSB_OK;
}

int ammsb_batch_init_state(ammsb_batch* b, const uint64_t* seeds) {
  if (!b) return AMMSB_ERR_ARG;
  for (int32_t i = 0; i < b->n; ++i) {
    ammsb_ctx* c = b->chains[i];
    const uint64_t seed = seeds ? seeds[i] : (uint64_t)(uint32_t)c->opt.random_seed;
    const int rc = ammsb_init_state(c, seed);
    if (rc != AMMSB_OK) return rc;
  }
  return AMMSB_OK;
}

int ammsb_batch_use_graph(ammsb_batch* b, int32_t enable) {
  if (!b) return AMMSB_ERR_ARG;
  b->use_graph = enable != 0;
  return AMMSB_OK;
}

int ammsb_batch_refill_in_order(ammsb_batch* b, int32_t enable) {
  if (!b) return AMMSB_ERR_ARG;
  b->refill_in_order = enable != 0;
  return AMMSB_OK;
}

int ammsb_batch_step(ammsb_batch* b, int64_t iters) {
  if (!b || iters < 0) return AMMSB_ERR_ARG;
  for (auto* c : b->chains)
    if (!c->state_ready) {
      set_error("ammsb_batch_step: state not initialised (ammsb_batch_init_state)");
      return AMMSB_ERR_STATE;
    }
  const int64_t S = batch_common_step(b);
  if (S < 0) {
    set_error("ammsb_batch_step: the chains are at different steps (ammsb_set_counters)");
    return AMMSB_ERR_STATE;
  }
  HIP_TRY(hipSetDevice(b->device));
  // the chains' own (single-context) work must be done before the batch's
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (!batch_ring_valid(b, S)) {
    const int rc = batch_ring_reset(b, S);
    if (rc != AMMSB_OK) return rc;
  }
  if (b->use_graph && !b->exec) {
    const int rc = batch_capture(b);
    if (rc != AMMSB_OK) return rc;
  }
  const int64_t R = b->R;
  for (int64_t t = S; t < S + iters; ++t) {
    const int64_t k = (t - b->ring_B) / R, pos = (t - b->ring_B) % R;
    if (pos == 0 && b->refill_in_order) {
      // block k+1 drawn in stream order, into the half block k-1 used
      // (block k was drawn before it, in stream order or by the event wait
      // of an overlapped refill)
      HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_fill[k % 2], 0));
      if (b->ring_hi == b->ring_B + (k + 1) * R) {
        const int rc = batch_enqueue_fill(b, k + 1, b->stream);
        if (rc != AMMSB_OK) return rc;
        HIP_TRY(hipEventRecord(b->ev_fill[(k + 1) % 2], b->stream));
        b->ring_hi += R;
      }
    } else if (pos == 0) {
      // block k's draws are complete; block k+1 is drawn on the sampler
      // stream while block k computes, into the half block k-1 used
      HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_fill[k % 2], 0));
      if (b->ring_hi == b->ring_B + (k + 1) * R) {
        HIP_TRY(hipStreamWaitEvent(b->fill_stream, b->ev_done[(k + 1) % 2], 0));
        const int rc = batch_enqueue_fill(b, k + 1, b->fill_stream);
        if (rc != AMMSB_OK) return rc;
        HIP_TRY(hipEventRecord(b->ev_fill[(k + 1) % 2], b->fill_stream));
        b->ring_hi += R;
      }
    }
    if (b->use_graph) {
      HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {
      const int rc = batch_enqueue_iteration(b);
      if (rc != AMMSB_OK) return rc;
    }
    if (pos == R - 1) HIP_TRY(hipEventRecord(b->ev_done[k % 2], b->stream));
    for (auto* c : b->chains) c->step_count++;
  }
  // nothing of the batch is in flight when it returns; a chain's own
  // (single-context) ring use refills it
  HIP_TRY(hipStreamSynchronize(b->fill_stream));
  for (auto* c : b->chains) c->ring_valid = false;
  return batch_check_states(b);
}

int ammsb_batch_perplexity(ammsb_batch* b, double* out) {
  if (!b || !out) return AMMSB_ERR_ARG;
  for (int32_t i = 0; i < b->n; ++i) {
    const int rc = ammsb_perplexity(b->chains[i], &out[i]);
    if (rc != AMMSB_OK) return rc;
  }
  return AMMSB_OK;
}

int ammsb_batch_set_shape(ammsb_batch* b, int32_t i, const ammsb_batch_shape* shape) {
  if (!b || !shape || i < 0 || i >= b->n) return AMMSB_ERR_ARG;
  const int cfg = shape_cfg(*shape);
  if (cfg < 0 || !shape_valid(b->chains[i], cfg, shape->nodes_per_block)) {
    set_error("ammsb_batch_set_shape: shape not available for this chain's K");
    return AMMSB_ERR_ARG;
  }
  b->shapes[i].cfg = cfg;
  b->shapes[i].npb = shape->nodes_per_block;
  b->tuned_us[i] = 0.0;
  return batch_upload_shapes(b);  // the graph reads shapes from device memory
}

int ammsb_batch_get_shape(ammsb_batch* b, int32_t i, ammsb_batch_shape* shape) {
  if (!b || !shape || i < 0 || i >= b->n) return AMMSB_ERR_ARG;
  shape->lanes_per_row = kBatchCfgLpr[b->shapes[i].cfg];
  shape->vectors_per_lane = kBatchCfgCh[b->shapes[i].cfg];
  shape->nodes_per_block = b->shapes[i].npb;
  shape->us = b->tuned_us[i];
  return AMMSB_OK;
}

int ammsb_batch_autotune(ammsb_batch* b, int32_t reps, ammsb_batch_shape* out) {
  if (!b || reps < 1) return AMMSB_ERR_ARG;
  for (auto* c : b->chains)
    if (!c->state_ready) {
      set_error("ammsb_batch_autotune: state not initialised");
      return AMMSB_ERR_STATE;
    }
  HIP_TRY(hipSetDevice(b->device));
  // this iteration's minibatches and neighbour sets in every chain's ring
  const int64_t S = batch_common_step(b);
  if (S < 0) {
    set_error("ammsb_batch_autotune: the chains are at different steps");
    return AMMSB_ERR_STATE;
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  int rc = batch_ring_reset(b, S);
  if (rc != AMMSB_OK) return rc;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  const BatchIndex tix = batch_index(b->d_tune_index, 1);
  for (int32_t i = 0; i < b->n && rc == AMMSB_OK; ++i) {
    ammsb_ctx* c = b->chains[i];
    double best = 1e300;
    BatchPhiShape pick = b->shapes[i];
    for (int cfg = 0; cfg < kBatchCfgs && rc == AMMSB_OK; ++cfg) {
      if (!batch_cfg_valid(cfg, c->K, c->prec)) continue;
      const int groups = kBatchWaves * (64 / kBatchCfgLpr[cfg]);
      for (int npb = 1; npb <= groups && rc == AMMSB_OK; npb *= 2) {
        b->shapes[i] = BatchPhiShape{cfg, npb};
        if ((rc = batch_upload_shapes(b)) != AMMSB_OK) break;
        // update_phi of chain i alone; it writes only the staging rows
        auto run = [&]() -> hipError_t {
          hipError_t e = launch_batch_index(b->d_pargs + i, b->d_shapes + i, b->d_beta_cfg + i, 1,
                                            tix, b->stream);
          if (e == hipSuccess)
            e = launch_update_phi_batch(b->d_pargs + i, b->d_shapes + i, 1, tix, b->prec,
                                        b->phi_grid, b->stream);
          return e;
        };
        hipError_t e = run();  // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
        for (int r = 0; r < reps && e == hipSuccess; ++r) e = run();
        if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          set_error(std::string("ammsb_batch_autotune: ") + hipGetErrorString(e));
          rc = AMMSB_ERR_HIP;
          break;
        }
        const double us = 1e3 * ms / reps;
        if (us < best) {
          best = us;
          pick = b->shapes[i];
        }
      }
    }
    b->shapes[i] = pick;
    b->tuned_us[i] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != AMMSB_OK) return rc;
  rc = batch_upload_shapes(b);
  if (rc != AMMSB_OK) return rc;
  if (out)
    for (int32_t i = 0; i < b->n; ++i) ammsb_batch_get_shape(b, i, &out[i]);
  return AMMSB_OK;
}

}  // extern "C"
