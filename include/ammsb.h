/*
 * ammsb.h -- C-ABI of the MI355X-native SGRLD a-MMSB sampler (libammsb.so).
 *
 * Drop-in boundary for the reference's sequential stochastic sampler
 *   mcmc::learning::MCMCSamplerStochastic
 *     (c++/src/mcmc/learning/mcmc_sampler_stochastic.h:72-115,
 *      c++/src/mcmc/learning/mcmc_sampler_stochastic.cc:9-327)
 * and its Learner base (c++/src/mcmc/learning/learner.h:24-172).
 *
 * State layout is the reference's distributed record layout
 *   pi[0..K) ++ phi_sum   per node  (mcmc_sampler_stochastic_distr.cc:743-751)
 * i.e. host arrays exchanged through this ABI are row-major pi[N][K+1] with
 * slot K holding phi_sum = sum_k phi_ik; theta is [K][2] double; beta is [K]
 * double.  pi is double for AMMSB_FP64 and float for AMMSB_FP32.
 *
 * Conventions: every function returns an int status (AMMSB_OK == 0, negative
 * on failure); nothing throws across the ABI; ammsb_last_error() returns a
 * thread-local message for the last failure.  One context owns one HIP stream
 * and all device buffers; calls on one context are not reentrant, separate
 * contexts are independent.  Host arrays are owned by the caller and copied.
 */
#ifndef AMMSB_H_
#define AMMSB_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (replace MCMCException, c++/src/mcmc/exception.h:15-84) */
#define AMMSB_OK 0
#define AMMSB_ERR_ARG (-1)       /* bad argument / unsupported size          */
#define AMMSB_ERR_OOM (-2)       /* host or device allocation failed         */
#define AMMSB_ERR_HIP (-3)       /* HIP runtime error                        */
#define AMMSB_ERR_INVARIANT (-4) /* sampler invariant violated (e.g. minibatch
                                    exceeds buffer, network.cc:399-404)      */
#define AMMSB_ERR_STATE (-5)     /* call order (e.g. step before init)       */

/* ---- sampling strategy (c++/src/mcmc/types.h:11-16) */
enum ammsb_strategy {
  AMMSB_STRATIFIED_RANDOM_NODE = 0,
  AMMSB_STRATIFIED_RANDOM_NODE_LINKS = 1,
  AMMSB_STRATIFIED_RANDOM_NODE_NONLINKS = 2,
  AMMSB_RANDOM_EDGE = 3  /* random_edge_sampling (network.cc:237-268); needs
                            2*num_edges < 2^31 */
};

/* ---- Float precision (c++/src/mcmc/config.h.in:14-20 picks float or double
 * for everything; AMMSB_FP32 here is the "precision mode": FP32 bulk
 * arithmetic with FP64 perplexity log-sum, theta/beta gradient accumulation
 * and pi renormalisation). */
enum ammsb_precision { AMMSB_FP64 = 0, AMMSB_FP32 = 1 };

/* ---- options: every --mcmc.* flag of c++/src/mcmc/options.h:139-212 that
 * the sequential sampler reads, with the same defaults.  Zero means "derive"
 * exactly where the reference derives:
 *   a      = b^-c            (options.h:257-259)
 *   alpha  = 1/K             (options.h:260-262)
 *   interval = 50, stats_print_interval = 65536
 *                            (mcmc_sampler_stochastic.cc:17-26)
 *   num_node_sample = N/50, mini_batch_size = N/10
 *                            (mcmc_sampler_stochastic.cc:35-44)
 *   held_out_ratio = 0.1     (learner.cc:87-91)                             */
typedef struct ammsb_options {
  double alpha;                 /* --mcmc.alpha (0)                */
  double eta0;                  /* --mcmc.eta0 (1)                 */
  double eta1;                  /* --mcmc.eta1 (1)                 */
  double epsilon;               /* --mcmc.epsilon (1e-7)           */
  double a;                     /* --mcmc.a (0 -> b^-c)            */
  double b;                     /* --mcmc.b (1024)                 */
  double c;                     /* --mcmc.c (0.5)                  */
  int64_t K;                    /* --mcmc.K (300)                  */
  int64_t mini_batch_size;      /* --mcmc.mini-batch-size (0)      */
  int64_t num_node_sample;      /* --mcmc.num-node-sample (0)      */
  int32_t strategy;             /* --mcmc.strategy                 */
  int64_t sampler_max_source;   /* --mcmc.sampler.max-source (1)   */
  int64_t max_iteration;        /* --mcmc.max-iteration (10000000) */
  int64_t interval;             /* --mcmc.interval (0 -> 50)       */
  int64_t stats_print_interval; /* --mcmc.stats (0 -> 65536)       */
  int64_t num_updates;          /* --mcmc.num-updates (1000)       */
  double held_out_ratio;        /* --mcmc.held-out-ratio (0 -> 0.1) */
  int32_t random_seed;          /* --mcmc.seed (42)                */
  double convergence_threshold; /* --mcmc.convergence (1e-12)      */
  int32_t precision;            /* ammsb_precision (AMMSB_FP64)    */
  int32_t device;               /* HIP device ordinal (0)          */
  int32_t verbose;              /* echo config to stderr like
                                   learner.cc:105-120 (0)          */
} ammsb_options;

/* Undirected graph as a symmetric CSR: row i lists the neighbours of i in
 * ascending order, no self loops, no duplicates; cols has 2*num_edges
 * entries.  Replaces NetworkGraph (c++/src/mcmc/data.h:135-313). */
typedef struct ammsb_graph {
  int32_t N;
  int64_t num_edges;
  const int64_t *row_ptr; /* N+1 */
  const int32_t *cols;    /* 2*num_edges */
} ammsb_graph;

/* Held-out and test sets (EdgeMap, data.h:315-317): edges (a,b) with a<b,
 * label 1 = link, 0 = non-link.  Replaces Network::held_out_map/test_map
 * (network.h:326-327). */
typedef struct ammsb_split {
  int64_t held_out_count;
  const int32_t *held_out_edges; /* [held_out_count][2] */
  const uint8_t *held_out_labels;
  int64_t test_count;
  const int32_t *test_edges; /* [test_count][2] */
  const uint8_t *test_labels;
} ammsb_split;

/* Derived quantities of a context (read-only view). */
typedef struct ammsb_info {
  int32_t N;
  int64_t K;
  int64_t num_edges;
  int64_t ld;              /* device row stride of pi ++ phi_sum (elements) */
  int64_t mini_batch_size; /* after derivation */
  int64_t num_node_sample; /* n; n+1 neighbours are drawn per node */
  int64_t interval;
  int64_t stats_print_interval;
  int64_t max_minibatch_edges;
  int64_t max_minibatch_nodes;
  int64_t held_out_count;
  int64_t test_count;
  int64_t step_count;    /* Learner::step_count (starts at 1) */
  int64_t average_count; /* Learner::average_count (starts at 1) */
  double alpha, a, b, c, epsilon, eta0, eta1;
  int32_t precision;
  int32_t strategy;
  int64_t sampler_max_source;
  int64_t max_iteration;
} ammsb_info;

typedef struct ammsb_ctx ammsb_ctx;

/* ---- library ----------------------------------------------------------- */
const char *ammsb_last_error(void);
const char *ammsb_version(void);
/* Fill with the reference defaults (options.h:139-212). */
int ammsb_options_default(ammsb_options *opts);

/* The input options of mcmc::Options (options.h:192-205). */
typedef struct ammsb_input {
  char filename[4096];   /* --mcmc.input.file, -f ("")            */
  char input_class[64];  /* --mcmc.input.class, -c ("relativity") */
  int32_t contiguous;    /* --mcmc.input.contiguous (0)           */
  int32_t compressed;    /* --mcmc.input.compressed (0)           */
} ammsb_input;

/* Replaces mcmc::Options::Parse (options.h:236-255, boost program_options
 * with allow_unregistered): argv[0 .. argc) (no program name) -> defaults
 * overridden by --mcmc.* flags (--name value or --name=value), the short
 * aliases -K -m -n -x -i -I -h -f -c (-K32 or -K 32), the input switches,
 * --verbose, --mcmc.precision fp64|fp32, --mcmc.device, and --config FILE
 * ("name = value" lines, [section] prefixes, '#' comments; the command line
 * wins).  Integers parse as the reference's parse<T> (options.h:39-57,
 * options.cc:5-67: 0x hex, leading-0 octal, a k/kb m/mb g/gb suffix for
 * 2^10/2^20/2^30).  Unrecognised arguments are skipped and counted in
 * *n_unrecognised (may be null).  a, alpha, interval, ... keep 0 where not
 * given; ammsb_create derives them.  A malformed value is AMMSB_ERR_ARG
 * (the reference exits 33 on an option error, main.cc:83-86). */
int ammsb_options_parse(int32_t argc, const char *const *argv, ammsb_options *opts,
                        ammsb_input *input, int32_t *n_unrecognised);

/* ---- host-side graph tools (no GPU needed) ----------------------------- */
/* Synthetic generators (see ammsb_gen_params).  The library allocates
 * *row_ptr (N+1) and *cols (2E); release both with ammsb_free(). */
enum ammsb_gen_kind {
  /* pi_a ~ Dir(gen_alpha), beta_k ~ U(beta_lo, beta_hi); pair (a,b) links
   * with prob sum_k pi_ak pi_bk beta_k + delta (1 - sum_k pi_ak pi_bk).
   * O(N^2 K): small N only. */
  AMMSB_GEN_PLANTED = 0,
  /* each node joins 1..max_comms communities (weights Dir(gen_alpha));
   * Erdos-Renyi inside each community with p set for mean_degree, plus
   * round(delta * N(N-1)/2) uniform background pairs.  O(N + E). */
  AMMSB_GEN_COMMUNITY = 1
};
typedef struct ammsb_gen_params {
  int32_t kind;
  int32_t N;
  int32_t K;
  double gen_alpha;   /* Dirichlet concentration */
  double beta_lo;     /* PLANTED */
  double beta_hi;     /* PLANTED */
  double delta;       /* background link probability (1e-7) */
  double mean_degree; /* COMMUNITY */
  int32_t max_comms;  /* COMMUNITY (3) */
  uint64_t seed;
} ammsb_gen_params;
int ammsb_graph_generate(const ammsb_gen_params *p, int64_t **row_ptr,
                         int32_t **cols, int64_t *num_edges);

/* Held-out/test split with the reference's semantics
 * (network.cc:503-614): held_out_size = ratio*E, p = held_out_size/2;
 * held-out = p random edges + p random non-edges; test = p + 1 random
 * edges not in held-out (init_test_set's `if (p < 0) break` after `p--`
 * admits one more, network.cc:566-580; p when only p remain) + p non-edges
 * not in either.  Arrays allocated by the library; release with
 * ammsb_free(). */
int ammsb_split_generate(const ammsb_graph *g, double held_out_ratio,
                         uint64_t seed, int32_t **held_out_edges,
                         uint8_t **held_out_labels, int64_t *held_out_count,
                         int32_t **test_edges, uint8_t **test_labels,
                         int64_t *test_count);
/* Edge-list reader of the reference's "relativity" input class
 * (preprocess/relativity.cc:20-193, data_factory.cc:7-45): skips the first 4
 * (header) lines, then one pair of integer ids per line; plain or gzip.
 * contiguous != 0: ids as given, N = max id + 1 (--mcmc.input.contiguous);
 * else ids renumbered 0..N-1 in order of first appearance.  Self-links and
 * repeated pairs are dropped and counted (self_links, duplicates may be
 * NULL).  *row_ptr (N+1) and *cols (2E, ascending rows) are allocated by the
 * library; release with ammsb_free(). */
int ammsb_graph_load_edge_list(const char *path, int32_t contiguous, int32_t *N,
                               int64_t *num_edges, int64_t **row_ptr, int32_t **cols,
                               int64_t *self_links, int64_t *duplicates);
void ammsb_free(void *p);
/* One Philox4x32-10 block of the library's counter-based stream:
 * out = philox(ctr[4], key[2]) (Salmon et al., SC'11).  The stream replaces
 * the reference's xorshift128+ generator (c++/src/mcmc/random.cc:136-176);
 * exported so callers can replay any draw of the samplers. */
void ammsb_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- context (Learner / MCMCSamplerStochastic) ------------------------- */
/* Replaces MCMCSamplerStochastic(const Options&) + LoadNetwork()
 * (mcmc_sampler_stochastic.cc:9-44, learner.cc:61-97): copies the graph,
 * builds the device CSR and the held-out∪test hash, derives defaults and
 * allocates the HBM state.  split may be NULL: the library then draws the
 * split itself with ammsb_split_generate(ratio, seed). */
int ammsb_create(const ammsb_options *opts, const ammsb_graph *g,
                 const ammsb_split *split, ammsb_ctx **out);
int ammsb_destroy(ammsb_ctx *ctx);
int ammsb_get_info(const ammsb_ctx *ctx, ammsb_info *info);

/* Replaces MCMCSamplerStochastic::init() state draw
 * (mcmc_sampler_stochastic.cc:54-76): theta ~ Gamma(eta0, eta1) per entry,
 * beta = theta[:,1]/rowsum, phi ~ Gamma(1,1), pi = phi/rowsum.  Counter-based
 * draws keyed by seed. */
int ammsb_init_state(ammsb_ctx *ctx, uint64_t seed);
/* Upload / download the state. pi: [N][K+1] (double or float per precision),
 * theta: [K][2] double.  beta is recomputed from theta on upload. */
int ammsb_set_state(ammsb_ctx *ctx, const void *pi, const double *theta);
int ammsb_get_state(ammsb_ctx *ctx, void *pi, double *theta, double *beta);
/* Rows of pi ++ phi_sum for a list of nodes ([n][K+1]). */
int ammsb_get_rows(ammsb_ctx *ctx, const int32_t *nodes, int64_t n,
                   void *out);
/* Learner::step_count / average_count setters (tests, resumption). */
int ammsb_set_counters(ammsb_ctx *ctx, int64_t step_count,
                       int64_t average_count);

/* ---- the loop ---------------------------------------------------------- */
/* iters full SGRLD iterations on the device, each exactly the body of
 * MCMCSamplerStochastic::run() (mcmc_sampler_stochastic.cc:131-174) without
 * the perplexity check: minibatch, nodes_in_batch, eps_t, neighbour
 * sampling + update_phi per node, update_pi, update_beta, step_count++.
 * Asynchronous w.r.t. the host except for error checks when sync != 0. */
int ammsb_step(ammsb_ctx *ctx, int64_t iters, int32_t sync);
/* Learner::cal_perplexity_held_out (learner.cc:126-183): updates the
 * running average per held-out edge, returns -mean log, increments
 * average_count. */
int ammsb_perplexity(ammsb_ctx *ctx, double *ppx);
/* MCMCSamplerStochastic::run() (mcmc_sampler_stochastic.cc:93-183): the
 * loop with perplexity every `interval` steps, the convergence test, the
 * reference's stdout line
 *   "average_count is: <c> step count: <t> time: <s> perplexity for hold
 *    out set: <x>"
 * and the per-phase timer table.  max_steps <= 0 means max_iteration.
 * If ppx_out/ppx_steps are non-NULL, up to cap checkpoints are stored and
 * *n_ppx receives how many were produced. */
int ammsb_run(ammsb_ctx *ctx, int64_t max_steps, int32_t print,
              double *ppx_out, int64_t *ppx_steps, int64_t cap,
              int64_t *n_ppx);
/* Per-phase accumulated device time in seconds and tick counts, in the
 * order of mcmc_sampler_stochastic.cc:95-102: outer, perplexity,
 * sample_mini_batch, nodes_in_mini_batch, sample_neighbor_nodes,
 * update_phi, update_pi, update_beta.  secs[8], ticks[8]. */
int ammsb_timers(ammsb_ctx *ctx, double *secs, int64_t *ticks);
/* on: 0 off, 1 every phase, 2 update_phi only (the fewest events in the
 * iteration's stream). */
int ammsb_enable_timers(ammsb_ctx *ctx, int32_t on);

/* ---- single operators with injected draws (parity) --------------------- */
/* update_phi for n_nodes distinct nodes followed by update_pi for the same
 * nodes: MCMCSamplerStochastic::update_phi (mcmc_sampler_stochastic.cc:
 * 269-313) in the pi ++ phi_sum form of update_phi_node
 * (mcmc_sampler_stochastic_distr.cc:1158-1227) + pi_from_phi (:743-751).
 * nbrs: [n_nodes][n_nb] neighbour ids (a neighbour equal to the node is
 * skipped as in :275-277); noise: [n_nodes][K] (same precision as pi) or
 * NULL for the Philox noise of iteration step_count. */
int ammsb_update_phi(ammsb_ctx *ctx, const int32_t *nodes, int64_t n_nodes,
                     const int32_t *nbrs, int64_t n_nb, const void *noise,
                     double eps_t);
/* MCMCSamplerStochastic::update_beta (mcmc_sampler_stochastic.cc:200-267):
 * edges [E][2], y [E] in {0,1}, noise [K][2] double (NULL = Philox).
 * Both operators take host arrays, check the ids on the host, copy them into
 * a device workspace kept with the context (grown when a call needs more)
 * and return after the device work has finished. */
int ammsb_update_beta(ammsb_ctx *ctx, const int32_t *edges, const uint8_t *y,
                      int64_t E, double scale, const double *noise,
                      double eps_t);

/* ---- device sampler outputs (bit-exact parity against a host replica) -- */
/* Runs the device minibatch sampler, nodes_in_batch and neighbour sampler
 * for step_count == step (state is not modified) and copies out:
 * flag (0 non-links, 1 links), scale, the minibatch edges [n_edges][2] in
 * sampler order, the unique nodes (unordered), and per node its n+1
 * neighbours [n_nodes][n+1] with bit 31 set when (node, nbr) is a link.
 * Buffers must hold max_minibatch_edges / max_minibatch_nodes entries. */
int ammsb_debug_sample(ammsb_ctx *ctx, int64_t step, int32_t *flag,
                       double *scale, int32_t *edges, int64_t *n_edges,
                       int32_t *nodes, int64_t *n_nodes, int32_t *nbrs);

/* Totals over every iteration run so far (synchronises the stream): nodes
 * whose phi/pi rows were updated and minibatch edges used by update_beta. */
int ammsb_get_totals(ammsb_ctx *ctx, int64_t *nodes_updated,
                     int64_t *minibatch_edges);

/* Launch-shape tuning (no counterpart in the reference, which runs on CPU):
 *   "phi_waves"  wavefronts per minibatch node in update_phi (2, 4, 8)
 *   "phi_unroll" neighbour rows in flight per row group (1, 2, 4)
 *   "ll_fwd_pairs", "ll_bwd_pairs"  pairs per LDS exchange of the
 *                K-chunked log-likelihood forward / adjoint kernels (rows
 *                wider than 2 KB): 2, 4, 8
 *   "ll_fwd_prefetch", "ll_bwd_prefetch"  1: the next pairs' rows are
 *                loaded before the exchange
 *   "ll_fwd_rows", "ll_bwd_rows"  1 (default): the forward / adjoint pass
 *                of rows wider than 2 KB streams whole rows per wavefront
 *                against the source row held in registers; 0: the
 *                K-chunked kernels above
 * Set phi_waves first; the (waves, unroll) pair is validated when
 * phi_unroll is set.  Results do not depend on the shape beyond
 * summation order (the log-likelihood shapes not even on that).  Defaults:
 * update_phi (8, 2) where that shape exists for the row (a row held by one
 * wavefront in 1, 2, 4 or 8 vectors per lane, e.g. FP32 K 1024), else
 * (4, 1); log-likelihood FP64 (4,0) both ways, FP32 forward (8,0), adjoint
 * (2,1).
 * ammsb_get_tuning also reads "sampler_lds" (read-only): bit 0 set when the
 * minibatch sampler's de-dup table lives in LDS, bit 1 when the neighbour
 * sampler's tables do (else they live in HBM); and "ring_slots"
 * (read-only): iterations of draws the sample-ahead ring holds (1 when one
 * iteration's draws exceed the ring's 1 GiB budget). */
int ammsb_set_tuning(ammsb_ctx *ctx, const char *name, int64_t value);
int ammsb_get_tuning(ammsb_ctx *ctx, const char *name, int64_t *value);

/* Bandwidth probe (no counterpart in the reference): `reps` contiguous reads
 * of the resident pi array; *seconds = average time of one read, *bytes =
 * bytes read.  Also a read of known size for calibrating HBM counters. */
int ammsb_probe_read(ammsb_ctx *ctx, int32_t reps, double *seconds, int64_t *bytes);
/* Bandwidth probe (no counterpart in the reference): the random row gather
 * of update_phi for iteration step_count -- its minibatch nodes' own rows
 * and n+1 neighbour rows, read with the context's update_phi launch shape
 * when it is 8 x 2, else 4 x 1 -- without the arithmetic and the row writes; `reps` launches, *seconds =
 * average time of one, *bytes = rows and neighbour ids read.  The state is
 * not changed; the sample-ahead ring is refilled before the next step. */
int ammsb_probe_gather(ammsb_ctx *ctx, int32_t reps, double *seconds, int64_t *bytes);

/* ---- differentiation --------------------------------------------------- */
/* Minibatch log-likelihood of M source nodes, all pointers on the device:
 *   L = sum_s sum_{p in [pair_ptr[s], pair_ptr[s+1])} log p(y_p | pi_s, pi_{b_p}, beta)
 *   p(y) = sum_k pi_sk (pi_bk f_k + e), f_k = beta_k - eps / eps - beta_k,
 *          e = eps / 1 - eps for y = 1 / 0
 * -- the per-pair likelihood of update_phi (mcmc_sampler_stochastic.cc:285-292)
 * and update_beta (:220-233).  pi_s = phi_s / sum(phi_s) from the caller's
 * rows phi [M][phi_ld] (precision of the context), neighbour rows b_p from the
 * resident state (constants, as in update_phi), beta_k = theta_k1 /
 * (theta_k0 + theta_k1) from theta [K][2] (NULL: the resident theta).  A pair
 * whose neighbour equals nodes[s] is skipped (:275-277).  P = number of pairs.
 * fwd: per_source [M] (may be NULL) and total (may be NULL), FP64.
 * bwd: the adjoint, recomputing each pair's likelihood:
 *   grad_phi [M][phi_ld] = gout * dL/dphi   (update_phi's grads, :293-295)
 *   grad_theta [K][2]    = gout * dL/dtheta (update_beta's G, :234-239; may be NULL)
 * gout points to one double.  stream: a HIP stream to run on (NULL: the
 * context's stream); the call returns after the work is done. */
int ammsb_loglik_fwd(ammsb_ctx *ctx, int64_t M, const int32_t *nodes, const void *phi,
                     int64_t phi_ld, const double *theta, const int64_t *pair_ptr,
                     const int32_t *pair_nbr, const uint8_t *pair_y, int64_t P,
                     double *per_source, double *total, void *stream);
int ammsb_loglik_bwd(ammsb_ctx *ctx, int64_t M, const int32_t *nodes, const void *phi,
                     int64_t phi_ld, const double *theta, const int64_t *pair_ptr,
                     const int32_t *pair_nbr, const uint8_t *pair_y, int64_t P,
                     const double *gout, void *grad_phi, double *grad_theta, void *stream);
/* ammsb_loglik_bwd with no device allocation.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * adjoint's scratch -- its beta partial rows, 8 * rows * K bytes, and an
 * error word -- lives in a file-scope device array of 32 MiB + 4 KiB per GPU,
 * held until the process ends; a call that needs more is AMMSB_ERR_ARG
 * (rows = the adjoint's workgroups, one partial row each: up to 1024 for
 * the K-chunked kernels, the resident workgroups for the whole-row one:
 * at K 2048 on an MI355X 8 MiB in FP64, 12 MiB in FP32).  Same arguments,
 * results and errors as
 * ammsb_loglik_bwd otherwise.  All calls on one GPU share the array: calls
 * that overlap on one GPU race. */
int ammsb_loglik_bwd_noalloc(ammsb_ctx *ctx, int64_t M, const int32_t *nodes, const void *phi,
                             int64_t phi_ld, const double *theta, const int64_t *pair_ptr,
                             const int32_t *pair_nbr, const uint8_t *pair_y, int64_t P,
                             const double *gout, void *grad_phi, double *grad_theta,
                             void *stream);

/* Device pointers of the resident state for interop (torch DLPack-free
 * views): pi ([N][ld] of the precision type), theta ([K][2] double),
 * beta ([K] double). */
int ammsb_device_ptrs(ammsb_ctx *ctx, void **pi, void **theta, void **beta);

/* ------------------------------------------------------------------------
 * Batches of independent chains (no counterpart in the reference, which
 * runs one sampler per process; each chain is the sampler of
 * mcmc_sampler_stochastic.cc:93-183 with its own graph, split, options,
 * K and seed).  All chains of a batch share one precision, one device and
 * one stream.  Every phase of an iteration is one launch over all chains
 * (update_phi a persistent kernel fed by a work queue), and the launch
 * sequence of an iteration is captured once as a hipGraph and replayed.
 * Limits: 1..1024 chains; K <= 256 (FP64) / 512 (FP32); stratified
 * minibatch strategies.  Chains draw their minibatch and neighbours every
 * iteration from the same Philox streams as a single context with the
 * chain's options and seed, so a chain's draws equal those of ammsb_step
 * on such a context. */
typedef struct ammsb_batch ammsb_batch;

/* opts, graphs: [n]; splits: [n] or NULL (no held-out set).  The arrays
 * are copied. */
int ammsb_batch_create(int32_t n, const ammsb_options *opts, const ammsb_graph *graphs,
                       const ammsb_split *splits, ammsb_batch **out);
int ammsb_batch_destroy(ammsb_batch *b);
/* Chain i as a context owned by the batch, for the single-context calls
 * (ammsb_get_info, ammsb_init_state, ammsb_set_state, ammsb_get_state,
 * ammsb_get_rows, ammsb_perplexity, ...).  ammsb_destroy refuses it. */
int ammsb_batch_chain(ammsb_batch *b, int32_t i, ammsb_ctx **chain);
/* init_state of every chain with seeds [n] (NULL: each chain's
 * options.random_seed). */
int ammsb_batch_init_state(ammsb_batch *b, const uint64_t *seeds);
/* `iters` iterations of every chain (graph replays unless disabled). */
int ammsb_batch_step(ammsb_batch *b, int64_t iters);
/* Held-out perplexity of every chain (cal_perplexity_held_out,
 * learner.cc:126-183) into out [n]. */
int ammsb_batch_perplexity(ammsb_batch *b, double *out);
/* 1: replay the captured hipGraph (default); 0: launch the kernels. */
int ammsb_batch_use_graph(ammsb_batch *b, int32_t enable);
/* Sampler refills of the batch's ring (no counterpart in the reference):
 * 0 (default) draws the next 16 iterations on a second stream while the
 * current 16 run; 1 draws them on the batch's stream, in order with the
 * iterations (64 chains on one MI355X: 0.492 against 0.518 ms per
 * iteration).  The draws are the same. */
int ammsb_batch_refill_in_order(ammsb_batch *b, int32_t enable);

/* Launch shape of chain i in the batched update_phi:
 *   lanes_per_row (4..64) x vectors_per_lane (1, 2) of 16 bytes = the
 *   K-chunk each lane owns; nodes_per_block = minibatch nodes a workgroup
 *   of 4 wavefronts processes together (power of two; nodes per wavefront
 *   = nodes_per_block / 4).  The shape changes summation order only. */
typedef struct ammsb_batch_shape {
  int32_t lanes_per_row;
  int32_t vectors_per_lane;
  int32_t nodes_per_block;
  double us;   /* autotune: measured update_phi time of the chain, else 0 */
} ammsb_batch_shape;
int ammsb_batch_set_shape(ammsb_batch *b, int32_t i, const ammsb_batch_shape *shape);
int ammsb_batch_get_shape(ammsb_batch *b, int32_t i, ammsb_batch_shape *shape);
/* Autotuner: for every chain, times each valid shape on the chain's
 * current minibatch (reps launches each, HIP events) and keeps the
 * fastest; chosen shapes into out [n] (may be NULL).  The state is not
 * changed.  Requires initialised state. */
int ammsb_batch_autotune(ammsb_batch *b, int32_t reps, ammsb_batch_shape *out);

#ifdef __cplusplus
}
#endif

#endif /* AMMSB_H_ */
