// Command-line driver: the sequential path of the reference's main
// (c++/apps/main/main.cc) on the MI355X sampler.  It echoes the command
// line, parses the --mcmc.* options (ammsb_options_parse, unrecognised
// arguments allowed), loads a SNAP edge list (input classes relativity,
// rc, rz, rcz: relativity.cc:20-193), draws the held-out/test split
// (network.cc:503-614), and runs MCMCSamplerStochastic::init() + run()
// (mcmc_sampler_stochastic.cc:54-183) with the reference's log line and
// timer table.  Option and input errors exit with 33, as main.cc:79-86.

#include <stdio.h>
#include <string.h>

#include "../../include/ammsb.h"

namespace {

void usage() {
  printf(
      "Options (mcmc::Options, options.h:139-263):\n"
      "  --config FILE                 name = value lines, [section] prefixes\n"
      "  --mcmc.alpha A (0: 1/K)       --mcmc.eta0 E (1)   --mcmc.eta1 E (1)\n"
      "  --mcmc.epsilon E (1e-7)       --mcmc.a A (0: b^-c)  --mcmc.b B (1024)  --mcmc.c C (0.5)\n"
      "  --mcmc.K,-K K (300)           --mcmc.mini-batch-size,-m M (0: N/10)\n"
      "  --mcmc.num-node-sample,-n N (0: N/50)\n"
      "  --mcmc.strategy S             stratified-random-node (default), "
      "stratified-random-node-links,\n"
      "                                stratified-random-node-nonlinks, random-edge\n"
      "  --mcmc.sampler.max-source S (1)   --mcmc.max-iteration,-x X (10000000)\n"
      "  --mcmc.interval,-i I (0: 50)  --mcmc.stats,-I S (0: 65536)  --mcmc.num-updates U (1000)\n"
      "  --mcmc.held-out-ratio,-h R (0: 0.1)  --mcmc.seed S (42)  --mcmc.convergence C (1e-12)\n"
      "  --mcmc.input.file,-f F        --mcmc.input.class,-c C (relativity; also rc, rz, rcz)\n"
      "  --mcmc.input.contiguous       --mcmc.input.compressed\n"
      "  --mcmc.precision fp64|fp32 (fp64)   --mcmc.device D (0)   --verbose\n");
}

bool edge_list_class(const char* c) {
  return !strcmp(c, "relativity") || !strcmp(c, "rc") || !strcmp(c, "rz") || !strcmp(c, "rcz");
}

}  // namespace

int main(int argc, char** argv) {
  printf("Command line: ");
  for (int i = 0; i < argc; ++i) printf("%s ", argv[i]);
  printf("\n");
  fflush(stdout);
  for (int i = 1; i < argc; ++i)
    if (!strcmp(argv[i], "--help")) {
      usage();
      return 0;
    }
  ammsb_options opts;
  ammsb_input input;
  int32_t unrecognised = 0;
  if (ammsb_options_parse(argc - 1, argv + 1, &opts, &input, &unrecognised) != AMMSB_OK) {
    fprintf(stderr, "Option error: %s\n", ammsb_last_error());
    return 33;
  }
  if (!edge_list_class(input.input_class)) {
    fprintf(stderr,
            "IO error: input class '%s' is not read by this driver (edge lists: relativity, "
            "rc, rz, rcz)\n",
            input.input_class);
    return 33;
  }
  if (!input.filename[0]) {
    fprintf(stderr, "IO error: no input file (--mcmc.input.file / -f)\n");
    return 33;
  }
  int32_t N = 0;
  int64_t E = 0, self_links = 0, duplicates = 0;
  int64_t* row_ptr = nullptr;
  int32_t* cols = nullptr;
  if (ammsb_graph_load_edge_list(input.filename, input.contiguous, &N, &E, &row_ptr, &cols,
                                 &self_links, &duplicates) != AMMSB_OK) {
    fprintf(stderr, "IO error: %s\n", ammsb_last_error());
    return 33;
  }
  printf("Loaded %s: N %d E %lld (%lld self-links, %lld repeated edges dropped)\n",
         input.filename, N, (long long)E, (long long)self_links, (long long)duplicates);
  ammsb_graph g = {N, E, row_ptr, cols};
  printf("start MCMC stochastical\n");
  ammsb_ctx* ctx = nullptr;
  int rc = ammsb_create(&opts, &g, nullptr, &ctx);  // null split: drawn as the reference does
  if (rc == AMMSB_OK) rc = ammsb_init_state(ctx, (uint64_t)(uint32_t)opts.random_seed);
  // every phase timed, as the reference's Timers are (its table is printed
  // every --mcmc.stats steps and at the end)
  if (rc == AMMSB_OK) rc = ammsb_enable_timers(ctx, 1);
  if (rc == AMMSB_OK) rc = ammsb_run(ctx, 0, 1, nullptr, nullptr, 0, nullptr);
  if (rc != AMMSB_OK) fprintf(stderr, "error %d: %s\n", rc, ammsb_last_error());
  if (ctx) ammsb_destroy(ctx);
  ammsb_free(row_ptr);
  ammsb_free(cols);
  return rc == AMMSB_OK ? 0 : 1;
}
