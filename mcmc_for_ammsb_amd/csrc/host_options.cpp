// The command-line surface of the reference's mcmc::Options
// (c++/src/mcmc/options.h:139-263, size suffixes options.cc:5-67), for C/C++
// hosts that replace its main (c++/apps/main/main.cc):
//   --name value | --name=value       every --mcmc.* option and --config
//   -K 32 | -K32                      the short aliases K m n x i I h f c
//   --mcmc.input.contiguous, --mcmc.input.compressed, --verbose   switches
//   --config FILE                     "name = value" lines, [section]
//                                     prefixes, '#' comments; a value given
//                                     on the command line wins
// Integers take the reference's forms: 0x hex, leading-0 octal, and a k/kb,
// m/mb or g/gb suffix (any case) for 2^10, 2^20, 2^30.
// Unrecognised arguments are counted and skipped (allow_unregistered).  The
// derived defaults (a = b^-c, alpha = 1/K, ...) are applied by ammsb_create,
// so zeros are kept here as given.

#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

#include <cctype>
#include <fstream>
#include <set>
#include <string>
#include <vector>

#include "../../include/ammsb.h"
#include "host_util.h"

namespace ammsb {
namespace {

enum Kind { K_FLOAT, K_SIZE, K_INT, K_STRATEGY, K_PRECISION, K_FILE, K_CLASS };

struct Flag {
  const char* name;   // long name without the leading dashes
  const char* alias;  // one-letter alias or nullptr
  Kind kind;
  size_t offset;      // into ammsb_options (numeric kinds)
};

#define OPT(field) offsetof(ammsb_options, field)
const Flag kFlags[] = {
    {"mcmc.alpha", nullptr, K_FLOAT, OPT(alpha)},
    {"mcmc.eta0", nullptr, K_FLOAT, OPT(eta0)},
    {"mcmc.eta1", nullptr, K_FLOAT, OPT(eta1)},
    {"mcmc.epsilon", nullptr, K_FLOAT, OPT(epsilon)},
    {"mcmc.a", nullptr, K_FLOAT, OPT(a)},
    {"mcmc.b", nullptr, K_FLOAT, OPT(b)},
    {"mcmc.c", nullptr, K_FLOAT, OPT(c)},
    {"mcmc.K", "K", K_SIZE, OPT(K)},
    {"mcmc.mini-batch-size", "m", K_SIZE, OPT(mini_batch_size)},
    {"mcmc.num-node-sample", "n", K_SIZE, OPT(num_node_sample)},
    {"mcmc.strategy", nullptr, K_STRATEGY, OPT(strategy)},
    {"mcmc.sampler.max-source", nullptr, K_SIZE, OPT(sampler_max_source)},
    {"mcmc.max-iteration", "x", K_SIZE, OPT(max_iteration)},
    {"mcmc.interval", "i", K_SIZE, OPT(interval)},
    {"mcmc.stats", "I", K_SIZE, OPT(stats_print_interval)},
    {"mcmc.num-updates", nullptr, K_SIZE, OPT(num_updates)},
    {"mcmc.held-out-ratio", "h", K_FLOAT, OPT(held_out_ratio)},
    {"mcmc.seed", nullptr, K_INT, OPT(random_seed)},
    {"mcmc.convergence", nullptr, K_FLOAT, OPT(convergence_threshold)},
    {"mcmc.precision", nullptr, K_PRECISION, OPT(precision)},
    {"mcmc.device", nullptr, K_INT, OPT(device)},
    {"mcmc.input.file", "f", K_FILE, 0},
    {"mcmc.input.class", "c", K_CLASS, 0},
};
#undef OPT

const char* const kSwitches[] = {"mcmc.input.contiguous", "mcmc.input.compressed", "verbose"};

const Flag* find_flag(const std::string& name, bool alias) {
  for (const Flag& f : kFlags)
    if (alias ? (f.alias && name == f.alias) : name == f.name) return &f;
  return nullptr;
}

bool is_switch(const std::string& name) {
  for (const char* s : kSwitches)
    if (name == s) return true;
  return false;
}

bool is_number(const std::string& t) {
  char* end = nullptr;
  strtod(t.c_str(), &end);
  return end && *end == '\0' && !t.empty();
}

// Integral options as the reference parses them (mcmc::parse<T>,
// options.h:39-57, over parse_size_t, options.cc:5-67): a leading '-'
// negates; "0x" selects hexadecimal and a leading '0' octal (whose digit
// test is isdigit, as there); the rest of the string may be one suffix
// k/kb, m/mb, g/gb (any case) for 2^10, 2^20, 2^30; the result must fit
// [lo, hi].
bool parse_size_t_ref(const char* arg, uint64_t* out) {
  uint64_t n = 0;
  int base = 10;
  if (strncmp(arg, "0x", 2) == 0 || strncmp(arg, "0X", 2) == 0) {
    base = 16;
    arg += 2;
  } else if (arg[0] == '0') {
    base = 8;
    arg++;
  }
  const uint64_t max = ~0ull;
  for (; *arg != '\0'; ++arg) {
    int d = -1;
    if (base == 16 && isxdigit((unsigned char)*arg))
      d = isdigit((unsigned char)*arg) ? *arg - '0' : tolower((unsigned char)*arg) - 'a' + 10;
    else if (base <= 10 && isdigit((unsigned char)*arg))
      d = *arg - '0';
    if (d >= 0) {
      if ((max - (uint64_t)d) / (uint64_t)base < n) return false;
      n = (uint64_t)d + n * (uint64_t)base;
      continue;
    }
    int shift = -1;
    if (strcasecmp(arg, "g") == 0 || strcasecmp(arg, "gb") == 0) shift = 30;
    else if (strcasecmp(arg, "m") == 0 || strcasecmp(arg, "mb") == 0) shift = 20;
    else if (strcasecmp(arg, "k") == 0 || strcasecmp(arg, "kb") == 0) shift = 10;
    if (shift < 0 || (max >> shift) < n) return false;
    n <<= shift;
    break;
  }
  *out = n;
  return true;
}

bool parse_integral(const std::string& s, int64_t lo, int64_t hi, int64_t* out) {
  const char* arg = s.c_str();
  uint64_t n;
  if (arg[0] == '-') {
    if (!parse_size_t_ref(arg + 1, &n) || n > (uint64_t)INT64_MAX) return false;
    const int64_t v = -(int64_t)n;
    if (v < lo) return false;
    *out = v;
    return true;
  }
  if (!parse_size_t_ref(arg, &n) || n > (uint64_t)hi) return false;
  *out = (int64_t)n;
  return true;
}

bool strategy_of(const std::string& v, int32_t* out) {
  // types.h:11-16 names, in the ABI's order
  static const char* const kNames[] = {"stratified-random-node", "stratified-random-node-links",
                                       "stratified-random-node-nonlinks", "random-edge"};
  for (int i = 0; i < 4; ++i)
    if (v == kNames[i]) {
      *out = i;
      return true;
    }
  return false;
}

struct Parser {
  ammsb_options* o;
  ammsb_input* in;
  std::string err;

  bool set(const Flag& f, const std::string& val) {
    char* base = reinterpret_cast<char*>(o);
    switch (f.kind) {
      case K_FLOAT: {  // as lexical_cast<double>: no blanks, hexadecimal or nan(...) forms
        bool plain = !val.empty();
        for (const char ch : val)
          if (isspace((unsigned char)ch) || ch == 'x' || ch == 'X' || ch == '(') plain = false;
        if (!plain) break;
        char* end = nullptr;
        const double v = strtod(val.c_str(), &end);
        if (!end || *end != '\0') break;
        memcpy(base + f.offset, &v, sizeof v);
        return true;
      }
      case K_SIZE: {  // size_t options: no negative values
        int64_t v;
        if (!parse_integral(val, 0, INT64_MAX, &v)) break;
        memcpy(base + f.offset, &v, sizeof v);
        return true;
      }
      case K_INT: {  // int options
        int64_t v;
        if (!parse_integral(val, INT32_MIN, INT32_MAX, &v)) break;
        const int32_t iv = (int32_t)v;
        memcpy(base + f.offset, &iv, sizeof iv);
        return true;
      }
      case K_STRATEGY: {
        int32_t v;
        if (!strategy_of(val, &v)) {
          err = "Unknown strategy '" + val + "'";
          return false;
        }
        o->strategy = v;
        return true;
      }
      case K_PRECISION:
        if (val == "fp64") {
          o->precision = AMMSB_FP64;
          return true;
        }
        if (val == "fp32") {
          o->precision = AMMSB_FP32;
          return true;
        }
        err = "precision must be fp64 or fp32";
        return false;
      case K_FILE:
        if (val.size() >= sizeof in->filename) break;
        strcpy(in->filename, val.c_str());
        return true;
      case K_CLASS:
        if (val.size() >= sizeof in->input_class) break;
        strcpy(in->input_class, val.c_str());
        return true;
    }
    err = "invalid value '" + val + "' for --" + f.name;
    return false;
  }

  void set_switch(const std::string& name, bool on) {
    if (name == "mcmc.input.contiguous") in->contiguous = on;
    if (name == "mcmc.input.compressed") in->compressed = on;
    if (name == "verbose") o->verbose = on;
  }
};

}  // namespace
}  // namespace ammsb

extern "C" int ammsb_options_parse(int32_t argc, const char* const* argv, ammsb_options* opts,
                                   ammsb_input* input, int32_t* n_unrecognised) {
  using namespace ammsb;
  if (!opts || !input || argc < 0 || (argc > 0 && !argv)) return AMMSB_ERR_ARG;
  ammsb_options_default(opts);
  memset(input, 0, sizeof(*input));
  strcpy(input->input_class, "relativity");
  Parser p{opts, input, {}};
  std::string config;
  std::set<std::string> given;  // long names given on the command line
  int32_t unknown = 0;
  for (int32_t i = 0; i < argc; ++i) {
    const std::string tok = argv[i] ? argv[i] : "";
    std::string name, val;
    bool has_val = false;
    const Flag* f = nullptr;
    if (tok.size() > 2 && tok.compare(0, 2, "--") == 0) {
      const size_t eq = tok.find('=');
      name = tok.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
      if (eq != std::string::npos) {
        val = tok.substr(eq + 1);
        has_val = true;
      }
      f = find_flag(name, false);
    } else if (tok.size() >= 2 && tok[0] == '-' && !is_number(tok)) {
      name = tok.substr(1, 1);
      if (tok.size() > 2) {
        val = tok.substr(2);
        has_val = true;
      }
      f = find_flag(name, true);
      if (f) name = f->name;
    } else {
      ++unknown;  // positional argument
      continue;
    }
    if (name == "config") {
      if (!has_val) {
        if (i + 1 >= argc) {
          set_error("option --config needs a value");
          return AMMSB_ERR_ARG;
        }
        val = argv[++i];
      }
      config = val;
      continue;
    }
    if (is_switch(name)) {  // a switch takes no value
      p.set_switch(name, true);
      given.insert(name);
      continue;
    }
    if (!f) {
      ++unknown;
      continue;
    }
    if (!has_val) {
      if (i + 1 >= argc) {
        set_error("option --" + std::string(f->name) + " needs a value");
        return AMMSB_ERR_ARG;
      }
      val = argv[++i];
    }
    if (!p.set(*f, val)) {
      set_error(p.err);
      return AMMSB_ERR_ARG;
    }
    given.insert(f->name);
  }
  if (!config.empty()) {
    std::ifstream file(config);
    if (!file) {
      set_error("Cannot open config file " + config);
      return AMMSB_ERR_ARG;
    }
    std::string line, section;
    while (std::getline(file, line)) {
      const size_t hash = line.find('#');
      if (hash != std::string::npos) line.resize(hash);
      auto trim = [](std::string s) {
        while (!s.empty() && isspace((unsigned char)s.back())) s.pop_back();
        size_t b = 0;
        while (b < s.size() && isspace((unsigned char)s[b])) ++b;
        return s.substr(b);
      };
      line = trim(line);
      if (line.empty()) continue;
      if (line.front() == '[' && line.back() == ']') {
        section = trim(line.substr(1, line.size() - 2)) + ".";
        continue;
      }
      const size_t eq = line.find('=');
      const std::string key = section + trim(line.substr(0, eq));
      const std::string val = eq == std::string::npos ? "" : trim(line.substr(eq + 1));
      if (given.count(key)) continue;  // the command line wins
      if (const Flag* f = find_flag(key, false)) {
        if (!p.set(*f, val)) {
          set_error(p.err);
          return AMMSB_ERR_ARG;
        }
      } else if (is_switch(key)) {
        std::string low = val;
        for (char& ch : low) ch = (char)tolower((unsigned char)ch);
        p.set_switch(key, low == "1" || low == "true" || low == "yes" || low == "on");
      }
    }
  }
  if (n_unrecognised) *n_unrecognised = unknown;
  return AMMSB_OK;
}
