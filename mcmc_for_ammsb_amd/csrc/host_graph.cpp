// Host-side graph tools: synthetic generators, symmetric CSR construction and
// the held-out / test split.
//
// The split follows the reference's Network::init_held_out_set /
// init_test_set (c++/src/mcmc/network.cc:503-614) and the non-link pair
// samplers (network.cc:677-719):
//   held_out_size = held_out_ratio * E (truncated), p = held_out_size / 2
//   held-out links    : p distinct edges drawn uniformly from the 2E directed
//                       adjacency entries (sample_random_edges, :448-501)
//   held-out non-links: p pairs a != b, not linked, not already held out
//   test links        : p edges not in held-out or test
//   test non-links    : p pairs not linked, not in held-out or test
// Every draw comes from the counter-based stream (tag TAG_SPLIT, iter =
// phase), so the split is a pure function of (graph, ratio, seed).

#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/ammsb.h"
#include "host_util.h"

namespace ammsb {

namespace {

// Build a symmetric CSR (sorted rows, no duplicates, no self loops) from an
// undirected pair list.
int build_csr(int32_t N, const std::vector<uint64_t>& pairs, int64_t** row_ptr_out,
              int32_t** cols_out, int64_t* E_out) {
  std::vector<int64_t> deg(N + 1, 0);
  for (uint64_t p : pairs) {
    const int32_t a = (int32_t)(p >> 32), b = (int32_t)(p & 0xFFFFFFFFu);
    if (a == b) continue;
    deg[a]++;
    deg[b]++;
  }
  std::vector<int64_t> start(N + 1, 0);
  for (int32_t i = 0; i < N; ++i) start[i + 1] = start[i] + deg[i];
  std::vector<int32_t> tmp(start[N]);
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  for (uint64_t p : pairs) {
    const int32_t a = (int32_t)(p >> 32), b = (int32_t)(p & 0xFFFFFFFFu);
    if (a == b) continue;
    tmp[fill[a]++] = b;
    tmp[fill[b]++] = a;
  }
  // sort + unique each row, in parallel
  std::vector<int64_t> newdeg(N, 0);
  parallel_for(
      N,
      [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i) {
          int32_t* b = tmp.data() + start[i];
          int32_t* e = tmp.data() + start[i + 1];
          std::sort(b, e);
          newdeg[i] = std::unique(b, e) - b;
        }
      },
      4096);
  int64_t* rp = (int64_t*)malloc(sizeof(int64_t) * (N + 1));
  if (!rp) return AMMSB_ERR_OOM;
  rp[0] = 0;
  for (int32_t i = 0; i < N; ++i) rp[i + 1] = rp[i] + newdeg[i];
  int32_t* cols = (int32_t*)malloc(sizeof(int32_t) * std::max<int64_t>(1, rp[N]));
  if (!cols) {
    free(rp);
    return AMMSB_ERR_OOM;
  }
  parallel_for(
      N,
      [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i)
          memcpy(cols + rp[i], tmp.data() + start[i], sizeof(int32_t) * newdeg[i]);
      },
      4096);
  *row_ptr_out = rp;
  *cols_out = cols;
  *E_out = rp[N] / 2;
  return AMMSB_OK;
}

inline uint64_t upair(int32_t a, int32_t b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

// Dirichlet(alpha * 1_K) row via log-gamma draws + log-sum-exp.
void dirichlet_row(Key key, double alpha, int K, uint32_t node, uint32_t iter,
                   double* out) {
  double mx = -INFINITY;
  for (int k = 0; k < K; ++k) {
    // one gamma stream per (node, k): id = node, iter encodes k
    out[k] = log_gamma_draw(key, alpha, node, iter + (uint32_t)k * 16u, TAG_GEN);
    mx = std::max(mx, out[k]);
  }
  double s = 0.0;
  for (int k = 0; k < K; ++k) {
    out[k] = std::exp(out[k] - mx);
    s += out[k];
  }
  for (int k = 0; k < K; ++k) out[k] /= s;
}

int gen_planted(const ammsb_gen_params* p, std::vector<uint64_t>* pairs) {
  const int N = p->N, K = p->K;
  const Key key = make_key(p->seed);
  std::vector<double> pi((size_t)N * K);
  parallel_for(N, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; ++i)
      dirichlet_row(key, p->gen_alpha, K, (uint32_t)i, 1000u, &pi[(size_t)i * K]);
  }, 64);
  std::vector<double> beta(K);
  for (int k = 0; k < K; ++k) {
    const double u = u01_from_u64(draw_u64(key, (uint32_t)k, 0, 2, TAG_GEN));
    beta[k] = p->beta_lo + (p->beta_hi - p->beta_lo) * u;
  }
  const int nt = host_threads();
  std::vector<std::vector<uint64_t>> local(nt);
  parallel_for(N, [&](int64_t lo, int64_t hi, int t) {
    auto& out = local[t];
    for (int64_t a = lo; a < hi; ++a) {
      const double* pa = &pi[(size_t)a * K];
      for (int64_t b = a + 1; b < N; ++b) {
        const double* pb = &pi[(size_t)b * K];
        double s = 0.0, tot = 0.0;
        for (int k = 0; k < K; ++k) {
          const double f = pa[k] * pb[k];
          s += f * beta[k];
          tot += f;
        }
        const double prob = s + p->delta * (1.0 - tot);
        const double u =
            u01_from_u64(draw_u64(key, (uint32_t)b, (uint32_t)a, 3, TAG_GEN));
        if (u < prob) out.push_back(upair((int32_t)a, (int32_t)b));
      }
    }
  }, 16);
  for (auto& v : local) pairs->insert(pairs->end(), v.begin(), v.end());
  return AMMSB_OK;
}

int gen_community(const ammsb_gen_params* p, std::vector<uint64_t>* pairs) {
  const int32_t N = p->N, K = p->K;
  const int maxc = std::max(1, std::min<int>(p->max_comms, K));
  const Key key = make_key(p->seed);
  // memberships: 1..maxc distinct communities per node
  std::vector<int32_t> comm((size_t)N * maxc, -1);
  std::vector<int32_t> ncomm(N);
  parallel_for(N, [&](int64_t lo, int64_t hi, int) {
    for (int64_t i = lo; i < hi; ++i) {
      const int c = 1 + (int)(draw_u64(key, 0, (uint32_t)i, 10, TAG_GEN) % maxc);
      int got = 0;
      for (uint32_t j = 1; got < c; ++j) {
        const int32_t k =
            (int32_t)(draw_u64(key, j, (uint32_t)i, 10, TAG_GEN) % (uint64_t)K);
        bool dup = false;
        for (int q = 0; q < got; ++q) dup |= comm[(size_t)i * maxc + q] == k;
        if (!dup) comm[(size_t)i * maxc + got++] = k;
      }
      ncomm[i] = c;
    }
  }, 4096);
  // members per community, ascending node id
  std::vector<int64_t> cstart(K + 1, 0);
  for (int32_t i = 0; i < N; ++i)
    for (int q = 0; q < ncomm[i]; ++q) cstart[comm[(size_t)i * maxc + q] + 1]++;
  for (int k = 0; k < K; ++k) cstart[k + 1] += cstart[k];
  std::vector<int32_t> members(cstart[K]);
  {
    std::vector<int64_t> f(cstart.begin(), cstart.end() - 1);
    for (int32_t i = 0; i < N; ++i)
      for (int q = 0; q < ncomm[i]; ++q) members[f[comm[(size_t)i * maxc + q]]++] = i;
  }
  double s2 = 0.0;
  for (int k = 0; k < K; ++k) {
    const double nk = (double)(cstart[k + 1] - cstart[k]);
    s2 += nk * (nk - 1.0);
  }
  const double prob = s2 > 0 ? std::min(1.0, (double)N * p->mean_degree / s2) : 0.0;
  std::vector<std::vector<uint64_t>> percomm(K);
  if (prob > 0.0) {
    const double lq = std::log1p(-std::min(prob, 1.0 - 1e-16));
    parallel_for(K, [&](int64_t lo, int64_t hi, int) {
      for (int64_t k = lo; k < hi; ++k) {
        const int64_t nk = cstart[k + 1] - cstart[k];
        const int32_t* mem = members.data() + cstart[k];
        const uint64_t P = (uint64_t)nk * (uint64_t)(nk - 1) / 2;
        auto& out = percomm[k];
        int64_t pos = -1;
        // geometric skipping over the lower-triangle pair index space
        for (uint64_t j = 0;; ++j) {
          const double u = u01_from_u64(
              draw_u64(key, (uint32_t)j, (uint32_t)k, 11u + (uint32_t)(j >> 32), TAG_GEN));
          const double skip = prob >= 1.0 ? 0.0 : std::floor(std::log(u) / lq);
          if (skip > 9e18) break;
          pos += 1 + (int64_t)skip;
          if (pos < 0 || (uint64_t)pos >= P) break;
          int64_t r = (int64_t)((1.0 + std::sqrt(1.0 + 8.0 * (double)pos)) / 2.0);
          while (r * (r - 1) / 2 > pos) --r;
          while ((r + 1) * r / 2 <= pos) ++r;
          const int64_t c = pos - r * (r - 1) / 2;
          out.push_back(upair(mem[r], mem[c]));
        }
      }
    }, 1);
  }
  for (auto& v : percomm) pairs->insert(pairs->end(), v.begin(), v.end());
  // background noise pairs
  const double npairs = (double)N * (double)(N - 1) / 2.0;
  const int64_t nbg = (int64_t)std::llround(p->delta * npairs);
  for (int64_t j = 0; j < nbg; ++j) {
    const u32x4 r = draw4(key, (uint32_t)j, (uint32_t)(j >> 32), 12, TAG_GEN);
    const int32_t a = (int32_t)(((uint64_t)r.x | ((uint64_t)r.y << 32)) % (uint64_t)N);
    const int32_t b = (int32_t)(((uint64_t)r.z | ((uint64_t)r.w << 32)) % (uint64_t)N);
    if (a != b) pairs->push_back(upair(a, b));
  }
  return AMMSB_OK;
}

}  // namespace

}  // namespace ammsb

using namespace ammsb;

extern "C" int ammsb_graph_generate(const ammsb_gen_params* p, int64_t** row_ptr,
                                    int32_t** cols, int64_t* num_edges) {
  if (!p || !row_ptr || !cols || !num_edges || p->N < 2 || p->K < 1) {
    set_error("ammsb_graph_generate: bad arguments");
    return AMMSB_ERR_ARG;
  }
  try {
    std::vector<uint64_t> pairs;
    int rc;
    if (p->kind == AMMSB_GEN_PLANTED) {
      if ((double)p->N * p->N * p->K > 5e11) {
        set_error("ammsb_graph_generate: PLANTED is O(N^2 K); N too large");
        return AMMSB_ERR_ARG;
      }
      rc = gen_planted(p, &pairs);
    } else if (p->kind == AMMSB_GEN_COMMUNITY) {
      rc = gen_community(p, &pairs);
    } else {
      set_error("ammsb_graph_generate: unknown kind");
      return AMMSB_ERR_ARG;
    }
    if (rc != AMMSB_OK) return rc;
    rc = build_csr(p->N, pairs, row_ptr, cols, num_edges);
    if (rc != AMMSB_OK) set_error("ammsb_graph_generate: out of memory");
    return rc;
  } catch (const std::bad_alloc&) {
    set_error("ammsb_graph_generate: out of memory");
    return AMMSB_ERR_OOM;
  }
}

extern "C" int ammsb_split_generate(const ammsb_graph* g, double ratio, uint64_t seed,
                                    int32_t** ho_edges, uint8_t** ho_labels,
                                    int64_t* ho_count, int32_t** te_edges,
                                    uint8_t** te_labels, int64_t* te_count) {
  if (!g || !g->row_ptr || !g->cols || g->N < 2 || !ho_edges || !ho_labels ||
      !ho_count || !te_edges || !te_labels || !te_count) {
    set_error("ammsb_split_generate: bad arguments");
    return AMMSB_ERR_ARG;
  }
  if (ratio == 0.0) ratio = 0.1;  // learner.cc:87-91
  const int64_t E = g->num_edges;
  const int64_t N = g->N;
  const int64_t hs = (int64_t)(ratio * (double)E);
  const int64_t p = hs / 2;
  if (E < 2 * p) {
    // network.cc:507-511 (held-out) and the test set need 2p distinct edges
    set_error("ammsb_split_generate: not enough linked edges; use a smaller held-out ratio");
    return AMMSB_ERR_ARG;
  }
  const double nonlinks = (double)N * (double)(N - 1) / 2.0 - (double)E;
  if (nonlinks < 2.0 * (double)p + 1.0) {
    set_error("ammsb_split_generate: not enough non-links for the split");
    return AMMSB_ERR_ARG;
  }
  try {
    const Key key = make_key(seed);
    EdgeHashHost chosen;
    chosen.init(4 * p + 16);
    std::vector<int32_t> ho, te;
    std::vector<uint8_t> hol, tel;
    ho.reserve(4 * p);
    te.reserve(4 * p);
    const uint64_t twoE = (uint64_t)g->row_ptr[N];
    auto edge_at = [&](uint64_t d, int32_t* a, int32_t* b) {
      const int64_t* it = std::upper_bound(g->row_ptr, g->row_ptr + N + 1, (int64_t)d);
      *a = (int32_t)(it - g->row_ptr - 1);
      *b = g->cols[d];
    };
    const uint64_t limit = 1000ull * (uint64_t)(p + 1000);
    // phase 0: held-out links
    uint64_t j = 0;
    for (int64_t got = 0; got < p; ++j) {
      if (j > limit) { set_error("split: attempt limit (held-out links)"); return AMMSB_ERR_INVARIANT; }
      int32_t a, b;
      edge_at(draw_u64(key, (uint32_t)j, (uint32_t)(j >> 32), 0, TAG_SPLIT) % twoE, &a, &b);
      if (a > b) std::swap(a, b);
      if (chosen.insert(edge_key(a, b), 1 | 4)) {
        ho.push_back(a); ho.push_back(b); hol.push_back(1);
        ++got;
      }
    }
    // phase 1: held-out non-links
    j = 0;
    for (int64_t got = 0; got < p; ++j) {
      if (j > limit) { set_error("split: attempt limit (held-out non-links)"); return AMMSB_ERR_INVARIANT; }
      const u32x4 r = draw4(key, (uint32_t)j, (uint32_t)(j >> 32), 1, TAG_SPLIT);
      int32_t a = (int32_t)(((uint64_t)r.x | ((uint64_t)r.y << 32)) % (uint64_t)N);
      int32_t b = (int32_t)(((uint64_t)r.z | ((uint64_t)r.w << 32)) % (uint64_t)N);
      if (a == b) continue;
      if (a > b) std::swap(a, b);
      if (csr_linked(g->row_ptr, g->cols, a, b)) continue;
      if (chosen.find(edge_key(a, b)) >= 0) continue;
      chosen.insert(edge_key(a, b), 1);
      ho.push_back(a); ho.push_back(b); hol.push_back(0);
      ++got;
    }
    // phase 2: test links.  init_test_set (network.cc:559-590) tests
    // `if (p < 0) break` after `p--`, so the edge that takes p from 0 to -1
    // is still added: the test set holds p + 1 links whenever the training
    // links allow it (its sampled batch of 2p edges nearly always does).
    j = 0;
    const int64_t test_links = E >= 2 * p + 1 ? p + 1 : p;
    for (int64_t got = 0; got < test_links; ++j) {
      if (j > limit) { set_error("split: attempt limit (test links)"); return AMMSB_ERR_INVARIANT; }
      int32_t a, b;
      edge_at(draw_u64(key, (uint32_t)j, (uint32_t)(j >> 32), 2, TAG_SPLIT) % twoE, &a, &b);
      if (a > b) std::swap(a, b);
      if (chosen.find(edge_key(a, b)) >= 0) continue;
      chosen.insert(edge_key(a, b), 2 | 4);
      te.push_back(a); te.push_back(b); tel.push_back(1);
      ++got;
    }
    // phase 3: test non-links
    j = 0;
    for (int64_t got = 0; got < p; ++j) {
      if (j > limit) { set_error("split: attempt limit (test non-links)"); return AMMSB_ERR_INVARIANT; }
      const u32x4 r = draw4(key, (uint32_t)j, (uint32_t)(j >> 32), 3, TAG_SPLIT);
      int32_t a = (int32_t)(((uint64_t)r.x | ((uint64_t)r.y << 32)) % (uint64_t)N);
      int32_t b = (int32_t)(((uint64_t)r.z | ((uint64_t)r.w << 32)) % (uint64_t)N);
      if (a == b) continue;
      if (a > b) std::swap(a, b);
      if (csr_linked(g->row_ptr, g->cols, a, b)) continue;
      if (chosen.find(edge_key(a, b)) >= 0) continue;
      chosen.insert(edge_key(a, b), 2);
      te.push_back(a); te.push_back(b); tel.push_back(0);
      ++got;
    }
    auto dup = [](const void* src, size_t bytes) -> void* {
      void* d = malloc(bytes ? bytes : 1);
      if (d && bytes) memcpy(d, src, bytes);
      return d;
    };
    *ho_edges = (int32_t*)dup(ho.data(), ho.size() * sizeof(int32_t));
    *ho_labels = (uint8_t*)dup(hol.data(), hol.size());
    *te_edges = (int32_t*)dup(te.data(), te.size() * sizeof(int32_t));
    *te_labels = (uint8_t*)dup(tel.data(), tel.size());
    *ho_count = (int64_t)hol.size();
    *te_count = (int64_t)tel.size();
    if (!*ho_edges || !*ho_labels || !*te_edges || !*te_labels) {
      set_error("ammsb_split_generate: out of memory");
      return AMMSB_ERR_OOM;
    }
    return AMMSB_OK;
  } catch (const std::bad_alloc&) {
    set_error("ammsb_split_generate: out of memory");
    return AMMSB_ERR_OOM;
  }
}

// SNAP edge list, the reference's Relativity reader (relativity.cc:20-193):
// the first 4 lines are the header; every further line holds two integer
// ids.  contiguous: ids are used as they are and N = max id + 1
// (relativity.cc:64-104); else ids are renumbered 0..N-1 in order of first
// appearance (the reference keeps its unordered_set's order,
// relativity.cc:106-150, which no other build reproduces).  Self-links and
// repeated pairs are dropped (relativity.cc:152-172).  gzip files are read
// through zlib (the reference's .gz inputs, relativity.cc:24-46).
extern "C" int ammsb_graph_load_edge_list(const char* path, int32_t contiguous, int32_t* N_out,
                                          int64_t* num_edges, int64_t** row_ptr_out,
                                          int32_t** cols_out, int64_t* self_links,
                                          int64_t* duplicates) {
  if (!path || !N_out || !num_edges || !row_ptr_out || !cols_out) {
    set_error("ammsb_graph_load_edge_list: null argument");
    return AMMSB_ERR_ARG;
  }
  *row_ptr_out = nullptr;
  *cols_out = nullptr;
  gzFile f = gzopen(path, "rb");
  if (!f) {
    set_error(std::string("Cannot open ") + path);
    return AMMSB_ERR_ARG;
  }
  try {
    std::vector<int64_t> raw;  // pairs of ids as read
    std::string line;
    char buf[4096];
    int64_t lineno = 0;
    bool more = true;
    while (more) {
      line.clear();
      while (true) {  // one line of any length
        if (!gzgets(f, buf, (int)sizeof buf)) {
          more = false;
          break;
        }
        line += buf;
        if (!line.empty() && line.back() == '\n') break;
      }
      if (!more && line.empty()) break;
      ++lineno;
      if (lineno <= 4) continue;  // header (relativity.cc:51-55)
      const char* p = line.c_str();
      char* end = nullptr;
      const long long a = strtoll(p, &end, 10);
      if (end == p) {
        gzclose(f);
        set_error("Fail to parse int (line " + std::to_string(lineno) + " of " + path + ")");
        return AMMSB_ERR_ARG;
      }
      p = end;
      const long long b = strtoll(p, &end, 10);
      if (end == p) {
        gzclose(f);
        set_error("Fail to parse int (line " + std::to_string(lineno) + " of " + path + ")");
        return AMMSB_ERR_ARG;
      }
      raw.push_back(a);
      raw.push_back(b);
    }
    gzclose(f);
    const int64_t npairs = (int64_t)raw.size() / 2;
    int64_t N = 0;
    if (contiguous) {
      long long mx = -1;
      for (int64_t i = 0; i < 2 * npairs; ++i) {
        if (raw[i] < 0) {
          set_error("ammsb_graph_load_edge_list: negative id in contiguous input");
          return AMMSB_ERR_ARG;
        }
        mx = std::max<long long>(mx, raw[i]);
      }
      N = mx + 1;
    } else {
      std::unordered_map<int64_t, int64_t> ids;
      ids.reserve((size_t)(2 * npairs));
      for (int64_t i = 0; i < 2 * npairs; ++i) {
        auto it = ids.find(raw[i]);
        if (it == ids.end()) it = ids.emplace(raw[i], (int64_t)ids.size()).first;
        raw[i] = it->second;
      }
      N = (int64_t)ids.size();
    }
    if (N < 1 || N > INT32_MAX) {
      set_error("ammsb_graph_load_edge_list: node count out of range");
      return AMMSB_ERR_ARG;
    }
    std::vector<std::pair<int32_t, int32_t>> e;
    e.reserve((size_t)npairs);
    int64_t self = 0;
    for (int64_t i = 0; i < npairs; ++i) {
      const int32_t a = (int32_t)raw[2 * i], b = (int32_t)raw[2 * i + 1];
      if (a == b) {
        ++self;
        continue;
      }
      e.emplace_back(std::min(a, b), std::max(a, b));
    }
    std::vector<int64_t>().swap(raw);
    std::sort(e.begin(), e.end());
    const int64_t before = (int64_t)e.size();
    e.erase(std::unique(e.begin(), e.end()), e.end());
    const int64_t E = (int64_t)e.size();
    int64_t* rp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(N + 1));
    int32_t* cols = (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, 2 * E));
    if (!rp || !cols) {
      free(rp);
      free(cols);
      set_error("ammsb_graph_load_edge_list: out of memory");
      return AMMSB_ERR_OOM;
    }
    std::fill(rp, rp + N + 1, 0);
    for (const auto& x : e) {
      rp[x.first + 1]++;
      rp[x.second + 1]++;
    }
    for (int64_t i = 0; i < N; ++i) rp[i + 1] += rp[i];
    std::vector<int64_t> fill(rp, rp + N);
    for (const auto& x : e) {
      cols[fill[x.first]++] = x.second;
      cols[fill[x.second]++] = x.first;
    }
    for (int64_t i = 0; i < N; ++i) std::sort(cols + rp[i], cols + rp[i + 1]);
    *N_out = (int32_t)N;
    *num_edges = E;
    *row_ptr_out = rp;
    *cols_out = cols;
    if (self_links) *self_links = self;
    if (duplicates) *duplicates = before - E;
    return AMMSB_OK;
  } catch (const std::bad_alloc&) {
    gzclose(f);
    set_error("ammsb_graph_load_edge_list: out of memory");
    return AMMSB_ERR_OOM;
  }
}

extern "C" void ammsb_free(void* p) { free(p); }

extern "C" void ammsb_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  u32x4 c;
  c.x = ctr[0];
  c.y = ctr[1];
  c.z = ctr[2];
  c.w = ctr[3];
  const u32x4 r = philox4x32_10(c, key[0], key[1]);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = r.w;
}
