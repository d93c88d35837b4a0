// Tuning of a graph handle: the run-time launch shapes and thresholds of the
// kernel families (olv_tuning), a search over them on the graph itself, a
// signature of the graph, and a JSON Lines cache keyed by that signature.
//
// The reference has no counterpart: Olive's programs pass a hand-picked
// GroupSize template argument per run (BFS.cu:127-135).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace olv {

int bfs_multisource(const olv_graph *g, const uint32_t *seeds, uint32_t k, bool seeds_host, int32_t *levels,
                    unsigned flags, olv_stats *st);

namespace {

void to_c(const Tuning &t, olv_tuning *o) {
  o->bfs_expand_per_cu = t.bfs_expand_per_cu;
  o->bfs_expand_per_cu_twin = t.bfs_expand_per_cu_twin;
  o->bfs_compact_per_cu = t.bfs_compact_per_cu;
  o->td_wave_thr = t.td_wave_thr;
  o->td_block_thr = t.td_block_thr;
  o->hub_thr = t.hub_thr;
  o->alpha = t.alpha;
  o->beta = t.beta;
  o->mbfs_pull_per_cu = t.mbfs_pull_per_cu;
  o->seg_grid_per_cu = t.seg_grid_per_cu;
  o->seg_unroll = t.seg_unroll;
  for (int i = 0; i < 3; ++i) o->seg_bounds[i] = t.seg_bounds[i];
  o->sssp_relax_per_cu = t.sssp_relax_per_cu;
  o->sssp_delta_num = t.sssp_delta_num;
  o->source = t.source;
  o->tune_ms = t.tune_ms;
}

int from_c(const olv_tuning *o, Tuning &t) {
  auto bad = [](const char *f) { return fail(OLV_EINVAL, "olv_tuning: %s out of range", f); };
  if (o->bfs_expand_per_cu < 1 || o->bfs_expand_per_cu > 64) return bad("bfs_expand_per_cu");
  if (o->bfs_expand_per_cu_twin < 1 || o->bfs_expand_per_cu_twin > 64) return bad("bfs_expand_per_cu_twin");
  if (o->bfs_compact_per_cu < 1 || o->bfs_compact_per_cu > 64) return bad("bfs_compact_per_cu");
  if (o->td_wave_thr < 2 || o->td_block_thr <= o->td_wave_thr) return bad("td_wave_thr / td_block_thr");
  if (o->hub_thr < kHubThrMin || o->hub_thr > (1u << 24)) return bad("hub_thr");
  if (!(o->alpha > 0.f) || !(o->beta > 0.f)) return bad("alpha / beta");
  if (o->mbfs_pull_per_cu < 1 || o->mbfs_pull_per_cu > 256) return bad("mbfs_pull_per_cu");
  if (o->seg_grid_per_cu > 256) return bad("seg_grid_per_cu");
  if (o->seg_unroll != 4 && o->seg_unroll != 8 && o->seg_unroll != 16) return bad("seg_unroll");
  if (o->seg_bounds[0] < 1 || o->seg_bounds[1] <= o->seg_bounds[0] || o->seg_bounds[2] <= o->seg_bounds[1] ||
      o->seg_bounds[2] > 8192)
    return bad("seg_bounds");
  if (o->sssp_relax_per_cu < 1 || o->sssp_relax_per_cu > 64) return bad("sssp_relax_per_cu");
  if (o->sssp_delta_num < 1 || o->sssp_delta_num > 1024) return bad("sssp_delta_num");
  t.bfs_expand_per_cu = o->bfs_expand_per_cu;
  t.bfs_expand_per_cu_twin = o->bfs_expand_per_cu_twin;
  t.bfs_compact_per_cu = o->bfs_compact_per_cu;
  t.td_wave_thr = o->td_wave_thr;
  t.td_block_thr = o->td_block_thr;
  t.hub_thr = o->hub_thr;
  t.alpha = o->alpha;
  t.beta = o->beta;
  t.mbfs_pull_per_cu = o->mbfs_pull_per_cu;
  t.seg_grid_per_cu = o->seg_grid_per_cu;
  t.seg_unroll = o->seg_unroll;
  for (int i = 0; i < 3; ++i) t.seg_bounds[i] = o->seg_bounds[i];
  t.sssp_relax_per_cu = o->sssp_relax_per_cu;
  t.sssp_delta_num = o->sssp_delta_num;
  return OLV_OK;
}

// FNV-1a over the vertex and edge counts and the log2 degree histograms.
std::string signature_of(uint32_t n, uint32_t m, const std::vector<uint32_t> &outdeg_hist,
                         const std::vector<uint32_t> &indeg_hist) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t x) {
    for (int b = 0; b < 8; ++b) {
      h ^= (x >> (8 * b)) & 0xffu;
      h *= 1099511628211ull;
    }
  };
  mix(n);
  mix(m);
  for (uint32_t c : outdeg_hist) mix(c);
  for (uint32_t c : indeg_hist) mix(c);
  char buf[96];
  std::snprintf(buf, sizeof(buf), "v%u-e%u-h%016llx", n, m, (unsigned long long)h);
  return buf;
}

inline uint32_t log2_bin(uint32_t d) { return d ? 32u - (uint32_t)__builtin_clz(d) : 0u; }  // 0..32

std::vector<uint32_t> hist_of_rp(const std::vector<uint32_t> &rp) {
  std::vector<uint32_t> h(33, 0);
  for (size_t v = 0; v + 1 < rp.size(); ++v) h[log2_bin(rp[v + 1] - rp[v])]++;
  return h;
}

// Fields of one cache line, in the order they are written.
struct Field {
  const char *name;
  bool is_float;
};
constexpr Field kFields[] = {
    {"bfs_expand_per_cu", false}, {"bfs_expand_per_cu_twin", false}, {"bfs_compact_per_cu", false},
    {"td_wave_thr", false},       {"td_block_thr", false},           {"hub_thr", false},
    {"alpha", true},              {"beta", true},                    {"mbfs_pull_per_cu", false},
    {"seg_grid_per_cu", false},   {"seg_unroll", false},             {"seg_bound0", false},
    {"seg_bound1", false},        {"seg_bound2", false},             {"sssp_relax_per_cu", false},
    {"sssp_delta_num", false},
};

void get_fields(const olv_tuning &t, double *v) {
  const double f[] = {(double)t.bfs_expand_per_cu, (double)t.bfs_expand_per_cu_twin, (double)t.bfs_compact_per_cu,
                      (double)t.td_wave_thr,       (double)t.td_block_thr,           (double)t.hub_thr,
                      (double)t.alpha,             (double)t.beta,                   (double)t.mbfs_pull_per_cu,
                      (double)t.seg_grid_per_cu,   (double)t.seg_unroll,             (double)t.seg_bounds[0],
                      (double)t.seg_bounds[1],     (double)t.seg_bounds[2],          (double)t.sssp_relax_per_cu,
                      (double)t.sssp_delta_num};
  for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); ++i) v[i] = f[i];
}

void set_fields(olv_tuning &t, const double *v) {
  t.bfs_expand_per_cu = (uint32_t)v[0];
  t.bfs_expand_per_cu_twin = (uint32_t)v[1];
  t.bfs_compact_per_cu = (uint32_t)v[2];
  t.td_wave_thr = (uint32_t)v[3];
  t.td_block_thr = (uint32_t)v[4];
  t.hub_thr = (uint32_t)v[5];
  t.alpha = (float)v[6];
  t.beta = (float)v[7];
  t.mbfs_pull_per_cu = (uint32_t)v[8];
  t.seg_grid_per_cu = (uint32_t)v[9];
  t.seg_unroll = (uint32_t)v[10];
  t.seg_bounds[0] = (uint32_t)v[11];
  t.seg_bounds[1] = (uint32_t)v[12];
  t.seg_bounds[2] = (uint32_t)v[13];
  t.sssp_relax_per_cu = (uint32_t)v[14];
  t.sssp_delta_num = (uint32_t)v[15];
}

constexpr int kNumFields = sizeof(kFields) / sizeof(kFields[0]);

std::string line_key(const std::string &line) {
  const char *k = std::strstr(line.c_str(), "\"key\": \"");
  if (!k) return "";
  k += 8;
  const char *e = std::strchr(k, '"');
  return e ? std::string(k, e) : std::string();
}

std::string format_line(const char *key, const olv_tuning &t) {
  double v[kNumFields];
  get_fields(t, v);
  std::ostringstream os;
  os << "{\"key\": \"" << key << "\"";
  for (int i = 0; i < kNumFields; ++i) {
    char num[48];
    if (kFields[i].is_float) std::snprintf(num, sizeof(num), "%.9g", v[i]);
    else std::snprintf(num, sizeof(num), "%u", (uint32_t)v[i]);
    os << ", \"" << kFields[i].name << "\": " << num;
  }
  os << "}";
  return os.str();
}

// Parses one line written by format_line; false when a field is missing.
bool parse_line(const std::string &line, olv_tuning &t) {
  double v[kNumFields];
  for (int i = 0; i < kNumFields; ++i) {
    const std::string pat = std::string("\"") + kFields[i].name + "\": ";
    const char *p = std::strstr(line.c_str(), pat.c_str());
    if (!p) return false;
    char *end = nullptr;
    v[i] = std::strtod(p + pat.size(), &end);
    if (end == p + pat.size()) return false;
  }
  set_fields(t, v);
  return true;
}

int graph_signature(const olv_graph *g, std::string *key) {
  std::vector<uint32_t> rp((size_t)g->n + 1);
  OLV_HIP(hipMemcpy(rp.data(), g->row_ptr, rp.size() * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> ih;
  if (g->in_row_ptr) {
    std::vector<uint32_t> irp((size_t)g->n + 1);
    OLV_HIP(hipMemcpy(irp.data(), g->in_row_ptr, irp.size() * 4, hipMemcpyDeviceToHost));
    ih = hist_of_rp(irp);
  } else {
    std::vector<uint32_t> col(std::max<uint32_t>(g->m, 1)), indeg(std::max<uint32_t>(g->n, 1), 0);
    if (g->m) OLV_HIP(hipMemcpy(col.data(), g->col, (size_t)g->m * 4, hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < g->m; ++k) indeg[col[k]]++;
    ih.assign(33, 0);
    for (uint32_t v = 0; v < g->n; ++v) ih[log2_bin(indeg[v])]++;
  }
  *key = signature_of(g->n, g->m, hist_of_rp(rp), ih);
  return OLV_OK;
}

// ------------------------------------------------------------------ search
using Clock = std::chrono::steady_clock;

struct Searcher {
  olv_graph *g;
  std::vector<uint32_t> srcs;
  DevBuf lv;
  int32_t *levels = nullptr;

  // Device time of plain BFS from the sample sources (best of two runs each).
  int time_bfs(double *ms) {
    double tot = 0;
    for (uint32_t s : srcs) {
      double best = 1e300;
      for (int r = 0; r < 2; ++r) {
        olv_stats st;
        OLV_TRY(bfs_multisource(g, &s, 1, true, levels, 0, &st));
        best = std::min(best, st.elapsed_ms);
      }
      tot += best;
    }
    *ms = tot;
    return OLV_OK;
  }
  int time_batched(double *ms) {
    olv_stats st;
    double best = 1e300;
    for (int r = 0; r < 2; ++r) {
      OLV_TRY(olv_bfs_batched(g, srcs.data(), (uint32_t)srcs.size(), levels, 0, &st));
      best = std::min(best, st.elapsed_ms);
    }
    *ms = best;
    return OLV_OK;
  }
  int time_pull(double *ms) {
    olv_stats st;
    double best = 1e300;
    for (int r = 0; r < 2; ++r) {
      OLV_TRY(olv_pagerank_f32(g, 0.85f, 5, 0.0, reinterpret_cast<float *>(levels), OLV_PR_ROW_PULL | OLV_PROFILE,
                               &st));
      best = std::min(best, st.kernel_ms / std::max<uint32_t>(st.kernel_launches, 1));
    }
    *ms = best;
    return OLV_OK;
  }
  int time_sssp(double *ms, unsigned flags) {  // the first two sample sources
    double tot = 0;
    for (size_t i = 0; i < srcs.size() && i < 2; ++i) {
      double best = 1e300;
      for (int r = 0; r < 2; ++r) {
        olv_stats st;
        OLV_TRY(olv_sssp(g, srcs[i], nullptr, levels, flags, &st));
        best = std::min(best, st.elapsed_ms);
      }
      tot += best;
    }
    *ms = tot;
    return OLV_OK;
  }

  // Coordinate descent: for each parameter in turn, try its candidates with
  // the others fixed and keep the fastest; two passes.
  template <class T>
  struct Param {
    T *field;
    std::vector<T> cand;
  };
  template <class Timef>
  int descend(std::vector<Param<uint32_t>> ups, std::vector<Param<float>> fps, Timef timef) {
    double best;
    OLV_TRY(timef(&best));
    for (int pass = 0; pass < 2; ++pass) {
      for (auto &p : ups) {
        const uint32_t keep = *p.field;
        uint32_t arg = keep;
        for (uint32_t c : p.cand) {
          if (c == keep) continue;
          *p.field = c;
          if (!valid()) continue;
          double t;
          OLV_TRY(timef(&t));
          if (t < best * 0.99) {
            best = t;
            arg = c;
          }
        }
        *p.field = arg;
      }
      for (auto &p : fps) {
        const float keep = *p.field;
        float arg = keep;
        for (float c : p.cand) {
          if (c == keep) continue;
          *p.field = c;
          double t;
          OLV_TRY(timef(&t));
          if (t < best * 0.99) {
            best = t;
            arg = c;
          }
        }
        *p.field = arg;
      }
    }
    return OLV_OK;
  }
  bool valid() const {
    olv_tuning c;
    to_c(g->tune, &c);
    Tuning t;
    return from_c(&c, t) == OLV_OK;
  }
};

}  // namespace

// A graph never tuned takes its signature's entry in the cache named by
// OLV_TUNE_CACHE, if there is one (used when plans are captured).
int apply_cached_tuning(olv_graph *g) {
  if (g->tune.tuned) return OLV_OK;
  const char *path = std::getenv("OLV_TUNE_CACHE");
  if (!path || !*path) return OLV_OK;
  std::string key;
  OLV_TRY(graph_signature(g, &key));
  olv_tuning t;
  int found = 0;
  OLV_TRY(olv_tuning_cache_load(path, key.c_str(), &t, &found));
  if (!found) return OLV_OK;
  OLV_TRY(from_c(&t, g->tune));
  g->tune.source = 2;
  g->tune.tune_ms = 0;
  g->tune.tuned = true;
  return OLV_OK;
}

}  // namespace olv

using namespace olv;

extern "C" {

int olv_get_tuning_ex(const olv_graph *g, olv_tuning *out) {
  if (!g || !out) return fail(OLV_EINVAL, "olv_get_tuning_ex: NULL argument");
  to_c(g->tune, out);
  return OLV_OK;
}

int olv_set_tuning_ex(olv_graph *g, const olv_tuning *t) {
  if (!g || !t) return fail(OLV_EINVAL, "olv_set_tuning_ex: NULL argument");
  Tuning n = g->tune;
  OLV_TRY(from_c(t, n));
  n.source = 3;
  n.tune_ms = 0;
  n.tuned = true;
  g->tune = n;
  return OLV_OK;
}

int olv_csr_signature(uint32_t n, uint32_t m, const uint32_t *row_ptr, const uint32_t *col, char *key, size_t len) {
  if (!row_ptr || (m && !col) || !key || len < 2) return fail(OLV_EINVAL, "olv_csr_signature: bad arguments");
  std::vector<uint32_t> rp(row_ptr, row_ptr + (size_t)n + 1), indeg(std::max<uint32_t>(n, 1), 0);
  if (rp[n] != m) return fail(OLV_EINVAL, "olv_csr_signature: row_ptr[n] = %u, m = %u", rp[n], m);
  for (uint32_t k = 0; k < m; ++k) {
    if (col[k] >= n) return fail(OLV_EINVAL, "olv_csr_signature: entry %u out of range", col[k]);
    indeg[col[k]]++;
  }
  std::vector<uint32_t> ih(33, 0);
  for (uint32_t v = 0; v < n; ++v) ih[log2_bin(indeg[v])]++;
  const std::string s = signature_of(n, m, hist_of_rp(rp), ih);
  if (s.size() + 1 > len) return fail(OLV_ERANGE, "olv_csr_signature: key needs %zu bytes", s.size() + 1);
  std::memcpy(key, s.c_str(), s.size() + 1);
  return OLV_OK;
}

int olv_graph_signature(const olv_graph *g, char *key, size_t len) {
  if (!g || !key || len < 2) return fail(OLV_EINVAL, "olv_graph_signature: bad arguments");
  std::string s;
  OLV_TRY(graph_signature(g, &s));
  if (s.size() + 1 > len) return fail(OLV_ERANGE, "olv_graph_signature: key needs %zu bytes", s.size() + 1);
  std::memcpy(key, s.c_str(), s.size() + 1);
  return OLV_OK;
}

int olv_tuning_cache_load(const char *path, const char *key, olv_tuning *out, int *found) {
  if (!path || !key || !out || !found) return fail(OLV_EINVAL, "olv_tuning_cache_load: NULL argument");
  *found = 0;
  std::ifstream f(path);
  if (!f) return OLV_OK;  // no cache yet
  std::string line;
  while (std::getline(f, line)) {
    if (line_key(line) != key) continue;
    olv_tuning t;
    to_c(Tuning(), &t);
    if (!parse_line(line, t)) return fail(OLV_EINVAL, "olv_tuning_cache_load: malformed entry for %s in %s", key, path);
    Tuning check;
    OLV_TRY(from_c(&t, check));
    t.source = 2;
    t.tune_ms = 0;
    *out = t;
    *found = 1;  // the last line of a key wins (stores replace in place)
  }
  return OLV_OK;
}

int olv_tuning_cache_store(const char *path, const char *key, const olv_tuning *t) {
  if (!path || !key || !t || !*key || std::strchr(key, '"')) return fail(OLV_EINVAL, "olv_tuning_cache_store: bad arguments");
  Tuning check;
  OLV_TRY(from_c(t, check));
  std::vector<std::string> lines;
  {
    std::ifstream f(path);
    std::string line;
    while (f && std::getline(f, line))
      if (!line.empty() && line_key(line) != key) lines.push_back(line);
  }
  lines.push_back(format_line(key, *t));
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream o(tmp, std::ios::trunc);
    if (!o) return fail(OLV_EINVAL, "olv_tuning_cache_store: cannot write %s", tmp.c_str());
    for (const auto &l : lines) o << l << "\n";
    if (!o) return fail(OLV_EINVAL, "olv_tuning_cache_store: write to %s failed", tmp.c_str());
  }
  if (std::rename(tmp.c_str(), path) != 0) return fail(OLV_EINVAL, "olv_tuning_cache_store: cannot replace %s", path);
  return OLV_OK;
}

int olv_autotune_ex(olv_graph *g, unsigned what, const char *cache_path, olv_tuning *out) {
  if (!g) return fail(OLV_EINVAL, "olv_autotune_ex: NULL graph");
  if (what & ~OLV_TUNE_ALL) return fail(OLV_EINVAL, "olv_autotune_ex: unknown bits %#x", what);
  const auto t0 = Clock::now();
  std::string key;
  if (cache_path) {
    OLV_TRY(graph_signature(g, &key));
    olv_tuning t;
    int found = 0;
    OLV_TRY(olv_tuning_cache_load(cache_path, key.c_str(), &t, &found));
    if (found) {
      OLV_TRY(from_c(&t, g->tune));
      g->tune.source = 2;
      g->tune.tune_ms = 0;
      g->tune.tuned = true;
      if (out) to_c(g->tune, out);
      return OLV_OK;
    }
  }
  if (g->n == 0 || g->m == 0) {  // nothing to measure: the defaults stand
    g->tune.source = 1;
    g->tune.tuned = true;
    if (out) to_c(g->tune, out);
    return OLV_OK;
  }
  Searcher S;
  S.g = g;
  // sample sources: the largest out-degree vertex and vertices with out-edges
  // spread over the id range (64 for the batched search)
  std::vector<uint32_t> rp((size_t)g->n + 1);
  OLV_HIP(hipMemcpy(rp.data(), g->row_ptr, rp.size() * 4, hipMemcpyDeviceToHost));
  uint32_t hub = 0;
  for (uint32_t v = 0; v < g->n; ++v)
    if (rp[v + 1] - rp[v] > rp[hub + 1] - rp[hub]) hub = v;
  S.srcs.push_back(hub);
  for (uint32_t i = 1; i < 64 && S.srcs.size() < 64; ++i) {
    uint32_t v = (uint32_t)(((uint64_t)g->n * i) / 64);
    for (uint32_t t = 0; t < g->n && rp[v + 1] == rp[v]; ++t) v = (v + 1) % g->n;
    S.srcs.push_back(v);
  }
  const size_t out_words = std::max<size_t>((size_t)g->n * ((what & OLV_TUNE_BATCHED) ? S.srcs.size() : 1), g->n);
  OLV_TRY(S.lv.alloc(out_words * 4));
  S.levels = S.lv.as<int32_t>();
  Tuning &T = g->tune;
  if (what & OLV_TUNE_BFS) {
    std::vector<uint32_t> keep(S.srcs);
    S.srcs.resize(3);
    double warm;
    OLV_TRY(S.time_bfs(&warm));  // first calls build the twin and work buffers
    OLV_TRY(S.descend({{&T.bfs_expand_per_cu_twin, {4, 6, 8, 12}},
                       {&T.bfs_compact_per_cu, {4, 8, 16}},
                       {&T.td_wave_thr, {32, 64, 128}},
                       {&T.td_block_thr, {256, 512, 1024}},
                       {&T.hub_thr, {512, 1024, 2048}}},
                      {{&T.alpha, {6.f, 10.f, 14.f, 20.f, 30.f}}, {&T.beta, {12.f, 24.f, 48.f}}},
                      [&](double *ms) { return S.time_bfs(ms); }));
    S.srcs = keep;
  }
  if ((what & OLV_TUNE_BATCHED) && g->in_row_ptr) {
    double warm;
    OLV_TRY(S.time_batched(&warm));
    OLV_TRY(S.descend({{&T.mbfs_pull_per_cu, {24, 32, 48, 64, 96}}}, {}, [&](double *ms) { return S.time_batched(ms); }));
  }
  if (what & OLV_TUNE_PULL) {
    double warm;
    OLV_TRY(S.time_pull(&warm));
    OLV_TRY(S.descend({{&T.seg_grid_per_cu, {16, 32, 48}},
                       {&T.seg_unroll, {4, 8, 16}},
                       {&T.seg_bounds[0], {4, 8, 16}},
                       {&T.seg_bounds[1], {48, 96, 192}},
                       {&T.seg_bounds[2], {512, 1024, 2048}}},
                      {}, [&](double *ms) { return S.time_pull(ms); }));
  }
  if ((what & OLV_TUNE_SSSP) && g->wdtype == OLV_W_I32) {
    // the step on the default (binned) path; the grid on the direct path,
    // the only one it shapes
    double warm;
    OLV_TRY(S.time_sssp(&warm, 0u));
    OLV_TRY(S.descend({{&T.sssp_delta_num, {16, 24, 32, 48, 64}}}, {},
                      [&](double *ms) { return S.time_sssp(ms, 0u); }));
    OLV_TRY(S.time_sssp(&warm, OLV_SSSP_DIRECT));
    OLV_TRY(S.descend({{&T.sssp_relax_per_cu, {4, 6, 8, 12}}}, {},
                      [&](double *ms) { return S.time_sssp(ms, OLV_SSSP_DIRECT); }));
  }
  T.source = 1;
  T.tuned = true;
  T.tune_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
  if (cache_path) {
    olv_tuning c;
    to_c(T, &c);
    OLV_TRY(olv_tuning_cache_store(cache_path, key.c_str(), &c));
  }
  if (out) to_c(T, out);
  return OLV_OK;
}

int olv_autotune(olv_graph *g, uint32_t *thread_max, uint32_t *wave_max, uint32_t *block_threads) {
  OLV_TRY(olv_autotune_ex(g, OLV_TUNE_BFS, nullptr, nullptr));
  if (thread_max) *thread_max = g->tune.td_wave_thr - 1;
  if (wave_max) *wave_max = g->tune.td_block_thr - 1;
  if (block_threads) *block_threads = g->tune.block_threads;
  return OLV_OK;
}

}  // extern "C"
