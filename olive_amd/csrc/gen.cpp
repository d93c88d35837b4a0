// Seeded synthetic graph generators (host side of libolive_hip.so).
//
// The reference has no generators: it reads edge-list / DIMACS files only
// (csrGraph.h:106-236).  These produce the same CSR layout that
// Oliver::readGraph consumes (row offsets `vertices[n+1]`, column ids
// `edges[m]`, csrGraph.h:61-63), with each row sorted by destination and
// duplicates kept.
//
// Every random draw is a pure function of (seed, draw index) through the
// SplitMix64 finaliser, so the output is identical for any thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "olive_hip.h"
#include "error.hpp"

namespace olv {

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

// Draw `idx` of stream `seed`.
static inline uint64_t draw(uint64_t seed, uint64_t idx) {
  return mix64(mix64(seed) + idx * kGolden);
}

int host_threads() {
  static int t = [] {
    if (const char *e = std::getenv("OLV_THREADS")) {
      int v = std::atoi(e);
      if (v > 0) return v;
    }
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 1;
    return (int)std::min(h, 16u);
  }();
  return t;
}

// Run fn(begin, end) over [0, total) split into contiguous chunks.
template <class Fn>
static void parallel_for(uint64_t total, Fn fn) {
  int nt = host_threads();
  if (total < 65536 || nt == 1) {
    fn((uint64_t)0, total);
    return;
  }
  std::vector<std::thread> th;
  uint64_t chunk = (total + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    uint64_t b = chunk * t, e = std::min(total, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] { fn(b, e); });
  }
  for (auto &x : th) x.join();
}

// COO (src,dst) -> CSR with rows sorted by dst.  Entries with keep[i]==0 are
// dropped (keep may be null).
static int coo_to_csr(uint32_t n, uint64_t m_in, const uint32_t *src,
                      const uint32_t *dst, bool drop_self_loops,
                      olv_host_csr *out) {
  std::vector<std::atomic<uint32_t>> deg(n);
  for (auto &d : deg) d.store(0, std::memory_order_relaxed);
  std::atomic<uint64_t> kept{0};
  parallel_for(m_in, [&](uint64_t b, uint64_t e) {
    uint64_t k = 0;
    for (uint64_t i = b; i < e; ++i) {
      if (drop_self_loops && src[i] == dst[i]) continue;
      deg[src[i]].fetch_add(1, std::memory_order_relaxed);
      ++k;
    }
    kept.fetch_add(k);
  });
  uint64_t m = kept.load();
  if (m >= (1ull << 32)) return fail(OLV_ERANGE, "edge count %llu exceeds uint32 EdgeId", (unsigned long long)m);
  uint32_t *row_ptr = (uint32_t *)std::malloc(sizeof(uint32_t) * ((size_t)n + 1));
  uint32_t *col = (uint32_t *)std::malloc(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
  if (!row_ptr || !col) {
    std::free(row_ptr);
    std::free(col);
    return fail(OLV_ENOMEM, "host allocation of CSR (%llu edges) failed", (unsigned long long)m);
  }
  uint64_t acc = 0;
  for (uint32_t v = 0; v < n; ++v) {
    row_ptr[v] = (uint32_t)acc;
    acc += deg[v].load(std::memory_order_relaxed);
  }
  row_ptr[n] = (uint32_t)acc;
  // scatter with per-row atomic cursors (order inside a row is fixed by the
  // sort below: rows of an unweighted multigraph are multisets of ids)
  for (uint32_t v = 0; v < n; ++v) deg[v].store(row_ptr[v], std::memory_order_relaxed);
  parallel_for(m_in, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      if (drop_self_loops && src[i] == dst[i]) continue;
      uint32_t pos = deg[src[i]].fetch_add(1, std::memory_order_relaxed);
      col[pos] = dst[i];
    }
  });
  parallel_for(n, [&](uint64_t b, uint64_t e) {
    for (uint64_t v = b; v < e; ++v) std::sort(col + row_ptr[v], col + row_ptr[v + 1]);
  });
  out->n = n;
  out->m = (uint32_t)m;
  out->row_ptr = row_ptr;
  out->col = col;
  out->weights = nullptr;
  out->weight_dtype = OLV_W_NONE;
  return OLV_OK;
}

}  // namespace olv

using namespace olv;

extern "C" {

void olv_host_csr_free(olv_host_csr *g) {
  if (!g) return;
  std::free(g->row_ptr);
  std::free(g->col);
  std::free(g->weights);
  g->row_ptr = nullptr;
  g->col = nullptr;
  g->weights = nullptr;
  g->n = g->m = 0;
  g->weight_dtype = OLV_W_NONE;
}

int olv_gen_grid(uint32_t rows, uint32_t cols, olv_host_csr *out) {
  if (!out) return fail(OLV_EINVAL, "olv_gen_grid: out is NULL");
  uint64_t n64 = (uint64_t)rows * cols;
  if (rows == 0 || cols == 0) return fail(OLV_EINVAL, "olv_gen_grid: empty grid");
  if (n64 >= (1ull << 32)) return fail(OLV_ERANGE, "olv_gen_grid: too many vertices");
  uint64_t m = 2ull * ((uint64_t)rows * (cols - 1) + (uint64_t)cols * (rows - 1));
  if (m >= (1ull << 32)) return fail(OLV_ERANGE, "olv_gen_grid: too many edges");
  uint32_t n = (uint32_t)n64;
  uint32_t *row_ptr = (uint32_t *)std::malloc(sizeof(uint32_t) * ((size_t)n + 1));
  uint32_t *col = (uint32_t *)std::malloc(sizeof(uint32_t) * std::max<uint64_t>(m, 1));
  if (!row_ptr || !col) {
    std::free(row_ptr);
    std::free(col);
    return fail(OLV_ENOMEM, "olv_gen_grid: host allocation failed");
  }
  uint32_t e = 0;
  for (uint32_t r = 0; r < rows; ++r)
    for (uint32_t c = 0; c < cols; ++c) {
      uint32_t id = r * cols + c;
      row_ptr[id] = e;
      // ascending ids: up, left, right, down
      if (r > 0) col[e++] = id - cols;
      if (c > 0) col[e++] = id - 1;
      if (c + 1 < cols) col[e++] = id + 1;
      if (r + 1 < rows) col[e++] = id + cols;
    }
  row_ptr[n] = e;
  out->n = n;
  out->m = e;
  out->row_ptr = row_ptr;
  out->col = col;
  out->weights = nullptr;
  out->weight_dtype = OLV_W_NONE;
  return OLV_OK;
}

int olv_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                 double c, uint64_t seed, int drop_self_loops,
                 olv_host_csr *out) {
  if (!out) return fail(OLV_EINVAL, "olv_gen_rmat: out is NULL");
  if (scale == 0 || scale > 31) return fail(OLV_EINVAL, "olv_gen_rmat: scale %u out of range", scale);
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0))
    return fail(OLV_EINVAL, "olv_gen_rmat: bad probabilities");
  uint64_t n = 1ull << scale;
  uint64_t m_gen = n * edge_factor;
  if (m_gen >= (1ull << 32)) return fail(OLV_ERANGE, "olv_gen_rmat: %llu edges exceed uint32", (unsigned long long)m_gen);
  // quadrant thresholds on a 32-bit uniform: q = #thresholds <= r
  auto thr = [](double p) -> uint64_t {
    double x = std::floor(p * 4294967296.0);
    return x >= 4294967296.0 ? 4294967296ull : (uint64_t)x;
  };
  const uint64_t ta = thr(a), tab = thr(a + b), tabc = thr(a + b + c);
  std::vector<uint32_t> src(m_gen), dst(m_gen);
  const uint32_t words = (scale + 1) / 2;  // one 64-bit draw per two levels
  parallel_for(m_gen, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) {
      uint32_t u = 0, v = 0;
      for (uint32_t w = 0; w < words; ++w) {
        uint64_t z = draw(seed, i * words + w);
        for (int half = 0; half < 2; ++half) {
          uint32_t level = 2 * w + half;
          if (level >= scale) break;
          uint64_t r = half == 0 ? (z >> 32) : (z & 0xffffffffull);
          uint32_t q = (uint32_t)(r >= ta) + (uint32_t)(r >= tab) + (uint32_t)(r >= tabc);
          uint32_t bit = scale - 1 - level;  // most significant bit first
          u |= (q >> 1) << bit;
          v |= (q & 1u) << bit;
        }
      }
      src[i] = u;
      dst[i] = v;
    }
  });
  return coo_to_csr((uint32_t)n, m_gen, src.data(), dst.data(), drop_self_loops != 0, out);
}

int olv_gen_uniform(uint32_t n, uint32_t m, uint64_t seed, olv_host_csr *out) {
  if (!out) return fail(OLV_EINVAL, "olv_gen_uniform: out is NULL");
  if (n == 0) return fail(OLV_EINVAL, "olv_gen_uniform: n == 0");
  std::vector<uint32_t> src(m), dst(m);
  parallel_for(m, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) {
      uint64_t z = draw(seed, i);
      src[i] = (uint32_t)(((z >> 32) * (uint64_t)n) >> 32);
      dst[i] = (uint32_t)(((z & 0xffffffffull) * (uint64_t)n) >> 32);
    }
  });
  return coo_to_csr(n, m, src.data(), dst.data(), false, out);
}

int olv_gen_weights_i32(olv_host_csr *g, int32_t lo, int32_t hi, uint64_t seed) {
  if (!g || !g->row_ptr) return fail(OLV_EINVAL, "olv_gen_weights_i32: no graph");
  if (hi < lo) return fail(OLV_EINVAL, "olv_gen_weights_i32: hi < lo");
  int32_t *w = (int32_t *)std::malloc(sizeof(int32_t) * std::max<uint32_t>(g->m, 1));
  if (!w) return fail(OLV_ENOMEM, "olv_gen_weights_i32: host allocation failed");
  uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo + 1);
  parallel_for(g->m, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      uint64_t z = draw(seed ^ 0x5eed0001ull, i);
      w[i] = (int32_t)((int64_t)lo + (int64_t)(((z >> 32) * range) >> 32));
    }
  });
  std::free(g->weights);
  g->weights = w;
  g->weight_dtype = OLV_W_I32;
  return OLV_OK;
}

int olv_gen_weights_f32_normalized(olv_host_csr *g, uint64_t seed) {
  if (!g || !g->row_ptr) return fail(OLV_EINVAL, "olv_gen_weights_f32_normalized: no graph");
  float *w = (float *)std::malloc(sizeof(float) * std::max<uint32_t>(g->m, 1));
  if (!w) return fail(OLV_ENOMEM, "olv_gen_weights_f32_normalized: host allocation failed");
  parallel_for(g->n, [&](uint64_t b, uint64_t e) {
    std::vector<double> raw;
    for (uint64_t v = b; v < e; ++v) {
      uint32_t s = g->row_ptr[v], t = g->row_ptr[v + 1];
      raw.resize(t - s);
      double sum = 0.0;
      for (uint32_t k = s; k < t; ++k) {
        uint64_t z = draw(seed ^ 0x5eed0002ull, k);
        double x = (double)((z >> 40) + 1) * (1.0 / 16777216.0);  // (0, 1]
        raw[k - s] = x;
        sum += x;
      }
      for (uint32_t k = s; k < t; ++k) w[k] = (float)(raw[k - s] / sum);
    }
  });
  std::free(g->weights);
  g->weights = w;
  g->weight_dtype = OLV_W_F32;
  return OLV_OK;
}

int olv_gen_dirichlet(uint32_t n, uint64_t seed, float *out) {
  if (!out || n == 0) return fail(OLV_EINVAL, "olv_gen_dirichlet: bad arguments");
  std::vector<double> e(n);
  parallel_for(n, [&](uint64_t b, uint64_t f) {
    for (uint64_t i = b; i < f; ++i) {
      uint64_t z = draw(seed ^ 0x5eed0003ull, i);
      double u = (double)((z >> 11) + 1) * (1.0 / 9007199254740992.0);  // (0, 1]
      e[i] = -std::log(u);
    }
  });
  double sum = 0.0;
  for (uint32_t i = 0; i < n; ++i) sum += e[i];
  for (uint32_t i = 0; i < n; ++i) out[i] = (float)(e[i] / sum);
  return OLV_OK;
}

uint32_t olv_gen_batch_scale(uint64_t seed, uint32_t i, uint32_t lo, uint32_t hi) {
  if (hi < lo) return lo;
  uint64_t z = draw(seed ^ 0x5eed0004ull, i);
  return lo + (uint32_t)(((z >> 32) * (uint64_t)(hi - lo + 1)) >> 32);
}

uint64_t olv_gen_batch_seed(uint64_t seed, uint32_t i) {
  return draw(seed ^ 0x5eed0005ull, i);
}

}  // extern "C"
