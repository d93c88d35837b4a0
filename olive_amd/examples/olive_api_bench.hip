// The reference BFS program's level loop (BFS.cu:124-148: edgeFilter, then
// vertexFilter, until the frontier is empty) run through the Olive C++ API
// on a generated graph, beside the engine's own BFS (olv_bfs, the C ABI fast
// path) on the same graph and sources.  Prints one JSON line: mean time per
// traversal of both paths, their ratio, and whether every traversal gave the
// same levels on both paths.
//
//   olive_api_bench rmat <scale> <sources> [seed]
//   olive_api_bench grid <rows> <cols>
#include <chrono>
#include <cstdlib>
#include <vector>

#include "oliver.h"

struct BfsVertex {
  int level;
};

struct PushLevel {  // BFS.cu:46-57: gather level + 1, store it
  __device__ inline int gather(BfsVertex src, EdgeId, Dump_Edge) { return src.level + 1; }
  __device__ inline void reduce(int &acc, int x) { acc = x; }
};

struct KeepUnreached {  // BFS.cu:59-71: update level, keep the unreached
  int inf;
  explicit KeepUnreached(int i) : inf(i) {}
  __device__ inline void update(BfsVertex &v, int acc) { v.level = acc; }
  __device__ inline bool cond(BfsVertex v, int) { return v.level == inf; }
};

struct SetLevel {  // BFS.cu:73-80
  int level;
  explicit SetLevel(int l) : level(l) {}
  __device__ inline void operator()(BfsVertex &v, int) { v.level = level; }
};

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s rmat <scale> <sources> [seed] | grid <rows> <cols>\n", argv[0]);
    return 2;
  }
  const std::string kind = argv[1];
  olv_host_csr h;
  uint32_t nsrc = 1;
  if (kind == "rmat") {
    const uint32_t scale = (uint32_t)atoi(argv[2]);
    nsrc = (uint32_t)atoi(argv[3]);
    const uint64_t seed = argc > 4 ? strtoull(argv[4], nullptr, 10) : 1;
    OLV_CHECK(olv_gen_rmat(scale, 16, 0.57, 0.19, 0.19, seed, 1, &h));
  } else {
    OLV_CHECK(olv_gen_grid((uint32_t)atoi(argv[2]), (uint32_t)atoi(argv[3]), &h));
  }
  CsrGraph<int, int> graph;
  graph.initGraph(h.n, h.m);
  std::copy(h.row_ptr, h.row_ptr + h.n + 1, graph.vertices);
  std::copy(h.col, h.col + h.m, graph.edges);
  // sources: vertex 0 for the grid; for RMAT, vertices with out-edges
  // spread over the id range
  std::vector<VertexId> srcs;
  if (kind == "grid") {
    srcs.push_back(0);
  } else {
    for (uint32_t i = 0; i < nsrc; ++i) {
      VertexId v = (VertexId)(((uint64_t)h.n * (2 * i + 1)) / (2 * nsrc));
      while (h.row_ptr[v + 1] == h.row_ptr[v]) v = (v + 1) % h.n;
      srcs.push_back(v);
    }
  }
  const int inf = 0x7fffffff;

  Oliver<BfsVertex, Dump_Edge, int> ol;
  ol.readGraph(graph);
  olv_graph *g = nullptr;
  OLV_CHECK(olv_graph_create(&g, h.n, h.m, h.row_ptr, h.col, nullptr, OLV_W_NONE, OLV_PTR_HOST | OLV_BUILD_TRANSPOSE));
  int32_t *d_levels = nullptr;
  OLV_CHECK(olv_malloc(reinterpret_cast<void **>(&d_levels), (size_t)h.n * 4));
  std::vector<BfsVertex> api(h.n);
  std::vector<int32_t> fast(h.n);

  double api_ms = 0, fast_ms = 0;
  bool exact = true;
  long levels_total = 0;
  for (int rep = -1; rep < (int)srcs.size(); ++rep) {  // rep -1: warm-up on the first source
    const VertexId s = srcs[rep < 0 ? 0 : rep];
    VertexSubset all(graph.vertexCount, true);
    ol.vertexMap<SetLevel>(all, SetLevel(inf));
    all.del();
    VertexSubset frontier(graph.vertexCount, s);
    ol.vertexMap<SetLevel>(frontier, SetLevel(0));
    VertexSubset edgeFrontier(graph.vertexCount, false);
    OLV_CHECK(olv_synchronize());
    const double t0 = now_ms();
    int iterations = 0;
    while (true) {  // BFS.cu:124-148 without the round cap
      const int size = frontier.size();
      ol.edgeFilter<PushLevel>(edgeFrontier, frontier, PushLevel());
      ol.vertexFilter<KeepUnreached>(frontier, edgeFrontier, KeepUnreached(inf));
      if (size == 0) break;
      iterations++;
    }
    OLV_CHECK(olv_synchronize());
    const double t1 = now_ms();
    frontier.del();
    edgeFrontier.del();
    // the engine's own BFS from the same source (wall clock, stream drained)
    OLV_CHECK(olv_bfs(g, s, d_levels, 0, nullptr));
    OLV_CHECK(olv_synchronize());
    const double t2 = now_ms();
    OLV_CHECK(olv_bfs(g, s, d_levels, 0, nullptr));
    OLV_CHECK(olv_synchronize());
    const double t3 = now_ms();
    if (rep < 0) continue;
    api_ms += t1 - t0;
    fast_ms += t3 - t2;
    levels_total += iterations;
    ol.copyVertices(api.data());
    OLV_CHECK(olv_memcpy_d2h(fast.data(), d_levels, (size_t)h.n * 4));
    for (uint32_t v = 0; v < h.n; ++v)
      if (api[v].level != fast[v]) {
        exact = false;
        break;
      }
  }
  const double k = (double)srcs.size();
  printf("{\"graph\": \"%s\", \"vertices\": %u, \"edges\": %u, \"sources\": %u, \"api_ms_mean\": %.4f, "
         "\"fast_path_ms_mean\": %.4f, \"api_over_fast\": %.3f, \"levels_mean\": %.2f, \"bit_exact\": %s}\n",
         kind == "rmat" ? ("rmat" + std::string(argv[2])).c_str() : ("grid" + std::string(argv[2]) + "x" + argv[3]).c_str(),
         h.n, h.m, (uint32_t)srcs.size(), api_ms / k, fast_ms / k, (api_ms / k) / (fast_ms / k), levels_total / k,
         exact ? "true" : "false");
  olv_free(d_levels);
  olv_graph_destroy(g);
  olv_host_csr_free(&h);
  return exact ? 0 : 1;
}
