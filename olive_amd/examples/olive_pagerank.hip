// fp64 PageRank written against the Olive C++ API: edgeMap scatters
// rank/outdeg into the accumulators, vertexMap applies the damping, and
// vertexReduce sums |delta| on the host.  Stops after `-max` updates or
// when the L1 change drops below `-eps` (the rule of testPageRank.cu:75-104).
//
//   olive_pagerank <graph> [-max 20] [-eps 1e-7] [-dimacs] [-o ranks.txt]
#include "oliver.h"

struct Rank {
  double rank;
  double delta;
  void reduce(double &r) { r += fabs(delta); }
};

// |delta| of a vertex, for the device reduction of the convergence check
struct AbsDelta {
  __device__ inline double operator()(const Rank &r) const { return fabs(r.delta); }
};

struct Scatter {
  __device__ inline double gather(Rank src, EdgeId outdeg, Dump_Edge) { return src.rank / outdeg; }
  __device__ inline void reduce(double &acc, double x) { atomicAdd(&acc, x); }
};

struct Apply {
  double d, base;
  Apply(double damping, double n) : d(damping), base((1.0 - damping) / n) {}
  __device__ inline void operator()(Rank &v, double acc) {
    const double r = d * acc + base;
    v.delta = r - v.rank;
    v.rank = r;
  }
};

struct Init {
  double r;
  explicit Init(double x) : r(x) {}
  __device__ inline void operator()(Rank &v, double) {
    v.rank = r;
    v.delta = r;
  }
};

int main(int argc, char **argv) {
  CommandLine cl(argc, argv, "<graph> [-max 20] [-eps 1e-7] [-dimacs] [-o ranks.txt]");
  const char *path = cl.getArgument(0);
  const int max_updates = cl.getOptionIntValue("-max", 20);
  const double eps = cl.getOptionDoubleValue("-eps", 1e-7);
  const char *out = cl.getOptionStringValue("-o", "olive_pagerank.txt");

  CsrGraph<int, int> graph;
  if (cl.getOption("-dimacs"))
    graph.fromDimacsFile(path);
  else
    graph.fromEdgeListFile(path);
  if (graph.vertexCount == 0) return 2;

  Oliver<Rank, Dump_Edge, double> ol;
  ol.readGraph(graph);
  VertexSubset all(graph.vertexCount, true);
  ol.vertexMap(all, Init(1.0 / graph.vertexCount));

  Stopwatch sw;
  sw.start();
  int updates = 0;
  double err = 0.0;
  while (updates < max_updates) {
    ol.edgeMap(all, Scatter());
    ol.vertexMap(all, Apply(0.85, (double)graph.vertexCount));
    updates++;
    err = ol.vertexReduce(AbsDelta());  // on the device; PageRank.cu:131 reduces on the host
    if (err < eps) break;
  }
  LOG(INFO) << "updates: " << updates << ", err: " << err << ", time: " << sw.getElapsedMillis() << "ms";

  FILE *f = fopen(out, "w");
  if (!f) return 3;
  std::vector<Rank> rv(graph.vertexCount);
  ol.copyVertices(rv.data());
  for (VertexId v = 0; v < graph.vertexCount; v++) fprintf(f, "%.17g\n", rv[v].rank);
  fclose(f);
  all.del();
  return 0;
}
