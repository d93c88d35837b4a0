// Frontier Bellman-Ford SSSP written against the Olive C++ API, on the
// int32 edge weights of the input file (third column of an edge list;
// 1 where a line has none).
//
//   olive_sssp <graph> [-s 0] [-dimacs] [-o dist.txt]
// Unreached vertices print as 2147483647; exit status 4 on a negative cycle.
#include <climits>

#include "oliver.h"

struct Dist {
  int d;
};

struct Weight {
  int w;
};

struct Relax {
  __device__ inline int gather(Dist src, EdgeId, Weight e) { return src.d + e.w; }
  __device__ inline void reduce(int &acc, int x) { atomicMin(&acc, x); }
};

struct Improved {
  __device__ inline bool cond(Dist v, int acc) { return acc < v.d; }
  __device__ inline void update(Dist &v, int acc) { v.d = acc; }
};

// edgeMap functor that writes weight 1 into every edge (SSSP.cu:90-103).
struct UnitWeight {
  __device__ inline int gather(Dist, EdgeId, Weight &e) {
    e.w = 1;
    return 0;
  }
  __device__ inline void reduce(int &, int) {}
};

struct SetDist {
  int d;
  explicit SetDist(int x) : d(x) {}
  __device__ inline void operator()(Dist &v, int) { v.d = d; }
};

int main(int argc, char **argv) {
  CommandLine cl(argc, argv, "<graph> [-s 0] [-dimacs] [-o dist.txt]");
  const char *path = cl.getArgument(0);
  const VertexId source = cl.getOptionIntValue("-s", 0);
  const char *out = cl.getOptionStringValue("-o", "olive_sssp.txt");

  CsrGraph<int, int> graph;
  if (cl.getOption("-dimacs"))
    graph.fromDimacsFile(path);
  else
    graph.fromEdgeListFile(path);
  if (graph.vertexCount == 0 || source >= graph.vertexCount) {
    fprintf(stderr, "source %u out of range\n", source);
    return 2;
  }
  if (!graph.hasEdgeValues) LOG(INFO) << "no weights in " << path << ": unit weights";

  Oliver<Dist, Weight, int> ol(INT_MAX);
  ol.readGraph(graph);
  if (!graph.hasEdgeValues) {
    VertexSubset all(graph.vertexCount, true);
    ol.edgeMap(all, UnitWeight());
    all.del();
  }
  VertexSubset all(graph.vertexCount, true);
  ol.vertexMap(all, SetDist(INT_MAX));
  all.del();

  VertexSubset frontier(graph.vertexCount, source);
  ol.vertexMap(frontier, SetDist(0));
  VertexSubset touched(graph.vertexCount, false);

  Stopwatch sw;
  sw.start();
  int rounds = 0;
  bool cycle = false;
  while (frontier.size() > 0) {
    if (rounds == (int)graph.vertexCount) {
      cycle = true;
      break;
    }
    ol.edgeFilter(touched, frontier, Relax());
    ol.vertexFilter(frontier, touched, Improved());
    rounds++;
  }
  LOG(INFO) << "rounds: " << rounds << ", time: " << sw.getElapsedMillis() << "ms";
  if (cycle) {
    LOG(INFO) << "negative cycle";
    return 4;
  }

  // "%d\n" per vertex, as SSSP.cu:41-43 prints, formatted on the device
  ol.writeVertices(out, offsetof(Dist, d), OLV_TXT_I32);
  frontier.del();
  touched.del();
  return 0;
}
