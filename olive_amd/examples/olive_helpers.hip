// The engine helpers of the Olive C++ API on a BFS: per level, the size of
// the dense frontier and of the sparse set of touched vertices
// (getWorkqueueSize, counted on the device), the stop test
// allVerticesInactive, and at the end a device vertexReduce and a host
// vertexTransform over the levels.
//
//   olive_helpers <graph> [-s 0] [-dimacs]
// Prints "level L frontier F touched T" per level, then
// "reached R levelsum S maxlevel M".
#include <cstdio>

#include "oliver.h"

struct Level {
  int level;
};

struct PushLevel {
  __device__ inline int gather(Level src, EdgeId, Dump_Edge) { return src.level + 1; }
  __device__ inline void reduce(int &acc, int x) { acc = x; }
};

struct KeepUnreached {
  __device__ inline bool cond(Level v, int) { return v.level < 0; }
  __device__ inline void update(Level &v, int acc) { v.level = acc; }
};

struct SetLevel {
  int level;
  explicit SetLevel(int l) : level(l) {}
  __device__ inline void operator()(Level &v, int) { v.level = level; }
};

// level of a reached vertex, 0 otherwise (summed on the device)
struct ReachedLevel {
  __device__ inline int operator()(const Level &v) const { return v.level > 0 ? v.level : 0; }
};

int main(int argc, char **argv) {
  CommandLine cl(argc, argv, "<graph> [-s 0] [-dimacs]");
  const char *path = cl.getArgument(0);
  const VertexId source = cl.getOptionIntValue("-s", 0);
  CsrGraph<int, int> graph;
  if (cl.getOption("-dimacs"))
    graph.fromDimacsFile(path);
  else
    graph.fromEdgeListFile(path);
  if (graph.vertexCount == 0 || source >= graph.vertexCount) return 2;

  Oliver<Level, Dump_Edge, int> ol;
  ol.readGraph(graph);
  VertexSubset all(graph.vertexCount, true);
  ol.vertexMap(all, SetLevel(-1));
  if (ol.getWorkqueueSize(all) != graph.vertexCount) return 4;
  all.del();

  VertexSubset frontier(graph.vertexCount, source);
  ol.vertexMap(frontier, SetLevel(0));
  VertexSubset touched(graph.vertexCount, false);
  int level = 0;
  while (!ol.allVerticesInactive(frontier)) {
    const VertexId f = ol.getWorkqueueSize(frontier);
    ol.edgeFilter(touched, frontier, PushLevel());
    printf("level %d frontier %u touched %u\n", level, f, ol.getWorkqueueSize(touched));
    ol.vertexFilter(frontier, touched, KeepUnreached());
    level++;
  }
  const int levelsum = ol.vertexReduce(ReachedLevel());
  long long reached = 0;
  int maxlevel = 0;
  ol.vertexTransform([&](VertexId, const Level &v) {
    if (v.level >= 0) {
      reached++;
      maxlevel = v.level > maxlevel ? v.level : maxlevel;
    }
  });
  printf("reached %lld levelsum %d maxlevel %d\n", reached, levelsum, maxlevel);
  frontier.del();
  touched.del();
  return 0;
}
