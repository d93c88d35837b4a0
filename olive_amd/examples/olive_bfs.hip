// Level-synchronous BFS written against the Olive C++ API: the frontier is
// a dense subset (queue), edgeFilter pushes level+1 to the neighbours and
// marks them, vertexFilter keeps the ones not yet reached.
//
//   olive_bfs <graph> [-s 0] [-dimacs] [-o levels.txt]
// Unreached vertices print as -1.
#include "oliver.h"

struct Level {
  int level;
};

struct PushLevel {
  __device__ inline int gather(Level src, EdgeId, Dump_Edge) { return src.level + 1; }
  // Every writer of one level stores the same value.
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

int main(int argc, char **argv) {
  CommandLine cl(argc, argv, "<graph> [-s 0] [-dimacs] [-o levels.txt]");
  const char *path = cl.getArgument(0);
  const VertexId source = cl.getOptionIntValue("-s", 0);
  const char *out = cl.getOptionStringValue("-o", "olive_bfs.txt");

  CsrGraph<int, int> graph;
  if (cl.getOption("-dimacs"))
    graph.fromDimacsFile(path);
  else
    graph.fromEdgeListFile(path);
  if (graph.vertexCount == 0 || source >= graph.vertexCount) {
    fprintf(stderr, "source %u out of range\n", source);
    return 2;
  }

  Oliver<Level, Dump_Edge, int> ol;
  ol.readGraph(graph);
  VertexSubset all(graph.vertexCount, true);
  ol.vertexMap(all, SetLevel(-1));
  all.del();

  VertexSubset frontier(graph.vertexCount, source);
  ol.vertexMap(frontier, SetLevel(0));
  VertexSubset touched(graph.vertexCount, false);

  Stopwatch sw;
  sw.start();
  int levels = 0;
  while (frontier.size() > 0) {
    ol.edgeFilter(touched, frontier, PushLevel());
    ol.vertexFilter(frontier, touched, KeepUnreached());
    levels++;
  }
  LOG(INFO) << "levels: " << levels << ", time: " << sw.getElapsedMillis() << "ms";

  // "%d\n" per vertex, as BFS.cu:41-43 prints, formatted on the device
  ol.writeVertices(out, offsetof(Level, level), OLV_TXT_I32);
  frontier.del();
  touched.del();
  return 0;
}
