// CSR graph and file loaders of the Olive C++ API (csrGraph.h:54-287 of the
// reference).  Layout as the reference: `vertices[n+1]` row offsets and
// `edges[m]` column ids (csrGraph.h:61-63).
//
// Differences, all deliberate:
//  * edge-list input need not be sorted: tuples are stable-sorted by source
//    (the sort is commented out at csrGraph.h:149, which corrupts unsorted
//    files such as data/disorderedGraph_9);
//  * a third column is kept as the edge value (dropped at csrGraph.h:139);
//  * arrays are released with delete[] (csrGraph.h:70-75 uses delete).
#ifndef OLIVE_CSR_GRAPH_H
#define OLIVE_CSR_GRAPH_H

#include <algorithm>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

#include "common.h"
#include "edgeTuple.h"
#include "logging.h"
#include "timer.h"
#include "utils.h"

template <typename VertexValue, typename EdgeValue>
class CsrGraph {
 public:
  VertexId vertexCount;
  EdgeId edgeCount;
  EdgeId *vertices;
  VertexId *edges;
  EdgeValue *edgeValues;
  VertexValue *vertexValues;
  bool hasEdgeValues;

  CsrGraph()
      : vertexCount(0), edgeCount(0), vertices(NULL), edges(NULL), edgeValues(NULL), vertexValues(NULL),
        hasEdgeValues(false) {}
  CsrGraph(const CsrGraph &) = delete;
  CsrGraph &operator=(const CsrGraph &) = delete;
  ~CsrGraph() { release(); }

  void release() {
    delete[] vertices;
    delete[] edges;
    delete[] edgeValues;
    delete[] vertexValues;
    vertices = NULL;
    edges = NULL;
    edgeValues = NULL;
    vertexValues = NULL;
    vertexCount = 0;
    edgeCount = 0;
    hasEdgeValues = false;
  }

  void initGraph(VertexId n, EdgeId m) {
    release();
    vertexCount = n;
    edgeCount = m;
    vertices = new EdgeId[(size_t)n + 1];
    edges = new VertexId[m ? m : 1];
    edgeValues = new EdgeValue[m ? m : 1]();
    vertexValues = new VertexValue[n ? n : 1]();
  }

  /**
   * Edge-list file: lines starting with '#' are comments; the first other
   * line is "n m"; then m lines "src dst [value]".  The file is parsed and
   * stable-sorted by source on the GPU (olv_graph_load_file) and the CSR
   * copied into this host graph; an unreadable or malformed file logs the
   * library's message and leaves the graph empty (csrGraph.h:107-111 of the
   * reference logs and returns too).
   */
  void fromEdgeListFile(const char *path) {
    Stopwatch sw;
    sw.start();
    olv_graph *g = NULL;
    if (olv_graph_load_file(&g, path, OLV_FMT_EDGE_LIST, 0, NULL) != OLV_OK) {
      LOG(ERROR) << olv_last_error();
      return;
    }
    uint32_t n = 0, m = 0;
    OLV_CHECK(olv_graph_counts(g, &n, &m));
    initGraph((VertexId)n, (EdgeId)m);
    std::vector<int32_t> w(m ? m : 1);
    int wd = OLV_W_NONE;
    OLV_CHECK(olv_graph_csr_to_host(g, reinterpret_cast<uint32_t *>(vertices), reinterpret_cast<uint32_t *>(edges),
                                    w.data(), &wd));
    OLV_CHECK(olv_graph_destroy(g));
    hasEdgeValues = std::is_arithmetic<EdgeValue>::value && wd == OLV_W_I32 && m > 0;
    if constexpr (std::is_arithmetic<EdgeValue>::value)
      if (hasEdgeValues)
        for (uint32_t e = 0; e < m; ++e) edgeValues[e] = (EdgeValue)w[e];
    LOG(INFO) << "Parsed " << path << ": " << n << " vertices, " << m << " edges in " << sw.getElapsedMillis()
              << "ms";
  }

  /**
   * DIMACS/METIS adjacency file: header "n m" (m undirected edges, 2m
   * arcs), then line i lists the 1-based neighbours of vertex i-1.  Parsed
   * on the GPU (olv_graph_load_file); a malformed file (the neighbour count
   * is not 2m: csrGraph.h:229-232 of the reference asserts) logs the
   * library's message and leaves the graph empty.
   */
  void fromDimacsFile(const char *path) {
    olv_graph *g = NULL;
    if (olv_graph_load_file(&g, path, OLV_FMT_DIMACS, 0, NULL) != OLV_OK) {
      LOG(ERROR) << olv_last_error();
      return;
    }
    uint32_t n = 0, m = 0;
    OLV_CHECK(olv_graph_counts(g, &n, &m));
    initGraph((VertexId)n, (EdgeId)m);
    OLV_CHECK(olv_graph_csr_to_host(g, reinterpret_cast<uint32_t *>(vertices), reinterpret_cast<uint32_t *>(edges),
                                    NULL, NULL));
    OLV_CHECK(olv_graph_destroy(g));
  }

  /** Degree histogram (and the adjacency when verbose). */
  void print(bool verbose = false) const {
    if (verbose) {
      for (VertexId v = 0; v < vertexCount; v++) {
        printf("[%u] ", v);
        for (EdgeId e = vertices[v]; e < vertices[v + 1]; e++) printf(" ->%u", edges[e]);
        printf("\n");
      }
    }
    size_t degLog[34] = {0};
    int slotMax = 0;
    for (VertexId v = 0; v < vertexCount; v++) {
      EdgeId deg = vertices[v + 1] - vertices[v];
      int slot = 0;
      while (deg > 0) {
        deg /= 2;
        slot++;
      }
      slotMax = std::max(slotMax, slot);
      degLog[slot]++;
    }
    printf("deg[0]: 0\t%zu\t%.2f%%\n", degLog[0], (float)degLog[0] / vertexCount * 100);
    for (int i = 1; i <= slotMax; i++)
      printf("deg[%d]: (%d, %d]\t%zu\t%.2f%%\n", i, i >= 2 ? 1 << (i - 2) : 0, 1 << (i - 1), degLog[i],
             (float)degLog[i] / vertexCount * 100);
  }
};

#endif  // OLIVE_CSR_GRAPH_H
