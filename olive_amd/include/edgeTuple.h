// COO edge tuple of the Olive C++ API (edgeTuple.h:42-67 of the reference).
#ifndef OLIVE_EDGE_TUPLE_H
#define OLIVE_EDGE_TUPLE_H

#include "common.h"

template <typename EdgeValue>
class EdgeTuple {
 public:
  VertexId srcId;
  VertexId dstId;
  EdgeValue value;
  EdgeTuple(VertexId src, VertexId dst, EdgeValue v) : srcId(src), dstId(dst), value(v) {}
};

template <typename EdgeTuple>
bool edgeTupleSrcCompare(EdgeTuple a, EdgeTuple b) {
  return a.srcId < b.srcId;
}

template <typename EdgeTuple>
bool edgeTupleDstCompare(EdgeTuple a, EdgeTuple b) {
  return a.dstId < b.dstId;
}

#endif  // OLIVE_EDGE_TUPLE_H
