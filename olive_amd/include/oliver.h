// `Oliver`, the class name the reference's applications instantiate
// (BFS.cu:103, SSSP.cu:125, PageRank.cu:110; oliver.h:46-49), is the same
// engine as Olive.
#ifndef OLIVER_H
#define OLIVER_H

#include "olive.h"

template <typename VertexValue, typename EdgeValue, typename AccumValue>
using Oliver = Olive<VertexValue, EdgeValue, AccumValue>;

#endif  // OLIVER_H
