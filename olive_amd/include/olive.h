// The Olive C++ API on MI355X: Olive<VertexValue, EdgeValue, AccumValue>
// with edgeFilter / vertexFilter / edgeMap / vertexMap / vertexReduce /
// readGraph, the interface of the reference's single-GPU engine
// (oliver.h:46-249, documented as `Olive` in README.md:44-97).
//
// Differences from the reference, all deliberate:
//  * edgeFilter / edgeMap schedule a frontier by out-degree bucket (one lane,
//    one wave64 or one workgroup per vertex) instead of a fixed GroupSize;
//    GroupSize is still accepted (and defaults to 0) so callers that pass it
//    (BFS.cu:127-135) or omit it (SSSP.cu:148) both compile.  UseScan
//    defaults to true for the same reason (oliver.h:92 has no default).
//  * vertexFilter hands cond() the vertex's accumulator, as update() gets
//    it; the reference passes the vertex id there (oliverKernel.h:112,135),
//    under which SSSP.cu's `v.distance > accum` compares against the id.
//  * readGraph accepts graphs with zero edges (grd.h:77 asserts), and copies
//    the loaded int edge values into EdgeValue when EdgeValue is a 4-byte
//    non-empty type (oliver.h:190-202 leaves them uninitialised).
//  * Every call is ordered on libolive_hip's stream; host reads (size(),
//    vertexReduce, printVertices) synchronise that stream instead of a
//    device-wide synchronise after every call (oliver.h:82,112,140,169).
#ifndef OLIVE_H
#define OLIVE_H

#include <algorithm>
#include <type_traits>

#include "commandLine.h"
#include "common.h"
#include "csrGraph.h"
#include "grd.h"
#include "logging.h"
#include "oliveKernel.h"
#include "timer.h"
#include "utils.h"
#include "vertexSubset.h"

template <typename VertexValue, typename EdgeValue, typename AccumValue>
class Olive {
 public:
  /**
   * Push every out-edge of the queued vertices of `src` (a dense subset):
   * acc[v] = reduce(acc[v], gather(val[u], outdeg(u), edge)), and set v in
   * the sparse subset `dst`.  Accumulators start at the default value.
   */
  template <typename F, int GroupSize = 0>
  void edgeFilter(VertexSubset &dst, const VertexSubset &src, F f) {
    assert(!dst.isDense);
    assert(src.isDense);
    dst.clear();
    accumulators.allTo(defaultAccumValue);
    expand<F, true>(src.queue, src.qSizeDevice, dst.flags.words, f);
  }

  /**
   * For every v of the sparse subset `src` with cond(val[v], acc[v]):
   * update(val[v], acc[v]) and append v to the dense subset `dst`.
   */
  template <typename F, bool UseScan = true>
  void vertexFilter(VertexSubset &dst, const VertexSubset &src, F f) {
    assert(!src.isDense);
    assert(dst.isDense);
    dst.clear();
    const uint32_t nw = src.flags.nwords;
    if (!nw) return;
    hipLaunchKernelGGL((olive_detail::filter_kernel<VertexValue, AccumValue, F>), dim3(grid(nw)),
                       dim3(olive_detail::kBlock), 0, stream(), src.flags.words, nw, vertexValues.elemsDevice,
                       accumulators.elemsDevice, dst.queue, dst.qSizeDevice, f);
    HIP_CHECK(hipGetLastError());
  }

  /**
   * Push every out-edge of the vertices of the sparse subset `src` (no
   * output subset).  The edge value reaches gather() by reference, so a
   * gather may write it (SSSP.cu:96-99).
   */
  template <typename F>
  void edgeMap(const VertexSubset &src, F f) {
    assert(!src.isDense);
    accumulators.allTo(defaultAccumValue);
    OLV_CHECK(olv_bitmap_to_queue(src.flags.words, vertexCount, mapQueue, mapCount));
    expand<F, false>(mapQueue, mapCount, NULL, f);
  }

  /** f(val[v], acc[v]) for every v of `src` (either representation). */
  template <typename F>
  void vertexMap(const VertexSubset &src, F f) {
    if (src.isDense) {
      hipLaunchKernelGGL((olive_detail::map_queue_kernel<VertexValue, AccumValue, F>), dim3(grid(vertexCount)),
                         dim3(olive_detail::kBlock), 0, stream(), src.queue, src.qSizeDevice,
                         vertexValues.elemsDevice, accumulators.elemsDevice, f);
    } else {
      const uint32_t nw = src.flags.nwords;
      if (!nw) return;
      hipLaunchKernelGGL((olive_detail::map_words_kernel<VertexValue, AccumValue, F>), dim3(grid(nw)),
                         dim3(olive_detail::kBlock), 0, stream(), src.flags.words, nw, vertexValues.elemsDevice,
                         accumulators.elemsDevice, f);
    }
    HIP_CHECK(hipGetLastError());
  }

  /** Host reduction of VertexValue::reduce over all vertices. */
  AccumValue vertexReduce() {
    vertexValues.persist();
    AccumValue r = (AccumValue)0;
    for (VertexId i = 0; i < vertexCount; i++) vertexValues.elemsHost[i].reduce(r);
    return r;
  }

  /**
   * Device map-reduce: the sum over all vertices of f(val[v]), where f is a
   * __device__ functor returning AccumValue.  The reduction runs on the GPU
   * and only the result crosses to the host (the host vertexReduce() above
   * copies every vertex value back, oliver.h:176-183).  The summation order
   * is fixed, so repeated calls give identical results.
   */
  template <typename F>
  AccumValue vertexReduce(F f) {
    if (!vertexCount) return AccumValue(0);
    const unsigned k = reduceGrid(vertexCount);
    AccumValue *part = reinterpret_cast<AccumValue *>(reduceScratch);
    hipLaunchKernelGGL((olive_detail::reduce_kernel<VertexValue, AccumValue, F>), dim3(k),
                       dim3(olive_detail::kBlock), 0, stream(), vertexValues.elemsDevice, vertexCount, f, part);
    hipLaunchKernelGGL(olive_detail::reduce_final_kernel<AccumValue>, dim3(1), dim3(olive_detail::kBlock), 0,
                       stream(), part, k, part + k);
    HIP_CHECK(hipGetLastError());
    AccumValue r;
    OLV_CHECK(olv_memcpy_d2h(&r, part + k, sizeof(AccumValue)));
    return r;
  }

  /**
   * Number of vertices in `s`: the queue length of a dense subset, the
   * population count (on the device) of a sparse one.  Replaces the
   * multi-GPU engine's getWorkqueueSize (olive.h:363-375), which copies
   * and adds one counter per partition.
   */
  VertexId getWorkqueueSize(const VertexSubset &s) {
    if (s.isDense) return s.size();
    const uint32_t nw = s.flags.nwords;
    if (!nw) return 0;
    const unsigned k = reduceGrid(nw);
    unsigned long long *part = reinterpret_cast<unsigned long long *>(reduceScratch);
    hipLaunchKernelGGL(olive_detail::popcount_kernel, dim3(k), dim3(olive_detail::kBlock), 0, stream(),
                       s.flags.words, nw, part);
    hipLaunchKernelGGL(olive_detail::reduce_final_kernel<unsigned long long>, dim3(1), dim3(olive_detail::kBlock), 0,
                       stream(), part, k, part + k);
    HIP_CHECK(hipGetLastError());
    unsigned long long c = 0;
    OLV_CHECK(olv_memcpy_d2h(&c, part + k, sizeof(c)));
    return (VertexId)c;
  }

  /** Whether `s` holds no vertex (olive.h:392-403, one flag per partition
   *  there; one device count here). */
  bool allVerticesInactive(const VertexSubset &s) { return getWorkqueueSize(s) == 0; }

  /**
   * f(globalId, value) for every vertex on the host, after copying the
   * values back (olive.h:414-426).  On one GPU the global id of a vertex is
   * its index.
   */
  template <typename F>
  void vertexTransform(F f) {
    vertexValues.persist();
    for (VertexId v = 0; v < vertexCount; v++) f(v, vertexValues.elemsHost[v]);
  }

  Olive(AccumValue accum) : defaultAccumValue(accum) { init(); }
  Olive() : defaultAccumValue(0) { init(); }

  /** Upload the CSR of `graph` and allocate the vertex and edge state. */
  void readGraph(const CsrGraph<int, int> &graph) {
    release();
    vertexCount = graph.vertexCount;
    edgeCount = graph.edgeCount;
    OLV_CHECK(olv_graph_create(&g, vertexCount, edgeCount, graph.vertices, graph.edges, NULL, OLV_W_NONE,
                               OLV_PTR_HOST));
    const void *w = NULL;
    OLV_CHECK(olv_graph_device_arrays(g, &rowPtr, &colIdx, &w, NULL, NULL, NULL));
    vertexValues.reserve(vertexCount);
    accumulators.reserve(vertexCount);
    vertexValues.clear();
    accumulators.clear();
    const bool perEdge = !std::is_empty<EdgeValue>::value;
    edgeValues.reserve(perEdge ? edgeCount : 1);
    edgeValues.clear();
    if (perEdge && sizeof(EdgeValue) == sizeof(int) && graph.hasEdgeValues && edgeCount) {
      memcpy(edgeValues.elemsHost, graph.edgeValues, sizeof(int) * (size_t)edgeCount);
      edgeValues.cache();
    }
    void *d = NULL;
    for (int b = 0; b < 3; b++) {
      OLV_CHECK(olv_malloc(&d, sizeof(uint32_t) * (size_t)(vertexCount ? vertexCount : 1)));
      buckets[b] = reinterpret_cast<uint32_t *>(d);
    }
    OLV_CHECK(olv_malloc(&d, 2 * sizeof(uint32_t) * olv_hub_capacity(edgeCount)));
    hubPairs = reinterpret_cast<uint32_t *>(d);
    OLV_CHECK(olv_malloc(&d, sizeof(uint32_t) * 4));
    bucketCounts = reinterpret_cast<uint32_t *>(d);
    OLV_CHECK(olv_malloc(&d, sizeof(uint32_t) * (size_t)(vertexCount ? vertexCount : 1)));
    mapQueue = reinterpret_cast<uint32_t *>(d);
    OLV_CHECK(olv_malloc(&d, sizeof(uint32_t)));
    mapCount = reinterpret_cast<uint32_t *>(d);
    const size_t slot = sizeof(AccumValue) > 8 ? sizeof(AccumValue) : 8;
    OLV_CHECK(olv_malloc(&reduceScratch, slot * ((size_t)cuCount * 4 + 1)));
  }

  inline void printVertices() {
    vertexValues.persist();
    vertexValues.print();
  }

  inline void printEdges() {
    edgeValues.persist();
    edgeValues.print();
  }

  /** The text printVertices() produces when VertexValue::print() is
   *  fprintf("%d\n" or "%f\n", field) (BFS.cu:41-43, PageRank.cu:42-44),
   *  formatted on the device: field_offset = offsetof(VertexValue, field),
   *  fmt = OLV_TXT_I32 / OLV_TXT_F32 / OLV_TXT_F64. */
  inline void writeVertices(const char *path, size_t field_offset, int fmt) {
    OLV_CHECK(olv_write_text(path, reinterpret_cast<const char *>(vertexValues.elemsDevice) + field_offset,
                             vertexCount, sizeof(VertexValue), fmt, 0, NULL));
  }

  /** Every vertex value as a binary dump (olv_write_binary). */
  inline void dumpVertices(const char *path) {
    OLV_CHECK(olv_write_binary(path, vertexValues.elemsDevice, vertexCount, sizeof(VertexValue), 0, 0));
  }

  /** Copy every vertex value to `out` (vertexCount elements, host). */
  inline void copyVertices(VertexValue *out) {
    vertexValues.persist();
    memcpy(out, vertexValues.elemsHost, sizeof(VertexValue) * (size_t)vertexCount);
  }

  inline VertexId getVertexCount() const { return vertexCount; }
  inline EdgeId getEdgeCount() const { return edgeCount; }

  /** The resident graph, for the C ABI algorithms (olv_bfs, olv_sssp ...). */
  inline olv_graph *graphHandle() const { return g; }

  ~Olive() { release(); }

 private:
  Olive(const Olive &) = delete;
  Olive &operator=(const Olive &) = delete;

  static hipStream_t stream() { return reinterpret_cast<hipStream_t>(olv_get_stream()); }

  unsigned reduceGrid(size_t items) const {
    const size_t want = (items + olive_detail::kBlock - 1) / olive_detail::kBlock;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(want, (size_t)cuCount * 4));
  }

  unsigned grid(size_t items) const {
    const size_t want = (items + olive_detail::kBlock - 1) / olive_detail::kBlock;
    return (unsigned)std::max<size_t>(1, std::min<size_t>(want, (size_t)cuCount * 8));
  }

  void init() {
    vertexCount = 0;
    edgeCount = 0;
    g = NULL;
    rowPtr = colIdx = NULL;
    buckets[0] = buckets[1] = buckets[2] = NULL;
    bucketCounts = mapQueue = mapCount = NULL;
    hubPairs = NULL;
    reduceScratch = NULL;
    OLV_CHECK(olv_device_info(&cuCount, NULL, NULL));
  }

  void release() {
    if (g) OLV_CHECK(olv_graph_destroy(g));
    g = NULL;
    for (int b = 0; b < 3; b++)
      if (buckets[b]) olv_free(buckets[b]);
    if (bucketCounts) olv_free(bucketCounts);
    if (mapQueue) olv_free(mapQueue);
    if (mapCount) olv_free(mapCount);
    if (reduceScratch) olv_free(reduceScratch);
    if (hubPairs) olv_free(hubPairs);
    hubPairs = NULL;
    buckets[0] = buckets[1] = buckets[2] = NULL;
    bucketCounts = mapQueue = mapCount = NULL;
    reduceScratch = NULL;
    vertexValues.del();
    accumulators.del();
    edgeValues.del();
  }

  // Split queue[0..*count) into low-degree entries and hub items, then push
  // both in one launch (olv::push_expand with the user's functors).  The
  // grid is sized from the CU count (the counts stay on the device); the
  // thresholds are the graph's tuning.
  template <typename F, bool MARK>
  void expand(const uint32_t *queue, const uint32_t *count, uint32_t *dstWords, F f) {
    if (!vertexCount) return;
    OLV_CHECK(olv_frontier_split(g, queue, count, buckets[0], bucketCounts, hubPairs));
    olv_tuning t;
    OLV_CHECK(olv_get_tuning_ex(g, &t));
    hipLaunchKernelGGL((olive_detail::push_kernel<VertexValue, AccumValue, EdgeValue, F, MARK>),
                       dim3((unsigned)cuCount * 8), dim3(olive_detail::kBlock), 0, stream(), buckets[0], bucketCounts,
                       reinterpret_cast<const uint2 *>(hubPairs), rowPtr, colIdx, vertexValues.elemsDevice,
                       accumulators.elemsDevice, edgeValues.elemsDevice, dstWords, f, t.td_wave_thr, t.td_block_thr,
                       t.hub_thr);
    HIP_CHECK(hipGetLastError());
  }

  VertexId vertexCount;
  EdgeId edgeCount;
  int cuCount;
  olv_graph *g;
  const uint32_t *rowPtr;
  const uint32_t *colIdx;
  uint32_t *buckets[3];
  uint32_t *bucketCounts;
  uint32_t *mapQueue;
  uint32_t *mapCount;
  uint32_t *hubPairs;   // hub items (vertex, chunk) of the push
  void *reduceScratch;  // partial sums of the device reductions + the result

  GRD<VertexValue> vertexValues;
  GRD<AccumValue> accumulators;
  GRD<EdgeValue> edgeValues;
  AccumValue defaultAccumValue;
};

#endif  // OLIVE_H
