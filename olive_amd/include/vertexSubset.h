// Vertex subsets of the Olive C++ API (vertexSubset.h:38-156 of the
// reference).  As in the reference, a subset is either "sparse" -- one flag
// per vertex, here a 1-bit-per-vertex device bitmap instead of one int
// (4V bytes) -- or "dense" -- a queue of ids with a device-side size
// (isDense == true).  Ids are compacted from bitmap to queue by
// vertexFilter with wave64 ballot/popcount and a workgroup prefix sum.
#ifndef OLIVE_VERTEX_SUBSET_H
#define OLIVE_VERTEX_SUBSET_H

#include "bitmap.h"
#include "common.h"

class VertexSubset {
 public:
  bool isDense;        // queue representation
  VertexId n;
  Bitmap flags;        // sparse representation
  uint32_t *queue;     // dense representation (device, n entries)
  uint32_t *qSizeDevice;
  uint32_t *qSize;     // pinned host mirror

  /** Empty (unallocated) subset. */
  VertexSubset() : isDense(false), n(0), queue(NULL), qSizeDevice(NULL), qSize(NULL) {}

  /** Sparse subset of n vertices, empty or universal. */
  VertexSubset(VertexId count, bool universal) : isDense(false), n(count), queue(NULL), qSizeDevice(NULL), qSize(NULL) {
    flags.reserve(count);
    if (universal)
      flags.fill();
    else
      flags.clear();
  }

  /** Empty dense subset (queue) of capacity n. */
  explicit VertexSubset(VertexId count) : isDense(true), n(count) {
    allocQueue();
    *qSize = 0;
    OLV_CHECK(olv_fill(qSizeDevice, 0, 1, 4));
  }

  /** Dense subset holding the single vertex v. */
  VertexSubset(VertexId count, VertexId v) : isDense(true), n(count) {
    assert(v < count);
    allocQueue();
    *qSize = 1;
    OLV_CHECK(olv_memcpy_h2d(queue, &v, 4));
    OLV_CHECK(olv_memcpy_h2d(qSizeDevice, qSize, 4));
  }

  /** Number of queued vertices (dense subsets; a device->host copy). */
  inline VertexId size() const {
    if (isDense) {
      OLV_CHECK(olv_memcpy_d2h(qSize, qSizeDevice, 4));
      return *qSize;
    }
    assert(0 && "size() of a sparse subset");
    return 0;
  }

  inline VertexId capacity() const { return n; }

  inline void clear() {
    if (isDense) {
      *qSize = 0;
      OLV_CHECK(olv_fill(qSizeDevice, 0, 1, 4));
    } else {
      flags.clear();
    }
  }

  inline void print() {
    if (isDense) {
      const VertexId k = size();
      uint32_t *h = (uint32_t *)malloc(sizeof(uint32_t) * (k ? k : 1));
      OLV_CHECK(olv_memcpy_d2h(h, queue, (size_t)k * 4));
      printf("dense: ");
      for (VertexId i = 0; i < k; i++) printf("%u ", h[i]);
      printf("\n");
      free(h);
    } else {
      uint32_t *h = (uint32_t *)malloc(sizeof(uint32_t) * (flags.nwords ? flags.nwords : 1));
      OLV_CHECK(olv_memcpy_d2h(h, flags.words, (size_t)flags.nwords * 4));
      printf("sparse: ");
      for (VertexId v = 0; v < n; v++)
        if ((h[v >> 5] >> (v & 31)) & 1u) printf("%u ", v);
      printf("\n");
      free(h);
    }
  }

  void del() {
    if (isDense) {
      if (queue) olv_free(queue);
      if (qSizeDevice) olv_free(qSizeDevice);
      if (qSize) olv_host_free(qSize);
      queue = qSizeDevice = qSize = NULL;
    } else {
      flags.del();
    }
  }

 private:
  void allocQueue() {
    void *d = NULL;
    OLV_CHECK(olv_malloc(&d, (size_t)(n ? n : 1) * 4));
    queue = reinterpret_cast<uint32_t *>(d);
    OLV_CHECK(olv_malloc(&d, 4));
    qSizeDevice = reinterpret_cast<uint32_t *>(d);
    void *h = NULL;
    OLV_CHECK(olv_host_malloc(&h, 4));
    qSize = reinterpret_cast<uint32_t *>(h);
  }
};

#endif  // OLIVE_VERTEX_SUBSET_H
