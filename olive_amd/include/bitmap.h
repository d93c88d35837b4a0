// Device bitmap (the reference's bitmap.h:44-199, which nothing uses there,
// kept managed 64-bit words; here 32-bit words in device memory, the layout
// the frontier kernels of libolive_hip.so work on).
#ifndef OLIVE_BITMAP_H
#define OLIVE_BITMAP_H

#include "common.h"

class Bitmap {
 public:
  uint32_t *words;  // device
  VertexId n;
  uint32_t nwords;

  Bitmap() : words(NULL), n(0), nwords(0) {}

  void reserve(VertexId count) {
    n = count;
    OLV_CHECK(olv_bitmap_words(count, &nwords));
    void *d = NULL;
    OLV_CHECK(olv_malloc(&d, (size_t)(nwords ? nwords : 1) * 4));
    words = reinterpret_cast<uint32_t *>(d);
  }

  void clear() {
    if (nwords) OLV_CHECK(olv_fill(words, 0, nwords, 4));
  }

  /** Every bit of [0, n) set, the padding bits of the last word clear. */
  void fill() {
    if (!nwords) return;
    OLV_CHECK(olv_fill(words, 0xffffffffu, nwords, 4));
    if (n % 32) {
      uint32_t last = (1u << (n % 32)) - 1u;
      OLV_CHECK(olv_memcpy_h2d(words + nwords - 1, &last, 4));
    }
  }

  void del() {
    if (words) olv_free(words);
    words = NULL;
    n = nwords = 0;
  }

  __device__ static inline bool get(const uint32_t *w, VertexId v) { return (w[v >> 5] >> (v & 31)) & 1u; }
  /** Sets bit v; true if this call set it. */
  __device__ static inline bool set(uint32_t *w, VertexId v) {
    const uint32_t bit = 1u << (v & 31);
    if (w[v >> 5] & bit) return false;
    return !(atomicOr(&w[v >> 5], bit) & bit);
  }
};

#endif  // OLIVE_BITMAP_H
