// Launch-shape helper of the Olive C++ API (utils.h:56-72 of the reference)
// and small utilities.  The reference caps the grid at 65,535 blocks and its
// kernels are not grid-stride (any vertex id >= 16,776,960 is skipped); the
// engine's kernels here are grid-stride, so the cap only bounds launch size.
#ifndef OLIVE_UTILS_H
#define OLIVE_UTILS_H

#include <utility>

#include "common.h"

namespace util {

/** (blocks, threads per block) for `threads` work items; (0, 0) if none. */
inline std::pair<int, int> kernelConfig(int threads, int threadsPerBlock = DEFAULT_THREADS_PER_BLOCK) {
  if (threads <= 0) return std::make_pair(0, 0);
  assert(threadsPerBlock <= MAX_THREADS_PER_BLOCK);
  if (threads < threadsPerBlock) threadsPerBlock = threads;
  int blocks = (threads + threadsPerBlock - 1) / threadsPerBlock;
  if (blocks > MAX_BLOCKS) blocks = MAX_BLOCKS;
  return std::make_pair(blocks, threadsPerBlock);
}

inline bool isNumeric(const char *str) {
  assert(str);
  for (; *str; ++str)
    if (!isdigit(*str)) return false;
  return true;
}

/** xorshift hash of a number. */
inline size_t hashCode(size_t a) {
  a ^= (a << 13);
  a ^= (a >> 17);
  a ^= (a << 5);
  return a;
}

}  // namespace util

#endif  // OLIVE_UTILS_H
