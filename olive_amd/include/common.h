// Types, launch constants and checks of the Olive C++ API (common.h:40-152
// of the reference), for HIP on MI355X.  Host code reaches the device
// through libolive_hip.so (include/olive_hip.h); only the functor-generic
// kernels are compiled into the application.
//
// HIP provides a native fp64 atomicAdd, so -- unlike the reference
// (common.h:141-150) -- none is defined here.
#ifndef OLIVE_COMMON_H
#define OLIVE_COMMON_H

#include <assert.h>
#include <ctype.h>
#include <inttypes.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime.h>

#include "olive_hip.h"

/** One word equals 64 bit. */
typedef uint64_t Word;
/** Vertex ids and edge ids are 32-bit (graphs up to 2^32 - 1 edges). */
typedef uint32_t VertexId;
typedef uint32_t EdgeId;
typedef uint32_t PartitionId;

/** Edge value for graphs whose edges carry nothing. */
struct Dump_Edge {};

/** Launch constants kept for source compatibility. */
const int DEFAULT_THREADS_PER_BLOCK = 256;
const int MAX_THREADS_PER_BLOCK = 1024;
const int MAX_BLOCKS = 65535;

#define BLOCK_INDEX blockIdx.x
#define THREAD_INDEX (threadIdx.x + blockDim.x * blockIdx.x)
#define NUM_BLOCKS gridDim.x
#define NUM_THREADS (blockDim.x * gridDim.x)

/** Abort with a message when a libolive_hip call fails. */
#define OLV_CHECK(call)                                                              \
  do {                                                                               \
    int olv_rc_ = (call);                                                            \
    if (olv_rc_ != OLV_OK) {                                                         \
      fprintf(stderr, "olive error in %s:%d: %s (%s)\n", __FILE__, __LINE__, #call,  \
              olv_last_error());                                                     \
      abort();                                                                       \
    }                                                                                \
  } while (0)

/** Abort with a message when a HIP call fails. */
#define HIP_CHECK(call)                                                              \
  do {                                                                               \
    hipError_t hip_rc_ = (call);                                                     \
    if (hip_rc_ != hipSuccess) {                                                     \
      fprintf(stderr, "HIP error in %s:%d: %s\n", __FILE__, __LINE__,                \
              hipGetErrorString(hip_rc_));                                           \
      abort();                                                                       \
    }                                                                                \
  } while (0)

#define H2D(dst, src, size) olv_memcpy_h2d(dst, src, size)
#define D2H(dst, src, size) olv_memcpy_d2h(dst, src, size)

#endif  // OLIVE_COMMON_H
