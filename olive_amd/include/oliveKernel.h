// Functor-generic kernels of the Olive C++ API, compiled into the
// application (they call the user's gather/reduce/cond/update/operator()
// functors).  Everything type-independent -- degree binning of a frontier,
// bitmap -> queue compaction, fills, device memory -- is in libolive_hip.so.
//
// Replaces the reference's K1-K5 (oliverKernel.h:45-221):
//  * K1/K3 (push expansion): the frontier queue is split into low-degree
//    entries and hub items of hub_thr edges (olv_frontier_split), and one
//    launch of the engine's push (olv::push_expand) runs the user's
//    gather/reduce: hub items a workgroup each, the rest by degree bucket --
//    a workgroup or a wave per vertex, and the short lists dealt to
//    consecutive lanes by a block-wide scan -- with four edge visits in
//    flight per lane; destination bits are set with a read first and a
//    non-returning atomic (the reference gives every vertex a fixed
//    GroupSize of lanes, and K3 one thread per vertex however large its
//    list).
//  * K2 (vertexFilter): one lane per 32-vertex bitmap word applies cond /
//    update and the kept ids are compacted into the queue with a wave64
//    inclusive scan, a workgroup prefix over the waves and one global atomic
//    per workgroup (the reference scans one int flag per vertex and takes
//    an atomic per vertex or per block-local queue of 1200).
//  * K4/K5 (vertexMap): over bitmap words or the queue, grid-stride.
#ifndef OLIVE_KERNEL_H
#define OLIVE_KERNEL_H

#include <type_traits>

#include "bitmap.h"
#include "common.h"
// the engine's workload-balanced push (olv::push_expand), shared with the
// library's own BFS and SSSP kernels
#include "../csrc/bfs_core.hpp"

namespace olive_detail {

constexpr int kBlock = 256;
constexpr int kWave = 64;

template <class T>
__global__ void fill_kernel(T *p, T v, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__device__ inline uint32_t lane() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Edge value handed to gather(): the edge's own value, or one shared dummy
// for value-less edge types (Dump_Edge).
template <class E>
__device__ inline E &edge_ref(E *ev, EdgeId e) {
  if constexpr (std::is_empty<E>::value)
    return ev[0];
  else
    return ev[e];
}

// Sets bit v of a flag bitmap (no result: a non-returning atomic, skipped
// when the bit is already set).
__device__ inline void mark(uint32_t *w, VertexId v) {
  const uint32_t bit = 1u << (v & 31);
  if (!(w[v >> 5] & bit)) __hip_atomic_fetch_or(&w[v >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Visitor of olv::push_expand for a user functor: the payload of a source u
// is (out-degree << 32 | u); an edge's target carries the gathered value,
// which apply() reduces into the accumulator (and marks the destination).
template <class V, class A, class E, class F, bool MARK>
struct PushVisitor {
  const EdgeId *row_ptr;
  const VertexId *col;
  V *vals;
  A *accs;
  E *evals;
  uint32_t *dst_words;
  F f;
  struct Tgt {
    VertexId v;
    A a;
  };
  __device__ inline uint32_t row(uint32_t q) const { return q; }
  __device__ inline unsigned long long load(uint32_t u) const {
    return ((unsigned long long)(row_ptr[u + 1] - row_ptr[u]) << 32) | u;
  }
  __device__ inline Tgt tgt(uint32_t e, unsigned long long pl) {
    const uint32_t u = (uint32_t)pl;
    return Tgt{col[e], f.gather(vals[u], (EdgeId)(pl >> 32), edge_ref(evals, e))};
  }
  __device__ inline int state(const Tgt &) const { return 0; }
  __device__ inline void apply(const Tgt &t, int, unsigned long long) {
    f.reduce(accs[t.v], t.a);
    if (MARK) mark(dst_words, t.v);
  }
};

// Push of a frontier split into a queue of low-degree vertices and hub
// items (olv_frontier_split; counts[0] queue entries, counts[1] items):
// hub items a workgroup each, the queue by the degree-bucketed,
// scan-balanced olv::push_expand with four edge visits in flight per lane.
template <class V, class A, class E, class F, bool MARK>
__global__ void __launch_bounds__(kBlock)
    push_kernel(const uint32_t *queue, const uint32_t *counts, const uint2 *hubs, const EdgeId *row_ptr,
                const VertexId *col, V *vals, A *accs, E *evals, uint32_t *dst_words, F f, uint32_t wave_thr,
                uint32_t block_thr, uint32_t hub_chunk) {
  PushVisitor<V, A, E, F, MARK> vis{row_ptr, col, vals, accs, evals, dst_words, f};
  unsigned long long insp = 0;
  olv::push_expand<kBlock>(queue, counts[0], hubs, counts[1], row_ptr, wave_thr, block_thr, hub_chunk, blockIdx.x,
                           gridDim.x, vis, insp);
}

// Exclusive prefix of `x` over the workgroup; *total gets the sum.
__device__ inline uint32_t block_scan(uint32_t x, uint32_t *total) {
  __shared__ uint32_t s_wave[kBlock / kWave + 1];
  const uint32_t l = lane(), wid = threadIdx.x / kWave;
  uint32_t inc = x;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o, kWave);
    if (l >= (uint32_t)o) inc += y;
  }
  if (l == kWave - 1) s_wave[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < kBlock / kWave; ++w) {
      const uint32_t t = s_wave[w];
      s_wave[w] = run;
      run += t;
    }
    s_wave[kBlock / kWave] = run;
  }
  __syncthreads();
  const uint32_t r = s_wave[wid] + inc - x;
  *total = s_wave[kBlock / kWave];
  __syncthreads();
  return r;
}

// vertexFilter: for every flagged v with cond(val[v], acc[v]): update and
// enqueue.  One lane per 32-vertex word.
template <class V, class A, class F>
__global__ void __launch_bounds__(kBlock)
    filter_kernel(const uint32_t *words, uint32_t nwords, V *vals, A *accs, uint32_t *queue, uint32_t *qcount, F f) {
  __shared__ uint32_t s_base;
  for (uint32_t w0 = blockIdx.x * kBlock; w0 < nwords; w0 += gridDim.x * kBlock) {
    const uint32_t w = w0 + threadIdx.x;
    uint32_t keep = 0;
    for (uint32_t x = w < nwords ? words[w] : 0u; x; x &= x - 1) {
      const uint32_t b = __ffs(x) - 1;
      const VertexId v = w * 32u + b;
      if (f.cond(vals[v], accs[v])) {
        f.update(vals[v], accs[v]);
        keep |= 1u << b;
      }
    }
    uint32_t total;
    const uint32_t off = block_scan((uint32_t)__popc(keep), &total);
    if (threadIdx.x == 0) s_base = total ? atomicAdd(qcount, total) : 0u;
    __syncthreads();
    uint32_t pos = s_base + off;
    for (uint32_t x = keep; x; x &= x - 1) queue[pos++] = w * 32u + (__ffs(x) - 1);
    __syncthreads();
  }
}

template <class V, class A, class F>
__global__ void map_queue_kernel(const uint32_t *queue, const uint32_t *qcount, V *vals, A *accs, F f) {
  const uint32_t cnt = *qcount;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += gridDim.x * blockDim.x) {
    const VertexId v = queue[i];
    f(vals[v], accs[v]);
  }
}

template <class V, class A, class F>
__global__ void map_words_kernel(const uint32_t *words, uint32_t nwords, V *vals, A *accs, F f) {
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += gridDim.x * blockDim.x)
    for (uint32_t x = words[w]; x; x &= x - 1) {
      const VertexId v = w * 32u + (__ffs(x) - 1);
      f(vals[v], accs[v]);
    }
}

// ---- device reductions (vertexReduce with a device functor, subset sizes)
// Partial sums per workgroup: grid-stride in a fixed order, then a tree in
// LDS; reduce_final_kernel adds the partials in a fixed order, so a result
// does not depend on scheduling.
template <class V, class A, class F>
__global__ void __launch_bounds__(kBlock) reduce_kernel(const V *__restrict__ vals, uint32_t n, F f,
                                                        A *__restrict__ partials) {
  __shared__ A s[kBlock];
  A acc = A(0);
  for (size_t v = (size_t)blockIdx.x * kBlock + threadIdx.x; v < n; v += (size_t)gridDim.x * kBlock)
    acc += f(vals[v]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

// Number of set bits of words[0..nw) per workgroup.
__global__ void __launch_bounds__(kBlock) popcount_kernel(const uint32_t *__restrict__ words, uint32_t nw,
                                                          unsigned long long *__restrict__ partials) {
  __shared__ unsigned long long s[kBlock];
  unsigned long long acc = 0;
  for (size_t w = (size_t)blockIdx.x * kBlock + threadIdx.x; w < nw; w += (size_t)gridDim.x * kBlock)
    acc += __popc(words[w]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

// out = partials[0] + ... + partials[k-1] (one workgroup).
template <class A>
__global__ void __launch_bounds__(kBlock) reduce_final_kernel(const A *__restrict__ partials, uint32_t k, A *out) {
  __shared__ A s[kBlock];
  A acc = A(0);
  for (uint32_t i = threadIdx.x; i < k; i += kBlock) acc += partials[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

}  // namespace olive_detail

#endif  // OLIVE_KERNEL_H
