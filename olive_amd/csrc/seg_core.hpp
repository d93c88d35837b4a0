// Bucketed segment sums over a CSR: one fused launch covers every vertex of
// a SegPlan.  Block roles by blockIdx (heavy roles first so they start
// first):
//   [hub chunks]  one workgroup per chunk of a row longer than plan.chunk,
//                 partial sum to chunk_partial[c] (combined later in a fixed
//                 order by the finalize kernel)
//   [bucket 3]    one workgroup per vertex
//   [bucket 2]    one wave per vertex
//   [bucket 1]    8 lanes per vertex
//   [bucket 0]    one lane per vertex
// Within a vertex, lane j sums terms j, j+G, j+2G, ... sequentially and the
// G partials are combined by a shuffle tree, so the fp32 summation order is
// fixed (bit-reproducible) and its error grows with deg/G + log2(G).
#pragma once
#include "device.hpp"
#include "internal.hpp"

namespace olv {

struct SegDev {
  const uint32_t *list;
  const uint32_t *hub_vertex;
  const uint32_t *hub_first;
  const uint32_t *chunk_hub;
  float *chunk_partial;
  uint32_t off[SegPlan::kBuckets + 1];
  uint32_t nhub, nchunks, chunk;
  // first block of each role: [0]=hub chunks, [1]=bucket3, [2]=bucket2,
  // [3]=bucket1, [4]=bucket0, [5]=end
  uint32_t role_start[6];
};

constexpr int kSegBlock = 256;

inline SegDev make_seg_dev(const SegPlan &p) {
  SegDev d;
  d.list = p.list;
  d.hub_vertex = p.hub_vertex;
  d.hub_first = p.hub_first;
  d.chunk_hub = p.chunk_hub;
  d.chunk_partial = p.chunk_partial;
  for (uint32_t i = 0; i <= SegPlan::kBuckets; ++i) d.off[i] = p.offsets[i];
  d.nhub = p.nhub;
  d.nchunks = p.nchunks;
  d.chunk = p.chunk;
  const uint32_t c3 = p.offsets[4] - p.offsets[3];
  const uint32_t c2 = p.offsets[3] - p.offsets[2];
  const uint32_t c1 = p.offsets[2] - p.offsets[1];
  const uint32_t c0 = p.offsets[1] - p.offsets[0];
  d.role_start[0] = 0;
  d.role_start[1] = d.role_start[0] + p.nchunks;
  d.role_start[2] = d.role_start[1] + c3;
  d.role_start[3] = d.role_start[2] + (c2 + kSegBlock / 64 - 1) / (kSegBlock / 64);
  d.role_start[4] = d.role_start[3] + (c1 + kSegBlock / 8 - 1) / (kSegBlock / 8);
  d.role_start[5] = d.role_start[4] + (c0 + kSegBlock - 1) / kSegBlock;
  return d;
}

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Op: float term(uint32_t v, uint32_t k)   -- summand of in/out slot k of v
//     void finish(uint32_t v, float sum, float &err)
// err is accumulated per thread and written to err_part[blockIdx.x].
template <class Op>
__global__ void __launch_bounds__(kSegBlock) seg_iter_kernel(SegDev d, const uint32_t *__restrict__ rp, Op op,
                                                             const uint32_t *__restrict__ done,
                                                             float *__restrict__ err_part) {
  if (done && __hip_atomic_load(done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  __shared__ float s_red[kSegBlock / kWave];
  const uint32_t b = blockIdx.x;
  float err = 0.f;
  if (b < d.role_start[1]) {
    // hub chunk: partial sum only
    const uint32_t c = b;
    const uint32_t h = d.chunk_hub[c];
    const uint32_t v = d.hub_vertex[h];
    const uint32_t k0 = c - d.hub_first[h];
    const uint32_t beg = rp[v] + k0 * d.chunk;
    const uint32_t end = min(rp[v + 1], beg + d.chunk);
    float s = 0.f;
    for (uint32_t k = beg + threadIdx.x; k < end; k += kSegBlock) s += op.term(v, k);
    s = block_sum<kSegBlock, float>(s, s_red);
    if (threadIdx.x == 0) d.chunk_partial[c] = s;
    if (threadIdx.x == 0) err_part[b] = 0.f;
    return;
  } else if (b < d.role_start[2]) {
    const uint32_t v = d.list[d.off[3] + (b - d.role_start[1])];
    float s = 0.f;
    for (uint32_t k = rp[v] + threadIdx.x; k < rp[v + 1]; k += kSegBlock) s += op.term(v, k);
    s = block_sum<kSegBlock, float>(s, s_red);
    if (threadIdx.x == 0) op.finish(v, s, err);
  } else if (b < d.role_start[3]) {
    const uint32_t idx = d.off[2] + (b - d.role_start[2]) * (kSegBlock / 64) + threadIdx.x / 64;
    const uint32_t j = threadIdx.x % 64;
    if (idx < d.off[3]) {
      const uint32_t v = d.list[idx];
      float s = 0.f;
      for (uint32_t k = rp[v] + j; k < rp[v + 1]; k += 64) s += op.term(v, k);
      s = group_sum<64>(s);
      if (j == 0) op.finish(v, s, err);
    }
  } else if (b < d.role_start[4]) {
    const uint32_t idx = d.off[1] + (b - d.role_start[3]) * (kSegBlock / 8) + threadIdx.x / 8;
    const uint32_t j = threadIdx.x % 8;
    const bool ok = idx < d.off[2];
    float s = 0.f;
    uint32_t v = 0;
    if (ok) {
      v = d.list[idx];
      for (uint32_t k = rp[v] + j; k < rp[v + 1]; k += 8) s += op.term(v, k);
    }
    s = group_sum<8>(s);
    if (ok && j == 0) op.finish(v, s, err);
  } else {
    const uint32_t idx = d.off[0] + (b - d.role_start[4]) * kSegBlock + threadIdx.x;
    if (idx < d.off[1]) {
      const uint32_t v = d.list[idx];
      float s = 0.f;
      for (uint32_t k = rp[v]; k < rp[v + 1]; ++k) s += op.term(v, k);
      op.finish(v, s, err);
    }
  }
  err = block_sum<kSegBlock, float>(err, s_red);
  if (threadIdx.x == 0) err_part[b] = err;
}

// Iteration control block (device).
struct IterCtl {
  uint32_t done;      // set once the residual drops below eps
  uint32_t iters;     // updates performed
  double eps;         // stop threshold on the L1 residual
  double residual;    // last residual
};

// Single-workgroup epilogue of one iteration: finish the chunked hub rows
// (partials summed in chunk order), total the residual, set `done`.
template <class Op>
__global__ void __launch_bounds__(1024) seg_finalize_kernel(SegDev d, Op op, const float *__restrict__ err_part,
                                                            uint32_t nparts, IterCtl *__restrict__ ctl) {
  if (__hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  __shared__ double s_red[1024 / kWave];
  float err = 0.f;
  for (uint32_t h = threadIdx.x; h < d.nhub; h += blockDim.x) {
    float s = 0.f;
    for (uint32_t c = d.hub_first[h]; c < d.hub_first[h + 1]; ++c) s += d.chunk_partial[c];
    op.finish(d.hub_vertex[h], s, err);
  }
  double tot = err;
  for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) tot += err_part[i];
  tot = block_sum<1024, double>(tot, s_red);
  if (threadIdx.x == 0) {
    ctl->residual = tot;
    ctl->iters += 1;
    if (tot < ctl->eps) ctl->done = 1u;
  }
}

}  // namespace olv
