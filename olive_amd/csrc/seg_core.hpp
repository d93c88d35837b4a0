// Bucketed segment sums over a CSR: one fused launch covers every vertex of
// a SegPlan.  Work is cut into virtual blocks by role (heavy roles first so
// they start first); a grid of a few thousand workgroups loops over them:
//   [hub chunks]  one workgroup per chunk of a row longer than plan.chunk,
//                 partial sum to chunk_partial[c] (combined later in chunk
//                 order)
//   [bucket 3]    one workgroup per vertex
//   [bucket 2]    one wave per vertex
//   [bucket 1]    8 lanes per vertex
//   [bucket 0]    one lane per vertex
// Within a vertex, lane j owns terms j, j+G, j+2G, ...; it loads U slot
// indices at once (predicated, no serial remainder), then the U gathered
// values, into U interleaved accumulators that are added pairwise; the G
// lane partials are combined by a shuffle tree.  So a lane has U gathers in
// flight and a chain of two dependent loads per U terms, and the fp32
// summation order is fixed (bit-reproducible), its error growing with
// deg/(U*G) + log2(U*G).
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
  uint32_t *hub_done;  // chunks of each chunked row finished in this launch
  uint32_t off[SegPlan::kBuckets + 1];
  uint32_t nhub, nchunks, chunk;
  // first virtual block of each role: [0]=hub chunks, [1]=bucket3,
  // [2]=bucket2, [3]=bucket1, [4]=bucket0, [5]=end
  uint32_t role_start[6];
  // end of the virtual blocks that hold a row with entries: the trailing
  // bucket-0 blocks hold only empty rows
  uint32_t live_end;
  bool contig;                          // list[i] = i + shift[bucket]
  uint32_t shift[SegPlan::kBuckets];
  uint32_t grid_per_cu;                 // workgroups per CU of a launch (host side)
  uint32_t unroll;                      // gathers in flight per lane: 4, 8 or 16 (host side)
  // Vertex at list position idx of bucket b (Contig: without the load).
  template <bool Contig>
  __device__ __forceinline__ uint32_t vertex(uint32_t idx, int b) const {
    return Contig ? idx + shift[b] : list[idx];
  }
};

constexpr int kSegBlock = 256;
// s_waitcnt operand (gfx9 encoding): vmcnt = 0, expcnt and lgkmcnt unbounded.
constexpr int kWaitVmcnt0 = (0x7 << 4) | (0xF << 8);

inline SegDev make_seg_dev(const SegPlan &p, const Tuning &t) {
  SegDev d;
  d.grid_per_cu = t.seg_grid_per_cu;
  d.unroll = t.seg_unroll;
  d.list = p.list;
  d.hub_vertex = p.hub_vertex;
  d.hub_first = p.hub_first;
  d.chunk_hub = p.chunk_hub;
  d.chunk_partial = p.chunk_partial;
  d.hub_done = p.hub_done;
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
  const uint32_t p0 = c0 - p.nempty;  // first empty row within bucket 0
  d.live_end = d.role_start[4] + (p0 + kSegBlock - 1) / kSegBlock;
  d.contig = p.contig;
  for (uint32_t i = 0; i < SegPlan::kBuckets; ++i) d.shift[i] = p.shift[i];
  return d;
}

// Workgroups of a segment launch: d.grid_per_cu per CU (0: one workgroup
// per virtual block), never more than there are virtual blocks.
inline uint32_t seg_grid(const SegDev &d) {
  const uint32_t nvb = d.role_start[5] ? d.role_start[5] : 1;
  if (d.grid_per_cu == 0) return nvb;
  const uint32_t cap = (uint32_t)cu_count() * d.grid_per_cu;
  return nvb < cap ? nvb : cap;
}

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Sum of op.val(v, k, op.idx(k)) for k = beg + j, beg + j + G, ... < end.
template <int G, int U, class Op>
__device__ __forceinline__ float lane_sum(const Op &op, uint32_t v, uint32_t beg, uint32_t end, uint32_t j) {
  float acc[U];
#pragma unroll
  for (int i = 0; i < U; ++i) acc[i] = 0.f;
  for (uint64_t k = (uint64_t)beg + j; k < end; k += (uint64_t)U * G) {
    uint32_t ix[U];
#pragma unroll
    for (int i = 0; i < U; ++i) ix[i] = k + (uint64_t)i * G < end ? op.idx((uint32_t)(k + (uint64_t)i * G)) : 0u;
#pragma unroll
    for (int i = 0; i < U; ++i)
      if (k + (uint64_t)i * G < end) acc[i] += op.val(v, (uint32_t)(k + (uint64_t)i * G), ix[i]);
  }
#pragma unroll
  for (int w = U / 2; w > 0; w >>= 1)
#pragma unroll
    for (int i = 0; i < w; ++i) acc[i] += acc[i + w];
  return acc[0];
}

// Iteration control block (device).  The L1 residual of iteration k is
// accumulated in errs[k % 3] -- kErrSlots copies on separate 128-byte lines,
// workgroup b adding to copy b % kErrSlots (device-scope atomics to one
// address serialise at the memory side).  Iteration k + 1 reads it at its
// start, so no separate reduction launch is needed between iterations.
constexpr int kErrSlots = 32;

struct alignas(128) ErrSlot {
  double v;
  double pad[15];
};

struct IterCtl {
  uint32_t done;      // set by the first iteration that finds the residual below eps
  uint32_t iters;     // updates performed, when done
  double eps;         // stop threshold on the L1 residual
  double residual;    // residual that stopped the iteration, when done
  uint32_t pad[26];
  ErrSlot errs[3][kErrSlots];
};

// Total of errs[slot] for every thread of the calling workgroup (all
// threads must call).
__device__ __forceinline__ double err_total(const IterCtl *ctl, uint32_t slot) {
  __shared__ double s_tot;
  if (threadIdx.x < kWave) {
    double x = threadIdx.x < kErrSlots ? ctl->errs[slot][threadIdx.x].v : 0.0;
    x = wave_sum(x);
    if (threadIdx.x == 0) s_tot = x;
  }
  __syncthreads();
  const double r = s_tot;
  __syncthreads();
  return r;
}

// Op: uint32_t idx(uint32_t k)                 -- index read of slot k
//     float val(uint32_t v, uint32_t k, uint32_t i) -- summand of slot k of row v
//     void finish(uint32_t v, float sum, float &err)
// Iteration k of a fixed point (ctl != nullptr): returns at once when done;
// stops (sets done) if iteration k - 1 left a residual below eps; else
// updates every vertex and adds its residual to errs[k % 3].  ctl ==
// nullptr: a pure edge pass (e.g. gradients).  Rows longer than d.chunk are
// cut into chunks; the workgroup that completes a row's last chunk sums the
// chunk partials in chunk order (fixed order: reproducible) and finishes it.
// Waves per SIMD the iteration kernel is compiled for (register budget):
// the kernel is bound by gather latency, so residency matters more than
// the few scalar registers spilled to fit it.
#ifndef OLV_SEG_WAVES_PER_EU
#define OLV_SEG_WAVES_PER_EU 8
#endif

template <class Op, bool Contig, int U>
__global__ void __launch_bounds__(kSegBlock) __attribute__((amdgpu_waves_per_eu(OLV_SEG_WAVES_PER_EU)))
seg_iter_kernel(SegDev d, const uint32_t *__restrict__ rp, Op op, IterCtl *ctl, uint32_t k) {
  if (ctl) {
    if (__hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    if (k > 0) {
      const double res = err_total(ctl, (k - 1) % 3);
      if (res < ctl->eps) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
          ctl->iters = k;
          ctl->residual = res;
          __hip_atomic_store(&ctl->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
      }
    }
    if (blockIdx.x == 0) {
      double *z = &ctl->errs[(k + 1) % 3][0].v;
      for (uint32_t i = threadIdx.x; i < kErrSlots * (sizeof(ErrSlot) / sizeof(double)); i += kSegBlock) z[i] = 0.0;
    }
  }
  __shared__ float s_red[kSegBlock / kWave];
  float err = 0.f;
  // An empty row's update is a constant: iterations 0 and 1 write it into
  // both ping-pong buffers, later iterations skip those rows (their
  // residual is 0).  An edge pass (ctl == nullptr) has nothing to do there.
  const uint32_t nvb = (!ctl || k >= 2) ? d.live_end : d.role_start[5];
  for (uint32_t b = blockIdx.x; b < nvb; b += gridDim.x) {
    if (b < d.role_start[1]) {
      // hub chunk: partial sum; the last chunk of the row to finish
      // completes the row
      const uint32_t c = b;
      const uint32_t h = d.chunk_hub[c];
      const uint32_t v = d.hub_vertex[h];
      const uint32_t c0 = d.hub_first[h], c1 = d.hub_first[h + 1];
      const uint32_t beg = rp[v] + (c - c0) * d.chunk;
      const uint32_t end = min(rp[v + 1], beg + d.chunk);
      float s = lane_sum<kSegBlock, U>(op, v, beg, end, threadIdx.x);
      s = block_sum<kSegBlock, float>(s, s_red);
      if (threadIdx.x == 0) {
        // The partials are device-coherent atomic stores and loads, so the
        // hand-off needs ordering, not cache maintenance: wait for the
        // partial's store to complete before counting in (an acq_rel
        // atomic would also write back and invalidate this XCD's whole L2,
        // once per chunk and iteration).
        __hip_atomic_store(&d.chunk_partial[c], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const uint32_t before =
            __hip_atomic_fetch_add(&d.hub_done[h], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (before == c1 - c0 - 1) {
          float t = 0.f;
          for (uint32_t q = c0; q < c1; ++q)
            t += __hip_atomic_load(&d.chunk_partial[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          op.finish(v, t, err);
          __hip_atomic_store(&d.hub_done[h], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
    } else if (b < d.role_start[2]) {
      const uint32_t v = d.template vertex<Contig>(d.off[3] + (b - d.role_start[1]), 3);
      float s = lane_sum<kSegBlock, U>(op, v, rp[v], rp[v + 1], threadIdx.x);
      s = block_sum<kSegBlock, float>(s, s_red);
      if (threadIdx.x == 0) op.finish(v, s, err);
    } else if (b < d.role_start[3]) {
      const uint32_t idx = d.off[2] + (b - d.role_start[2]) * (kSegBlock / 64) + threadIdx.x / 64;
      const uint32_t j = threadIdx.x % 64;
      if (idx < d.off[3]) {
        const uint32_t v = d.template vertex<Contig>(idx, 2);
        float s = lane_sum<64, U>(op, v, rp[v], rp[v + 1], j);
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
        v = d.template vertex<Contig>(idx, 1);
        s = lane_sum<8, U>(op, v, rp[v], rp[v + 1], j);
      }
      s = group_sum<8>(s);
      if (ok && j == 0) op.finish(v, s, err);
    } else {
      const uint32_t idx = d.off[0] + (b - d.role_start[4]) * kSegBlock + threadIdx.x;
      if (idx < d.off[1]) {
        const uint32_t v = d.template vertex<Contig>(idx, 0);
        const float s = lane_sum<1, U>(op, v, rp[v], rp[v + 1], 0u);
        op.finish(v, s, err);
      }
    }
  }
  if (!ctl) return;
  err = block_sum<kSegBlock, float>(err, s_red);
  if (threadIdx.x == 0 && err != 0.f)
    atomicAdd(&ctl->errs[k % 3][blockIdx.x % kErrSlots].v, (double)err);
}

// One launch of seg_iter_kernel with the descriptor's unroll depth and list
// form (on the library stream).
template <class Op>
inline void launch_seg(const SegDev &d, const uint32_t *rp, const Op &op, IterCtl *ctl, uint32_t k) {
  const dim3 grid(seg_grid(d)), block(kSegBlock);
#define OLV_SEG_LAUNCH(C, U) hipLaunchKernelGGL((seg_iter_kernel<Op, C, U>), grid, block, 0, stream(), d, rp, op, ctl, k)
  if (d.contig) {
    if (d.unroll == 4) OLV_SEG_LAUNCH(true, 4);
    else if (d.unroll == 16) OLV_SEG_LAUNCH(true, 16);
    else OLV_SEG_LAUNCH(true, 8);
  } else {
    if (d.unroll == 4) OLV_SEG_LAUNCH(false, 4);
    else if (d.unroll == 16) OLV_SEG_LAUNCH(false, 16);
    else OLV_SEG_LAUNCH(false, 8);
  }
#undef OLV_SEG_LAUNCH
}

}  // namespace olv
