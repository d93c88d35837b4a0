// Device-side building blocks for gfx950 (CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace olv {

constexpr int kWave = 64;
constexpr int32_t kInf = 0x7fffffff;

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Loads of arrays that are read once per pass (CSR indices, edge weights).
// Non-temporal loads were measured slower on gfx950 for the PageRank pull
// (110 us vs 82 us per RMAT-20 iteration), so they are off by default.
#ifndef OLV_STREAM_NT
#define OLV_STREAM_NT 0
#endif
template <class T>
__device__ __forceinline__ T ld_stream(const T *p) {
#if OLV_STREAM_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

// Random gather from a vector (PageRank contributions, ranks, adjoints).
#ifndef OLV_GATHER_NT
#define OLV_GATHER_NT 0
#endif
template <class T>
__device__ __forceinline__ T ld_gather(const T *p) {
#if OLV_GATHER_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T x = __shfl_xor(v, o, kWave);
    v = x > v ? x : v;
  }
  return v;
}

template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T x = __shfl_xor(v, o, kWave);
    v = x < v ? x : v;
  }
  return v;
}

// Inclusive prefix sum over the wave.
template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T x = __shfl_up(v, o, kWave);
    if (lane >= (uint32_t)o) v += x;
  }
  return v;
}

// Block-wide exclusive scan for blockDim.x == BLOCK (multiple of 64).
// `smem` needs BLOCK/64 + 1 entries.  Returns the exclusive prefix and
// writes the block total to *total.  Contains __syncthreads().
template <int BLOCK, class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T *smem, T *total) {
  constexpr int W = BLOCK / kWave;
  const int wid = threadIdx.x / kWave;
  const uint32_t lane = lane_id();
  T inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) smem[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      T x = smem[i];
      smem[i] = run;
      run += x;
    }
    smem[W] = run;
  }
  __syncthreads();
  T res = smem[wid] + inc - v;
  *total = smem[W];
  __syncthreads();
  return res;
}

template <int BLOCK, class T>
__device__ __forceinline__ T block_sum(T v, T *smem) {
  constexpr int W = BLOCK / kWave;
  v = wave_sum(v);
  const int wid = threadIdx.x / kWave;
  if (lane_id() == 0) smem[wid] = v;
  __syncthreads();
  T r = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) r += smem[i];
  __syncthreads();
  return r;
}

__device__ __forceinline__ bool bm_test(const uint32_t *bm, uint32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

// XCC (XCD) the calling wave runs on: hardware register XCC_ID (20), bits 3:0.
__device__ __forceinline__ uint32_t xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15u; }

// Grid-wide barrier for persistent (cooperatively launched) kernels, XCD
// aware.  An agent-scope release fence on gfx950 writes back the whole L2 of
// the XCD (buffer_wbl2), so instead of one per workgroup there is one per
// XCD: workgroups first meet on a per-XCD counter, the last one on each XCD
// releases (L2 write-back) and arrives at the global counter; after the
// global generation moves on, every workgroup acquires (L2/L1 invalidate).
// Measured on MI355X with 256 workgroups: about 6 us per barrier, against
// 15-18 us for a fence in every workgroup.
//
// The state must be zeroed before launch.  Spins are bounded: on a timeout
// the error flag is raised and every caller returns false, so every wave
// still exits.
// Arrival counters only ever grow (barrier k completes at the k-th multiple
// of the arrivals it expects), so no reset store can race with an early
// arrival at the next barrier.
struct GridBar {
  uint32_t count, gen, err, icount;
  uint32_t local[16 * 16];  // per-XCD arrivals, 64 bytes apart
  uint32_t xcnt[16 * 16];   // workgroups per XCD
};

struct GridBarLocal {
  uint32_t gen = 0, xcc = 0, xsize = 0, nxcc = 0;
};

__device__ __forceinline__ bool gbar_wait(GridBar *b, uint32_t gen, uint32_t spin_limit) {
  uint32_t spins = 0;
  while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
    if (__hip_atomic_load(&b->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || ++spins > spin_limit) {
      __hip_atomic_store(&b->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// Every workgroup registers on its XCD, then a flat barrier makes the
// per-XCD sizes visible.  Returns false on timeout.
__device__ __forceinline__ bool gbar_init(GridBar *b, GridBarLocal &l, uint32_t nblocks, uint32_t spin_limit) {
  __shared__ uint32_t s_ok, s_xsize, s_nxcc, s_xcc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ok = 1;
    const uint32_t x = xcc_id();
    __hip_atomic_fetch_add(&b->xcnt[x * 16], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // release: this workgroup's xcnt arrival is ordered before its count
    // arrival; the last arriver's acquire sees every xcnt before bumping gen
    const uint32_t a = __hip_atomic_fetch_add(&b->icount, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (a == nblocks - 1) __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    ok = gbar_wait(b, 0u, spin_limit) ? 1u : 0u;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    uint32_t nx = 0;
    for (uint32_t i = 0; i < 16; ++i)
      nx += __hip_atomic_load(&b->xcnt[i * 16], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 1u : 0u;
    s_ok = ok;
    s_xcc = x;
    s_xsize = __hip_atomic_load(&b->xcnt[x * 16], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_nxcc = nx;
  }
  __syncthreads();
  l.gen = 1;
  l.xcc = s_xcc;
  l.xsize = s_xsize;
  l.nxcc = s_nxcc;
  const bool ok = s_ok != 0;
  __syncthreads();
  return ok;
}

__device__ __forceinline__ bool grid_barrier(GridBar *b, GridBarLocal &l, uint32_t spin_limit) {
  __shared__ uint32_t s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    // barrier l.gen (1, 2, ...): its arrivals are the l.gen-th block of
    // xsize arrivals on this XCD and of nxcc arrivals on count
    const uint32_t p = __hip_atomic_fetch_add(&b->local[l.xcc * 16], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (p == l.gen * l.xsize - 1) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      const uint32_t a = __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (a == l.gen * l.nxcc - 1) __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_ok = gbar_wait(b, l.gen, spin_limit) ? 1u : 0u;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  l.gen += 1;
  const bool ok = s_ok != 0;
  __syncthreads();
  return ok;
}

}  // namespace olv
