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

// Grid-wide barrier for persistent (cooperatively launched) kernels.
// bar[0] = arrival count, bar[1] = generation, bar[2] = error flag.
// The spin is bounded: if the barrier does not complete within `spin_limit`
// polls the error flag is raised and every caller returns false, so every
// wave still exits.
__device__ __forceinline__ bool grid_barrier(uint32_t *bar, uint32_t nblocks,
                                             uint32_t &gen,
                                             uint32_t spin_limit) {
  __shared__ uint32_t ok_s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t ok = 1;
    __threadfence();
    uint32_t arrived = atomicAdd(&bar[0], 1u);
    if (arrived == nblocks - 1) {
      atomicExch(&bar[0], 0u);
      __threadfence();
      atomicAdd(&bar[1], 1u);
    } else {
      uint32_t spins = 0;
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        if (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
            ++spins > spin_limit) {
          atomicOr(&bar[2], 1u);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();
    gen += 1;
    ok_s = ok;
  }
  __syncthreads();
  bool ok = ok_s != 0;
  __syncthreads();
  return ok;
}

}  // namespace olv
