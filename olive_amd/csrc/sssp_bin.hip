// SSSP by destination binning: each round's relaxations are scattered into
// per-partition bins, then applied partition by partition in LDS.
//
// Same fixed point as the frontier relaxation of SSSP.cu:146-163 (and the
// direct path in sssp.hip): every vertex whose distance drops is expanded
// again with its new distance until nothing drops, and shortest distances
// are unique, so the order of relaxations changes no output.  What changes is
// where the min-reductions happen.  A relaxation in the direct path is a
// random 4-byte probe of dist and, when it improves, two device-scope atomics
// (atomicMin on dist, atomicOr on a bitmap).  On MI355X those atomics are
// performed at the memory side, and binning the probes by the XCD that owns
// the destination does not help them: 20M random relaxations into a 16 MB
// array cost about 735 us either way.  Here a round is two kernels:
//
//   expand (A): the frontier's out-edges become (target, candidate) pairs,
//     staged in LDS by destination partition (2^sb consecutive virtual
//     vertices) and written to that partition's bin as contiguous runs; one
//     device-scope atomic per (tile, partition) reserves the run.  No dist
//     probe, no per-edge atomic.
//   apply (B): one workgroup per partition loads the partition's dist slice
//     and dirty bits into LDS, takes the min of every binned candidate with
//     LDS atomics, writes back the 256-byte wave spans that changed, and
//     builds the next frontier from its own slice.
//
// 20M pairs: A about 90 us, B about 35 us, against 735 us for the same
// relaxations done directly (measured with synthetic pairs).
//
// Near-far order (Davidson et al., IPDPS'14) is kept with one "dirty" bit per
// vertex: the distance dropped since the vertex was last expanded.  With
// split T, the frontier is {dirty, dist < T}; B builds it for the next round
// and records the smallest dirty distance >= T.  When the frontier is empty
// the next round re-splits at (smallest dirty far distance) + delta -- no
// empty promotion, the split always lands past at least one vertex -- and
// when nothing is dirty the run is done.  A vertex leaves the dirty set when
// it is expanded: B of round R+1 clears the bits of the frontier A of round
// R+1 expanded, recomputed from the slice state B of round R left.
//
// k sources run as k instances over virtual vertices s*n + v (as in
// sssp.hip); partitions cover virtual ranges, so one launch serves all.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "bfs_core.hpp"
#include "internal.hpp"

namespace olv {

constexpr int kBinA = 512;                           // expand workgroup (~53 KB LDS: 3 per CU)
constexpr int kBinSlots = 8;                         // pairs per thread per tile
constexpr uint32_t kBinTile = kBinA * kBinSlots;     // 4096 pairs per tile
constexpr uint32_t kBinPartMax = 1024;               // partitions
constexpr int kBinB = 1024;                          // apply workgroup
constexpr uint32_t kFillStride = 32;                 // one fill counter per 128-byte line
// slice: 2^14 vertices (64 KB of dist; 16 pairs per partition run in a
// 4096-pair tile of a 4M-vertex graph) up to 2^15 (128 KB)
constexpr uint32_t kBinSbMin = 14, kBinSbMax = 15;

enum : uint32_t { kModeNear = 0, kModeResplit = 1, kModeDone = 2 };

// Smallest dirty far distance, one copy per 128-byte line (workgroup b of
// the apply kernel updates copy b % kCtrSlots).
struct alignas(128) MinSlot {
  int32_t v;
  uint32_t pad[31];
};

struct BinCtl {
  uint32_t round;    // next round to expand (written by apply)
  uint32_t round_b;  // round being applied (written by expand)
  uint32_t done, rounds, capped, err, resplits, pad0;
  uint32_t mode[2];  // mode of round R, slot R & 1
  int32_t split[2];  // split[R & 1]: the split the frontier of round R was built with
  uint32_t qn[2];    // frontier queue length of round R
  uint32_t hn[2];    // frontier hub items of round R
  unsigned long long relax;  // result: edges read
  MinSlot minfar[2][kCtrSlots];
  CtrSet cnt[2];     // edges read by the expand of round R (slot R & 1)
};

struct BinArgs {
  const uint32_t *row_ptr;
  const uint32_t *col;
  const int32_t *w;
  int32_t *dist;       // nv
  uint32_t *dirty;     // nwp words
  const uint32_t *hub_bm;  // rows with out-degree >= hub_thr (n bits)
  uint32_t *queue[2];  // frontier entries (virtual ids) of round R in queue[R & 1]
  uint2 *hubs[2];      // frontier hub items of round R
  uint2 *bins;         // pairs (virtual target, candidate)
  const uint32_t *pbase;  // nparts + 1: bin offsets (capacities = in-edges of the partition)
  uint32_t *fill;      // nparts x kFillStride: pairs in the partition's bin this round
  BinCtl *ctl;
  uint32_t n, k, nv, nw, nparts, sb;
  uint32_t hub_thr;
  int32_t delta;
  uint32_t max_rounds;
};

__device__ __forceinline__ uint32_t bin_row(const BinArgs &a, uint32_t x) { return a.k > 1 ? x % a.n : x; }

// Virtual id -> CSR row.
struct BinRow {
  uint32_t n, k;
  __device__ __forceinline__ uint32_t operator()(uint32_t x) const { return k > 1 ? x % n : x; }
};

__device__ __forceinline__ int32_t add_clamp(int32_t t, int32_t d) { return t > kInf - d ? kInf : t + d; }

// ---------------------------------------------------------------- expand
__global__ void __launch_bounds__(kBinA) ssspb_expand_kernel(BinArgs a) {
  BinCtl *ctl = a.ctl;
  if (__hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const uint32_t R = ctl->round;
  const uint32_t t = threadIdx.x;
  const bool lead = blockIdx.x == 0;
  if (R >= a.max_rounds) {
    if (lead && t == 0) {
      ctl->rounds = R;
      ctl->capped = 1u;
      ctl->done = 1u;
    }
    return;
  }
  const uint32_t qn = ctl->qn[R & 1], hn = ctl->hn[R & 1];
  const int32_t split = ctl->split[R & 1];
  __shared__ int32_t s_mf;
  if (t < kWave) {
    int32_t mf = t < kCtrSlots ? ctl->minfar[R & 1][t].v : kInf;
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const int32_t y = __shfl_xor(mf, o, kWave);
      mf = y < mf ? y : mf;
    }
    if (t == 0) s_mf = mf;
  }
  __syncthreads();
  const int32_t minfar = s_mf;
  const uint32_t mode = (qn + hn) ? kModeNear : (minfar < kInf ? kModeResplit : kModeDone);
  if (lead) {
    if (t == 0) {
      ctl->round_b = R;
      ctl->mode[R & 1] = mode;
      ctl->split[(R + 1) & 1] = mode == kModeNear ? split : add_clamp(minfar, a.delta);
      ctl->qn[(R + 1) & 1] = 0;
      ctl->hn[(R + 1) & 1] = 0;
      if (mode == kModeResplit) ctl->resplits += 1;
      if (mode == kModeDone) {
        ctl->rounds = R;
        ctl->done = 1u;
      }
    }
    if (t < kCtrSlots) ctl->minfar[(R + 1) & 1][t].v = kInf;
  }
  if (mode != kModeNear) return;

  __shared__ uint32_t s_off[kBinA + 1], s_beg[kBinA], s_xb[kBinA];
  __shared__ int32_t s_du[kBinA];
  __shared__ uint32_t s_scan[kBinA / kWave + 1];
  __shared__ uint32_t s_cnt[kBinPartMax], s_po[kBinPartMax], s_g[kBinPartMax];
  __shared__ uint2 s_st[kBinTile];
  constexpr uint32_t kPP = kBinPartMax / kBinA;  // partitions per thread (scan)
  for (uint32_t j = 0; j < kPP; ++j) s_cnt[t * kPP + j] = 0;

  const uint32_t *queue = a.queue[R & 1];
  const uint2 *hubs = a.hubs[R & 1];
  const uint32_t nitems = qn + hn;
  const uint32_t G = gridDim.x;
  const uint32_t per = (nitems + G - 1) / G;
  const uint32_t i0 = blockIdx.x * per;
  const uint32_t i1 = min(nitems, i0 + per);
  unsigned long long insp = 0;
  for (uint32_t ib = i0; ib < i1; ib += kBinA) {
    const uint32_t i = ib + t;
    uint32_t beg = 0, deg = 0, xb = 0;
    int32_t du = 0;
    if (i < i1) {
      uint32_t x, rb, re;
      if (i < qn) {
        x = queue[i];
        const uint32_t row = bin_row(a, x);
        rb = a.row_ptr[row];
        re = a.row_ptr[row + 1];
        beg = rb;
        deg = re - rb;
        xb = x - row;
      } else {
        const uint2 it = hubs[i - qn];
        x = it.x;
        const uint32_t row = bin_row(a, x);
        rb = a.row_ptr[row];
        re = a.row_ptr[row + 1];
        beg = rb + it.y * a.hub_thr;
        deg = beg < re ? min(a.hub_thr, re - beg) : 0u;
        xb = x - row;
      }
      du = a.dist[x];
    }
    uint32_t E;
    const uint32_t off = block_exclusive_scan<kBinA>(deg, s_scan, &E);
    s_off[t] = i < i1 ? off : E;
    s_beg[t] = beg;
    s_xb[t] = xb;
    s_du[t] = du;
    if (t == 0) s_off[kBinA] = E;
    insp += deg;
    __syncthreads();
    for (uint32_t tb = 0; tb < E; tb += kBinTile) {
      const uint32_t tn = min(kBinTile, E - tb);
      uint2 pr[kBinSlots];
      uint32_t rk[kBinSlots];
#pragma unroll
      for (int j = 0; j < kBinSlots; ++j) {
        const uint32_t kk = j * kBinA + t;
        pr[j] = make_uint2(0xffffffffu, 0u);
        if (kk < tn) {
          const uint32_t slot = tb + kk;
          // owner: the last item whose offset is <= slot
          uint32_t lo = 0, hi = kBinA;  // s_off[lo] <= slot < s_off[hi]
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s_off[mid] <= slot)
              lo = mid;
            else
              hi = mid;
          }
          const uint32_t e = s_beg[lo] + (slot - s_off[lo]);
          const uint32_t v = s_xb[lo] + a.col[e];
          pr[j] = make_uint2(v, (uint32_t)s_du[lo] + (uint32_t)a.w[e]);
        }
      }
#pragma unroll
      for (int j = 0; j < kBinSlots; ++j)
        if (pr[j].x != 0xffffffffu) rk[j] = atomicAdd(&s_cnt[pr[j].x >> a.sb], 1u);
      __syncthreads();
      // partition offsets in the tile, and one reservation per partition
      uint32_t c[kPP], sum = 0;
#pragma unroll
      for (uint32_t j = 0; j < kPP; ++j) {
        c[j] = s_cnt[t * kPP + j];
        sum += c[j];
      }
      uint32_t tot;
      uint32_t run = block_exclusive_scan<kBinA>(sum, s_scan, &tot);
#pragma unroll
      for (uint32_t j = 0; j < kPP; ++j) {
        const uint32_t p = t * kPP + j;
        s_po[p] = run;
        uint32_t g = 0xffffffffu;
        if (c[j]) {
          const uint32_t old = atomicAdd(&a.fill[p * kFillStride], c[j]);
          const uint32_t cap = a.pbase[p + 1] - a.pbase[p];
          if (old + c[j] <= cap)
            g = a.pbase[p] + old;
          else
            __hip_atomic_store(&ctl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_g[p] = g;
        run += c[j];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kBinSlots; ++j)
        if (pr[j].x != 0xffffffffu) s_st[s_po[pr[j].x >> a.sb] + rk[j]] = pr[j];
      __syncthreads();
      for (uint32_t q = t; q < tn; q += kBinA) {
        const uint2 e = s_st[q];
        const uint32_t p = e.x >> a.sb;
        const uint32_t g = s_g[p];
        if (g != 0xffffffffu) a.bins[g + (q - s_po[p])] = e;
      }
      __syncthreads();
#pragma unroll
      for (uint32_t j = 0; j < kPP; ++j) s_cnt[t * kPP + j] = 0;
      __syncthreads();
    }
  }
  // edges read: the sum of every lane's item degrees
  flush_counters<kBinA>(&ctl->cnt[R & 1], 0u, 0ull, insp);
}

// ----------------------------------------------------------------- apply
template <uint32_t SB>
__global__ void __launch_bounds__(kBinB) ssspb_apply_kernel(BinArgs a) {
  constexpr uint32_t S = 1u << SB;
  constexpr uint32_t J = S / kBinB;
  // the pre-apply values stay in registers up to 16 per lane; above that
  // they are read again from dist (this workgroup's own slice, unchanged)
  constexpr bool kKeep = J <= 16;
  extern __shared__ uint32_t s_dyn[];
  int32_t *s_d = reinterpret_cast<int32_t *>(s_dyn);  // S values
  uint32_t *s_dw = s_dyn + S;                         // S / 32 dirty words
  __shared__ uint32_t s_scan[kBinB / kWave + 1];
  __shared__ uint32_t s_base;
  __shared__ int32_t s_mf[kBinB / kWave];
  BinCtl *ctl = a.ctl;
  if (__hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const uint32_t R = ctl->round_b;
  const uint32_t t = threadIdx.x;
  const bool near_mode = ctl->mode[R & 1] == kModeNear;
  const int32_t tc = ctl->split[R & 1];       // split of the frontier just expanded
  const int32_t tb = ctl->split[(R + 1) & 1];  // split of the next frontier
  if (blockIdx.x == 0) {
    const CtrSum c = ctr_total(&ctl->cnt[R & 1]);
    if (t == 0) {
      ctl->relax += c.inspected;
      ctl->round = R + 1;
    }
    ctr_clear(&ctl->cnt[(R + 1) & 1]);
  }
  uint32_t *nq = a.queue[(R + 1) & 1];
  uint2 *nh = a.hubs[(R + 1) & 1];
  const uint32_t lane = lane_id();
  for (uint32_t p = blockIdx.x; p < a.nparts; p += gridDim.x) {
    const uint32_t x0 = p << SB;
    const uint32_t cv = min(S, a.nv - x0);  // vertices in this partition
    int32_t old[kKeep ? J : 1];
#pragma unroll
    for (uint32_t j = 0; j < J; ++j) {
      const uint32_t x = j * kBinB + t;
      const int32_t o = x < cv ? a.dist[x0 + x] : kInf;
      if (kKeep) old[kKeep ? j : 0] = o;
      s_d[x] = o;
    }
    const uint32_t w0 = x0 >> 5;
    for (uint32_t w = t; w < S / 32; w += kBinB) s_dw[w] = w0 + w < a.nw ? a.dirty[w0 + w] : 0u;
    __syncthreads();
    const uint32_t np = a.fill[p * kFillStride];
    const uint32_t cap = a.pbase[p + 1] - a.pbase[p];
    const uint32_t cnt = np <= cap ? np : cap;
    const uint2 *b = a.bins + a.pbase[p];
    // (every binned target lies in this slice; the mask only keeps a
    // corrupted bin from reaching past the LDS array)
    uint32_t i = t;
    for (; i + 3 * kBinB < cnt; i += 4 * kBinB) {
      const uint2 q0 = b[i], q1 = b[i + kBinB], q2 = b[i + 2 * kBinB], q3 = b[i + 3 * kBinB];
      atomicMin(&s_d[(q0.x - x0) & (S - 1)], (int32_t)q0.y);
      atomicMin(&s_d[(q1.x - x0) & (S - 1)], (int32_t)q1.y);
      atomicMin(&s_d[(q2.x - x0) & (S - 1)], (int32_t)q2.y);
      atomicMin(&s_d[(q3.x - x0) & (S - 1)], (int32_t)q3.y);
    }
    for (; i < cnt; i += kBinB) {
      const uint2 q = b[i];
      atomicMin(&s_d[(q.x - x0) & (S - 1)], (int32_t)q.y);
    }
    __syncthreads();
    if (t == 0) a.fill[p * kFillStride] = 0u;
    // classify: dirty' = (dirty & !expanded) | improved; next frontier =
    // dirty' & dist < tb; far minimum over dirty' & dist >= tb
    uint32_t nnear = 0;
    uint32_t nearbits = 0;  // bit j: vertex j * kBinB + t joins the queue
    uint32_t wasbits = 0;   // bit j: vertex j * kBinB + t was dirty
#pragma unroll
    for (uint32_t j = 0; j < J; ++j) {
      const uint32_t x = j * kBinB + t;
      wasbits |= ((s_dw[x >> 5] >> (x & 31)) & 1u) << j;
    }
    __syncthreads();  // every dirty word is read before any is replaced
    int32_t mf = kInf;
#pragma unroll
    for (uint32_t j = 0; j < J; ++j) {
      const uint32_t x = j * kBinB + t;
      const int32_t nv = s_d[x];
      const bool inb = x < cv;
      const int32_t o = kKeep ? old[kKeep ? j : 0] : (inb ? a.dist[x0 + x] : kInf);
      const bool improved = inb && nv < o;
      const bool was = (wasbits >> j) & 1u;
      const bool expanded = near_mode && was && o < tc;
      const bool dirty = (was && !expanded) || improved;
      const uint64_t any = ballot(improved);
      if (any && inb) a.dist[x0 + x] = nv;  // whole wave span: no partial lines
      const uint64_t db = ballot(dirty);
      if (lane == 0) s_dw[x >> 5] = (uint32_t)db;
      if (lane == 32) s_dw[x >> 5] = (uint32_t)(db >> 32);
      if (dirty && nv < tb) {
        bool hub = false;
        if (a.hub_bm) {
          const uint32_t row = bin_row(a, x0 + x);
          hub = (a.hub_bm[row >> 5] >> (row & 31)) & 1u;
        }
        if (hub) {
          const uint32_t row = bin_row(a, x0 + x);
          const uint32_t d = a.row_ptr[row + 1] - a.row_ptr[row];
          const uint32_t kk = (d + a.hub_thr - 1) / a.hub_thr;
          const uint32_t hb = atomicAdd(&ctl->hn[(R + 1) & 1], kk);
          for (uint32_t c = 0; c < kk; ++c) nh[hb + c] = make_uint2(x0 + x, c);
        } else {
          nearbits |= 1u << j;
          ++nnear;
        }
      } else if (dirty) {
        mf = nv < mf ? nv : mf;
      }
    }
    __syncthreads();
    for (uint32_t w = t; w < S / 32; w += kBinB)
      if (w0 + w < a.nw) a.dirty[w0 + w] = s_dw[w];
    uint32_t total;
    const uint32_t off = block_exclusive_scan<kBinB>(nnear, s_scan, &total);
    if (t == 0) s_base = total ? atomicAdd(&ctl->qn[(R + 1) & 1], total) : 0u;
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const int32_t y = __shfl_xor(mf, o, kWave);
      mf = y < mf ? y : mf;
    }
    if (lane == 0) s_mf[t / kWave] = mf;
    __syncthreads();
    uint32_t pos = s_base + off;
#pragma unroll
    for (uint32_t j = 0; j < J; ++j)
      if ((nearbits >> j) & 1u) nq[pos++] = x0 + j * kBinB + t;
    if (t == 0) {
      int32_t m = kInf;
      for (int q = 0; q < kBinB / kWave; ++q) m = s_mf[q] < m ? s_mf[q] : m;
      if (m < kInf) atomicMin(&ctl->minfar[(R + 1) & 1][blockIdx.x % kCtrSlots].v, m);
    }
    __syncthreads();
  }
}

// dist = INF, dirty = 0, fill = 0, control zero (minfar INF), split = delta.
__global__ void ssspb_reset_kernel(BinArgs a) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = tid; i < a.nv; i += stride) a.dist[i] = kInf;
  for (size_t i = tid; i < a.nw; i += stride) a.dirty[i] = 0u;
  for (size_t i = tid; i < (size_t)a.nparts * kFillStride; i += stride) a.fill[i] = 0u;
  uint32_t *c = reinterpret_cast<uint32_t *>(a.ctl);
  for (size_t i = tid; i < sizeof(BinCtl) / 4; i += stride) c[i] = 0u;
  if (tid == 0) a.ctl->split[0] = a.delta;
}

// One workgroup, after the reset: minfar = INF, then the k seeds.
__global__ void ssspb_seed_kernel(const uint32_t *srcs, BinArgs a) {
  if (threadIdx.x < kCtrSlots) {
    a.ctl->minfar[0][threadIdx.x].v = kInf;
    a.ctl->minfar[1][threadIdx.x].v = kInf;
  }
  auto claim = [&](uint32_t s) -> uint32_t {
    const uint32_t x = s * a.n + srcs[s];
    a.dist[x] = 0;
    atomicOr(&a.dirty[x >> 5], 1u << (x & 31));
    return x;
  };
  unsigned long long deg;
  seed_frontier(srcs, a.k, a.row_ptr, a.hub_thr, a.queue[0], &a.ctl->qn[0], a.hubs[0], &a.ctl->hn[0], claim,
                BinRow{a.n, a.k}, &deg);
}

// Capacity of each partition's bin: in-edges of its virtual vertices, i.e.
// over the k instances, the edges whose target s*n + col[e] falls in it.
__global__ void ssspb_cap_kernel(const uint32_t *col, uint32_t m, uint32_t n, uint32_t k, uint32_t sb,
                                 uint32_t nparts, uint32_t *cap) {
  __shared__ uint32_t s_h[kBinPartMax];
  for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x) s_h[p] = 0;
  __syncthreads();
  const uint64_t total = (uint64_t)k * m;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = (uint32_t)(i / m), e = (uint32_t)(i % m);
    atomicAdd(&s_h[(s * n + col[e]) >> sb], 1u);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < nparts; p += blockDim.x)
    if (s_h[p]) atomicAdd(&cap[p], s_h[p]);
}

// Rows with out-degree >= thr.
__global__ void ssspb_hub_kernel(const uint32_t *rp, uint32_t n, uint32_t thr, uint32_t *bm) {
  const uint32_t lane = lane_id();
  for (uint32_t v0 = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; v0 < n; v0 += gridDim.x * blockDim.x) {
    const uint32_t v = v0 + lane;
    const bool h = v < n && rp[v + 1] - rp[v] >= thr;
    const uint64_t b = ballot(h);
    if (lane == 0) bm[v0 >> 5] = (uint32_t)b;
    if (lane == 32 && v0 + 32 < n) bm[(v0 >> 5) + 1] = (uint32_t)(b >> 32);
  }
}

// ------------------------------------------------------------ fused rounds
// One kernel per round for graphs without hub vertices: workgroup q owns
// partition q (2^14 virtual vertices) and, in launch R,
//   1. applies the pairs that every partition sent it in launch R-1 (region
//      p -> q of the pair store, one count per region) to its dist slice in
//      LDS,
//   2. updates its dirty bits and picks its frontier {dirty, dist < T},
//   3. expands that frontier at once: the pairs for partition q' go to region
//      q -> q' of the other pair store, at offsets this workgroup counts in
//      LDS -- each region has a single writer, so no device-scope atomic is
//      needed to place them (region capacities are the edge counts between
//      the two slices, fixed per graph).
// The split T of launch R comes from the previous launch's totals (striped
// counters): frontier non-empty -> same split; empty -> re-split at the
// smallest dirty far distance + delta; nothing dirty -> done.  Half the
// kernel boundaries of the expand/apply pair above, and none of its
// reservation atomics.
// Slices of 2^14 vertices (64 KB of dist; ~135 KB of LDS, one workgroup per
// CU) up to 16M virtual vertices.  (2^13-vertex slices fit two workgroups per
// CU by LDS, but the second ones started up to 120 us late on MI355X and a
// 4M-vertex run took 1.78 ms against 1.30 ms.)
constexpr uint32_t kFParts = 1024;               // partitions at most (16M virtual vertices)
constexpr int kFB = 1024;                        // workgroup

struct alignas(128) FusedSlot {
  uint32_t nnear;  // frontier vertices chosen (and expanded) in the launch
  int32_t mf;      // smallest dirty distance left at or above the split
  unsigned long long edges;
  uint32_t pad[28];
};

struct FusedCtl {
  uint32_t done, rounds, capped, err, resplits, pad[3];
  int32_t split[2];  // split of launch R in slot R & 1
  unsigned long long relax;
  FusedSlot slot[3][kCtrSlots];  // totals of launch R in slot R % 3
};

struct FusedArgs {
  const uint32_t *row_ptr;
  const uint32_t *col;
  const int32_t *w;
  int32_t *dist;
  uint32_t *dirty;
  uint2 *pairs[2];        // pair store written by launch R in pairs[(R + 1) & 1]
  const uint32_t *base;   // [q * nparts + p]: start of region p -> q; base[nparts^2] = total
  uint32_t *cnt[2];       // [q * nparts + p]: pairs in region p -> q (written by p, read as a row by q)
  FusedCtl *ctl;
  uint32_t n, k, nv, nw, nparts;
  int32_t delta;
  uint32_t max_rounds;
};

// Exclusive prefix over the workgroup with one barrier: every wave reads the
// kFB / 64 wave totals itself (no serial pass by one lane).  `wt` holds
// kFB / 64 words and must not be reused before the next barrier.
__device__ __forceinline__ uint32_t fused_scan(uint32_t v, uint32_t *wt, uint32_t *total) {
  const uint32_t lane = lane_id(), wid = threadIdx.x / kWave;
  const uint32_t inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) wt[wid] = inc;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (uint32_t i = 0; i < kFB / kWave; ++i) {
    const uint32_t x = wt[i];
    before += i < wid ? x : 0u;
    tot += x;
  }
  *total = tot;
  return before + inc - v;
}

template <uint32_t kFSb>
__global__ void __launch_bounds__(kFB) ssspf_round_kernel(FusedArgs a, uint32_t R) {
  constexpr uint32_t kFS = 1u << kFSb;
  constexpr uint32_t kFV = kFS / kFB;  // consecutive vertices per lane
  FusedCtl *ctl = a.ctl;
  const uint32_t t = threadIdx.x, lane = lane_id();
  const uint32_t q = blockIdx.x;
  const uint32_t P = a.nparts;
  __shared__ int4 s_d4[kFS / 4];
  int32_t *s_d = reinterpret_cast<int32_t *>(s_d4);
  __shared__ uint16_t s_near[kFS];
  __shared__ uint32_t s_in[kFParts + 1], s_base[kFParts];  // incoming regions: scan, start
  __shared__ uint32_t s_obase[kFParts], s_ocap[kFParts];   // outgoing regions: start, capacity
  __shared__ uint32_t s_run[kFParts];  // pairs placed in region q -> q'
  constexpr uint32_t kNW = kFB / kWave;
  __shared__ uint32_t s_wo[kNW][kWave + 1], s_wb[kNW][kWave], s_wx[kNW][kWave];  // per-wave items
  __shared__ int32_t s_wd[kNW][kWave];
  __shared__ uint32_t s_scan[2][kNW];
  __shared__ int32_t s_mfw[kFB / kWave];
  __shared__ int32_t s_T;
  __shared__ uint32_t s_stop;
  // Everything this launch reads that does not depend on the round decision
  // is requested first: the slice (16 consecutive values per lane), its dirty
  // word, the region counts and offsets of this partition.
  const uint32_t x0 = q << kFSb;
  const uint32_t cv = min(kFS, a.nv - x0);
  const uint32_t o0 = t * kFV;  // this lane's first vertex in the slice
  int32_t old[kFV];
  if (o0 + kFV <= cv) {
    const int4 *src = reinterpret_cast<const int4 *>(a.dist + x0 + o0);
#pragma unroll
    for (uint32_t i = 0; i < kFV / 4; ++i) {
      const int4 v = src[i];
      old[4 * i] = v.x;
      old[4 * i + 1] = v.y;
      old[4 * i + 2] = v.z;
      old[4 * i + 3] = v.w;
    }
  } else {
#pragma unroll
    for (uint32_t i = 0; i < kFV; ++i) old[i] = o0 + i < cv ? a.dist[x0 + o0 + i] : kInf;
  }
  const uint32_t wd = (x0 + o0) >> 5;  // dirty word holding this lane's 16 bits
  const uint32_t dw = o0 < cv ? a.dirty[wd] : 0u;
  uint32_t cin = 0, bin = 0, ob = 0, oe = 0;
  if (t < P) {
    if (R > 0) cin = a.cnt[R & 1][q * P + t];
    bin = a.base[q * P + t];
    ob = a.base[t * P + q];
    oe = a.base[t * P + q + 1];
  }
  // round decision (identical in every workgroup)
  if (t < kWave) {
    uint32_t nn = 0;
    int32_t mf = kInf;
    unsigned long long ed = 0;
    if (R > 0 && t < kCtrSlots) {
      const FusedSlot &sl = ctl->slot[(R - 1) % 3][t];
      nn = sl.nnear;
      mf = sl.mf;
      ed = sl.edges;
    }
    nn = wave_sum(nn);
    ed = wave_sum(ed);
    for (int o = kWave / 2; o > 0; o >>= 1) {
      const int32_t y = __shfl_xor(mf, o, kWave);
      mf = y < mf ? y : mf;
    }
    if (t == 0) {
      int32_t T;
      uint32_t stop = __hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ? 3u : 0u;
      if (R == 0) {
        T = a.delta;
      } else if (nn > 0) {
        T = ctl->split[(R - 1) & 1];
      } else if (mf < kInf) {
        T = add_clamp(mf, a.delta);
      } else {
        T = kInf;
        if (!stop) stop = 1;  // nothing dirty: done
      }
      if (!stop && R >= a.max_rounds) stop = 2;
      s_T = T;
      s_stop = stop;
      if (q == 0 && stop != 3) {
        ctl->relax += ed;
        if (stop) {
          ctl->rounds = R;
          ctl->capped = stop == 2;
          ctl->done = 1u;
        } else {
          ctl->split[R & 1] = T;
          if (R > 0 && nn == 0) ctl->resplits += 1;
        }
      }
    }
  }
  __syncthreads();
  if (s_stop) return;
  if (q == 0 && t < kCtrSlots) {
    FusedSlot &z = ctl->slot[(R + 1) % 3][t];
    z.nnear = 0;
    z.mf = kInf;
    z.edges = 0;
  }
  const int32_t T = s_T;
#pragma unroll
  for (uint32_t i = 0; i < kFV / 4; ++i)
    s_d4[t * (kFV / 4) + i] = make_int4(old[4 * i], old[4 * i + 1], old[4 * i + 2], old[4 * i + 3]);
  // 1. incoming pairs
  uint32_t tin;
  const uint32_t ein = fused_scan(cin, s_scan[0], &tin);
  if (t < P) {
    s_in[t] = ein;
    s_base[t] = bin;
    s_obase[t] = ob;
    s_ocap[t] = oe - ob;
    s_run[t] = 0;
  }
  if (t == 0) s_in[P] = tin;
  __syncthreads();
  {
    const uint2 *src = a.pairs[R & 1];
    auto pair_at = [&](uint32_t i) -> uint2 {
      uint32_t lo = 0, hi = P;  // s_in[lo] <= i < s_in[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_in[mid] <= i)
          lo = mid;
        else
          hi = mid;
      }
      return src[s_base[lo] + (i - s_in[lo])];
    };
    constexpr int UA = 8;  // pairs in flight per lane
    for (uint32_t i0 = 0; i0 < tin; i0 += UA * kFB) {
      uint2 pp[UA];
#pragma unroll
      for (int u = 0; u < UA; ++u) {
        const uint32_t i = i0 + u * kFB + t;
        pp[u] = i < tin ? pair_at(i) : make_uint2(0xffffffffu, 0u);
      }
#pragma unroll
      for (int u = 0; u < UA; ++u)
        if (pp[u].x != 0xffffffffu) atomicMin(&s_d[(pp[u].x - x0) & (kFS - 1)], (int32_t)pp[u].y);
    }
  }
  __syncthreads();
  // 2. dirty' = dirty | improved; frontier = dirty' & dist < T (expanded
  // now, so it leaves the dirty set); far minimum over the rest
  int32_t nv[kFV];
#pragma unroll
  for (uint32_t i = 0; i < kFV / 4; ++i) {
    const int4 v = s_d4[t * (kFV / 4) + i];
    nv[4 * i] = v.x;
    nv[4 * i + 1] = v.y;
    nv[4 * i + 2] = v.z;
    nv[4 * i + 3] = v.w;
  }
  const uint32_t was = (dw >> (o0 & 31)) & ((1u << kFV) - 1u);
  uint32_t imp = 0, rest = 0, nr = 0;
  int32_t mf = kInf;
#pragma unroll
  for (uint32_t i = 0; i < kFV; ++i) {
    const bool improved = nv[i] < old[i];  // (vertices past cv hold INF on both sides)
    const bool dirty = ((was >> i) & 1u) || improved;
    const bool near = dirty && nv[i] < T;
    imp |= (uint32_t)improved << i;
    nr |= (uint32_t)near << i;
    rest |= (uint32_t)(dirty && !near) << i;
    if (dirty && !near) mf = nv[i] < mf ? nv[i] : mf;
  }
  // write back the 128-byte lines (lane pairs) that changed
  uint32_t lim = imp;
#pragma unroll
  for (uint32_t o = 1; o < 32 / kFV; o <<= 1) lim |= __shfl_xor(lim, o, kWave);
  const bool line = lim != 0;
  if (line) {
    if (o0 + kFV <= cv) {
      int4 *dst = reinterpret_cast<int4 *>(a.dist + x0 + o0);
#pragma unroll
      for (uint32_t i = 0; i < kFV / 4; ++i) dst[i] = make_int4(nv[4 * i], nv[4 * i + 1], nv[4 * i + 2], nv[4 * i + 3]);
    } else {
#pragma unroll
      for (uint32_t i = 0; i < kFV; ++i)
        if (o0 + i < cv) a.dist[x0 + o0 + i] = nv[i];
    }
  }
  // lanes sharing a dirty word combine their bits; the first one stores it
  constexpr uint32_t kLW = 32 / kFV;  // lanes per word
  uint32_t word = rest << (o0 & 31);
#pragma unroll
  for (uint32_t o = 1; o < kLW; o <<= 1) word |= __shfl_xor(word, o, kWave);
  if ((t % kLW) == 0 && o0 < cv) a.dirty[wd] = word;
  uint32_t N;
  const uint32_t nb = fused_scan((uint32_t)__popc(nr), s_scan[1], &N);
  {
    uint32_t pos = nb;
    for (uint32_t x = nr; x; x &= x - 1) s_near[pos++] = (uint16_t)(o0 + __ffs(x) - 1);
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const int32_t y = __shfl_xor(mf, o, kWave);
    mf = y < mf ? y : mf;
  }
  if (lane == 0) s_mfw[t / kWave] = mf;
  __syncthreads();
  // 3. expand the frontier into the regions q -> q'.  Each wave takes 64
  // frontier vertices at a time and deals their out-edges to its lanes (a
  // wave-local prefix of the degrees); a pair's place in its region is an
  // LDS atomic on the region's counter (one writer per region: this
  // workgroup), so no barrier stands between reading an edge and storing its
  // pair.  Lines of a region fill from consecutive claims and merge in L2.
  uint2 *dst = a.pairs[(R + 1) & 1];
  unsigned long long edges = 0;
  {
    const uint32_t wv = t / kWave;
    constexpr int U = 8;  // edges in flight per lane
    // a lane's frontier vertex of chunk cb: out-list start, degree, instance
    // base, distance (the next chunk's are requested before this one's edges)
    auto item = [&](uint32_t i, uint32_t &beg, uint32_t &deg, uint32_t &xb, int32_t &du) {
      beg = deg = xb = 0;
      du = 0;
      if (i < N) {
        const uint32_t o = s_near[i];
        const uint32_t x = x0 + o;
        const uint32_t row = a.k > 1 ? x % a.n : x;
        beg = a.row_ptr[row];
        deg = a.row_ptr[row + 1] - beg;
        xb = x - row;
        du = s_d[o];
      }
    };
    uint32_t beg, deg, xb;
    int32_t du;
    item(wv * kWave + lane, beg, deg, xb, du);
    for (uint32_t cb = wv * kWave; cb < N; cb += kNW * kWave) {
      uint32_t nbeg, ndeg, nxb;
      int32_t ndu;
      item(cb + kNW * kWave + lane, nbeg, ndeg, nxb, ndu);
      const uint32_t inc = wave_inclusive_scan(deg);
      const uint32_t E = __shfl(inc, kWave - 1, kWave);
      s_wo[wv][lane] = inc - deg;
      s_wb[wv][lane] = beg;
      s_wx[wv][lane] = xb;
      s_wd[wv][lane] = du;
      if (lane == 0) s_wo[wv][kWave] = E;
      edges += deg;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      for (uint32_t s0 = 0; s0 < E; s0 += kWave * U) {
        uint32_t e[U], xbv[U];
        int32_t duv[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t slot = s0 + u * kWave + lane;
          ok[u] = slot < E;
          uint32_t lo = 0, hi = kWave;  // s_wo[lo] <= slot < s_wo[hi]
          if (ok[u])
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) >> 1;
              if (s_wo[wv][mid] <= slot)
                lo = mid;
              else
                hi = mid;
            }
          e[u] = ok[u] ? s_wb[wv][lo] + (slot - s_wo[wv][lo]) : 0u;
          xbv[u] = s_wx[wv][lo];
          duv[u] = s_wd[wv][lo];
        }
        uint32_t cv_[U], wv_[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          cv_[u] = ok[u] ? a.col[e[u]] : 0u;
          wv_[u] = ok[u] ? (uint32_t)a.w[e[u]] : 0u;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (ok[u]) {
            const uint32_t v = xbv[u] + cv_[u];
            const uint32_t p = v >> kFSb;
            const uint32_t pos = atomicAdd(&s_run[p], 1u);
            if (pos < s_ocap[p]) dst[s_obase[p] + pos] = make_uint2(v, (uint32_t)duv[u] + wv_[u]);
          }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();  // the wave's item arrays are rewritten next
      beg = nbeg;
      deg = ndeg;
      xb = nxb;
      du = ndu;
    }
  }
  __syncthreads();
  if (t < P && s_run[t] > s_ocap[t]) __hip_atomic_store(&ctl->err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // counts of the regions q -> q' filled in this launch (zeros included)
  if (t < P) a.cnt[(R + 1) & 1][t * P + q] = min(s_run[t], s_ocap[t]);
  // totals
  __shared__ unsigned long long s_red[kFB / kWave];
  const unsigned long long te = block_sum<kFB, unsigned long long>(edges, s_red);
  if (t == 0) {
    int32_t m = kInf;
    for (int i = 0; i < kFB / kWave; ++i) m = s_mfw[i] < m ? s_mfw[i] : m;
    FusedSlot &sl = ctl->slot[R % 3][q % kCtrSlots];
    if (N) atomicAdd(&sl.nnear, N);
    if (m < kInf) atomicMin(&sl.mf, m);
    if (te) atomicAdd(&sl.edges, te);
  }
}

// Edges from slice p to slice q (virtual): cap[q * nparts + p]; one
// workgroup per source slice.
__global__ void ssspf_cap_kernel(const uint32_t *rp, const uint32_t *col, uint32_t n, uint32_t k, uint32_t nv,
                                 uint32_t nparts, uint32_t sb, uint32_t *cap) {
  __shared__ uint32_t s_h[kFParts];
  const uint32_t p = blockIdx.x;
  for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) s_h[i] = 0;
  __syncthreads();
  const uint32_t x0 = p << sb, x1 = min(nv, x0 + (1u << sb));
  // one wave per source vertex
  const uint32_t lane = lane_id(), wv = threadIdx.x / kWave, nwv = blockDim.x / kWave;
  for (uint32_t x = x0 + wv; x < x1; x += nwv) {
    const uint32_t row = k > 1 ? x % n : x, xb = x - row;
    for (uint32_t e = rp[row] + lane; e < rp[row + 1]; e += kWave) atomicAdd(&s_h[(xb + col[e]) >> sb], 1u);
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x)
    if (s_h[i]) cap[i * nparts + p] = s_h[i];
}

// dist = INF, dirty = 0, counts of parity 0 = 0, control zero (slots: mf =
// INF), then the seeds: dist 0 and dirty.
__global__ void ssspf_reset_kernel(FusedArgs a) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = tid; i < a.nv; i += stride) a.dist[i] = kInf;
  for (size_t i = tid; i < a.nw; i += stride) a.dirty[i] = 0u;
  for (size_t i = tid; i < (size_t)a.nparts * a.nparts; i += stride) a.cnt[0][i] = 0u;
  uint32_t *c = reinterpret_cast<uint32_t *>(a.ctl);
  for (size_t i = tid; i < sizeof(FusedCtl) / 4; i += stride) c[i] = 0u;
}

__global__ void ssspf_seed_kernel(const uint32_t *srcs, FusedArgs a) {
  if (threadIdx.x < 3 * kCtrSlots) a.ctl->slot[threadIdx.x / kCtrSlots][threadIdx.x % kCtrSlots].mf = kInf;
  for (uint32_t s = threadIdx.x; s < a.k; s += blockDim.x) {
    const uint32_t x = s * a.n + srcs[s];
    a.dist[x] = 0;
    atomicOr(&a.dirty[x >> 5], 1u << (x & 31));
  }
}

// --------------------------------------------------------------------- host
struct SsspBinWork {
  bool fused = false;
  uint32_t k = 0, hub_thr = 0, sb = 0, nparts = 0, nw = 0, nwp = 0;
  uint64_t nv = 0;
  uint32_t *pbase = nullptr, *fill = nullptr, *dirty = nullptr, *hub_bm = nullptr;
  uint32_t *queue[2] = {nullptr, nullptr};
  uint2 *hubs[2] = {nullptr, nullptr};
  uint2 *bins = nullptr;
  BinCtl *ctl = nullptr, *h_ctl = nullptr;
  // fused rounds
  uint2 *pairs[2] = {nullptr, nullptr};
  uint32_t *base = nullptr, *cnt[2] = {nullptr, nullptr};
  FusedCtl *fctl = nullptr, *h_fctl = nullptr;
  ~SsspBinWork() {
    void *ps[] = {pbase, fill, dirty, hub_bm, queue[0], queue[1], hubs[0], hubs[1], bins, ctl,
                  pairs[0], pairs[1], base, cnt[0], cnt[1], fctl};
    for (void *q : ps)
      if (q) (void)hipFree(q);
    if (h_ctl) (void)hipHostFree(h_ctl);
    if (h_fctl) (void)hipHostFree(h_fctl);
  }
};

void free_sssp_bin(olv_graph *g) {
  delete g->sssp_bin;
  g->sssp_bin = nullptr;
}

static unsigned bgrid(size_t items, size_t per) {
  size_t b = (items + per - 1) / per;
  return (unsigned)std::max<size_t>(1, std::min(b, (size_t)cu_count() * 8));
}

template <uint32_t SB>
constexpr size_t apply_lds() {
  return ((size_t)1 << SB) * 4 + ((size_t)1 << SB) / 8;  // dist slice + dirty words
}

template <uint32_t SB>
static void launch_apply(const BinArgs &a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(ssspb_apply_kernel<SB>, dim3(grid), dim3(kBinB), apply_lds<SB>(), s, a);
}

// Allow the apply kernel's dynamic LDS (up to 132 KB) on the current device.
static int allow_apply_lds(uint32_t sb) {
  if (sb == 13)
    OLV_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(ssspb_apply_kernel<13>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)apply_lds<13>()));
  else if (sb == 14)
    OLV_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(ssspb_apply_kernel<14>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)apply_lds<14>()));
  else
    OLV_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(ssspb_apply_kernel<15>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)apply_lds<15>()));
  return OLV_OK;
}

// Work arrays for k instances on g (kept on the graph for the next call
// with the same k and hub threshold).
static int bin_work(olv_graph *g, uint32_t k, uint32_t sb, SsspBinWork **out) {
  SsspBinWork *w = g->sssp_bin;
  const uint32_t hub_thr = g->tune.hub_thr;
  if (w && !w->fused && w->k == k && w->hub_thr == hub_thr && w->sb == sb) {
    *out = w;
    return OLV_OK;
  }
  free_sssp_bin(g);
  w = new SsspBinWork();
  g->sssp_bin = w;
  w->k = k;
  w->hub_thr = hub_thr;
  w->sb = sb;
  w->nv = (uint64_t)k * g->n;
  w->nparts = (uint32_t)((w->nv + (1ull << sb) - 1) >> sb);
  w->nw = (uint32_t)((w->nv + 31) / 32);
  w->nwp = (w->nw + 63) / 64 * 64;
  auto fail_alloc = [&]() {
    free_sssp_bin(g);
    return fail(OLV_ENOMEM, "olv_sssp: device allocation for the binned relaxation failed");
  };
  const size_t hubcap = hub_capacity(g->m) * k;
  if (hipMalloc(reinterpret_cast<void **>(&w->pbase), (size_t)(w->nparts + 1) * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->fill), (size_t)w->nparts * kFillStride * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->dirty), (size_t)w->nwp * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->queue[0]), w->nv * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->queue[1]), w->nv * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->hubs[0]), hubcap * sizeof(uint2)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->hubs[1]), hubcap * sizeof(uint2)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->bins), std::max<uint64_t>((uint64_t)k * g->m, 1) * sizeof(uint2)) !=
          hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->ctl), sizeof(BinCtl)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&w->h_ctl), sizeof(BinCtl), hipHostMallocDefault) != hipSuccess)
    return fail_alloc();
  if (g->max_outdeg >= hub_thr) {
    const uint32_t hw = (g->n + 31) / 32;
    if (hipMalloc(reinterpret_cast<void **>(&w->hub_bm), (size_t)hw * 4) != hipSuccess) return fail_alloc();
    hipLaunchKernelGGL(ssspb_hub_kernel, dim3(bgrid(g->n, 256)), dim3(256), 0, stream(), g->row_ptr, g->n, hub_thr,
                       w->hub_bm);
    OLV_HIP(hipGetLastError());
  }
  // bin offsets: exclusive scan of the partition capacities (host, <= 1024)
  DevBuf cap;
  OLV_TRY(cap.alloc((size_t)w->nparts * 4));
  OLV_HIP(hipMemsetAsync(cap.p, 0, (size_t)w->nparts * 4, stream()));
  hipLaunchKernelGGL(ssspb_cap_kernel, dim3(bgrid((size_t)k * g->m, 1024)), dim3(1024), 0, stream(), g->col, g->m,
                     g->n, k, sb, w->nparts, cap.as<uint32_t>());
  OLV_HIP(hipGetLastError());
  std::vector<uint32_t> hc(w->nparts), hb(w->nparts + 1);
  OLV_HIP(hipMemcpyAsync(hc.data(), cap.p, (size_t)w->nparts * 4, hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  uint64_t run = 0;
  for (uint32_t p = 0; p < w->nparts; ++p) {
    hb[p] = (uint32_t)run;
    run += hc[p];
  }
  hb[w->nparts] = (uint32_t)run;
  if (run != (uint64_t)k * g->m) return fail(OLV_EINTERNAL, "olv_sssp: bin capacities %llu != %llu edges",
                                              (unsigned long long)run, (unsigned long long)k * g->m);
  OLV_HIP(hipMemcpy(w->pbase, hb.data(), (size_t)(w->nparts + 1) * 4, hipMemcpyHostToDevice));
  OLV_TRY(allow_apply_lds(sb));
  *out = w;
  return OLV_OK;
}

// Work for the fused rounds: region capacities (edges between every pair of
// slices), their offsets, two pair stores and two count matrices.
static int fused_work(olv_graph *g, uint32_t k, uint32_t sb, SsspBinWork **out) {
  SsspBinWork *w = g->sssp_bin;
  if (w && w->fused && w->k == k && w->sb == sb) {
    *out = w;
    return OLV_OK;
  }
  free_sssp_bin(g);
  w = new SsspBinWork();
  g->sssp_bin = w;
  w->fused = true;
  w->k = k;
  w->sb = sb;
  w->nv = (uint64_t)k * g->n;
  w->nparts = (uint32_t)((w->nv + (1u << sb) - 1) >> sb);
  w->nw = (uint32_t)((w->nv + 31) / 32);
  w->nwp = (w->nw + 63) / 64 * 64;
  const size_t P2 = (size_t)w->nparts * w->nparts;
  const size_t npairs = std::max<uint64_t>((uint64_t)k * g->m, 1);
  if (hipMalloc(reinterpret_cast<void **>(&w->dirty), (size_t)w->nwp * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->pairs[0]), npairs * sizeof(uint2)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->pairs[1]), npairs * sizeof(uint2)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->base), (P2 + 1) * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->cnt[0]), P2 * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->cnt[1]), P2 * 4) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&w->fctl), sizeof(FusedCtl)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void **>(&w->h_fctl), sizeof(FusedCtl), hipHostMallocDefault) != hipSuccess) {
    free_sssp_bin(g);
    return fail(OLV_ENOMEM, "olv_sssp: device allocation for the fused rounds failed");
  }
  DevBuf cap;
  OLV_TRY(cap.alloc(P2 * 4));
  OLV_HIP(hipMemsetAsync(cap.p, 0, P2 * 4, stream()));
  hipLaunchKernelGGL(ssspf_cap_kernel, dim3(w->nparts), dim3(1024), 0, stream(), g->row_ptr, g->col, g->n, k,
                     (uint32_t)w->nv, w->nparts, sb, cap.as<uint32_t>());
  OLV_HIP(hipGetLastError());
  std::vector<uint32_t> hc(P2), hb(P2 + 1);
  OLV_HIP(hipMemcpyAsync(hc.data(), cap.p, P2 * 4, hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  uint64_t run = 0;
  for (size_t i = 0; i < P2; ++i) {
    hb[i] = (uint32_t)run;
    run += hc[i];
  }
  hb[P2] = (uint32_t)run;
  if (run != (uint64_t)k * g->m)
    return fail(OLV_EINTERNAL, "olv_sssp: region capacities %llu != %llu edges", (unsigned long long)run,
                (unsigned long long)k * g->m);
  OLV_HIP(hipMemcpy(w->base, hb.data(), (P2 + 1) * 4, hipMemcpyHostToDevice));
  *out = w;
  return OLV_OK;
}

static int fused_run(const olv_graph *g, const uint32_t *d_srcs, uint32_t k, const int32_t *weights, int32_t *dist,
                     int32_t delta, unsigned flags, olv_stats *st) {
  olv_graph *gm = const_cast<olv_graph *>(g);
  SsspBinWork *w = nullptr;
  const uint64_t nv = (uint64_t)k * g->n;
  const uint32_t sb = 14;
  (void)nv;
  OLV_TRY(fused_work(gm, k, sb, &w));
  FusedArgs a;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.w = weights;
  a.dist = dist;
  a.dirty = w->dirty;
  a.pairs[0] = w->pairs[0];
  a.pairs[1] = w->pairs[1];
  a.base = w->base;
  a.cnt[0] = w->cnt[0];
  a.cnt[1] = w->cnt[1];
  a.ctl = w->fctl;
  a.n = g->n;
  a.k = k;
  a.nv = (uint32_t)w->nv;
  a.nw = w->nw;
  a.nparts = w->nparts;
  a.delta = delta;
  a.max_rounds = (uint32_t)std::min<uint64_t>(4ull * g->n + 1024, 0xffffffffull);
  hipStream_t s = stream();
  Timer t;
  t.start();
  hipLaunchKernelGGL(ssspf_reset_kernel,
                     dim3(bgrid(std::max<uint64_t>(w->nv, (uint64_t)w->nparts * w->nparts), 256)), dim3(256), 0, s, a);
  hipLaunchKernelGGL(ssspf_seed_kernel, dim3(1), dim3(256), 0, s, d_srcs, a);
  OLV_HIP(hipGetLastError());
  uint32_t hist = 8;
  for (uint32_t r : gm->sssp_rounds) hist = std::max(hist, std::min(r + 1, 4096u));
  uint32_t chunk = (flags & OLV_HOST_LOOP) ? 1u : hist;
  uint64_t enq = 0;
  for (;;) {
    for (uint32_t i = 0; i < chunk; ++i)
      hipLaunchKernelGGL(ssspf_round_kernel<14>, dim3(w->nparts), dim3(kFB), 0, s, a, (uint32_t)(enq + i));
    OLV_HIP(hipGetLastError());
    t.stop();
    enq += chunk;
    OLV_HIP(hipMemcpyAsync(w->h_fctl, w->fctl, sizeof(FusedCtl), hipMemcpyDeviceToHost, s));
    OLV_HIP(hipStreamSynchronize(s));
    if (w->h_fctl->done) break;
    if (enq > (uint64_t)a.max_rounds + 2) return fail(OLV_EINTERNAL, "olv_sssp: round loop did not terminate");
    if (!(flags & OLV_HOST_LOOP)) chunk = std::min(chunk * 2, 256u);
  }
  const FusedCtl &c = *w->h_fctl;
  for (int i = 3; i > 0; --i) gm->sssp_rounds[i] = gm->sssp_rounds[i - 1];
  gm->sssp_rounds[0] = c.rounds;
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = t.ms();
    st->iterations = c.rounds;
    st->edges_inspected = c.relax;
  }
  if (c.err) return fail(OLV_EINTERNAL, "olv_sssp: a pair region overflowed its capacity");
  if (c.capped) return fail(OLV_ENOCONV, "olv_sssp: still relaxing after %u rounds", c.rounds);
  return OLV_OK;
}

// Binned SSSP for k instances (sources d_srcs[0..k), on the device) with
// near-far step delta > 0.  *used = false (and nothing run) when the graph
// does not fit the scheme: more than 1024 partitions of 2^15 virtual
// vertices, or k x m >= 2^32 bin entries.
int sssp_bin_run(const olv_graph *g, const uint32_t *d_srcs, uint32_t k, const int32_t *weights, int32_t *dist,
                 int32_t delta, unsigned flags, olv_stats *st, bool *used) {
  *used = false;
  if (delta <= 0 || g->m == 0 || g->n == 0) return OLV_OK;
  const uint64_t nv = (uint64_t)k * g->n;
  if ((uint64_t)k * g->m >= (1ull << 32) || nv >= (1ull << 32)) return OLV_OK;
  // one kernel per round when no out-list needs splitting (hub items);
  // more instances than the fused kernel's 16M virtual vertices run in
  // groups of as many as fit
  const uint64_t fmax = (uint64_t)kFParts << 14;
  if (g->max_outdeg < g->tune.hub_thr && g->n <= fmax) {
    *used = true;
    const uint32_t gk = (uint32_t)std::min<uint64_t>(k, fmax / g->n);
    if (gk == k) return fused_run(g, d_srcs, k, weights, dist, delta, flags, st);
    olv_stats tot;
    std::memset(&tot, 0, sizeof(tot));
    for (uint32_t s0 = 0; s0 < k; s0 += gk) {
      const uint32_t kk = std::min(gk, k - s0);
      olv_stats one;
      OLV_TRY(fused_run(g, d_srcs + s0, kk, weights, dist + (size_t)s0 * g->n, delta, flags, &one));
      tot.elapsed_ms += one.elapsed_ms;
      tot.iterations += one.iterations;
      tot.edges_inspected += one.edges_inspected;
    }
    if (st) *st = tot;
    return OLV_OK;
  }
  uint32_t sb = kBinSbMin;
  while (sb < kBinSbMax && ((nv + (1ull << sb) - 1) >> sb) > kBinPartMax) ++sb;
  if (((nv + (1ull << sb) - 1) >> sb) > kBinPartMax) return OLV_OK;
  olv_graph *gm = const_cast<olv_graph *>(g);
  SsspBinWork *w = nullptr;
  OLV_TRY(bin_work(gm, k, sb, &w));
  *used = true;
  BinArgs a;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.w = weights;
  a.dist = dist;
  a.dirty = w->dirty;
  a.hub_bm = w->hub_bm;
  a.queue[0] = w->queue[0];
  a.queue[1] = w->queue[1];
  a.hubs[0] = w->hubs[0];
  a.hubs[1] = w->hubs[1];
  a.bins = w->bins;
  a.pbase = w->pbase;
  a.fill = w->fill;
  a.ctl = w->ctl;
  a.n = g->n;
  a.k = k;
  a.nv = (uint32_t)nv;
  a.nw = w->nw;
  a.nparts = w->nparts;
  a.sb = sb;
  a.hub_thr = w->hub_thr;
  a.delta = delta;
  a.max_rounds = (uint32_t)std::min<uint64_t>(4ull * g->n + 1024, 0xffffffffull);
  hipStream_t s = stream();
  const unsigned ga = (unsigned)cu_count() * 2;  // two expand workgroups per CU (VGPR-bound)
  const unsigned gb = std::min<uint32_t>(w->nparts, (uint32_t)cu_count() * 2);
  Timer t;
  t.start();
  hipLaunchKernelGGL(ssspb_reset_kernel, dim3(bgrid(std::max<uint64_t>(nv, w->nparts * kFillStride), 256)),
                     dim3(256), 0, s, a);
  hipLaunchKernelGGL(ssspb_seed_kernel, dim3(1), dim3(256), 0, s, d_srcs, a);
  OLV_HIP(hipGetLastError());
  // first chunk: the longest of the last runs (a run of r rounds takes r + 1
  // launch pairs: the last one finds nothing left)
  uint32_t hist = 8;
  for (uint32_t r : gm->sssp_rounds) hist = std::max(hist, std::min(r + 1, 4096u));
  uint32_t chunk = (flags & OLV_HOST_LOOP) ? 1u : hist;
  uint64_t enq = 0;
  for (;;) {
    for (uint32_t i = 0; i < chunk; ++i) {
      hipLaunchKernelGGL(ssspb_expand_kernel, dim3(ga), dim3(kBinA), 0, s, a);
      if (sb == 13)
        launch_apply<13>(a, gb, s);
      else if (sb == 14)
        launch_apply<14>(a, gb, s);
      else
        launch_apply<15>(a, gb, s);
    }
    OLV_HIP(hipGetLastError());
    t.stop();
    enq += chunk;
    OLV_HIP(hipMemcpyAsync(w->h_ctl, w->ctl, sizeof(BinCtl), hipMemcpyDeviceToHost, s));
    OLV_HIP(hipStreamSynchronize(s));
    if (w->h_ctl->done) break;
    if (enq > (uint64_t)a.max_rounds + 2) return fail(OLV_EINTERNAL, "olv_sssp: round loop did not terminate");
    if (!(flags & OLV_HOST_LOOP)) chunk = std::min(chunk * 2, 256u);
  }
  const BinCtl &c = *w->h_ctl;
  for (int i = 3; i > 0; --i) gm->sssp_rounds[i] = gm->sssp_rounds[i - 1];
  gm->sssp_rounds[0] = c.rounds;
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = t.ms();
    st->iterations = c.rounds;
    st->edges_inspected = c.relax;
  }
  if (c.err) return fail(OLV_EINTERNAL, "olv_sssp: a bin overflowed its capacity");
  if (c.capped) return fail(OLV_ENOCONV, "olv_sssp: still relaxing after %u rounds", c.rounds);
  return OLV_OK;
}

}  // namespace olv
