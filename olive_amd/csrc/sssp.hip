// Single- and multi-source shortest paths over int32 weights: frontier
// relaxation.
//
// The reference SSSP (SSSP.cu:146-163) alternates edgeFilter (reset every
// accumulator to INF, gather dist[u] + w, atomicMin into acc[v], flag v;
// oliver.h:56-83) and vertexFilter (if dist[v] > acc[v]: dist[v] = acc[v],
// enqueue v; oliver.h:92-113) until the queue is empty, capped at n rounds
// (negative-cycle guard, SSSP.cu:153-157).  Shortest distances are unique,
// so relaxing straight into dist with atomicMin -- no accumulator array, no
// per-round O(V) reset -- reaches the same fixed point: a vertex whose
// distance dropped this round is marked in a bitmap and expanded next round.
//
// Near-far ordering (Davidson et al., IPDPS'14): with a split distance T, an
// improved vertex joins the next round's frontier only if its new distance
// is below T; otherwise its bit goes to a far bitmap.  When the near
// frontier runs dry, T grows by delta and the far vertices now below T (read
// from dist) form the next frontier.  Vertices are then mostly expanded once
// with their final distance instead of once per improvement, and the fixed
// point -- hence every distance -- is the same.  Negative weights turn the
// ordering off (plain rounds, so the reference's n-round negative-cycle rule
// applies unchanged).
//
// k sources run as k instances over one shared CSR: instance s owns virtual
// vertices s*n .. s*n+n-1 (dist row s), so one launch per round relaxes all
// k frontiers together.  Rounds are a relax kernel and a compact kernel
// whose control lives on the device (as for BFS); OLV_PERSISTENT runs all
// rounds (plain, without the near-far split) in one cooperative kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <type_traits>
#include <vector>

#include "bfs_core.hpp"
#include "internal.hpp"

namespace olv {

constexpr int kSBlock = 256;
constexpr uint32_t kSSpin = 1u << 24;

template <bool Multi>
struct SsspVisitor {
  using P = typename std::conditional<Multi, unsigned long long, int32_t>::type;
  const uint32_t *col;
  const int32_t *w;
  int32_t *dist;
  uint32_t *next_bm;
  uint32_t *far_bm;  // nullptr: plain rounds
  int32_t split;     // near-far split distance T
  uint32_t n;
  uint32_t nf = 0;
  __device__ SsspVisitor(const uint32_t *c, const int32_t *ww, int32_t *d, uint32_t *nb, uint32_t *fb, int32_t t,
                         uint32_t nn)
      : col(c), w(ww), dist(d), next_bm(nb), far_bm(fb), split(t), n(nn) {}
  __device__ __forceinline__ uint32_t row(uint32_t q) const { return Multi ? q % n : q; }
  __device__ __forceinline__ P load(uint32_t q) const {
    if constexpr (Multi)
      return ((unsigned long long)(q - q % n) << 32) | (unsigned long long)(uint32_t)dist[q];
    else
      return dist[q];
  }
  // target of edge e: (vertex, candidate distance)
  __device__ __forceinline__ uint2 tgt(uint32_t e, P pl) const {
    int32_t du;
    uint32_t base;
    if constexpr (Multi) {
      du = (int32_t)(uint32_t)pl;
      base = (uint32_t)(pl >> 32);
    } else {
      du = pl;
      base = 0;
    }
    return make_uint2(base + col[e], (uint32_t)du + (uint32_t)w[e]);
  }
  __device__ __forceinline__ int32_t state(uint2 t) const { return dist[t.x]; }
  // Relax: lower dist[v] and mark v.  Both atomics are non-returning (no
  // lane waits on them): a relaxation that loses a race to a lower distance
  // still marks v, which only re-expands v with its final distance, and nf
  // may count a vertex more than once -- it only has to be non-zero exactly
  // when something improved.
  __device__ __forceinline__ void apply(uint2 t, int32_t cur, P) {
    const uint32_t v = t.x;
    const int32_t cand = (int32_t)t.y;
    if (cand >= cur) return;
    __hip_atomic_fetch_min(&dist[v], cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t bit = 1u << (v & 31);
    uint32_t *bm = (far_bm && cand >= split) ? far_bm : next_bm;
    __hip_atomic_fetch_or(&bm[v >> 5], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (bm == next_bm) nf += 1;
  }
};

struct SsspCtl {
  uint32_t round;     // next round to relax (written by compact)
  uint32_t round_e;   // round being compacted (written by relax)
  uint32_t done;
  uint32_t rounds;    // result
  uint32_t capped;    // result: n rounds without convergence
  uint32_t pad;
  uint32_t qfill[2];  // frontier queue length of round R, slot R & 1
  uint32_t hfill[2];  // hub items of round R, slot R & 1
  unsigned long long relax;  // result: edges read
  int32_t split[2];   // near-far split T of round R, slot R & 1
  uint32_t promo[2];  // round R's frontier came from the far bitmap
  uint32_t kept[2];   // far bits left by the promotion that built round R
  uint32_t phases;    // result: near-far phases
  uint32_t pad2;
  CtrSet cnt[3];
};

struct SsspParams {
  const uint32_t *row_ptr;
  const uint32_t *col;
  const int32_t *w;
  int32_t *dist;     // ninst x n
  uint32_t *queue;   // virtual ids
  uint2 *hubs;       // hub items (virtual id, chunk)
  uint32_t *bms;     // 2 x nwp
  SsspCtl *ctl;
  GridBar *bar;
  uint32_t *far_bm;  // nwp words, or nullptr for plain rounds
  int32_t delta;     // near-far step (0: plain rounds)
  uint32_t n, ninst, nw, nwp;
  uint32_t wave_thr, block_thr, hub_thr;
  uint32_t max_rounds;
};

// Virtual id -> CSR row.
struct ModRow {
  uint32_t n;
  __device__ __forceinline__ uint32_t operator()(uint32_t v) const { return v % n; }
};

template <bool Multi>
__device__ __forceinline__ void sssp_round(const SsspParams &p, uint32_t qsize, uint32_t nhubs, uint32_t *nextb,
                                           int32_t split, uint32_t vb, uint32_t nvb, uint32_t &nf,
                                           unsigned long long &insp) {
  SsspVisitor<Multi> vis(p.col, p.w, p.dist, nextb, p.far_bm, split, p.n);
  push_expand<kSBlock>(p.queue, qsize, p.hubs, nhubs, p.row_ptr, p.wave_thr, p.block_thr, p.hub_thr, vb, nvb, vis,
                       insp);
  nf += vis.nf;
}

template <bool Multi>
__global__ void __launch_bounds__(kSBlock) sssp_relax_kernel(SsspParams p) {
  SsspCtl *ctl = p.ctl;
  if (__hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const uint32_t R = ctl->round;
  const uint32_t qsize = ctl->qfill[R & 1], nhubs = ctl->hfill[R & 1];
  uint32_t *nextb = p.bms + (size_t)(R & 1) * p.nwp;
  uint32_t *clearb = p.bms + (size_t)((R + 1) & 1) * p.nwp;
  for (uint32_t i = blockIdx.x * kSBlock + threadIdx.x; i < p.nwp; i += gridDim.x * kSBlock) clearb[i] = 0u;
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      ctl->round_e = R;
      ctl->qfill[(R + 1) & 1] = 0;
      ctl->hfill[(R + 1) & 1] = 0;
      ctl->kept[(R + 1) & 1] = 0;
    }
    ctr_clear(&ctl->cnt[(R + 1) % 3]);
  }
  uint32_t nf = 0;
  unsigned long long insp = 0;
  sssp_round<Multi>(p, qsize, nhubs, nextb, ctl->split[R & 1], blockIdx.x, gridDim.x, nf, insp);
  flush_counters<kSBlock>(&ctl->cnt[R % 3], nf, 0ull, insp);
}

__global__ void __launch_bounds__(kSBlock) sssp_compact_kernel(SsspParams p) {
  SsspCtl *ctl = p.ctl;
  if (__hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const uint32_t R = ctl->round_e;
  const CtrSum c = ctr_total(&ctl->cnt[R % 3]);
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (lead) ctl->relax += c.inspected;
  const int32_t split = ctl->split[R & 1];
  // near frontier dry: finished in plain mode, or when the far bitmap was
  // found empty by the promotion that built this round's (empty) frontier
  const bool dry = c.nf == 0;
  const bool finished =
      dry && (p.delta == 0 || (ctl->promo[R & 1] && ctl->qfill[R & 1] == 0 && ctl->hfill[R & 1] == 0 &&
                               ctl->kept[R & 1] == 0));
  if (finished || R + 1 >= p.max_rounds) {
    if (lead) {
      ctl->rounds = R + 1;
      ctl->capped = !finished;
      ctl->done = 1u;
    }
    return;
  }
  if (!dry) {
    if (lead) {
      ctl->split[(R + 1) & 1] = split;
      ctl->promo[(R + 1) & 1] = 0;
      ctl->round = R + 1;
    }
    compact_step<kSBlock>(p.bms + (size_t)(R & 1) * p.nwp, p.nw, p.row_ptr, p.hub_thr, p.queue,
                          &ctl->qfill[(R + 1) & 1], p.hubs, &ctl->hfill[(R + 1) & 1], blockIdx.x, gridDim.x,
                          ModRow{p.n});
    return;
  }
  // promotion: T += delta; far vertices now below T become the frontier
  const int32_t nsplit = split > kInf - p.delta ? kInf : split + p.delta;
  if (lead) {
    ctl->split[(R + 1) & 1] = nsplit;
    ctl->promo[(R + 1) & 1] = 1;
    ctl->phases += 1;
    ctl->round = R + 1;
  }
  uint32_t kept = 0;
  const int32_t *dist = p.dist;
  uint32_t *far = p.far_bm;
  auto take = [&](uint32_t w) -> uint32_t {
    const uint32_t word = far[w];
    uint32_t sel = 0;
    for (uint32_t x = word; x; x &= x - 1) {
      const uint32_t b = __ffs(x) - 1;
      if (dist[w * 32u + b] < nsplit) sel |= 1u << b;
    }
    if (sel) far[w] = word & ~sel;
    kept += __popc(word & ~sel);
    return sel;
  };
  compact_words<kSBlock>(p.nw, take, p.row_ptr, p.hub_thr, p.queue, &ctl->qfill[(R + 1) & 1], p.hubs,
                         &ctl->hfill[(R + 1) & 1], blockIdx.x, gridDim.x, ModRow{p.n});
  __shared__ unsigned long long s_red[kSBlock / kWave];
  const unsigned long long tk = block_sum<kSBlock, unsigned long long>(kept, s_red);
  if (threadIdx.x == 0 && tk) atomicAdd(&ctl->kept[(R + 1) & 1], (uint32_t)tk);
}

template <bool Multi>
__global__ void __launch_bounds__(kSBlock) sssp_persistent_kernel(SsspParams p) {
  GridBarLocal gb;
  if (!gbar_init(p.bar, gb, gridDim.x, kSSpin)) return;
  SsspCtl *ctl = p.ctl;
  const uint32_t vb = blockIdx.x, nvb = gridDim.x;
  const bool lead = vb == 0 && threadIdx.x == 0;
  unsigned long long insp_all = 0;
  for (uint32_t R = 0;; ++R) {
    uint32_t *nextb = p.bms + (size_t)(R & 1) * p.nwp;
    uint32_t *clearb = p.bms + (size_t)((R + 1) & 1) * p.nwp;
    const uint32_t qsize = __hip_atomic_load(&ctl->qfill[R & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t nhubs = __hip_atomic_load(&ctl->hfill[R & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (uint32_t i = vb * kSBlock + threadIdx.x; i < p.nwp; i += nvb * kSBlock) clearb[i] = 0u;
    if (vb == 0) {
      if (threadIdx.x == 0) {
        ctl->qfill[(R + 1) & 1] = 0;
        ctl->hfill[(R + 1) & 1] = 0;
      }
      ctr_clear(&ctl->cnt[(R + 1) % 3]);
    }
    uint32_t nf = 0;
    unsigned long long insp = 0;
    sssp_round<Multi>(p, qsize, nhubs, nextb, kInf, vb, nvb, nf, insp);
    flush_counters<kSBlock>(&ctl->cnt[R % 3], nf, 0ull, insp);
    if (!grid_barrier(p.bar, gb, kSSpin)) break;
    const CtrSum c = ctr_total(&ctl->cnt[R % 3]);
    insp_all += c.inspected;
    if (c.nf == 0 || R + 1 >= p.max_rounds) {
      if (lead) {
        ctl->rounds = R + 1;
        ctl->capped = c.nf != 0;
        ctl->relax = insp_all;
        ctl->done = 1u;
      }
      break;
    }
    compact_step<kSBlock>(nextb, p.nw, p.row_ptr, p.hub_thr, p.queue, &ctl->qfill[(R + 1) & 1], p.hubs,
                          &ctl->hfill[(R + 1) & 1], vb, nvb, ModRow{p.n});
    if (!grid_barrier(p.bar, gb, kSSpin)) break;
  }
}

// dist = INF, bitmaps (near x2, far)/control/barrier = 0, split = delta.
__global__ void sssp_reset_kernel(int32_t *__restrict__ dist, size_t total, uint32_t *__restrict__ bms,
                                  uint32_t nwp, SsspCtl *__restrict__ ctl, GridBar *__restrict__ bar, int32_t delta) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = tid; i < total; i += stride) dist[i] = kInf;
  for (size_t i = tid; i < (size_t)nwp * 3; i += stride) bms[i] = 0u;
  for (size_t i = tid; i < sizeof(SsspCtl) / 4; i += stride) reinterpret_cast<uint32_t *>(ctl)[i] = 0u;
  for (size_t i = tid; i < sizeof(GridBar) / 4; i += stride) reinterpret_cast<uint32_t *>(bar)[i] = 0u;
  if (tid == 0) ctl->split[0] = delta > 0 ? delta : kInf;
}

// One workgroup: instance s starts from virtual vertex s*n + srcs[s].
__global__ void sssp_seed_kernel(const uint32_t *__restrict__ srcs, uint32_t k, SsspParams p) {
  auto claim = [&](uint32_t s) -> uint32_t {
    const uint32_t x = s * p.n + srcs[s];
    p.dist[x] = 0;
    return x;
  };
  unsigned long long deg;
  seed_frontier(srcs, k, p.row_ptr, p.hub_thr, p.queue, &p.ctl->qfill[0], p.hubs, &p.ctl->hfill[0], claim,
                ModRow{p.n}, &deg);
}

struct SsspWork {
  uint32_t nw = 0, nwp = 0;
  uint32_t *queue = nullptr, *bms = nullptr;
  uint2 *hubs = nullptr;
  SsspCtl *ctl = nullptr, *h_ctl = nullptr;
  GridBar *bar = nullptr;
  ~SsspWork() {
    void *ps[] = {queue, bms, hubs, ctl, bar};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    if (h_ctl) (void)hipHostFree(h_ctl);
  }
  int alloc(uint64_t nv, uint32_t k, uint64_t m) {
    nw = (uint32_t)((nv + 31) / 32);
    nwp = (nw + 63) / 64 * 64;
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&queue), (size_t)nv * 4));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&hubs), hub_capacity(m) * k * sizeof(uint2)));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&bms), (size_t)nwp * 12));  // near x2, far
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&ctl), sizeof(SsspCtl)));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&bar), sizeof(GridBar)));
    OLV_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_ctl), sizeof(SsspCtl), hipHostMallocDefault));
    return OLV_OK;
  }
};

static unsigned sgrid(size_t items, size_t per) {
  size_t b = (items + per - 1) / per;
  return (unsigned)std::max<size_t>(1, std::min(b, (size_t)cu_count() * 8));
}

// Minimum and mean of the int32 weights (one reduction, read back).
__global__ void wstat_kernel(const int32_t *__restrict__ w, uint32_t m, long long *__restrict__ out) {
  __shared__ long long s_red[256 / kWave];
  long long mn = LLONG_MAX, sum = 0;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) {
    const long long x = w[e];
    mn = x < mn ? x : mn;
    sum += x;
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const long long y = __shfl_xor(mn, o, kWave);
    mn = y < mn ? y : mn;
  }
  if (lane_id() == 0) s_red[threadIdx.x / kWave] = mn;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 256 / kWave; ++i) mn = s_red[i] < mn ? s_red[i] : mn;
    atomicMin(&out[0], mn);
  }
  __syncthreads();
  sum = block_sum<256, long long>(sum, s_red);
  if (threadIdx.x == 0) atomicAdd(reinterpret_cast<unsigned long long *>(&out[1]), (unsigned long long)sum);
}

static int weight_stats(const int32_t *w, uint32_t m, int32_t *wmin, double *wmean) {
  long long h[2] = {LLONG_MAX, 0};
  DevBuf b;
  OLV_TRY(b.alloc(sizeof(h)));
  OLV_HIP(hipMemcpyAsync(b.p, h, sizeof(h), hipMemcpyHostToDevice, stream()));
  hipLaunchKernelGGL(wstat_kernel, dim3(sgrid(m, 256)), dim3(256), 0, stream(), w, m, b.as<long long>());
  OLV_HIP(hipMemcpyAsync(h, b.p, sizeof(h), hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  *wmin = m ? (int32_t)h[0] : 0;
  *wmean = m ? (double)h[1] / m : 0.0;
  return OLV_OK;
}

// Near-far step: mean weight x the graph's sssp_delta_num (32 unless tuned)
// / mean out-degree; 0 (plain rounds) when a weight is negative.
static int sssp_delta(const olv_graph *g, const int32_t *w, int32_t *delta) {
  *delta = 0;
  if (!g->m || !g->n) return OLV_OK;
  int32_t wmin = 0;
  double wmean = 0.0;
  olv_graph *mg = const_cast<olv_graph *>(g);
  if (w == g->weights && mg->wstat_ready) {
    wmin = mg->wmin;
    wmean = mg->wmean;
  } else {
    OLV_TRY(weight_stats(w, g->m, &wmin, &wmean));
    if (w == g->weights) {
      mg->wmin = wmin;
      mg->wmean = wmean;
      mg->wstat_ready = true;
    }
  }
  if (wmin < 0) return OLV_OK;
  const double deg = (double)g->m / g->n;
  const double d = wmean * g->tune.sssp_delta_num / (deg > 1.0 ? deg : 1.0);
  *delta = d < 1.0 ? 1 : (d > 1e9 ? 1000000000 : (int32_t)(d + 0.5));
  return OLV_OK;
}

static int sssp_read(SsspWork &w) {
  OLV_HIP(hipMemcpyAsync(w.h_ctl, w.ctl, sizeof(SsspCtl), hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  return OLV_OK;
}

// k instances from device sources d_srcs[0..k); dist is k x n.
static int sssp_run(const olv_graph *g, const uint32_t *d_srcs, uint32_t k, const int32_t *weights, int32_t *dist,
                    unsigned flags, olv_stats *st) {
  const uint64_t nv = (uint64_t)k * g->n;
  if (nv >= (1ull << 32)) return fail(OLV_ERANGE, "sssp: %u sources x %u vertices exceed 2^32 ids", k, g->n);
  const bool multi = k > 1;
  int32_t delta = 0;
  if (!(flags & OLV_PERSISTENT)) OLV_TRY(sssp_delta(g, weights, &delta));
  // destination binning (sssp_bin.hip) unless asked to relax directly
  if (delta > 0 && !(flags & (OLV_PERSISTENT | OLV_SSSP_DIRECT))) {
    bool used = false;
    OLV_TRY(sssp_bin_run(g, d_srcs, k, weights, dist, delta, flags, st, &used));
    if (used) return OLV_OK;
  }
  SsspWork w;
  OLV_TRY(w.alloc(nv, k, g->m));
  SsspParams p;
  p.row_ptr = g->row_ptr;
  p.col = g->col;
  p.w = weights;
  p.dist = dist;
  p.queue = w.queue;
  p.hubs = w.hubs;
  p.bms = w.bms;
  p.ctl = w.ctl;
  p.bar = w.bar;
  p.n = g->n;
  p.ninst = k;
  p.nw = w.nw;
  p.nwp = w.nwp;
  p.wave_thr = g->tune.td_wave_thr;
  p.block_thr = g->tune.td_block_thr;
  p.hub_thr = g->tune.hub_thr;
  p.delta = delta;
  p.far_bm = delta ? w.bms + (size_t)2 * w.nwp : nullptr;
  // plain rounds keep the reference's negative-cycle rule (n rounds); with
  // non-negative weights the cap only guards against a runaway loop
  p.max_rounds = delta ? (uint32_t)std::min<uint64_t>(4ull * g->n + 1024, 0xffffffffull) : g->n;
  hipStream_t s = stream();
  Timer t;
  t.start();
  hipLaunchKernelGGL(sssp_reset_kernel, dim3(sgrid(std::max<size_t>(nv, (size_t)w.nwp * 3), 256)), dim3(256), 0, s,
                     dist, (size_t)nv, w.bms, w.nwp, w.ctl, w.bar, delta);
  hipLaunchKernelGGL(sssp_seed_kernel, dim3(1), dim3(256), 0, s, d_srcs, k, p);
  OLV_HIP(hipGetLastError());
  if (flags & OLV_PERSISTENT) {
    const void *fn = multi ? reinterpret_cast<const void *>(sssp_persistent_kernel<true>)
                           : reinterpret_cast<const void *>(sssp_persistent_kernel<false>);
    int per_cu = 0;
    OLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kSBlock, 0));
    if (per_cu <= 0) return fail(OLV_EUNSUPPORTED, "persistent SSSP kernel does not fit on a CU");
    void *args[] = {&p};
    OLV_HIP(hipLaunchCooperativeKernel(fn, dim3(cu_count()), dim3(kSBlock), args, 0, s));
    t.stop();
    OLV_TRY(sssp_read(w));
    GridBar hb;
    OLV_HIP(hipMemcpy(&hb, w.bar, sizeof(hb), hipMemcpyDeviceToHost));
    if (hb.err || !w.h_ctl->done) return fail(OLV_EINTERNAL, "olv_sssp: grid barrier timed out (co-residency lost)");
  } else {
    // relax workgroups per CU: the graph's tuning (uniform 2^22 x 2^25 graph:
    // 6 gives 1.99 ms; 4: 2.00, 8: 2.03, 12: 2.01, 16: 2.05)
    const unsigned gr = (unsigned)cu_count() * g->tune.sssp_relax_per_cu;
    const unsigned gc = sgrid(w.nw, kSBlock);
    // first chunk: the longest of the last few runs on this graph (8 at
    // first); longer runs continue in doubling chunks after a host check
    olv_graph *gm = const_cast<olv_graph *>(g);
    uint32_t hist = 8;
    for (uint32_t r : gm->sssp_rounds) hist = std::max(hist, std::min(r, 4096u));
    uint32_t chunk = (flags & OLV_HOST_LOOP) ? 1u : hist;
    uint64_t enq = 0;
    for (;;) {
      for (uint32_t i = 0; i < chunk; ++i) {
        if (multi)
          hipLaunchKernelGGL(sssp_relax_kernel<true>, dim3(gr), dim3(kSBlock), 0, s, p);
        else
          hipLaunchKernelGGL(sssp_relax_kernel<false>, dim3(gr), dim3(kSBlock), 0, s, p);
        hipLaunchKernelGGL(sssp_compact_kernel, dim3(gc), dim3(kSBlock), 0, s, p);
      }
      OLV_HIP(hipGetLastError());
      t.stop();  // ahead of the read-back: host latency after the last round is not timed
      enq += chunk;
      OLV_TRY(sssp_read(w));
      if (w.h_ctl->done) break;
      // the device stops by itself after p.max_rounds rounds (capped)
      if (enq > (uint64_t)p.max_rounds + 2) return fail(OLV_EINTERNAL, "olv_sssp: round loop did not terminate");
      if (!(flags & OLV_HOST_LOOP)) chunk = std::min(chunk * 2, 256u);
    }
  }
  const double ms = t.ms();
  {
    olv_graph *gm = const_cast<olv_graph *>(g);
    for (int i = 3; i > 0; --i) gm->sssp_rounds[i] = gm->sssp_rounds[i - 1];
    gm->sssp_rounds[0] = w.h_ctl->rounds;
  }
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = w.h_ctl->rounds;
    st->edges_inspected = w.h_ctl->relax;
  }
  if (w.h_ctl->capped) return fail(OLV_ENOCONV, "olv_sssp: still relaxing after n=%u rounds (negative cycle)", g->n);
  return OLV_OK;
}

__global__ void reach_stat_kernel(const int32_t *__restrict__ d, const uint32_t *__restrict__ rp, uint32_t n,
                                  size_t total, unsigned long long *__restrict__ out) {
  __shared__ unsigned long long s_red[256 / kWave];
  unsigned long long vr = 0, er = 0;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x)
    if (d[x] != kInf) {
      const uint32_t v = (uint32_t)(x % n);
      vr += 1;
      er += rp[v + 1] - rp[v];
    }
  vr = block_sum<256, unsigned long long>(vr, s_red);
  er = block_sum<256, unsigned long long>(er, s_red);
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], vr);
    atomicAdd(&out[1], er);
  }
}

// Reached (vertex, source) pairs and the sum of their out-degrees.
int reach_stats(const olv_graph *g, const int32_t *d, uint32_t k, olv_stats *st) {
  DevBuf b;
  OLV_TRY(b.alloc(16));
  OLV_HIP(hipMemsetAsync(b.p, 0, 16, stream()));
  const size_t total = (size_t)k * g->n;
  hipLaunchKernelGGL(reach_stat_kernel, dim3(sgrid(total, 256)), dim3(256), 0, stream(), d, g->row_ptr, g->n, total,
                     b.as<unsigned long long>());
  unsigned long long h[2];
  OLV_HIP(hipMemcpyAsync(h, b.p, 16, hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  st->vertices_reached = h[0];
  st->edges_traversed = h[1];
  return OLV_OK;
}

static int sssp_entry(const olv_graph *g, const uint32_t *srcs, uint32_t k, const int32_t *weights, int32_t *dist,
                      unsigned flags, olv_stats *st) {
  if (!g || !dist || !srcs) return fail(OLV_EINVAL, "olv_sssp: NULL argument");
  if (k == 0) return fail(OLV_EINVAL, "olv_sssp: no source");
  const bool host = (flags & OLV_PTR_HOST) != 0;
  for (uint32_t i = 0; i < k; ++i)
    if (srcs[i] >= g->n) return fail(OLV_EINVAL, "olv_sssp: source %u out of range (n=%u)", srcs[i], g->n);
  DevBuf dw, dd, ds;
  const int32_t *w = weights;
  if (!w) {
    if (g->wdtype != OLV_W_I32) return fail(OLV_EINVAL, "olv_sssp: graph has no int32 weights and none were given");
    w = static_cast<const int32_t *>(g->weights);
  } else if (host) {
    OLV_TRY(dw.alloc((size_t)g->m * 4));
    if (g->m) OLV_HIP(hipMemcpy(dw.p, weights, (size_t)g->m * 4, hipMemcpyHostToDevice));
    w = dw.as<int32_t>();
  }
  const size_t total = (size_t)k * g->n;
  int32_t *out = dist;
  if (host) {
    OLV_TRY(dd.alloc(total * 4));
    out = dd.as<int32_t>();
  }
  OLV_TRY(ds.alloc((size_t)k * 4));
  OLV_HIP(hipMemcpy(ds.p, srcs, (size_t)k * 4, hipMemcpyHostToDevice));
  OLV_TRY(sssp_run(g, ds.as<uint32_t>(), k, w, out, flags, st));
  if (st) OLV_TRY(reach_stats(g, out, k, st));
  if (host) OLV_HIP(hipMemcpy(dist, out, total * 4, hipMemcpyDeviceToHost));
  return OLV_OK;
}

}  // namespace olv

using namespace olv;

extern "C" int olv_sssp(const olv_graph *g, uint32_t src, const int32_t *weights, int32_t *dist, unsigned flags,
                        olv_stats *st) {
  if (!g) return fail(OLV_EINVAL, "olv_sssp: NULL graph");
  return sssp_entry(g, &src, 1, weights, dist, flags, st);
}

extern "C" int olv_sssp_batched(const olv_graph *g, const uint32_t *srcs, uint32_t k, const int32_t *weights,
                                int32_t *dist, unsigned flags, olv_stats *st) {
  return sssp_entry(g, srcs, k, weights, dist, flags, st);
}
