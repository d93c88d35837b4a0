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
// k sources run as k instances over one shared CSR: instance s owns virtual
// vertices s*n .. s*n+n-1 (dist row s), so one persistent launch relaxes all
// k frontiers together and each CSR row is read once per round for every
// instance that has it on its frontier.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bfs_core.hpp"
#include "internal.hpp"

namespace olv {

constexpr int kSBlock = 256;
constexpr uint32_t kSSpin = 1u << 26;

template <bool Multi>
struct SsspVisitor {
  using P = typename std::conditional<Multi, unsigned long long, int32_t>::type;
  const uint32_t *col;
  const int32_t *w;
  int32_t *dist;
  uint32_t *next_bm;
  uint32_t n;
  uint32_t nf = 0;
  __device__ SsspVisitor(const uint32_t *c, const int32_t *ww, int32_t *d, uint32_t *nb, uint32_t nn)
      : col(c), w(ww), dist(d), next_bm(nb), n(nn) {}
  __device__ __forceinline__ uint32_t row(uint32_t q) const { return Multi ? q % n : q; }
  __device__ __forceinline__ P load(uint32_t q) const {
    if constexpr (Multi)
      return ((unsigned long long)(q - q % n) << 32) | (unsigned long long)(uint32_t)dist[q];
    else
      return dist[q];
  }
  __device__ __forceinline__ void visit(uint32_t e, P pl) {
    int32_t du;
    uint32_t base;
    if constexpr (Multi) {
      du = (int32_t)(uint32_t)pl;
      base = (uint32_t)(pl >> 32);
    } else {
      du = pl;
      base = 0;
    }
    const uint32_t v = base + col[e];
    const int32_t cand = (int32_t)((uint32_t)du + (uint32_t)w[e]);
    if (cand >= dist[v]) return;
    const int32_t old = atomicMin(&dist[v], cand);
    if (cand >= old) return;
    const uint32_t bit = 1u << (v & 31);
    if (!(atomicOr(&next_bm[v >> 5], bit) & bit)) nf += 1;
  }
};

struct SsspParams {
  const uint32_t *row_ptr;
  const uint32_t *col;
  const int32_t *w;
  int32_t *dist;     // ninst x n
  uint32_t *queue;   // virtual ids
  uint32_t *bms;     // 2 x nwp
  BfsCtr *ctr;       // 3 slots
  uint32_t *bar;     // [0..2] barrier, [4] rounds, [6] hit cap
  unsigned long long *relax;  // edges read
  uint32_t n, ninst, nw, nwp;
  uint32_t wave_thr, block_thr;
  uint32_t max_rounds;
};

template <bool Multi>
__device__ __forceinline__ void sssp_round(const SsspParams &p, uint32_t qsize, uint32_t *nextb, uint32_t vb,
                                           uint32_t nvb, uint32_t &nf, unsigned long long &insp) {
  SsspVisitor<Multi> vis(p.col, p.w, p.dist, nextb, p.n);
  push_expand<kSBlock>(p.queue, qsize, p.row_ptr, p.wave_thr, p.block_thr, vb, nvb, vis, insp);
  nf += vis.nf;
}

template <bool Multi>
__global__ void __launch_bounds__(kSBlock) sssp_persistent_kernel(SsspParams p) {
  uint32_t gen = 0;
  const uint32_t vb = blockIdx.x, nvb = gridDim.x;
  uint32_t R = 0;
  unsigned long long insp_all = 0;
  for (;;) {
    BfsCtr *cur = &p.ctr[R % 3];
    BfsCtr *nxt = &p.ctr[(R + 1) % 3];
    uint32_t *nextb = p.bms + (size_t)(R & 1) * p.nwp;
    uint32_t *clearb = p.bms + (size_t)((R + 1) & 1) * p.nwp;
    if (vb == 0 && threadIdx.x < sizeof(BfsCtr) / 4) reinterpret_cast<uint32_t *>(nxt)[threadIdx.x] = 0u;
    for (uint32_t i = vb * kSBlock + threadIdx.x; i < p.nwp; i += nvb * kSBlock) clearb[i] = 0u;
    uint32_t nf = 0;
    unsigned long long insp = 0;
    const uint32_t qsize = __hip_atomic_load(&cur->qsize, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sssp_round<Multi>(p, qsize, nextb, vb, nvb, nf, insp);
    insp_all += insp;
    flush_counters<kSBlock>(cur, nf, 0ull, 0ull);
    if (!grid_barrier(p.bar, nvb, gen, kSSpin)) break;
    const uint32_t tnf = __hip_atomic_load(&cur->nf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ++R;
    if (tnf == 0) break;
    if (R >= p.max_rounds) {
      if (vb == 0 && threadIdx.x == 0) p.bar[6] = 1u;
      break;
    }
    compact_step<kSBlock>(nextb, p.nw, p.queue, &nxt->qsize, vb, nvb);
    if (!grid_barrier(p.bar, nvb, gen, kSSpin)) break;
  }
  __shared__ unsigned long long s_red[kSBlock / kWave];
  insp_all = block_sum<kSBlock, unsigned long long>(insp_all, s_red);
  if (threadIdx.x == 0 && insp_all) atomicAdd(p.relax, insp_all);
  if (vb == 0 && threadIdx.x == 0) p.bar[4] = R;
}

template <bool Multi>
__global__ void __launch_bounds__(kSBlock) sssp_round_kernel(SsspParams p, uint32_t qsize, uint32_t *nextb,
                                                             BfsCtr *ctr) {
  uint32_t nf = 0;
  unsigned long long insp = 0;
  sssp_round<Multi>(p, qsize, nextb, blockIdx.x, gridDim.x, nf, insp);
  flush_counters<kSBlock>(ctr, nf, 0ull, insp);
}

__global__ void __launch_bounds__(kSBlock) sssp_compact_kernel(const uint32_t *bm, uint32_t nw, uint32_t *queue,
                                                               uint32_t *count) {
  compact_step<kSBlock>(bm, nw, queue, count, blockIdx.x, gridDim.x);
}

// dist = INF except the sources, bitmaps/counters/barrier = 0, queue =
// the distinct virtual source ids.
__global__ void sssp_reset_kernel(int32_t *__restrict__ dist, size_t total, uint32_t *__restrict__ bms,
                                  uint32_t nwp, BfsCtr *__restrict__ ctr, uint32_t *__restrict__ bar,
                                  unsigned long long *__restrict__ relax) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = tid; i < total; i += stride) dist[i] = kInf;
  for (size_t i = tid; i < (size_t)nwp * 2; i += stride) bms[i] = 0u;
  if (tid < sizeof(BfsCtr) * 3 / 4) reinterpret_cast<uint32_t *>(ctr)[tid] = 0u;
  if (tid < 16) bar[tid] = 0u;
  if (tid == 0) *relax = 0ull;
}

__global__ void sssp_seed_kernel(const uint32_t *__restrict__ srcs, uint32_t k, uint32_t n, int32_t *__restrict__ dist,
                                 uint32_t *__restrict__ queue, BfsCtr *__restrict__ ctr) {
  for (uint32_t s = threadIdx.x; s < k; s += blockDim.x) {
    const uint32_t x = s * n + srcs[s];
    dist[x] = 0;
    queue[s] = x;
  }
  if (threadIdx.x == 0) ctr[0].qsize = k;
}

__global__ void sssp_clear_kernel(BfsCtr *c, uint32_t *bm, uint32_t nwp) {
  if (blockIdx.x == 0 && threadIdx.x < sizeof(BfsCtr) / 4) reinterpret_cast<uint32_t *>(c)[threadIdx.x] = 0u;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nwp; i += gridDim.x * blockDim.x) bm[i] = 0u;
}

struct SsspWork {
  uint32_t nw = 0, nwp = 0;
  uint32_t *queue = nullptr, *bms = nullptr, *bar = nullptr, *h_bar = nullptr;
  BfsCtr *ctr = nullptr, *h_ctr = nullptr;
  unsigned long long *relax = nullptr;
  ~SsspWork() {
    void *ps[] = {queue, bms, bar, ctr, relax};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    if (h_bar) (void)hipHostFree(h_bar);
    if (h_ctr) (void)hipHostFree(h_ctr);
  }
  int alloc(uint64_t nv) {
    nw = (uint32_t)((nv + 31) / 32);
    nwp = (nw + 63) / 64 * 64;
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&queue), (size_t)nv * 4));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&bms), (size_t)nwp * 8));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&bar), 64));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&ctr), sizeof(BfsCtr) * 3));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&relax), 8));
    OLV_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_bar), 64, hipHostMallocDefault));
    OLV_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_ctr), sizeof(BfsCtr) * 3, hipHostMallocDefault));
    return OLV_OK;
  }
};

static unsigned sgrid(size_t items, size_t per) {
  size_t b = (items + per - 1) / per;
  return (unsigned)std::max<size_t>(1, std::min(b, (size_t)cu_count() * 8));
}

// k instances from device sources d_srcs[0..k); dist is k x n.
static int sssp_run(const olv_graph *g, const uint32_t *d_srcs, uint32_t k, const int32_t *weights, int32_t *dist,
                    unsigned flags, olv_stats *st) {
  const uint64_t nv = (uint64_t)k * g->n;
  if (nv >= (1ull << 32)) return fail(OLV_ERANGE, "sssp: %u sources x %u vertices exceed 2^32 ids", k, g->n);
  const bool multi = k > 1;
  SsspWork w;
  OLV_TRY(w.alloc(nv));
  SsspParams p;
  p.row_ptr = g->row_ptr;
  p.col = g->col;
  p.w = weights;
  p.dist = dist;
  p.queue = w.queue;
  p.bms = w.bms;
  p.ctr = w.ctr;
  p.bar = w.bar;
  p.relax = w.relax;
  p.n = g->n;
  p.ninst = k;
  p.nw = w.nw;
  p.nwp = w.nwp;
  p.wave_thr = g->tune.td_wave_thr;
  p.block_thr = g->tune.td_block_thr;
  p.max_rounds = g->n;
  hipStream_t s = stream();
  Timer t;
  t.start();
  hipLaunchKernelGGL(sssp_reset_kernel, dim3(sgrid(std::max<size_t>(nv, (size_t)w.nwp * 2), 256)), dim3(256), 0, s,
                     dist, (size_t)nv, w.bms, w.nwp, w.ctr, w.bar, w.relax);
  hipLaunchKernelGGL(sssp_seed_kernel, dim3(1), dim3(256), 0, s, d_srcs, k, g->n, dist, w.queue, w.ctr);
  OLV_HIP(hipGetLastError());
  uint32_t rounds = 0;
  bool capped = false;
  unsigned long long relax = 0;
  if (flags & OLV_HOST_LOOP) {
    uint32_t qsize = k;
    for (;;) {
      uint32_t *nextb = w.bms + (size_t)(rounds & 1) * w.nwp;
      hipLaunchKernelGGL(sssp_clear_kernel, dim3(sgrid(w.nwp, 256)), dim3(256), 0, s, w.ctr,
                         w.bms + (size_t)((rounds + 1) & 1) * w.nwp, w.nwp);
      if (multi)
        hipLaunchKernelGGL(sssp_round_kernel<true>, dim3(sgrid(qsize, kSBlock)), dim3(kSBlock), 0, s, p, qsize, nextb,
                           w.ctr);
      else
        hipLaunchKernelGGL(sssp_round_kernel<false>, dim3(sgrid(qsize, kSBlock)), dim3(kSBlock), 0, s, p, qsize,
                           nextb, w.ctr);
      OLV_HIP(hipGetLastError());
      OLV_HIP(hipMemcpyAsync(w.h_ctr, w.ctr, sizeof(BfsCtr), hipMemcpyDeviceToHost, s));
      OLV_HIP(hipStreamSynchronize(s));
      const uint32_t nf = w.h_ctr[0].nf;
      relax += w.h_ctr[0].inspected;
      ++rounds;
      if (nf == 0) break;
      if (rounds >= p.max_rounds) {
        capped = true;
        break;
      }
      hipLaunchKernelGGL(sssp_clear_kernel, dim3(1), dim3(64), 0, s, w.ctr, (uint32_t *)nullptr, 0u);
      hipLaunchKernelGGL(sssp_compact_kernel, dim3(sgrid(w.nw, kSBlock)), dim3(kSBlock), 0, s, (const uint32_t *)nextb,
                         w.nw, w.queue, &w.ctr[0].qsize);
      qsize = nf;
    }
    t.stop();
  } else {
    const void *fn = multi ? reinterpret_cast<const void *>(sssp_persistent_kernel<true>)
                           : reinterpret_cast<const void *>(sssp_persistent_kernel<false>);
    int per_cu = 0;
    OLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kSBlock, 0));
    if (per_cu <= 0) return fail(OLV_EUNSUPPORTED, "persistent SSSP kernel does not fit on a CU");
    const int grid = std::min(per_cu, 2) * cu_count();
    void *args[] = {&p};
    OLV_HIP(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kSBlock), args, 0, s));
    t.stop();
    OLV_HIP(hipMemcpyAsync(w.h_bar, w.bar, 64, hipMemcpyDeviceToHost, s));
    OLV_HIP(hipMemcpyAsync(&relax, w.relax, 8, hipMemcpyDeviceToHost, s));
    OLV_HIP(hipStreamSynchronize(s));
    if (w.h_bar[2]) return fail(OLV_EINTERNAL, "olv_sssp: grid barrier timed out (co-residency lost)");
    rounds = w.h_bar[4];
    capped = w.h_bar[6] != 0;
  }
  const double ms = t.ms();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = rounds;
    st->edges_inspected = relax;
  }
  if (capped) return fail(OLV_ENOCONV, "olv_sssp: still relaxing after n=%u rounds (negative cycle)", g->n);
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

static int sssp_entry(const olv_graph *g, const uint32_t *srcs, uint32_t k, bool srcs_host, const int32_t *weights,
                      int32_t *dist, unsigned flags, olv_stats *st) {
  if (!g || !dist || !srcs) return fail(OLV_EINVAL, "olv_sssp: NULL argument");
  if (k == 0) return fail(OLV_EINVAL, "olv_sssp: no source");
  const bool host = (flags & OLV_PTR_HOST) != 0;
  std::vector<uint32_t> hs(k);
  if (srcs_host) std::memcpy(hs.data(), srcs, (size_t)k * 4);
  else OLV_HIP(hipMemcpy(hs.data(), srcs, (size_t)k * 4, hipMemcpyDeviceToHost));
  for (uint32_t s : hs)
    if (s >= g->n) return fail(OLV_EINVAL, "olv_sssp: source %u out of range (n=%u)", s, g->n);
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
  OLV_HIP(hipMemcpy(ds.p, hs.data(), (size_t)k * 4, hipMemcpyHostToDevice));
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
  return sssp_entry(g, &src, 1, true, weights, dist, flags, st);
}

extern "C" int olv_sssp_batched(const olv_graph *g, const uint32_t *srcs, uint32_t k, const int32_t *weights,
                                int32_t *dist, unsigned flags, olv_stats *st) {
  return sssp_entry(g, srcs, k, true, weights, dist, flags, st);
}
