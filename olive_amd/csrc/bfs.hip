// Plain and multi-source BFS over one resident CSR.
//
// Replaces the BFS.cu:124-148 level loop, in which every level costs an
// O(V) host fill + H2D of the accumulators (oliver.h:66, grd.h:90-97), a
// 4V-byte memset of the flag frontier, the push kernel K1
// (oliverKernel.h:45-83), a full O(V) scan in K2 (oliverKernel.h:85-141) and
// two device syncs.  Here a level is one step of a persistent cooperative
// kernel (or, with OLV_HOST_LOOP, one launch group) and the frontier is a
// bitmap that is compacted into a queue only when the next step is top-down.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bfs_core.hpp"
#include "internal.hpp"

namespace olv {

constexpr int kBlock = 256;
constexpr uint32_t kSpinLimit = 1u << 26;

struct BfsWork {
  uint32_t n = 0, nw = 0, nwp = 0;  // vertices, bitmap words, padded words
  uint32_t *visited = nullptr;
  uint32_t *bms = nullptr;          // 3 bitmaps of nwp words
  uint32_t *queue = nullptr;
  BfsCtr *ctr = nullptr;            // 3 counter slots
  BfsCtr *h_ctr = nullptr;          // pinned mirror
  uint32_t *bar = nullptr;          // barrier + result words (16)
  uint32_t *h_res = nullptr;        // pinned mirror of result words
  unsigned long long *stat = nullptr;  // reached vertices, reached out-degree sum
};

void free_bfs_work(olv_graph *g) {
  BfsWork *w = g->bfs;
  if (!w) return;
  void *ps[] = {w->visited, w->bms, w->queue, w->ctr, w->bar, w->stat};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  if (w->h_ctr) (void)hipHostFree(w->h_ctr);
  if (w->h_res) (void)hipHostFree(w->h_res);
  delete w;
  g->bfs = nullptr;
}

static int get_work(const olv_graph *cg, BfsWork **out) {
  olv_graph *g = const_cast<olv_graph *>(cg);
  if (!g->bfs) {
    BfsWork *w = new BfsWork();
    g->bfs = w;
    w->n = g->n;
    w->nw = (g->n + 31) / 32;
    w->nwp = (w->nw + 63) / 64 * 64;
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->visited), (size_t)w->nwp * 4));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->bms), (size_t)w->nwp * 4 * 3));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->queue), (size_t)g->n * 4));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->ctr), sizeof(BfsCtr) * 3));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->bar), 16 * 4));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->stat), 2 * sizeof(unsigned long long)));
    OLV_HIP(hipHostMalloc(reinterpret_cast<void **>(&w->h_ctr), sizeof(BfsCtr) * 3, hipHostMallocDefault));
    OLV_HIP(hipHostMalloc(reinterpret_cast<void **>(&w->h_res), 16 * 4, hipHostMallocDefault));
  }
  *out = g->bfs;
  return OLV_OK;
}

// ------------------------------------------------------------- kernels
// levels = INF, bitmaps/counters/barrier = 0, then the seeds: level 0,
// visited, queue[0..k) = seeds, ctr[0].qsize = #distinct seeds.
__global__ void bfs_reset_kernel(int32_t *__restrict__ levels, uint32_t n, uint32_t *__restrict__ visited,
                                 uint32_t *__restrict__ bms, uint32_t nwp, BfsCtr *__restrict__ ctr,
                                 uint32_t *__restrict__ bar) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = tid; i < n; i += stride) levels[i] = kInf;
  for (size_t i = tid; i < nwp; i += stride) visited[i] = 0u;
  for (size_t i = tid; i < (size_t)nwp * 3; i += stride) bms[i] = 0u;
  if (tid < sizeof(BfsCtr) * 3 / 4) reinterpret_cast<uint32_t *>(ctr)[tid] = 0u;
  if (tid < 16) bar[tid] = 0u;
}

// Single workgroup: seeds are few (one source, or one per graph of a batch).
__global__ void bfs_seed_kernel(const uint32_t *__restrict__ seeds, uint32_t k, int32_t *__restrict__ levels,
                                uint32_t *__restrict__ visited, uint32_t *__restrict__ queue,
                                BfsCtr *__restrict__ ctr) {
  for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
    const uint32_t s = seeds[i];
    const uint32_t old = atomicOr(&visited[s >> 5], 1u << (s & 31));
    if (!(old & (1u << (s & 31)))) {
      levels[s] = 0;
      const uint32_t pos = atomicAdd(&ctr[0].qsize, 1u);
      queue[pos] = s;
    }
  }
}

__global__ void __launch_bounds__(kBlock) td_kernel(BfsArgs a, uint32_t qsize, int32_t nlevel,
                                                    uint32_t *next_bm, BfsCtr *ctr) {
  uint32_t nf = 0;
  unsigned long long mf = 0, insp = 0;
  td_step<kBlock>(a, qsize, nlevel, next_bm, blockIdx.x, gridDim.x, nf, mf, insp);
  flush_counters<kBlock>(ctr, nf, mf, insp);
}

__global__ void __launch_bounds__(kBlock) bu_kernel(BfsArgs a, const uint32_t *front_bm, int32_t nlevel,
                                                    uint32_t *next_bm, BfsCtr *ctr) {
  uint32_t nf = 0;
  unsigned long long mf = 0, insp = 0;
  bu_step<kBlock>(a, front_bm, nlevel, next_bm, blockIdx.x, gridDim.x, nf, mf, insp);
  flush_counters<kBlock>(ctr, nf, mf, insp);
}

__global__ void __launch_bounds__(kBlock) compact_kernel(const uint32_t *bm, uint32_t nwords, uint32_t *queue,
                                                         uint32_t *count) {
  compact_step<kBlock>(bm, nwords, queue, count, blockIdx.x, gridDim.x);
}

__global__ void clear_ctr_kernel(BfsCtr *c) {
  if (threadIdx.x < sizeof(BfsCtr) / 4) reinterpret_cast<uint32_t *>(c)[threadIdx.x] = 0u;
}

__global__ void clear_words_kernel(uint32_t *p, uint32_t nw) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += gridDim.x * blockDim.x) p[i] = 0u;
}

struct BfsPersist {
  BfsArgs a;
  uint32_t *bms;       // 3 x nwp words
  BfsCtr *ctr;         // 3 slots
  uint32_t *bar;       // [0..2] barrier, [4] levels run, [5] bottom-up levels, [6] error
  unsigned long long m;
  uint32_t nw, nwp;
  uint32_t can_bu;
  float alpha, beta;
};

// All levels in one launch.  Level L: expand (top-down from the queue or
// bottom-up against bitmap (L+2)%3) into bitmap L%3, clear bitmap (L+1)%3,
// barrier; every block then reads the same counters and takes the same
// decision; if the next step is top-down, compact bitmap L%3 into the queue
// and barrier again.
__global__ void __launch_bounds__(kBlock) bfs_persistent_kernel(BfsPersist p) {
  uint32_t gen = 0;
  const uint32_t vb = blockIdx.x, nvb = gridDim.x;
  // unexplored out-edges: m minus the out-degrees of the seeds
  unsigned long long mu = p.m;
  const uint32_t nseed = __hip_atomic_load(&p.ctr[0].qsize, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t i = 0; i < nseed; ++i) {
    const uint32_t s = p.a.queue[i];
    mu -= p.a.row_ptr[s + 1] - p.a.row_ptr[s];
  }
  bool bu = false;
  uint32_t nbu = 0, L = 0;
  for (;;) {
    BfsCtr *cur = &p.ctr[L % 3];
    BfsCtr *nxt = &p.ctr[(L + 1) % 3];
    uint32_t *nextb = p.bms + (size_t)(L % 3) * p.nwp;
    const uint32_t *front = p.bms + (size_t)((L + 2) % 3) * p.nwp;
    uint32_t *clearb = p.bms + (size_t)((L + 1) % 3) * p.nwp;
    if (vb == 0 && threadIdx.x < sizeof(BfsCtr) / 4) reinterpret_cast<uint32_t *>(nxt)[threadIdx.x] = 0u;
    for (uint32_t i = vb * kBlock + threadIdx.x; i < p.nwp; i += nvb * kBlock) clearb[i] = 0u;
    uint32_t nf = 0;
    unsigned long long mf = 0, insp = 0;
    if (!bu) {
      const uint32_t qsize = __hip_atomic_load(&cur->qsize, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      td_step<kBlock>(p.a, qsize, (int32_t)(L + 1), nextb, vb, nvb, nf, mf, insp);
    } else {
      bu_step<kBlock>(p.a, front, (int32_t)(L + 1), nextb, vb, nvb, nf, mf, insp);
    }
    flush_counters<kBlock>(cur, nf, mf, insp);
    if (!grid_barrier(p.bar, nvb, gen, kSpinLimit)) break;
    const uint32_t tnf = __hip_atomic_load(&cur->nf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long tmf = __hip_atomic_load(&cur->mf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (bu) ++nbu;
    ++L;
    if (tnf == 0 || L > p.a.n) break;
    mu = mu > tmf ? mu - tmf : 0ull;
    const bool next_bu = choose_bottom_up(bu, p.can_bu != 0, tnf, tmf, mu, p.a.n, p.alpha, p.beta);
    if (!next_bu) {
      compact_step<kBlock>(nextb, p.nw, p.a.queue, &nxt->qsize, vb, nvb);
      if (!grid_barrier(p.bar, nvb, gen, kSpinLimit)) break;
    }
    bu = next_bu;
  }
  if (vb == 0 && threadIdx.x == 0) {
    p.bar[4] = L;
    p.bar[5] = nbu;
  }
}

// Reached vertices and the sum of their out-degrees (E_r of the byte model).
__global__ void bfs_stat_kernel(const int32_t *__restrict__ levels, const uint32_t *__restrict__ row_ptr,
                                uint32_t n, unsigned long long *__restrict__ stat) {
  __shared__ unsigned long long s_red[kBlock / kWave];
  unsigned long long vr = 0, er = 0;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    if (levels[v] != kInf) {
      vr += 1;
      er += row_ptr[v + 1] - row_ptr[v];
    }
  vr = block_sum<kBlock, unsigned long long>(vr, s_red);
  er = block_sum<kBlock, unsigned long long>(er, s_red);
  if (threadIdx.x == 0) {
    atomicAdd(&stat[0], vr);
    atomicAdd(&stat[1], er);
  }
}

// ------------------------------------------------------------- drivers
static int persistent_grid(int *grid) {
  int per_cu = 0;
  OLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bfs_persistent_kernel, kBlock, 0));
  if (per_cu <= 0) return fail(OLV_EUNSUPPORTED, "persistent BFS kernel does not fit on a CU");
  per_cu = std::min(per_cu, 2);
  *grid = per_cu * cu_count();
  return OLV_OK;
}

static BfsArgs make_args(const olv_graph *g, BfsWork *w, int32_t *levels) {
  BfsArgs a;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.in_row_ptr = g->in_row_ptr;
  a.in_col = g->in_col;
  a.levels = levels;
  a.visited = w->visited;
  a.queue = w->queue;
  a.n = g->n;
  a.wave_thr = g->tune.td_wave_thr;
  a.block_thr = g->tune.td_block_thr;
  return a;
}

static unsigned grid_for(size_t items, size_t per_block) {
  size_t b = (items + per_block - 1) / per_block;
  size_t cap = (size_t)cu_count() * 8;
  return (unsigned)std::max<size_t>(1, std::min(b, cap));
}

// Enqueue reset + seeding (stream-ordered, capturable: no memset nodes).
static int enqueue_init(const olv_graph *g, BfsWork *w, const uint32_t *d_seeds, uint32_t k, int32_t *levels) {
  hipLaunchKernelGGL(bfs_reset_kernel, dim3(grid_for(std::max<size_t>(g->n, (size_t)w->nwp * 3), 256)), dim3(256),
                     0, stream(), levels, g->n, w->visited, w->bms, w->nwp, w->ctr, w->bar);
  hipLaunchKernelGGL(bfs_seed_kernel, dim3(1), dim3(256), 0, stream(), d_seeds, k, levels, w->visited, w->queue,
                     w->ctr);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

static int enqueue_persistent(const olv_graph *g, BfsWork *w, int32_t *levels, bool can_bu) {
  int grid = 0;
  OLV_TRY(persistent_grid(&grid));
  BfsPersist p;
  p.a = make_args(g, w, levels);
  p.bms = w->bms;
  p.ctr = w->ctr;
  p.bar = w->bar;
  p.m = g->m;
  p.nw = w->nw;
  p.nwp = w->nwp;
  p.can_bu = can_bu ? 1u : 0u;
  p.alpha = g->tune.alpha;
  p.beta = g->tune.beta;
  void *args[] = {&p};
  OLV_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(bfs_persistent_kernel), dim3(grid),
                                     dim3(kBlock), args, 0, stream()));
  return OLV_OK;
}

// Host-driven levels: one launch group per level, counters read back.
static int run_host_loop(const olv_graph *g, BfsWork *w, int32_t *levels, const std::vector<uint32_t> &h_seeds,
                         bool can_bu, uint32_t *levels_run, uint32_t *bu_levels, unsigned long long *inspected) {
  hipStream_t s = stream();
  BfsArgs a = make_args(g, w, levels);
  // seed statistics from the host copy of the seeds
  OLV_HIP(hipMemcpyAsync(w->h_ctr, w->ctr, sizeof(BfsCtr), hipMemcpyDeviceToHost, s));
  OLV_HIP(hipStreamSynchronize(s));
  uint32_t qsize = w->h_ctr[0].qsize;
  unsigned long long mu = g->m;
  {
    std::vector<uint32_t> rp(2);
    for (uint32_t sd : h_seeds) {
      OLV_HIP(hipMemcpy(rp.data(), g->row_ptr + sd, 8, hipMemcpyDeviceToHost));
      mu -= rp[1] - rp[0];
    }
  }
  uint32_t *bm[3] = {w->bms, w->bms + w->nwp, w->bms + 2 * (size_t)w->nwp};
  uint32_t *front = bm[2], *next = bm[0], *spare = bm[1];
  bool bu = false;
  uint32_t L = 0, nbu = 0;
  unsigned long long insp_total = 0;
  for (;;) {
    hipLaunchKernelGGL(clear_ctr_kernel, dim3(1), dim3(64), 0, s, w->ctr);
    if (!bu)
      hipLaunchKernelGGL(td_kernel, dim3(grid_for(qsize, kBlock)), dim3(kBlock), 0, s, a, qsize, (int32_t)(L + 1),
                         next, w->ctr);
    else
      hipLaunchKernelGGL(bu_kernel, dim3(grid_for(g->n, kBlock)), dim3(kBlock), 0, s, a, (const uint32_t *)front,
                         (int32_t)(L + 1), next, w->ctr);
    OLV_HIP(hipGetLastError());
    OLV_HIP(hipMemcpyAsync(w->h_ctr, w->ctr, sizeof(BfsCtr), hipMemcpyDeviceToHost, s));
    OLV_HIP(hipStreamSynchronize(s));
    const uint32_t nf = w->h_ctr[0].nf;
    const unsigned long long mf = w->h_ctr[0].mf;
    insp_total += w->h_ctr[0].inspected;
    if (bu) ++nbu;
    ++L;
    if (nf == 0 || L > g->n) break;
    mu = mu > mf ? mu - mf : 0ull;
    const bool next_bu = choose_bottom_up(bu, can_bu, nf, mf, mu, g->n, g->tune.alpha, g->tune.beta);
    if (!next_bu) {
      hipLaunchKernelGGL(clear_ctr_kernel, dim3(1), dim3(64), 0, s, w->ctr);
      hipLaunchKernelGGL(compact_kernel, dim3(grid_for(w->nw, kBlock)), dim3(kBlock), 0, s, (const uint32_t *)next,
                         w->nw, w->queue, &w->ctr[0].qsize);
      qsize = nf;
    }
    // rotate: next becomes the frontier bitmap, the old frontier is cleared
    uint32_t *old_front = front;
    front = next;
    next = spare;
    spare = old_front;
    hipLaunchKernelGGL(clear_words_kernel, dim3(grid_for(w->nwp, 256)), dim3(256), 0, s, next, w->nwp);
    bu = next_bu;
  }
  *levels_run = L;
  *bu_levels = nbu;
  *inspected = insp_total;
  return OLV_OK;
}

static int collect_stats(const olv_graph *g, BfsWork *w, const int32_t *levels, olv_stats *st) {
  if (!st) return OLV_OK;
  OLV_HIP(hipMemsetAsync(w->stat, 0, 2 * sizeof(unsigned long long), stream()));
  hipLaunchKernelGGL(bfs_stat_kernel, dim3(grid_for(g->n, kBlock)), dim3(kBlock), 0, stream(), levels, g->row_ptr,
                     g->n, w->stat);
  unsigned long long h[2];
  OLV_HIP(hipMemcpyAsync(h, w->stat, sizeof(h), hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  st->vertices_reached = h[0];
  st->edges_traversed = h[1];
  return OLV_OK;
}

// Transpose (for bottom-up steps) and work buffers of g.
int bfs_prepare(const olv_graph *g, bool can_bu, BfsWork **w) {
  if (can_bu && !g->in_row_ptr) OLV_TRY(olv_graph_build_transpose(const_cast<olv_graph *>(g)));
  return get_work(g, w);
}

// Stream-ordered, capturable: reset, seeding and the persistent kernel.
int bfs_enqueue(const olv_graph *g, BfsWork *w, const uint32_t *d_seeds, uint32_t k, int32_t *levels, bool can_bu) {
  OLV_TRY(enqueue_init(g, w, d_seeds, k, levels));
  return enqueue_persistent(g, w, levels, can_bu && g->in_row_ptr != nullptr);
}

// Wait for the stream, then read the level count of the last persistent run.
int bfs_result(BfsWork *w, uint32_t *levels_run, uint32_t *bu_levels) {
  OLV_HIP(hipMemcpyAsync(w->h_res, w->bar, 16 * 4, hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  if (w->h_res[2]) return fail(OLV_EINTERNAL, "olv_bfs: grid barrier timed out (co-residency lost)");
  *levels_run = w->h_res[4];
  *bu_levels = w->h_res[5];
  return OLV_OK;
}

uint32_t *bfs_result_words(BfsWork *w) { return w->bar; }

// Multi-source BFS on g (levels = distance to the nearest seed).
// seeds_host: seeds is a host array (levels follow OLV_PTR_HOST in flags).
int bfs_multisource(const olv_graph *g, const uint32_t *seeds, uint32_t k, bool seeds_host, int32_t *levels,
                    unsigned flags, olv_stats *st) {
  if (!g || !seeds || !levels) return fail(OLV_EINVAL, "olv_bfs: NULL argument");
  if (k == 0) return fail(OLV_EINVAL, "olv_bfs: no source");
  const bool host = (flags & OLV_PTR_HOST) != 0;
  std::vector<uint32_t> h_seeds(k);
  if (seeds_host) std::memcpy(h_seeds.data(), seeds, (size_t)k * 4);
  else OLV_HIP(hipMemcpy(h_seeds.data(), seeds, (size_t)k * 4, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < k; ++i)
    if (h_seeds[i] >= g->n) return fail(OLV_EINVAL, "olv_bfs: source %u out of range (n=%u)", h_seeds[i], g->n);
  const bool can_bu = !(flags & OLV_NO_DIRECTION_OPT);
  BfsWork *w = nullptr;
  OLV_TRY(bfs_prepare(g, can_bu, &w));
  DevBuf d_seeds, d_levels;
  OLV_TRY(d_seeds.alloc((size_t)k * 4));
  OLV_HIP(hipMemcpyAsync(d_seeds.p, h_seeds.data(), (size_t)k * 4, hipMemcpyHostToDevice, stream()));
  int32_t *out = levels;
  if (host) {
    OLV_TRY(d_levels.alloc((size_t)g->n * 4));
    out = d_levels.as<int32_t>();
  }
  Timer t;
  t.start();
  OLV_TRY(enqueue_init(g, w, d_seeds.as<uint32_t>(), k, out));
  uint32_t levels_run = 0, bu_levels = 0;
  unsigned long long insp = 0;
  if (flags & OLV_HOST_LOOP) {
    OLV_TRY(run_host_loop(g, w, out, h_seeds, can_bu, &levels_run, &bu_levels, &insp));
    t.stop();
  } else {
    OLV_TRY(enqueue_persistent(g, w, out, can_bu));
    t.stop();
    OLV_TRY(bfs_result(w, &levels_run, &bu_levels));
  }
  const double ms = t.ms();
  if (host) OLV_HIP(hipMemcpy(levels, out, (size_t)g->n * 4, hipMemcpyDeviceToHost));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = levels_run;
    st->bottom_up_levels = bu_levels;
    if (flags & OLV_HOST_LOOP) st->edges_inspected = insp;
    OLV_TRY(collect_stats(g, w, out, st));
  }
  return OLV_OK;
}

}  // namespace olv

using namespace olv;

extern "C" int olv_bfs(const olv_graph *g, uint32_t src, int32_t *levels, unsigned flags, olv_stats *st) {
  if (!g) return fail(OLV_EINVAL, "olv_bfs: NULL graph");
  if (src >= g->n) return fail(OLV_EINVAL, "olv_bfs: source %u out of range (n=%u)", src, g->n);
  return bfs_multisource(g, &src, 1, true, levels, flags, st);
}
