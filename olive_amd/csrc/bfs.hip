// Plain and multi-source BFS over one resident CSR.
//
// Replaces the BFS.cu:124-148 level loop, in which every level costs an
// O(V) host fill + H2D of the accumulators (oliver.h:66, grd.h:90-97), a
// 4V-byte memset of the flag frontier, the push kernel K1
// (oliverKernel.h:45-83), a full O(V) scan in K2 (oliverKernel.h:85-141) and
// two device syncs.
//
// Here a level is two kernels, expand and compact, and the loop control
// lives on the device (BfsCtl): expand reads the level number, direction and
// queue length the previous compact wrote; compact reads the level's
// counters, takes the direction decision, writes the next level's state and
// -- when the next step is top-down -- turns the next-frontier bitmap into
// the queue.  The host enqueues levels in chunks without waiting and checks
// the done flag once per chunk; surplus levels return at once.  The whole
// sequence is capturable in a hipGraph (olv_plan_*).  A kernel boundary costs
// about 3 us on MI355X (1.8 us inside a hipGraph), less than a grid barrier
// (6-34 us), so this beats one persistent kernel; the persistent variant is
// kept behind OLV_PERSISTENT for comparison.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "bfs_core.hpp"
#include "internal.hpp"

namespace olv {

constexpr int kBlock = 256;
constexpr uint32_t kSpinLimit = 1u << 24;
// Workgroups per CU of the per-level kernels come from the graph's tuning
// (Tuning::bfs_*_per_cu).  Defaults: expansion in the graph's own numbering
// 8 (full occupancy; 4 and 16 measured slower); on the renumbered twin 6
// (RMAT-22 0.148 ms; 7: 0.148, 5: 0.150, 4: 0.154, 8: 0.160, 3: 0.165);
// compaction 8 (insensitive).

struct BfsState {              // per level parity
  unsigned long long mu;       // unexplored out-edges before this level
  uint32_t mode;               // 0 = top-down, 1 = bottom-up
  uint32_t nbu;                // bottom-up levels before this one
};

struct BfsCtl {
  uint32_t level;      // next level to expand (written by compact)
  uint32_t level_e;    // level being compacted (written by expand)
  uint32_t done;       // traversal finished
  uint32_t levels;     // result: levels run, the final empty one included
  uint32_t nbu;        // result: bottom-up levels
  uint32_t overflow;   // result: level cap (n + 1) reached
  uint32_t qfill[2];   // frontier queue length of level L, slot L & 1
  uint32_t hfill[2];   // hub items of level L, slot L & 1
  uint32_t pad[2];
  unsigned long long inspected;  // result: edges read
  BfsState st[2];
  CtrSet cnt[3];       // level counters, slot L % 3
};

struct BfsWork {
  uint32_t n = 0, nw = 0, nwp = 0;  // vertices, bitmap words, padded words
  uint8_t *visited = nullptr;       // one byte per vertex, nwp * 32 bytes
  uint32_t *snap = nullptr;         // claims as of the last frontier extraction (bitmap)
  uint32_t *bms = nullptr;          // 2 frontier bitmaps of nwp words
  uint32_t *queue = nullptr;
  uint2 *hubs = nullptr;
  BfsCtl *ctl = nullptr;
  BfsCtl *h_ctl = nullptr;          // pinned mirror
  GridBar *bar = nullptr;           // persistent mode
  unsigned long long *stat = nullptr;
  uint32_t last_levels = 0;         // depth of the previous run (plan capture depth)
  uint32_t recent[4] = {0, 0, 0, 0};  // depths of the last runs (first chunk = their maximum)
  bool force_bu = false;            // OLV_FORCE_BOTTOM_UP of the current run
};

static void release_bfs_work(BfsWork *w) {
  void *ps[] = {w->visited, w->snap, w->bms, w->queue, w->hubs, w->ctl, w->bar, w->stat};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  if (w->h_ctl) (void)hipHostFree(w->h_ctl);
  delete w;
}

void free_bfs_work(olv_graph *g) {
  if (!g->bfs) return;
  release_bfs_work(g->bfs);
  g->bfs = nullptr;
}

static int alloc_bfs_work(BfsWork *w, const olv_graph *g) {
  w->n = g->n;
  w->nw = (g->n + 31) / 32;
  w->nwp = (w->nw + 63) / 64 * 64;
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->visited), (size_t)w->nwp * 32));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->snap), (size_t)w->nwp * 4));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->bms), (size_t)w->nwp * 4 * 2));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->queue), (size_t)g->n * 4));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->hubs), hub_capacity(g->m) * sizeof(uint2)));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->ctl), sizeof(BfsCtl)));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->bar), sizeof(GridBar)));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->stat), 2 * sizeof(unsigned long long)));
  OLV_HIP(hipHostMalloc(reinterpret_cast<void **>(&w->h_ctl), sizeof(BfsCtl), hipHostMallocDefault));
  return OLV_OK;
}

// The graph's work buffers, allocated on first use; published to the graph
// only once every allocation has succeeded.
static int get_work(const olv_graph *cg, BfsWork **out) {
  olv_graph *g = const_cast<olv_graph *>(cg);
  if (!g->bfs) {
    BfsWork *w = new BfsWork();
    const int rc = alloc_bfs_work(w, g);
    if (rc != OLV_OK) {
      release_bfs_work(w);
      return rc;
    }
    g->bfs = w;
  }
  *out = g->bfs;
  return OLV_OK;
}

// ------------------------------------------------------------- kernels
// levels = INF, bitmaps/control/barrier = 0; the vertices without in-edges
// are pre-claimed (visited byte and snapshot bit) when bottom-up steps may
// run -- no parent can claim them, so bottom-up steps skip them;
// bfs_seed_kernel releases the seeds among them -- else nothing is.
// With a single seed (`seed1`, device) the seeding is done here as well:
// the seed's level, visited byte and snapshot bit are written by the lanes
// that own them, and workgroup 0 builds the level-0 frontier after clearing
// the control block -- one launch instead of two.
__global__ void bfs_reset_kernel(int32_t *__restrict__ levels, uint32_t n, uint8_t *__restrict__ visited,
                                 uint32_t *__restrict__ snap, const uint32_t *__restrict__ no_in,
                                 uint32_t *__restrict__ bms, uint32_t nwp, BfsCtl *__restrict__ ctl,
                                 GridBar *__restrict__ bar, const uint32_t *__restrict__ seed1, BfsArgs a,
                                 unsigned long long m) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t nw = ((size_t)n + 31) / 32;
  const uint32_t s = seed1 ? *seed1 : 0xffffffffu;
  const uint32_t so = (seed1 && a.omap) ? a.omap[s] : s;  // the seed's levels index
  for (size_t i = tid; i < n; i += stride) levels[i] = i == so ? 0 : kInf;
  for (size_t i = tid; i < nwp; i += stride) {
    uint32_t d = (no_in && i < nw) ? no_in[i] : 0u;
    if (s != 0xffffffffu && i == (s >> 5)) d |= 1u << (s & 31);  // the seed is claimed
    snap[i] = d;
    uint32_t b[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const uint32_t nib = (d >> (4 * k)) & 0xfu;
      b[k] = (nib & 1u) | ((nib & 2u) << 7) | ((nib & 4u) << 14) | ((nib & 8u) << 21);
    }
    uint4 *o = reinterpret_cast<uint4 *>(visited + i * 32);
    o[0] = make_uint4(b[0], b[1], b[2], b[3]);
    o[1] = make_uint4(b[4], b[5], b[6], b[7]);
  }
  for (size_t i = tid; i < (size_t)nwp * 2; i += stride) bms[i] = 0u;
  if (blockIdx.x != 0) return;
  for (uint32_t i = threadIdx.x; i < sizeof(BfsCtl) / 4; i += blockDim.x) reinterpret_cast<uint32_t *>(ctl)[i] = 0u;
  for (uint32_t i = threadIdx.x; i < sizeof(GridBar) / 4; i += blockDim.x) reinterpret_cast<uint32_t *>(bar)[i] = 0u;
  if (!seed1) return;
  __syncthreads();
  unsigned long long deg = 0;
  seed_frontier(seed1, 1u, a.row_ptr, a.hub_thr, a.queue, &ctl->qfill[0], a.hubs, &ctl->hfill[0],
                [&](uint32_t) -> uint32_t { return s; }, IdRow(), &deg);
  if (threadIdx.x == 0) {
    ctl->st[0].mu = m > deg ? m - deg : 0ull;
    ctl->st[0].mode = 0;
    ctl->st[0].nbu = 0;
  }
}

// One workgroup: distinct seeds become level 0 and the level-0 frontier.
__global__ void bfs_seed_kernel(const uint32_t *__restrict__ seeds, uint32_t k, BfsArgs a, unsigned long long m,
                                BfsCtl *__restrict__ ctl) {
  for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
    const uint32_t s = seeds[i];
    atomicAnd(&a.snap[s >> 5], ~(1u << (s & 31)));
  }
  __syncthreads();
  // the snapshot bit decides between duplicate seeds (level 0 is not a next
  // frontier, so seeds stay in the snapshot)
  auto claim = [&](uint32_t i) -> uint32_t {
    const uint32_t s = seeds[i];
    const uint32_t old = atomicOr(&a.snap[s >> 5], 1u << (s & 31));
    if (old & (1u << (s & 31))) return 0xffffffffu;
    a.visited[s] = 1;
    if (a.omap) a.set_level<true>(s, 0);
    else a.set_level<false>(s, 0);
    return s;
  };
  unsigned long long deg = 0;
  seed_frontier(seeds, k, a.row_ptr, a.hub_thr, a.queue, &ctl->qfill[0], a.hubs, &ctl->hfill[0], claim, IdRow(),
                &deg);
  if (threadIdx.x == 0) {
    ctl->st[0].mu = m > deg ? m - deg : 0ull;
    ctl->st[0].mode = 0;
    ctl->st[0].nbu = 0;
  }
}

struct BfsDir {
  uint32_t can_bu;
  uint32_t force_bu;  // every level after the first bottom-up (OLV_FORCE_BOTTOM_UP)
  float alpha, beta;
};

// Level L = ctl->level: claim the vertices of level L + 1, top-down from
// the level's queue and hub items or bottom-up against frontier bitmap
// L & 1; reset the next level's fill counters.
template <bool Mapped>
__global__ void __launch_bounds__(kBlock) bfs_expand_kernel(BfsArgs a, BfsCtl *ctl, uint32_t *bms, uint32_t nwp) {
  if (__hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const uint32_t L = ctl->level;
  const BfsState st = ctl->st[L & 1];
  const uint32_t qsize = ctl->qfill[L & 1], nhubs = ctl->hfill[L & 1];
  const uint32_t *front = bms + (size_t)(L & 1) * nwp;
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      ctl->level_e = L;
      ctl->qfill[(L + 1) & 1] = 0;
      ctl->hfill[(L + 1) & 1] = 0;
    }
    ctr_clear(&ctl->cnt[(L + 1) % 3]);
  }
  uint32_t nf = 0;
  unsigned long long mf = 0, insp = 0;
  if (st.mode == 0)
    td_step<kBlock, Mapped>(a, qsize, nhubs, (int32_t)(L + 1), blockIdx.x, gridDim.x, nf, mf, insp);
  else
    bu_step<kBlock, Mapped>(a, front, (int32_t)(L + 1), blockIdx.x, gridDim.x, nf, mf, insp);
  flush_counters<kBlock>(&ctl->cnt[L % 3], nf, mf, insp);
}

// Level L = ctl->level_e: decide the next direction and extract level
// L + 1's frontier (bitmap (L + 1) & 1; also queue and hub items if the
// next step is top-down).  Workgroup 0 publishes the next level's state.
__global__ void __launch_bounds__(kBlock) bfs_compact_kernel(BfsArgs a, BfsCtl *ctl, uint32_t *bms, uint32_t nw,
                                                             uint32_t nwp, unsigned long long m, BfsDir dir) {
  const uint32_t w0 = blockIdx.x * kBlock + threadIdx.x;
  const WordLoad pre = w0 < nw ? load_word(a, w0) : WordLoad{};
  if (__hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const uint32_t L = ctl->level_e;
  const CtrSum c = ctr_total(&ctl->cnt[L % 3]);
  const BfsState st = ctl->st[L & 1];
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  const uint32_t nbu = st.nbu + st.mode;
  if (lead) ctl->inspected += c.inspected;
  if (c.nf == 0 || L + 1 > a.n) {
    if (lead) {
      ctl->levels = L + 1;
      ctl->nbu = nbu;
      ctl->overflow = c.nf != 0;
      ctl->done = 1u;
    }
    return;
  }
  const unsigned long long mu = st.mu > c.mf ? st.mu - c.mf : 0ull;
  const bool next_bu =
      dir.force_bu != 0 || choose_bottom_up(st.mode != 0, dir.can_bu != 0, c.nf, c.mf, mu, a.n, dir.alpha, dir.beta);
  if (lead) {
    BfsState ns;
    ns.mu = mu;
    ns.mode = next_bu ? 1u : 0u;
    ns.nbu = nbu;
    ctl->st[(L + 1) & 1] = ns;
    ctl->level = L + 1;
  }
  frontier_words<kBlock>(a, bms + (size_t)((L + 1) & 1) * nwp, nw, !next_bu, &ctl->qfill[(L + 1) & 1],
                         &ctl->hfill[(L + 1) & 1], blockIdx.x, gridDim.x, pre);
}

// All levels in one cooperative launch with the XCD-aware grid barrier
// (OLV_PERSISTENT).  Same steps and decisions as the pipeline.
template <bool Mapped>
__global__ void __launch_bounds__(kBlock) bfs_persistent_kernel(BfsArgs a, BfsCtl *ctl, uint32_t *bms, uint32_t nw,
                                                                uint32_t nwp, BfsDir dir, GridBar *bar) {
  GridBarLocal gb;
  if (!gbar_init(bar, gb, gridDim.x, kSpinLimit)) return;
  const uint32_t vb = blockIdx.x, nvb = gridDim.x;
  const bool lead = vb == 0 && threadIdx.x == 0;
  BfsState st = ctl->st[0];
  unsigned long long insp_all = 0;
  uint32_t L = 0;
  for (;;) {
    const uint32_t *front = bms + (size_t)(L & 1) * nwp;
    const uint32_t qsize = __hip_atomic_load(&ctl->qfill[L & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t nhubs = __hip_atomic_load(&ctl->hfill[L & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (vb == 0) {
      if (threadIdx.x == 0) {
        ctl->qfill[(L + 1) & 1] = 0;
        ctl->hfill[(L + 1) & 1] = 0;
      }
      ctr_clear(&ctl->cnt[(L + 1) % 3]);
    }
    uint32_t nf = 0;
    unsigned long long mf = 0, insp = 0;
    if (st.mode == 0)
      td_step<kBlock, Mapped>(a, qsize, nhubs, (int32_t)(L + 1), vb, nvb, nf, mf, insp);
    else
      bu_step<kBlock, Mapped>(a, front, (int32_t)(L + 1), vb, nvb, nf, mf, insp);
    flush_counters<kBlock>(&ctl->cnt[L % 3], nf, mf, insp);
    if (!grid_barrier(bar, gb, kSpinLimit)) break;
    const CtrSum c = ctr_total(&ctl->cnt[L % 3]);
    insp_all += c.inspected;
    const uint32_t nbu = st.nbu + st.mode;
    if (c.nf == 0 || L + 1 > a.n) {
      if (lead) {
        ctl->levels = L + 1;
        ctl->nbu = nbu;
        ctl->overflow = c.nf != 0;
        ctl->inspected = insp_all;
        ctl->done = 1u;
      }
      break;
    }
    const unsigned long long mu = st.mu > c.mf ? st.mu - c.mf : 0ull;
    const bool next_bu =
        dir.force_bu != 0 || choose_bottom_up(st.mode != 0, dir.can_bu != 0, c.nf, c.mf, mu, a.n, dir.alpha, dir.beta);
    st.mu = mu;
    st.mode = next_bu ? 1u : 0u;
    st.nbu = nbu;
    const uint32_t wp = vb * kBlock + threadIdx.x;
    frontier_words<kBlock>(a, bms + (size_t)((L + 1) & 1) * nwp, nw, !next_bu, &ctl->qfill[(L + 1) & 1],
                           &ctl->hfill[(L + 1) & 1], vb, nvb, wp < nw ? load_word(a, wp) : WordLoad{});
    if (!grid_barrier(bar, gb, kSpinLimit)) break;
    ++L;
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
static BfsArgs make_args(const olv_graph *g, BfsWork *w, int32_t *levels) {
  BfsArgs a;
  a.row_ptr = g->row_ptr;
  a.col = g->col;
  a.in_row_ptr = g->in_row_ptr;
  a.in_col = g->in_col;
  a.in_head = g->in_head;
  a.levels = levels;
  a.omap = g->out_map;
  a.visited = w->visited;
  a.snap = w->snap;
  a.queue = w->queue;
  a.hubs = w->hubs;
  a.n = g->n;
  a.wave_thr = g->tune.td_wave_thr;
  a.block_thr = g->tune.td_block_thr;
  a.hub_thr = g->tune.hub_thr;
  return a;
}

static unsigned grid_for(size_t items, size_t per_block, size_t per_cu) {
  size_t b = (items + per_block - 1) / per_block;
  size_t cap = (size_t)cu_count() * per_cu;
  return (unsigned)std::max<size_t>(1, std::min(b, cap));
}

static BfsDir make_dir(const olv_graph *g, const BfsWork *w, bool can_bu) {
  BfsDir d;
  d.can_bu = (can_bu && g->in_row_ptr) ? 1u : 0u;
  d.force_bu = (d.can_bu && w->force_bu) ? 1u : 0u;
  d.alpha = g->tune.alpha;
  d.beta = g->tune.beta;
  return d;
}

// Transpose (for bottom-up steps) and work buffers of g.
int bfs_prepare(const olv_graph *g, bool can_bu, BfsWork **w) {
  if (can_bu && !g->in_row_ptr) OLV_TRY(olv_graph_build_transpose(const_cast<olv_graph *>(g)));
  if (can_bu) OLV_TRY(ensure_in_head(const_cast<olv_graph *>(g)));
  return get_work(g, w);
}

// Stream-ordered, capturable: reset and seeding.
int bfs_enqueue_init(const olv_graph *g, BfsWork *w, const uint32_t *d_seeds, uint32_t k, int32_t *levels,
                     bool can_bu) {
  const uint32_t *no_in = can_bu ? g->no_in_bm : nullptr;
  const BfsArgs a = make_args(g, w, levels);
  hipLaunchKernelGGL(bfs_reset_kernel, dim3(grid_for(std::max<size_t>(g->n, (size_t)w->nwp * 2), 256, 8)), dim3(256),
                     0, stream(), levels, g->n, w->visited, w->snap, no_in, w->bms, w->nwp, w->ctl, w->bar,
                     k == 1 ? d_seeds : nullptr, a, (unsigned long long)g->m);
  if (k != 1)
    hipLaunchKernelGGL(bfs_seed_kernel, dim3(1), dim3(256), 0, stream(), d_seeds, k, a, (unsigned long long)g->m,
                       w->ctl);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

// Stream-ordered, capturable: `count` levels (expand + compact each).
int bfs_enqueue_levels(const olv_graph *g, BfsWork *w, int32_t *levels, bool can_bu, uint32_t count) {
  const BfsArgs a = make_args(g, w, levels);
  const BfsDir dir = make_dir(g, w, can_bu);
  const unsigned ge = (unsigned)cu_count() * (a.omap ? g->tune.bfs_expand_per_cu_twin : g->tune.bfs_expand_per_cu);
  const unsigned gc = grid_for(w->nw, kBlock, g->tune.bfs_compact_per_cu);
  for (uint32_t i = 0; i < count; ++i) {
    if (a.omap)
      hipLaunchKernelGGL(bfs_expand_kernel<true>, dim3(ge), dim3(kBlock), 0, stream(), a, w->ctl, w->bms, w->nwp);
    else
      hipLaunchKernelGGL(bfs_expand_kernel<false>, dim3(ge), dim3(kBlock), 0, stream(), a, w->ctl, w->bms, w->nwp);
    hipLaunchKernelGGL(bfs_compact_kernel, dim3(gc), dim3(kBlock), 0, stream(), a, w->ctl, w->bms, w->nw, w->nwp,
                       (unsigned long long)g->m, dir);
  }
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

static int enqueue_persistent(const olv_graph *g, BfsWork *w, int32_t *levels, bool can_bu) {
  BfsArgs a = make_args(g, w, levels);
  const void *kern = a.omap ? reinterpret_cast<const void *>(bfs_persistent_kernel<true>)
                            : reinterpret_cast<const void *>(bfs_persistent_kernel<false>);
  int per_cu = 0;
  OLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0));
  if (per_cu <= 0) return fail(OLV_EUNSUPPORTED, "persistent BFS kernel does not fit on a CU");
  const int grid = cu_count();  // one workgroup per CU: cheapest barrier (measured)
  BfsCtl *ctl = w->ctl;
  uint32_t *bms = w->bms;
  uint32_t nw = w->nw, nwp = w->nwp;
  BfsDir dir = make_dir(g, w, can_bu);
  GridBar *bar = w->bar;
  void *args[] = {&a, &ctl, &bms, &nw, &nwp, &dir, &bar};
  OLV_HIP(hipLaunchCooperativeKernel(kern, dim3(grid), dim3(kBlock), args, 0, stream()));
  return OLV_OK;
}

// Wait for the stream, then read the control block.
int bfs_result(BfsWork *w, uint32_t *levels_run, uint32_t *bu_levels) {
  OLV_HIP(hipMemcpyAsync(w->h_ctl, w->ctl, sizeof(BfsCtl), hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  if (levels_run) *levels_run = w->h_ctl->levels;
  if (bu_levels) *bu_levels = w->h_ctl->nbu;
  return OLV_OK;
}

bool bfs_done(BfsWork *w) { return w->h_ctl->done != 0; }
uint32_t bfs_depth_hint(BfsWork *w) { return w->last_levels; }
void bfs_set_depth_hint(BfsWork *w, uint32_t d) { w->last_levels = d; }
uint32_t *bfs_ctl_words(BfsWork *w) { return reinterpret_cast<uint32_t *>(w->ctl); }
GridBar *bfs_bar(BfsWork *w) { return w->bar; }

// Enqueue levels in growing chunks until the device reports done.
// chunk0 == 1 gives one host round trip per level (OLV_HOST_LOOP).  The
// stop event of `t` follows the last enqueued level, ahead of the control
// block read-back, so host latency after the last kernel is not timed (the
// gaps between chunks are).
static int run_pipeline(const olv_graph *g, BfsWork *w, int32_t *levels, bool can_bu, uint32_t chunk0, Timer *t) {
  uint32_t chunk = std::max(1u, chunk0);
  uint64_t enq = 0;
  for (;;) {
    OLV_TRY(bfs_enqueue_levels(g, w, levels, can_bu, chunk));
    if (t) t->stop();
    enq += chunk;
    OLV_TRY(bfs_result(w, nullptr, nullptr));
    if (w->h_ctl->done) break;
    if (enq > (uint64_t)g->n + 2) return fail(OLV_EINTERNAL, "olv_bfs: level loop did not terminate");
    if (chunk0 > 1) chunk = std::min(chunk * 2, 256u);
  }
  return OLV_OK;
}

static int collect_stats(const olv_graph *g, BfsWork *w, const int32_t *levels, olv_stats *st) {
  if (!st) return OLV_OK;
  OLV_HIP(hipMemsetAsync(w->stat, 0, 2 * sizeof(unsigned long long), stream()));
  hipLaunchKernelGGL(bfs_stat_kernel, dim3(grid_for(g->n, kBlock, 8)), dim3(kBlock), 0, stream(), levels, g->row_ptr,
                     g->n, w->stat);
  unsigned long long h[2];
  OLV_HIP(hipMemcpyAsync(h, w->stat, sizeof(h), hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  st->vertices_reached = h[0];
  st->edges_traversed = h[1];
  return OLV_OK;
}

// ------------------------------------------------------------- small graphs
// A graph of at most 16,384 vertices whose out-lists are short runs its
// whole traversal in one workgroup: the row offsets, the visited bits and
// both frontier queues sit in LDS (~130 KB), so a level costs the column
// loads (L2 hits after the first level), LDS claims and one barrier -- not
// two kernel launches.  The 128 x 128 grid (255 levels of at most 128
// vertices) was launch-bound at ~11 us per level in the pipeline.
constexpr uint32_t kSmallN = 16384;
constexpr uint32_t kSmallMaxDeg = 256;
constexpr int kSmallBlock = 1024;

// out[0] = levels run (the deepest level + 1), out[1] = edges read (= the
// out-degree sum of reached vertices), out[2] = vertices reached.
__global__ void __launch_bounds__(kSmallBlock) bfs_small_kernel(const uint32_t *__restrict__ rp,
                                                                const uint32_t *__restrict__ col, uint32_t n,
                                                                const uint32_t *__restrict__ seeds, uint32_t k,
                                                                int32_t *__restrict__ levels,
                                                                unsigned long long *__restrict__ out) {
  __shared__ uint32_t s_rp[kSmallN + 1];
  __shared__ uint32_t s_vis[kSmallN / 32];
  __shared__ uint16_t s_q[2][kSmallN];
  __shared__ uint32_t s_cnt[2];
  __shared__ unsigned long long s_red[kSmallBlock / kWave];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i <= n; i += kSmallBlock) s_rp[i] = rp[i];
  for (uint32_t i = t; i < kSmallN / 32; i += kSmallBlock) s_vis[i] = 0u;
  for (uint32_t v = t; v < n; v += kSmallBlock) levels[v] = kInf;
  if (t < 2) s_cnt[t] = 0u;
  __syncthreads();
  for (uint32_t i = t; i < k; i += kSmallBlock) {
    const uint32_t v = seeds[i], bit = 1u << (v & 31);
    if (!(atomicOr(&s_vis[v >> 5], bit) & bit)) {
      s_q[0][atomicAdd(&s_cnt[0], 1u)] = (uint16_t)v;
      levels[v] = 0;
    }
  }
  __syncthreads();
  uint32_t cur = 0, L = 0;
  unsigned long long insp = 0, reached = 0;
  for (;;) {
    const uint32_t qs = s_cnt[cur];
    if (qs == 0 || L > n) break;
    reached += t < qs ? (qs - t + kSmallBlock - 1) / kSmallBlock : 0u;
    if (t == 0) s_cnt[cur ^ 1] = 0u;
    __syncthreads();
    const uint32_t nxt = cur ^ 1;
    for (uint32_t i = t; i < qs; i += kSmallBlock) {
      const uint32_t u = s_q[cur][i];
      const uint32_t b = s_rp[u], e = s_rp[u + 1];
      insp += e - b;
      for (uint32_t x = b; x < e; x += 4) {
        uint32_t v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = x + j < e ? col[x + j] : 0xffffffffu;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (v[j] == 0xffffffffu) continue;
          const uint32_t bit = 1u << (v[j] & 31);
          if (s_vis[v[j] >> 5] & bit) continue;
          if (!(atomicOr(&s_vis[v[j] >> 5], bit) & bit)) {
            levels[v[j]] = (int32_t)(L + 1);
            s_q[nxt][atomicAdd(&s_cnt[nxt], 1u)] = (uint16_t)v[j];
          }
        }
      }
    }
    __syncthreads();
    cur = nxt;
    ++L;
  }
  const unsigned long long ti = block_sum<kSmallBlock, unsigned long long>(insp, s_red);
  const unsigned long long tr = block_sum<kSmallBlock, unsigned long long>(reached, s_red);
  if (t == 0) {
    out[0] = L;
    out[1] = ti;
    out[2] = tr;
  }
}

static bool small_bfs_fits(const olv_graph *g, unsigned flags) {
  return g->n <= kSmallN && g->max_outdeg <= kSmallMaxDeg &&
         !(flags & (OLV_PERSISTENT | OLV_HOST_LOOP | OLV_FORCE_BOTTOM_UP));
}

static int bfs_small(const olv_graph *g, const std::vector<uint32_t> &seeds, int32_t *levels, bool host,
                     olv_stats *st) {
  const uint32_t k = (uint32_t)seeds.size();
  DevBuf d_seeds, d_out, d_levels;
  OLV_TRY(d_seeds.alloc((size_t)k * 4));
  OLV_TRY(d_out.alloc(3 * sizeof(unsigned long long)));
  OLV_HIP(hipMemcpyAsync(d_seeds.p, seeds.data(), (size_t)k * 4, hipMemcpyHostToDevice, stream()));
  int32_t *outl = levels;
  if (host) {
    OLV_TRY(d_levels.alloc((size_t)g->n * 4));
    outl = d_levels.as<int32_t>();
  }
  Timer t;
  t.start();
  hipLaunchKernelGGL(bfs_small_kernel, dim3(1), dim3(kSmallBlock), 0, stream(), g->row_ptr, g->col, g->n,
                     d_seeds.as<uint32_t>(), k, outl, d_out.as<unsigned long long>());
  OLV_HIP(hipGetLastError());
  t.stop();
  unsigned long long h[3];
  OLV_HIP(hipMemcpyAsync(h, d_out.p, sizeof(h), hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  if (h[0] > g->n) return fail(OLV_EINTERNAL, "olv_bfs: more than n levels");
  if (host) OLV_HIP(hipMemcpy(levels, outl, (size_t)g->n * 4, hipMemcpyDeviceToHost));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = t.ms();
    st->iterations = (uint32_t)h[0];
    st->edges_inspected = h[1];
    st->edges_traversed = h[1];
    st->vertices_reached = h[2];
  }
  return OLV_OK;
}

// Multi-source BFS on g (levels = distance to the nearest seed).
// seeds_host: seeds is a host array (levels follow OLV_PTR_HOST in flags).
// Direction-optimising runs go to g's twin -- g renumbered by descending
// in-degree, so hubs and their visited bytes sit together, and in-lists
// start with the hubs -- whose level stores go through the twin's out_map
// straight to g's numbering.  BFS levels do not depend on the numbering, so
// they are the same values.  OLV_GRAPH_ORDER (or OLV_NO_DIRECTION_OPT) runs
// on g itself.
int bfs_multisource(const olv_graph *g0, const uint32_t *seeds, uint32_t k, bool seeds_host, int32_t *levels,
                    unsigned flags, olv_stats *st) {
  if (!g0 || !seeds || !levels) return fail(OLV_EINVAL, "olv_bfs: NULL argument");
  if (k == 0) return fail(OLV_EINVAL, "olv_bfs: no source");
  const bool host = (flags & OLV_PTR_HOST) != 0;
  std::vector<uint32_t> h_seeds(k);
  if (seeds_host) std::memcpy(h_seeds.data(), seeds, (size_t)k * 4);
  else OLV_HIP(hipMemcpy(h_seeds.data(), seeds, (size_t)k * 4, hipMemcpyDeviceToHost));
  for (uint32_t i = 0; i < k; ++i)
    if (h_seeds[i] >= g0->n) return fail(OLV_EINVAL, "olv_bfs: source %u out of range (n=%u)", h_seeds[i], g0->n);
  const bool can_bu = !(flags & OLV_NO_DIRECTION_OPT);
  if (!can_bu && (flags & OLV_FORCE_BOTTOM_UP))
    return fail(OLV_EINVAL, "olv_bfs: OLV_FORCE_BOTTOM_UP and OLV_NO_DIRECTION_OPT together");
  if (small_bfs_fits(g0, flags)) return bfs_small(g0, h_seeds, levels, host, st);
  const bool twin = can_bu && !(flags & OLV_GRAPH_ORDER);
  const olv_graph *g = g0;
  if (twin) {
    olv_graph *gm = const_cast<olv_graph *>(g0);
    OLV_TRY(ensure_twin(gm));
    gm->twin->tune = gm->tune;
    for (uint32_t &s : h_seeds) s = gm->h_rl_pi[s];
    g = gm->twin;
  }
  BfsWork *w = nullptr;
  OLV_TRY(bfs_prepare(g, can_bu, &w));
  w->force_bu = (flags & OLV_FORCE_BOTTOM_UP) != 0;
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
  OLV_TRY(bfs_enqueue_init(g, w, d_seeds.as<uint32_t>(), k, out, can_bu));
  if (flags & OLV_PERSISTENT) {
    OLV_TRY(enqueue_persistent(g, w, out, can_bu));
    t.stop();
  } else {
    // first chunk: the deepest of the last few runs (traversals of one graph
    // from different sources have similar depths); deeper runs continue in
    // doubling chunks after one host check
    const uint32_t recent_max = std::max(std::max(w->recent[0], w->recent[1]), std::max(w->recent[2], w->recent[3]));
    const uint32_t chunk0 = (flags & OLV_HOST_LOOP) ? 1u : std::max(4u, recent_max);
    OLV_TRY(run_pipeline(g, w, out, can_bu, chunk0, &t));
  }
  const double ms = t.ms();
  uint32_t levels_run = 0, bu_levels = 0;
  OLV_TRY(bfs_result(w, &levels_run, &bu_levels));
  if (flags & OLV_PERSISTENT) {
    GridBar hb;
    OLV_HIP(hipMemcpy(&hb, w->bar, sizeof(hb), hipMemcpyDeviceToHost));
    if (hb.err || !w->h_ctl->done) return fail(OLV_EINTERNAL, "olv_bfs: grid barrier timed out (co-residency lost)");
  }
  if (w->h_ctl->overflow) return fail(OLV_EINTERNAL, "olv_bfs: more than n levels");
  w->last_levels = levels_run;
  w->recent[3] = w->recent[2];
  w->recent[2] = w->recent[1];
  w->recent[1] = w->recent[0];
  w->recent[0] = levels_run;
  if (host) OLV_HIP(hipMemcpy(levels, out, (size_t)g->n * 4, hipMemcpyDeviceToHost));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = levels_run;
    st->bottom_up_levels = bu_levels;
    st->edges_inspected = w->h_ctl->inspected;
    OLV_TRY(collect_stats(g0, w, out, st));  // levels are in g0's numbering
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

// ------------------------------------------------ allocation-free entry point
// olv_bfs_noalloc keeps every work buffer of a traversal (visited bytes,
// snapshot and frontier bitmaps, queue, hub items, control block, barrier,
// statistics, the seed) in one static device array per GPU instead of
// buffers allocated with the graph.  The host code only ever uses the array's
// device address, obtained once per device with hipGetSymbolAddress.
namespace olv {
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
}  // namespace olv

static __device__ __align__(256) unsigned char scratch[olv::kNoAllocScratchBytes];

namespace olv {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

struct NoAllocLayout {
  size_t visited, snap, bms, queue, hubs, ctl, bar, stat, seed, total;
  uint32_t nw, nwp;
};

NoAllocLayout noalloc_layout(uint32_t n, uint64_t m) {
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  NoAllocLayout L;
  L.nw = (n + 31) / 32;
  L.nwp = (L.nw + 63) / 64 * 64;
  size_t o = 0;
  L.visited = o;
  o += al((size_t)L.nwp * 32);
  L.snap = o;
  o += al((size_t)L.nwp * 4);
  L.bms = o;
  o += al((size_t)L.nwp * 8);
  L.queue = o;
  o += al(std::max<size_t>(n, 1) * 4);
  L.hubs = o;
  o += al(hub_capacity(m) * sizeof(uint2));
  L.ctl = o;
  o += al(sizeof(BfsCtl));
  L.bar = o;
  o += al(sizeof(GridBar));
  L.stat = o;
  o += al(2 * sizeof(unsigned long long));
  L.seed = o;
  o += al(sizeof(uint32_t));
  L.total = o;
  return L;
}

__global__ void store_word_kernel(uint32_t *__restrict__ p, uint32_t v) { *p = v; }
}  // namespace

// The current device's copy of `scratch` (its address looked up once per
// device, with that device current).
int noalloc_scratch(unsigned char **out) {
  const int dev = current_device();
  if (dev < 0 || dev >= kMaxDevices) return fail(OLV_EUNSUPPORTED, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(OLV_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return OLV_OK;
}

}  // namespace olv

extern "C" int olv_bfs_noalloc(const olv_graph *g, uint32_t src, int32_t *levels, unsigned flags, olv_stats *st) {
  if (!g || !levels) return fail(OLV_EINVAL, "olv_bfs_noalloc: NULL argument");
  if (src >= g->n) return fail(OLV_EINVAL, "olv_bfs_noalloc: source %u out of range (n=%u)", src, g->n);
  if (flags & OLV_PTR_HOST) return fail(OLV_EINVAL, "olv_bfs_noalloc: levels must be a device pointer");
  if (flags & OLV_PERSISTENT) return fail(OLV_EINVAL, "olv_bfs_noalloc: OLV_PERSISTENT is not supported");
  const bool can_bu = !(flags & OLV_NO_DIRECTION_OPT);
  if (!can_bu && (flags & OLV_FORCE_BOTTOM_UP))
    return fail(OLV_EINVAL, "olv_bfs_noalloc: OLV_FORCE_BOTTOM_UP and OLV_NO_DIRECTION_OPT together");
  if (can_bu && (!g->in_row_ptr || !g->in_head))
    return fail(OLV_EINVAL,
                "olv_bfs_noalloc: the graph has no in-edge CSR with in-list heads (create it with "
                "OLV_BUILD_TRANSPOSE and run olv_bfs on it once, or pass OLV_NO_DIRECTION_OPT)");
  const NoAllocLayout L = noalloc_layout(g->n, g->m);
  if (L.total > kNoAllocScratchBytes)
    return fail(OLV_ERANGE, "olv_bfs_noalloc: the graph needs %zu bytes of work memory, the static scratch holds %zu "
                            "(n=%u, m=%u)",
                L.total, kNoAllocScratchBytes, g->n, g->m);
  unsigned char *base = nullptr;
  OLV_TRY(noalloc_scratch(&base));
  BfsWork w;  // views into the scratch array; nothing here is freed
  BfsCtl hctl;
  w.n = g->n;
  w.nw = L.nw;
  w.nwp = L.nwp;
  w.visited = base + L.visited;
  w.snap = reinterpret_cast<uint32_t *>(base + L.snap);
  w.bms = reinterpret_cast<uint32_t *>(base + L.bms);
  w.queue = reinterpret_cast<uint32_t *>(base + L.queue);
  w.hubs = reinterpret_cast<uint2 *>(base + L.hubs);
  w.ctl = reinterpret_cast<BfsCtl *>(base + L.ctl);
  w.bar = reinterpret_cast<GridBar *>(base + L.bar);
  w.stat = reinterpret_cast<unsigned long long *>(base + L.stat);
  w.h_ctl = &hctl;
  w.force_bu = (flags & OLV_FORCE_BOTTOM_UP) != 0;
  uint32_t *seed = reinterpret_cast<uint32_t *>(base + L.seed);
  const auto t0 = std::chrono::steady_clock::now();
  hipLaunchKernelGGL(store_word_kernel, dim3(1), dim3(1), 0, stream(), seed, src);
  OLV_TRY(bfs_enqueue_init(g, &w, seed, 1, levels, can_bu));
  OLV_TRY(run_pipeline(g, &w, levels, can_bu, (flags & OLV_HOST_LOOP) ? 1u : 8u, nullptr));
  uint32_t levels_run = 0, bu_levels = 0;
  OLV_TRY(bfs_result(&w, &levels_run, &bu_levels));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (hctl.overflow) return fail(OLV_EINTERNAL, "olv_bfs_noalloc: more than n levels");
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = levels_run;
    st->bottom_up_levels = bu_levels;
    st->edges_inspected = hctl.inspected;
    OLV_TRY(collect_stats(g, &w, levels, st));
  }
  return OLV_OK;
}
