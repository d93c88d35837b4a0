// Batched traversals (no reference counterpart; each result equals repeated
// runs of the plain engine):
//  * olv_bfs_batched: up to 256 sources per sweep, bit-parallel.  Every
//    vertex carries a 256-bit visited mask V, a frontier mask F and (for
//    push levels) an accumulator N; one adjacency sweep per level serves all
//    sources.  Push (top-down) levels OR F[u] into N[v] with atomics;
//    pull (bottom-up) levels OR the F of in-neighbours into v and stop once
//    v has every source.  Source s's level of v is written when bit s of
//    V[v] first turns on.
//  * olv_bfs_multisource / olv_graph_union / olv_bfs_multigraph: many
//    graphs as one block-diagonal union traversed by one multi-source BFS.
//  * olv_plan_*: the whole traversal (reset, seeding, persistent kernel)
//    captured once in a hipGraph and replayed.
//  * olv_autotune: picks degree-bucket thresholds and direction-switch
//    parameters for a graph by timing the plain BFS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bfs_core.hpp"
#include "internal.hpp"

namespace olv {

int bfs_multisource(const olv_graph *g, const uint32_t *seeds, uint32_t k, bool seeds_host, int32_t *levels,
                    unsigned flags, olv_stats *st);
int bfs_prepare(const olv_graph *g, bool can_bu, BfsWork **w);
int bfs_enqueue(const olv_graph *g, BfsWork *w, const uint32_t *d_seeds, uint32_t k, int32_t *levels, bool can_bu);
int bfs_result(BfsWork *w, uint32_t *levels_run, uint32_t *bu_levels);
int reach_stats(const olv_graph *g, const int32_t *d, uint32_t k, olv_stats *st);
uint32_t *bfs_result_words(BfsWork *w);

constexpr int kMB = 256;          // workgroup size
constexpr int kMW = 8;            // 32-bit words per vertex mask
constexpr uint32_t kMSpin = 1u << 26;

struct MbfsParams {
  const uint32_t *row_ptr, *col, *in_row_ptr, *in_col;
  uint32_t n, nw, nwp;
  uint32_t k;
  uint32_t full[kMW];
  uint32_t *V;     // n * kMW
  uint32_t *F;     // 2 * n * kMW
  uint32_t *N;     // n * kMW
  uint32_t *A;     // 3 * nwp
  uint32_t *T;     // 2 * nwp
  uint32_t *D;     // nwp
  uint32_t *queue; // n
  int32_t *levels; // k * n
  BfsCtr *ctr;     // 3
  uint32_t *bar;   // 16
  unsigned long long m;
  uint32_t wave_thr, block_thr;
  float alpha, beta;
  uint32_t can_bu;
};

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
  return v;
}

struct MbfsTdVisitor {
  const MbfsParams &p;
  const uint32_t *Ff;
  uint32_t *T;
  __device__ MbfsTdVisitor(const MbfsParams &pp, const uint32_t *f, uint32_t *t) : p(pp), Ff(f), T(t) {}
  __device__ __forceinline__ uint32_t row(uint32_t q) const { return q; }
  __device__ __forceinline__ uint32_t load(uint32_t q) const { return q; }
  __device__ __forceinline__ void visit(uint32_t e, uint32_t u) {
    const uint32_t v = p.col[e];
    if (bm_test(p.D, v)) return;
    bool any = false;
#pragma unroll
    for (int j = 0; j < kMW; ++j) {
      const uint32_t f = Ff[(size_t)u * kMW + j];
      if (!f) continue;
      const uint32_t nv = f & ~p.V[(size_t)v * kMW + j];
      if (nv) {
        atomicOr(&p.N[(size_t)v * kMW + j], nv);
        any = true;
      }
    }
    if (any) {
      const uint32_t bit = 1u << (v & 31);
      if (!(T[v >> 5] & bit)) atomicOr(&T[v >> 5], bit);
    }
  }
};

// Apply new bits of v: V |= nw, F_next = nw, levels for each new source.
__device__ __forceinline__ bool mbfs_apply(const MbfsParams &p, uint32_t v, const uint32_t *nwv, uint32_t *vv,
                                           uint32_t *Fn, int32_t nlevel) {
  bool full = true;
#pragma unroll
  for (int j = 0; j < kMW; ++j) {
    Fn[(size_t)v * kMW + j] = nwv[j];
    if (nwv[j]) p.V[(size_t)v * kMW + j] = vv[j] | nwv[j];
    full = full && ((vv[j] | nwv[j]) == p.full[j]);
  }
  return full;
}

template <int BLOCK>
__device__ void mbfs_contract(const MbfsParams &p, const uint32_t *T, uint32_t *An, uint32_t *Fn, int32_t nlevel,
                              uint32_t vb, uint32_t nvb, uint32_t &nf, unsigned long long &mf) {
  for (uint32_t w = vb * BLOCK + threadIdx.x; w < p.nw; w += nvb * BLOCK) {
    uint32_t tw = T[w];
    uint32_t anext = 0, dnew = 0;
    while (tw) {
      const uint32_t b = __ffs(tw) - 1;
      tw &= tw - 1;
      const uint32_t v = w * 32u + b;
      uint32_t nwv[kMW], vv[kMW];
      bool any = false;
#pragma unroll
      for (int j = 0; j < kMW; ++j) {
        const uint32_t nj = p.N[(size_t)v * kMW + j];
        if (nj) p.N[(size_t)v * kMW + j] = 0u;
        vv[j] = p.V[(size_t)v * kMW + j];
        nwv[j] = nj & ~vv[j];
        any = any || nwv[j];
      }
      if (!any) continue;
      if (mbfs_apply(p, v, nwv, vv, Fn, nlevel)) dnew |= 1u << b;
      anext |= 1u << b;
      nf += 1;
      mf += p.row_ptr[v + 1] - p.row_ptr[v];
#pragma unroll
      for (int j = 0; j < kMW; ++j) {
        uint32_t x = nwv[j];
        while (x) {
          const uint32_t s = j * 32 + (__ffs(x) - 1);
          x &= x - 1;
          p.levels[(size_t)s * p.n + v] = nlevel;
        }
      }
    }
    if (anext) An[w] = anext;
    if (dnew) p.D[w] |= dnew;
  }
}

template <int BLOCK>
__device__ void mbfs_bu(const MbfsParams &p, const uint32_t *Af, const uint32_t *Ff, uint32_t *An, uint32_t *Fn,
                        int32_t nlevel, uint32_t vb, uint32_t nvb, uint32_t &nf, unsigned long long &mf,
                        unsigned long long &insp) {
  const uint32_t lane = lane_id();
  const uint32_t waves = nvb * (BLOCK / kWave);
  const uint32_t gw = vb * (BLOCK / kWave) + threadIdx.x / kWave;
  for (uint64_t base = (uint64_t)gw * kWave; base < p.n; base += (uint64_t)waves * kWave) {
    const uint32_t v = (uint32_t)base + lane;
    const bool inb = v < p.n;
    const uint32_t w = v >> 5;
    const uint32_t dw = inb ? p.D[w] : 0xffffffffu;
    const bool todo = inb && !((dw >> (v & 31)) & 1u);
    if (!ballot(todo)) continue;
    uint32_t acc[kMW], vv[kMW];
#pragma unroll
    for (int j = 0; j < kMW; ++j) {
      acc[j] = 0u;
      vv[j] = todo ? p.V[(size_t)v * kMW + j] : 0xffffffffu;
    }
    if (todo) {
      uint32_t k = 0;
      for (uint32_t e = p.in_row_ptr[v]; e < p.in_row_ptr[v + 1]; ++e) {
        ++k;
        const uint32_t u = p.in_col[e];
        if (!bm_test(Af, u)) continue;
        bool done = true;
#pragma unroll
        for (int j = 0; j < kMW; ++j) {
          acc[j] |= Ff[(size_t)u * kMW + j];
          done = done && ((acc[j] | vv[j]) == p.full[j]);
        }
        if (done) break;
      }
      insp += k;
    }
    uint32_t nwv[kMW];
    bool any = false;
#pragma unroll
    for (int j = 0; j < kMW; ++j) {
      nwv[j] = acc[j] & ~vv[j];
      any = any || nwv[j];
    }
    bool isfull = false;
    if (any) {
      isfull = mbfs_apply(p, v, nwv, vv, Fn, nlevel);
      nf += 1;
      mf += p.row_ptr[v + 1] - p.row_ptr[v];
    }
    // coalesced level writes: walk the sources present anywhere in the wave
#pragma unroll
    for (int j = 0; j < kMW; ++j) {
      uint32_t wm = wave_or(nwv[j]);
      while (wm) {
        const uint32_t b = __ffs(wm) - 1;
        wm &= wm - 1;
        if ((nwv[j] >> b) & 1u) p.levels[(size_t)(j * 32 + b) * p.n + v] = nlevel;
      }
    }
    const uint64_t am = ballot(any), fm = ballot(isfull);
    const uint32_t amine = lane < 32 ? (uint32_t)am : (uint32_t)(am >> 32);
    const uint32_t fmine = lane < 32 ? (uint32_t)fm : (uint32_t)(fm >> 32);
    if (lane == 0 || lane == 32) {
      if (amine) An[w] = amine;
      if (fmine) p.D[w] = dw | fmine;
    }
  }
}

__global__ void __launch_bounds__(kMB) mbfs_persistent_kernel(MbfsParams p) {
  uint32_t gen = 0;
  const uint32_t vb = blockIdx.x, nvb = gridDim.x;
  unsigned long long mu = p.m;
  const uint32_t nseed = __hip_atomic_load(&p.ctr[0].qsize, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t i = 0; i < nseed; ++i) {
    const uint32_t s = p.queue[i];
    mu -= p.row_ptr[s + 1] - p.row_ptr[s];
  }
  const size_t fstride = (size_t)p.n * kMW;
  bool bu = false;
  uint32_t L = 0, nbu = 0;
  for (;;) {
    BfsCtr *cur = &p.ctr[L % 3];
    BfsCtr *nxt = &p.ctr[(L + 1) % 3];
    const uint32_t *Ff = p.F + (size_t)((L + 1) & 1) * fstride;
    uint32_t *Fn = p.F + (size_t)(L & 1) * fstride;
    const uint32_t *Af = p.A + (size_t)((L + 2) % 3) * p.nwp;
    uint32_t *An = p.A + (size_t)(L % 3) * p.nwp;
    uint32_t *Ac = p.A + (size_t)((L + 1) % 3) * p.nwp;
    uint32_t *Tc = p.T + (size_t)(L & 1) * p.nwp;
    uint32_t *Tx = p.T + (size_t)((L + 1) & 1) * p.nwp;
    if (vb == 0 && threadIdx.x < sizeof(BfsCtr) / 4) reinterpret_cast<uint32_t *>(nxt)[threadIdx.x] = 0u;
    for (uint32_t i = vb * kMB + threadIdx.x; i < p.nwp; i += nvb * kMB) {
      Ac[i] = 0u;
      Tx[i] = 0u;
    }
    uint32_t nf = 0;
    unsigned long long mf = 0, insp = 0;
    const int32_t nlevel = (int32_t)(L + 1);
    if (!bu) {
      const uint32_t qsize = __hip_atomic_load(&cur->qsize, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      MbfsTdVisitor vis(p, Ff, Tc);
      push_expand<kMB>(p.queue, qsize, p.row_ptr, p.wave_thr, p.block_thr, vb, nvb, vis, insp);
      if (!grid_barrier(p.bar, nvb, gen, kMSpin)) break;
      mbfs_contract<kMB>(p, Tc, An, Fn, nlevel, vb, nvb, nf, mf);
    } else {
      mbfs_bu<kMB>(p, Af, Ff, An, Fn, nlevel, vb, nvb, nf, mf, insp);
    }
    flush_counters<kMB>(cur, nf, mf, insp);
    if (!grid_barrier(p.bar, nvb, gen, kMSpin)) break;
    const uint32_t tnf = __hip_atomic_load(&cur->nf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long tmf = __hip_atomic_load(&cur->mf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (bu) ++nbu;
    ++L;
    if (tnf == 0 || L > p.n) break;
    mu = mu > tmf ? mu - tmf : 0ull;
    const bool next_bu = choose_bottom_up(bu, p.can_bu != 0, tnf, tmf, mu, p.n, p.alpha, p.beta);
    if (!next_bu) {
      compact_step<kMB>(An, p.nw, p.queue, &nxt->qsize, vb, nvb);
      if (!grid_barrier(p.bar, nvb, gen, kMSpin)) break;
    }
    bu = next_bu;
  }
  if (vb == 0 && threadIdx.x == 0) {
    p.bar[4] = L;
    p.bar[5] = nbu;
  }
}

// Reset state for one group of sources (grid-stride).
__global__ void mbfs_reset_kernel(MbfsParams p) {
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t nm = (size_t)p.n * kMW;
  for (size_t i = tid; i < nm; i += stride) {
    p.V[i] = 0u;
    p.N[i] = 0u;
  }
  for (size_t i = tid; i < (size_t)p.nwp * 3; i += stride) p.A[i] = 0u;
  for (size_t i = tid; i < (size_t)p.nwp * 2; i += stride) p.T[i] = 0u;
  for (size_t i = tid; i < (size_t)p.nwp; i += stride) p.D[i] = 0u;
  if (tid < sizeof(BfsCtr) * 3 / 4) reinterpret_cast<uint32_t *>(p.ctr)[tid] = 0u;
  if (tid < 16) p.bar[tid] = 0u;
}

// Level-0 state from the sources (one workgroup): V and the level-0
// frontier F (buffer 1) get bit s at srcs[s]; queue = distinct sources.
__global__ void mbfs_seed_kernel(MbfsParams p, const uint32_t *__restrict__ srcs) {
  const size_t fstride = (size_t)p.n * kMW;
  uint32_t *F0 = p.F + fstride;  // frontier of level 0 is buffer (0+1)&1
  uint32_t *A0 = p.A + (size_t)2 * p.nwp;
  // clear the seeds' frontier masks first
  for (uint32_t s = threadIdx.x; s < p.k; s += blockDim.x)
    for (int j = 0; j < kMW; ++j) F0[(size_t)srcs[s] * kMW + j] = 0u;
  __syncthreads();
  for (uint32_t s = threadIdx.x; s < p.k; s += blockDim.x) {
    const uint32_t v = srcs[s];
    atomicOr(&p.V[(size_t)v * kMW + s / 32], 1u << (s % 32));
    atomicOr(&F0[(size_t)v * kMW + s / 32], 1u << (s % 32));
    p.levels[(size_t)s * p.n + v] = 0;
    const uint32_t old = atomicOr(&A0[v >> 5], 1u << (v & 31));
    if (!(old & (1u << (v & 31)))) {
      const uint32_t pos = atomicAdd(&p.ctr[0].qsize, 1u);
      p.queue[pos] = v;
    }
  }
  __syncthreads();
  // vertices whose mask is already full (only possible when k sources share it)
  for (uint32_t s = threadIdx.x; s < p.k; s += blockDim.x) {
    const uint32_t v = srcs[s];
    bool full = true;
    for (int j = 0; j < kMW; ++j) full = full && (p.V[(size_t)v * kMW + j] == p.full[j]);
    if (full) atomicOr(&p.D[v >> 5], 1u << (v & 31));
  }
}

// Fill with 16-byte stores on the 16-byte-aligned body, scalar head/tail.
__global__ void fill_i32_kernel(int32_t *__restrict__ p, int32_t v, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const uintptr_t addr = reinterpret_cast<uintptr_t>(p);
  size_t head = ((16 - (addr & 15)) & 15) / 4;
  if (head > count) head = count;
  if (i < head) p[i] = v;
  const size_t n4 = (count - head) / 4;
  int4 *q = reinterpret_cast<int4 *>(p + head);
  const int4 val = make_int4(v, v, v, v);
  for (size_t j = i; j < n4; j += stride) q[j] = val;
  for (size_t j = head + n4 * 4 + i; j < count; j += stride) p[j] = v;
}

struct MbfsWork {
  uint32_t *V = nullptr, *F = nullptr, *N = nullptr, *A = nullptr, *T = nullptr, *D = nullptr, *queue = nullptr;
  BfsCtr *ctr = nullptr;
  uint32_t *bar = nullptr, *h_bar = nullptr;
  ~MbfsWork() {
    void *ps[] = {V, F, N, A, T, D, queue, ctr, bar};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    if (h_bar) (void)hipHostFree(h_bar);
  }
  int alloc(uint32_t n, uint32_t nwp) {
    const size_t nm = (size_t)n * kMW * 4;
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&V), nm));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&F), nm * 2));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&N), nm));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&A), (size_t)nwp * 4 * 3));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&T), (size_t)nwp * 4 * 2));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&D), (size_t)nwp * 4));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&queue), (size_t)n * 4));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&ctr), sizeof(BfsCtr) * 3));
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&bar), 64));
    OLV_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_bar), 64, hipHostMallocDefault));
    return OLV_OK;
  }
};

static unsigned mgrid(size_t items, size_t per) {
  size_t b = (items + per - 1) / per;
  return (unsigned)std::max<size_t>(1, std::min(b, (size_t)cu_count() * 8));
}

static int mbfs_grid(int *grid) {
  int per_cu = 0;
  OLV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, mbfs_persistent_kernel, kMB, 0));
  if (per_cu <= 0) return fail(OLV_EUNSUPPORTED, "batched BFS kernel does not fit on a CU");
  *grid = std::min(per_cu, 2) * cu_count();
  return OLV_OK;
}

static MbfsParams mbfs_params(const olv_graph *g, MbfsWork &w, uint32_t k, int32_t *levels) {
  MbfsParams p;
  std::memset(&p, 0, sizeof(p));
  p.row_ptr = g->row_ptr;
  p.col = g->col;
  p.in_row_ptr = g->in_row_ptr;
  p.in_col = g->in_col;
  p.n = g->n;
  p.nw = (g->n + 31) / 32;
  p.nwp = (p.nw + 63) / 64 * 64;
  p.k = k;
  for (int j = 0; j < kMW; ++j) {
    const int lo = j * 32;
    const int bits = std::max(0, std::min(32, (int)k - lo));
    p.full[j] = bits == 32 ? 0xffffffffu : ((1u << bits) - 1u);
  }
  p.V = w.V;
  p.F = w.F;
  p.N = w.N;
  p.A = w.A;
  p.T = w.T;
  p.D = w.D;
  p.queue = w.queue;
  p.levels = levels;
  p.ctr = w.ctr;
  p.bar = w.bar;
  p.m = g->m;
  p.wave_thr = g->tune.td_wave_thr;
  p.block_thr = g->tune.td_block_thr;
  p.alpha = g->tune.alpha;
  p.beta = g->tune.beta;
  p.can_bu = g->in_row_ptr ? 1u : 0u;
  return p;
}

// Enqueue one group (<= 256 sources): reset, level fill, seed, persistent.
static int mbfs_enqueue_group(const MbfsParams &p, const uint32_t *d_srcs, int grid) {
  hipStream_t s = stream();
  hipLaunchKernelGGL(mbfs_reset_kernel, dim3(mgrid((size_t)p.n * kMW, 256)), dim3(256), 0, s, p);
  hipLaunchKernelGGL(fill_i32_kernel, dim3(mgrid((size_t)p.k * p.n / 4 + 1, 256)), dim3(256), 0, s, p.levels, kInf,
                     (size_t)p.k * p.n);
  hipLaunchKernelGGL(mbfs_seed_kernel, dim3(1), dim3(256), 0, s, p, d_srcs);
  OLV_HIP(hipGetLastError());
  MbfsParams pp = p;
  void *args[] = {&pp};
  OLV_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(mbfs_persistent_kernel), dim3(grid), dim3(kMB),
                                     args, 0, s));
  return OLV_OK;
}

}  // namespace olv

// --------------------------------------------------------------- plans
struct olv_plan {
  int kind = 0;  // 1 = multi-source BFS, 2 = batched BFS
  const olv_graph *g = nullptr;
  olv::DevBuf seeds;
  uint32_t k = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  olv::BfsWork *bw = nullptr;
  olv::MbfsWork *mw = nullptr;
  uint32_t *result = nullptr;  // device words [4]=levels [5]=bu levels [2]=error
  uint32_t *h_result = nullptr;
  ~olv_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    delete mw;
    if (h_result) (void)hipHostFree(h_result);
  }
};

namespace olv {

// Capture `enqueue` (stream-ordered work) into plan->graph/exec.
template <class Fn>
static int capture(olv_plan *plan, Fn enqueue) {
  hipStream_t user = stream();
  hipStream_t cap = user;
  bool own = false;
  if (!cap) {
    OLV_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    own = true;
  }
  OLV_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
  olv_set_stream(cap);
  int st = enqueue();
  olv_set_stream(user);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(cap, &graph);
  if (own) (void)hipStreamDestroy(cap);
  if (st != OLV_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) return fail(OLV_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
  plan->graph = graph;
  e = hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) return fail(OLV_EHIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
  return OLV_OK;
}

}  // namespace olv

using namespace olv;

extern "C" {

int olv_bfs_multisource(const olv_graph *g, const uint32_t *srcs, uint32_t k, int32_t *levels, unsigned flags,
                        olv_stats *st) {
  return bfs_multisource(g, srcs, k, true, levels, flags, st);
}

int olv_bfs_batched(const olv_graph *cg, const uint32_t *srcs, uint32_t k, int32_t *levels, unsigned flags,
                    olv_stats *st) {
  if (!cg || !srcs || !levels) return fail(OLV_EINVAL, "olv_bfs_batched: NULL argument");
  if (k == 0) return fail(OLV_EINVAL, "olv_bfs_batched: no source");
  olv_graph *g = const_cast<olv_graph *>(cg);
  for (uint32_t i = 0; i < k; ++i)
    if (srcs[i] >= g->n) return fail(OLV_EINVAL, "olv_bfs_batched: source %u out of range (n=%u)", srcs[i], g->n);
  if ((uint64_t)k * g->n > (1ull << 40)) return fail(OLV_ERANGE, "olv_bfs_batched: output too large");
  if (!(flags & OLV_NO_DIRECTION_OPT) && !g->in_row_ptr) OLV_TRY(olv_graph_build_transpose(g));
  const bool host = (flags & OLV_PTR_HOST) != 0;
  MbfsWork w;
  OLV_TRY(w.alloc(g->n, ((g->n + 31) / 32 + 63) / 64 * 64));
  DevBuf d_srcs, d_out;
  OLV_TRY(d_srcs.alloc((size_t)k * 4));
  OLV_HIP(hipMemcpy(d_srcs.p, srcs, (size_t)k * 4, hipMemcpyHostToDevice));
  int32_t *out = levels;
  if (host) {
    OLV_TRY(d_out.alloc((size_t)k * g->n * 4));
    out = d_out.as<int32_t>();
  }
  int grid = 0;
  OLV_TRY(mbfs_grid(&grid));
  Timer t;
  t.start();
  uint32_t max_levels = 0, bu_levels = 0;
  for (uint32_t g0 = 0; g0 < k; g0 += 256) {
    const uint32_t kg = std::min(256u, k - g0);
    MbfsParams p = mbfs_params(g, w, kg, out + (size_t)g0 * g->n);
    if (flags & OLV_NO_DIRECTION_OPT) p.can_bu = 0;
    OLV_TRY(mbfs_enqueue_group(p, d_srcs.as<uint32_t>() + g0, grid));
    OLV_HIP(hipMemcpyAsync(w.h_bar, w.bar, 64, hipMemcpyDeviceToHost, stream()));
    OLV_HIP(hipStreamSynchronize(stream()));
    if (w.h_bar[2]) return fail(OLV_EINTERNAL, "olv_bfs_batched: grid barrier timed out");
    max_levels = std::max(max_levels, w.h_bar[4]);
    bu_levels += w.h_bar[5];
  }
  t.stop();
  const double ms = t.ms();
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = max_levels;
    st->bottom_up_levels = bu_levels;
    OLV_TRY(reach_stats(g, out, k, st));
  }
  if (host) OLV_HIP(hipMemcpy(levels, out, (size_t)k * g->n * 4, hipMemcpyDeviceToHost));
  return OLV_OK;
}

}  // extern "C"

// ------------------------------------------------------------ graph union
namespace olv {

__global__ void union_rows_kernel(const uint32_t *__restrict__ rp, uint32_t n, uint32_t voff, uint32_t eoff,
                                  uint32_t *__restrict__ out) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    out[voff + v] = rp[v] + eoff;
}

__global__ void union_cols_kernel(const uint32_t *__restrict__ c, uint32_t m, uint32_t add,
                                  uint32_t *__restrict__ out) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x) out[e] = c[e] + add;
}

__global__ void set_word_kernel(uint32_t *p, uint32_t v) { *p = v; }

static int concat_rows(const uint32_t *const *rps, const uint32_t *ns, const uint32_t *voff, const uint32_t *eoff,
                       uint32_t k, uint32_t n_total, uint32_t m_total, uint32_t **out) {
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(out), ((size_t)n_total + 1) * 4));
  for (uint32_t i = 0; i < k; ++i)
    hipLaunchKernelGGL(union_rows_kernel, dim3(mgrid(ns[i], 256)), dim3(256), 0, stream(), rps[i], ns[i], voff[i],
                       eoff[i], *out);
  hipLaunchKernelGGL(set_word_kernel, dim3(1), dim3(1), 0, stream(), *out + n_total, m_total);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

static int concat_cols(const uint32_t *const *cs, const uint32_t *ms, const uint32_t *eoff, const uint32_t *adds,
                       uint32_t k, uint32_t m_total, uint32_t **out) {
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(out), std::max<size_t>(m_total, 1) * 4));
  for (uint32_t i = 0; i < k; ++i)
    if (ms[i])
      hipLaunchKernelGGL(union_cols_kernel, dim3(mgrid(ms[i], 256)), dim3(256), 0, stream(), cs[i], ms[i], adds[i],
                         *out + eoff[i]);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

static int graph_union(olv_graph *const *gs, uint32_t k, olv_graph **out, std::vector<uint32_t> &voff) {
  if (!gs || !out || k == 0) return fail(OLV_EINVAL, "olv_graph_union: bad arguments");
  voff.assign(k + 1, 0);
  std::vector<uint32_t> eoff(k + 1, 0), ns(k), ms(k);
  bool all_t = true;
  uint64_t nt = 0, mt = 0;
  for (uint32_t i = 0; i < k; ++i) {
    if (!gs[i]) return fail(OLV_EINVAL, "olv_graph_union: graph %u is NULL", i);
    ns[i] = gs[i]->n;
    ms[i] = gs[i]->m;
    voff[i] = (uint32_t)nt;
    eoff[i] = (uint32_t)mt;
    nt += gs[i]->n;
    mt += gs[i]->m;
    all_t = all_t && gs[i]->in_row_ptr;
    if (nt >= (1ull << 32) || mt >= (1ull << 32)) return fail(OLV_ERANGE, "olv_graph_union: union too large");
  }
  voff[k] = (uint32_t)nt;
  eoff[k] = (uint32_t)mt;
  olv_graph *u = new olv_graph();
  u->device = current_device();
  u->n = (uint32_t)nt;
  u->m = (uint32_t)mt;
  std::vector<const uint32_t *> rps(k), cs(k), irps(k), ics(k), ies(k);
  for (uint32_t i = 0; i < k; ++i) {
    rps[i] = gs[i]->row_ptr;
    cs[i] = gs[i]->col;
    irps[i] = gs[i]->in_row_ptr;
    ics[i] = gs[i]->in_col;
    ies[i] = gs[i]->in_eid;
    u->max_outdeg = std::max(u->max_outdeg, gs[i]->max_outdeg);
    u->max_indeg = std::max(u->max_indeg, gs[i]->max_indeg);
  }
  int st = OLV_OK;
  do {
    if ((st = concat_rows(rps.data(), ns.data(), voff.data(), eoff.data(), k, u->n, u->m, &u->row_ptr))) break;
    if ((st = concat_cols(cs.data(), ms.data(), eoff.data(), voff.data(), k, u->m, &u->col))) break;
    if (hipMalloc(reinterpret_cast<void **>(&u->inv_outdeg), (size_t)u->n * 4) != hipSuccess) {
      st = fail(OLV_ENOMEM, "olv_graph_union: allocation failed");
      break;
    }
    for (uint32_t i = 0; i < k; ++i)
      if (hipMemcpyAsync(u->inv_outdeg + voff[i], gs[i]->inv_outdeg, (size_t)ns[i] * 4, hipMemcpyDeviceToDevice,
                         stream()) != hipSuccess) {
        st = fail(OLV_EHIP, "olv_graph_union: copy failed");
        break;
      }
    if (st) break;
    if (all_t) {
      if ((st = concat_rows(irps.data(), ns.data(), voff.data(), eoff.data(), k, u->n, u->m, &u->in_row_ptr))) break;
      if ((st = concat_cols(ics.data(), ms.data(), eoff.data(), voff.data(), k, u->m, &u->in_col))) break;
      if ((st = concat_cols(ies.data(), ms.data(), eoff.data(), eoff.data(), k, u->m, &u->in_eid))) break;
    }
    if (hipStreamSynchronize(stream()) != hipSuccess) st = fail(OLV_EHIP, "olv_graph_union: build failed");
  } while (0);
  if (st != OLV_OK) {
    olv_graph_destroy(u);
    return st;
  }
  u->tune = gs[0]->tune;
  *out = u;
  return OLV_OK;
}

}  // namespace olv

extern "C" {

int olv_graph_union(olv_graph *const *gs, uint32_t k, olv_graph **out, uint32_t *vertex_offsets) {
  std::vector<uint32_t> voff;
  OLV_TRY(graph_union(gs, k, out, voff));
  if (vertex_offsets) std::memcpy(vertex_offsets, voff.data(), voff.size() * 4);
  return OLV_OK;
}

int olv_bfs_multigraph(olv_graph *const *gs, uint32_t k, const uint32_t *srcs, int32_t *const *levels,
                       unsigned flags, olv_stats *st) {
  if (!gs || !srcs || !levels || k == 0) return fail(OLV_EINVAL, "olv_bfs_multigraph: bad arguments");
  for (uint32_t i = 0; i < k; ++i) {
    if (!gs[i] || !levels[i]) return fail(OLV_EINVAL, "olv_bfs_multigraph: NULL graph or output %u", i);
    if (srcs[i] >= gs[i]->n) return fail(OLV_EINVAL, "olv_bfs_multigraph: source %u out of range in graph %u", srcs[i], i);
  }
  olv_graph *u = nullptr;
  std::vector<uint32_t> voff;
  OLV_TRY(graph_union(gs, k, &u, voff));
  std::vector<uint32_t> seeds(k);
  for (uint32_t i = 0; i < k; ++i) seeds[i] = voff[i] + srcs[i];
  DevBuf lv;
  int s = lv.alloc((size_t)u->n * 4);
  if (s == OLV_OK) s = bfs_multisource(u, seeds.data(), k, true, lv.as<int32_t>(), flags & ~OLV_PTR_HOST, st);
  if (s == OLV_OK) {
    const bool host = (flags & OLV_PTR_HOST) != 0;
    for (uint32_t i = 0; i < k && s == OLV_OK; ++i) {
      const size_t bytes = (size_t)gs[i]->n * 4;
      hipError_t e = host ? hipMemcpy(levels[i], lv.as<int32_t>() + voff[i], bytes, hipMemcpyDeviceToHost)
                          : hipMemcpyAsync(levels[i], lv.as<int32_t>() + voff[i], bytes, hipMemcpyDeviceToDevice, stream());
      if (e != hipSuccess) s = fail(OLV_EHIP, "olv_bfs_multigraph: output copy failed");
    }
    if (s == OLV_OK && hipStreamSynchronize(stream()) != hipSuccess) s = fail(OLV_EHIP, "olv_bfs_multigraph: sync failed");
  }
  olv_graph_destroy(u);
  return s;
}

int olv_plan_bfs(const olv_graph *g, const uint32_t *srcs, uint32_t k, int32_t *levels, unsigned flags,
                 olv_plan **out) {
  if (!g || !srcs || !levels || !out || k == 0) return fail(OLV_EINVAL, "olv_plan_bfs: bad arguments");
  *out = nullptr;
  for (uint32_t i = 0; i < k; ++i)
    if (srcs[i] >= g->n) return fail(OLV_EINVAL, "olv_plan_bfs: source %u out of range", srcs[i]);
  const bool can_bu = !(flags & OLV_NO_DIRECTION_OPT);
  olv_plan *p = new olv_plan();
  p->kind = 1;
  p->g = g;
  p->k = k;
  int s = bfs_prepare(g, can_bu, &p->bw);
  if (s == OLV_OK) s = p->seeds.alloc((size_t)k * 4);
  if (s == OLV_OK && hipMemcpy(p->seeds.p, srcs, (size_t)k * 4, hipMemcpyHostToDevice) != hipSuccess)
    s = fail(OLV_EHIP, "olv_plan_bfs: seed copy failed");
  if (s == OLV_OK)
    s = capture(p, [&] { return bfs_enqueue(g, p->bw, p->seeds.as<uint32_t>(), k, levels, can_bu); });
  if (s != OLV_OK) {
    delete p;
    return s;
  }
  p->result = bfs_result_words(p->bw);
  *out = p;
  return OLV_OK;
}

int olv_plan_bfs_batched(const olv_graph *cg, const uint32_t *srcs, uint32_t k, int32_t *levels, unsigned flags,
                         olv_plan **out) {
  if (!cg || !srcs || !levels || !out || k == 0) return fail(OLV_EINVAL, "olv_plan_bfs_batched: bad arguments");
  *out = nullptr;
  olv_graph *g = const_cast<olv_graph *>(cg);
  for (uint32_t i = 0; i < k; ++i)
    if (srcs[i] >= g->n) return fail(OLV_EINVAL, "olv_plan_bfs_batched: source %u out of range", srcs[i]);
  if (!(flags & OLV_NO_DIRECTION_OPT) && !g->in_row_ptr) OLV_TRY(olv_graph_build_transpose(g));
  olv_plan *p = new olv_plan();
  p->kind = 2;
  p->g = g;
  p->k = k;
  p->mw = new MbfsWork();
  int grid = 0;
  int s = p->mw->alloc(g->n, ((g->n + 31) / 32 + 63) / 64 * 64);
  if (s == OLV_OK) s = mbfs_grid(&grid);
  if (s == OLV_OK) s = p->seeds.alloc((size_t)k * 4);
  if (s == OLV_OK && hipMemcpy(p->seeds.p, srcs, (size_t)k * 4, hipMemcpyHostToDevice) != hipSuccess)
    s = fail(OLV_EHIP, "olv_plan_bfs_batched: seed copy failed");
  if (s == OLV_OK)
    s = capture(p, [&]() -> int {
      for (uint32_t g0 = 0; g0 < k; g0 += 256) {
        const uint32_t kg = std::min(256u, k - g0);
        MbfsParams prm = mbfs_params(g, *p->mw, kg, levels + (size_t)g0 * g->n);
        if (flags & OLV_NO_DIRECTION_OPT) prm.can_bu = 0;
        OLV_TRY(mbfs_enqueue_group(prm, p->seeds.as<uint32_t>() + g0, grid));
      }
      return OLV_OK;
    });
  if (s != OLV_OK) {
    delete p;
    return s;
  }
  p->result = p->mw->bar;
  *out = p;
  return OLV_OK;
}

int olv_plan_launch(olv_plan *p, olv_stats *st) {
  if (!p || !p->exec) return fail(OLV_EINVAL, "olv_plan_launch: bad plan");
  Timer t;
  t.start();
  OLV_HIP(hipGraphLaunch(p->exec, stream()));
  t.stop();
  if (st) {
    const double ms = t.ms();
    uint32_t res[16];
    OLV_HIP(hipMemcpyAsync(res, p->result, sizeof(res), hipMemcpyDeviceToHost, stream()));
    OLV_HIP(hipStreamSynchronize(stream()));
    if (res[2]) return fail(OLV_EINTERNAL, "olv_plan_launch: grid barrier timed out");
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = res[4];
    st->bottom_up_levels = res[5];
  }
  return OLV_OK;
}

int olv_plan_destroy(olv_plan *p) {
  if (p) {
    (void)hipStreamSynchronize(stream());
    delete p;
  }
  return OLV_OK;
}

int olv_set_direction(olv_graph *g, float alpha, float beta) {
  if (!g || !(alpha > 0) || !(beta > 0)) return fail(OLV_EINVAL, "olv_set_direction: bad arguments");
  g->tune.alpha = alpha;
  g->tune.beta = beta;
  return OLV_OK;
}

int olv_get_tuning(const olv_graph *g, uint32_t *wave_thr, uint32_t *block_thr, float *alpha, float *beta) {
  if (!g) return fail(OLV_EINVAL, "olv_get_tuning: NULL graph");
  if (wave_thr) *wave_thr = g->tune.td_wave_thr;
  if (block_thr) *block_thr = g->tune.td_block_thr;
  if (alpha) *alpha = g->tune.alpha;
  if (beta) *beta = g->tune.beta;
  return OLV_OK;
}

// Time the plain BFS over a grid of (wave threshold, workgroup threshold,
// alpha) on sources picked from the graph (its largest out-degree vertex and
// two more with out-edges) and keep the fastest setting.
int olv_autotune(olv_graph *g, uint32_t *thread_max, uint32_t *wave_max, uint32_t *block_threads) {
  if (!g) return fail(OLV_EINVAL, "olv_autotune: NULL graph");
  std::vector<uint32_t> rp((size_t)g->n + 1);
  OLV_HIP(hipMemcpy(rp.data(), g->row_ptr, rp.size() * 4, hipMemcpyDeviceToHost));
  std::vector<uint32_t> srcs;
  uint32_t best_v = 0;
  for (uint32_t v = 0; v < g->n; ++v)
    if (rp[v + 1] - rp[v] > rp[best_v + 1] - rp[best_v]) best_v = v;
  srcs.push_back(best_v);
  for (uint32_t step = 1; step <= 2; ++step) {
    uint32_t v = (uint32_t)(((uint64_t)g->n * step) / 3);
    for (uint32_t t = 0; t < g->n && rp[v + 1] == rp[v]; ++t) v = (v + 1) % g->n;
    srcs.push_back(v);
  }
  DevBuf lv;
  OLV_TRY(lv.alloc((size_t)g->n * 4));
  const uint32_t waves[] = {16, 32, 64};
  const uint32_t blocks[] = {256, 1024, 4096};
  const float alphas[] = {4.f, 14.f, 40.f};
  Tuning best = g->tune;
  double best_ms = 1e300;
  for (uint32_t wt : waves)
    for (uint32_t bt : blocks)
      for (float al : alphas) {
        g->tune.td_wave_thr = wt;
        g->tune.td_block_thr = bt;
        g->tune.alpha = al;
        double tot = 0;
        for (uint32_t s : srcs) {
          olv_stats st;
          // warm once, then time
          OLV_TRY(bfs_multisource(g, &s, 1, true, lv.as<int32_t>(), 0, nullptr));
          OLV_TRY(bfs_multisource(g, &s, 1, true, lv.as<int32_t>(), 0, &st));
          tot += st.elapsed_ms;
        }
        if (tot < best_ms) {
          best_ms = tot;
          best = g->tune;
        }
      }
  best.tuned = true;
  g->tune = best;
  if (thread_max) *thread_max = g->tune.td_wave_thr - 1;
  if (wave_max) *wave_max = g->tune.td_block_thr - 1;
  if (block_threads) *block_threads = g->tune.block_threads;
  return OLV_OK;
}

}  // extern "C"
