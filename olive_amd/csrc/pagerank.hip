// PageRank in fp32 (pull over the in-edge CSR), the weighted/teleported
// fixed point, and its adjoint / vector-Jacobian product.
//
// The reference runs PageRank as edgeMap (thread per source vertex, fp64
// atomicAdd per edge, oliverKernel.h:148-178 + PageRank.cu:51-61), vertexMap
// (PageRank.cu:63-75) and a host-side vertexReduce (full D2H copy,
// oliver.h:176-183) per iteration.  Here one launch per iteration computes
// every vertex's in-sum with a fixed summation order, applies the update,
// writes the next contribution vector and adds its workgroup's residual to
// one of 32 counter slots; the next iteration totals those slots at its start
// and, once the residual is below epsilon, raises a device flag that turns
// the remaining pre-enqueued iterations into no-ops.  Index and weight
// arrays are read with plain loads (non-temporal loads measured slower on
// MI355X).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "blocked.hpp"
#include "internal.hpp"
#include "seg_core.hpp"

namespace olv {

// The iterations run over renumbered copies of the CSRs: vertices numbered
// by descending in-degree (ties by id), each row's entries mapped to the new
// numbers in their original slot order (the in-CSR for the pulls, the
// out-CSR for the adjoint).  On a skewed graph the vertices most gathered
// from (large out-degree, which follows in-degree there) then sit together
// in a few cache lines of the gathered vector, and the rows of one degree
// bucket are consecutive.  Every vertex keeps its bucket and its summation
// order, so every iterate is bit-identical to the iteration in the graph's
// own numbering (OLV_GRAPH_ORDER).  The numbering itself, g->rl_pi /
// g->rl_order, belongs to the graph.
struct RelabelSide {
  uint32_t *rp = nullptr;   // n + 1: row offsets, rows in new-number order
  uint32_t *col = nullptr;  // m: new numbers of the row's entries
  uint32_t *eid = nullptr;  // m: edge id (index into w) of each slot
  SegPlan plan;
  bool ready = false;
  void release() {
    void *ps[] = {rp, col, eid};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    rp = col = eid = nullptr;
    plan.release();
    ready = false;
  }
};

struct PrRelabel {
  RelabelSide in, out;
  BlockedPlan blk_in;    // source-segmented pull over the renumbered in-CSR
  // the same for the weighted solves: [0] over the renumbered in-CSR
  // (forward), [1] over the out-CSR (adjoint); their lslot holds each
  // entry's edge id (the slot mapped through the side's eid when built)
  BlockedPlan blk_w[2];
  float *inv = nullptr;  // n: 1/outdeg by new number
  void release() {
    in.release();
    out.release();
    blk_in.release();
    blk_w[0].release();
    blk_w[1].release();
    if (inv) (void)hipFree(inv);
    inv = nullptr;
  }
};

struct PrWork {
  uint32_t n = 0, m = 0;
  PrRelabel rl;
  float *buf[4] = {nullptr, nullptr, nullptr, nullptr};  // n floats each
  // Edge weights in a solver's slot order (w[eid[k]] at slot k): [0] for the
  // pulls over the in-CSR, [1] for the adjoint over the out-CSR; each with
  // the weights and slot map it was made from, reused by calls that pass
  // OLV_W_UNCHANGED with the same weights.
  float *wslot[2] = {nullptr, nullptr};  // m floats each
  const float *wslot_src[2] = {nullptr, nullptr};
  const uint32_t *wslot_eid[2] = {nullptr, nullptr};
  // The same in the entry order of the segmented pulls (rl.blk_w[side]):
  // w[eid[lslot[e]]] at entry e, 0 in the alignment gaps.
  float *wpiece[2] = {nullptr, nullptr};  // rl.blk_w[side].nent floats each
  const float *wpiece_src[2] = {nullptr, nullptr};
  const uint32_t *wpiece_eid[2] = {nullptr, nullptr};
  IterCtl *ctl = nullptr;
  IterCtl *h_ctl = nullptr;
  double *part = nullptr;  // per-workgroup partial sums (kPartials)
};

constexpr uint32_t kPartials = 2048;

void free_pr_work(olv_graph *g) {
  PrWork *w = g->pr;
  if (!w) return;
  for (float *p : w->buf)
    if (p) (void)hipFree(p);
  for (float *p : w->wslot)
    if (p) (void)hipFree(p);
  for (float *p : w->wpiece)
    if (p) (void)hipFree(p);
  if (w->ctl) (void)hipFree(w->ctl);
  if (w->h_ctl) (void)hipHostFree(w->h_ctl);
  if (w->part) (void)hipFree(w->part);
  w->rl.release();
  delete w;
  g->pr = nullptr;
}

static int alloc_pr_work(PrWork *w, const olv_graph *g) {
  w->n = g->n;
  w->m = g->m;
  for (auto &b : w->buf) OLV_HIP(hipMalloc(reinterpret_cast<void **>(&b), (size_t)g->n * 4));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->ctl), sizeof(IterCtl)));
  OLV_HIP(hipHostMalloc(reinterpret_cast<void **>(&w->h_ctl), sizeof(IterCtl), hipHostMallocDefault));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->part), kPartials * sizeof(double)));
  return OLV_OK;
}

// The graph's PageRank buffers, allocated on first use and published to the
// graph only once every allocation has succeeded.
static int get_pr_work(olv_graph *g, PrWork **out) {
  if (!g->pr) {
    PrWork *w = new PrWork();
    g->pr = w;
    const int rc = alloc_pr_work(w, g);
    if (rc != OLV_OK) {
      free_pr_work(g);
      return rc;
    }
  }
  *out = g->pr;
  return OLV_OK;
}

static int ensure_plan(olv_graph *g, bool in_edges) {
  if (in_edges) {
    if (!g->in_row_ptr) OLV_TRY(olv_graph_build_transpose(g));
    if (!g->plan_in_ready || seg_plan_stale(g->plan_in, g->tune)) {
      OLV_TRY(build_seg_plan(g, true, g->plan_in));
      g->plan_in_ready = true;
    }
  } else if (!g->plan_out_ready || seg_plan_stale(g->plan_out, g->tune)) {
    OLV_TRY(build_seg_plan(g, false, g->plan_out));
    g->plan_out_ready = true;
  }
  return OLV_OK;
}

// Row i of a renumbered CSR is row order[i] of the source CSR with every
// entry u replaced by pi[u]; eid (if not null) receives each slot's edge id,
// src_eid[slot] or, with src_eid null, the source slot itself.  One wave per
// row.
__global__ void relabel_rows_kernel(const uint32_t *__restrict__ order, const uint32_t *__restrict__ pi,
                                    const uint32_t *__restrict__ src_rp, const uint32_t *__restrict__ src_col,
                                    const uint32_t *__restrict__ src_eid, const uint32_t *__restrict__ rp, uint32_t n,
                                    uint32_t *__restrict__ col, uint32_t *__restrict__ eid) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t lane = threadIdx.x % kWave;
  for (uint32_t i = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; i < n; i += waves) {
    const uint32_t v = order[i];
    const uint32_t b = src_rp[v], len = src_rp[v + 1] - b, o = rp[i];
    for (uint32_t j0 = lane; j0 < len; j0 += 4 * kWave) {  // four loads in flight per lane
      uint32_t c[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) c[q] = j0 + q * kWave < len ? src_col[b + j0 + q * kWave] : 0u;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (j0 + q * kWave < len) {
          const uint32_t k = b + j0 + q * kWave;
          col[o + j0 + q * kWave] = pi[c[q]];
          if (eid) eid[o + j0 + q * kWave] = src_eid ? src_eid[k] : k;
        }
    }
  }
}

// out[i] = in[idx[i]]
template <class T>
__global__ void gather_kernel(const T *__restrict__ in, const uint32_t *__restrict__ idx, uint32_t n,
                              T *__restrict__ out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = in[idx[i]];
}

static unsigned ew_grid(size_t n);

// Renumbered copy of the in-CSR (in_edges) or the out-CSR, with the slot
// edge ids when with_eid (weighted solves).
static int ensure_side(olv_graph *g, PrWork *w, bool in_edges, bool with_eid) {
  RelabelSide &sd = in_edges ? w->rl.in : w->rl.out;
  if (sd.ready && (sd.eid || !with_eid)) {
    if (seg_plan_stale(sd.plan, g->tune)) {  // the bucket bounds changed: new plan, same arrays
      std::vector<uint32_t> rp((size_t)g->n + 1);
      OLV_HIP(hipMemcpy(rp.data(), sd.rp, rp.size() * 4, hipMemcpyDeviceToHost));
      OLV_TRY(build_seg_plan_rp(rp, g->tune.seg_bounds, sd.plan));
    }
    return OLV_OK;
  }
  sd.release();
  const uint32_t n = g->n, m = g->m;
  if (in_edges && !g->in_row_ptr) OLV_TRY(olv_graph_build_transpose(g));
  std::vector<uint32_t> order;
  OLV_TRY(indeg_order(g, &order));
  const uint32_t *d_src_rp = in_edges ? g->in_row_ptr : g->row_ptr;
  std::vector<uint32_t> src_rp((size_t)n + 1), rp((size_t)n + 1);
  OLV_HIP(hipMemcpy(src_rp.data(), d_src_rp, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
  rp[0] = 0;
  for (uint32_t i = 0; i < n; ++i) rp[i + 1] = rp[i] + (src_rp[order[i] + 1] - src_rp[order[i]]);
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&sd.rp), ((size_t)n + 1) * 4));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&sd.col), std::max<size_t>(m, 1) * 4));
  if (with_eid) OLV_HIP(hipMalloc(reinterpret_cast<void **>(&sd.eid), std::max<size_t>(m, 1) * 4));
  OLV_HIP(hipMemcpy(sd.rp, rp.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice));
  if (!w->rl.inv) {
    OLV_HIP(hipMalloc(reinterpret_cast<void **>(&w->rl.inv), std::max<size_t>(n, 1) * 4));
    if (n)
      hipLaunchKernelGGL(gather_kernel<float>, dim3(ew_grid(n)), dim3(256), 0, stream(),
                         (const float *)g->inv_outdeg, g->rl_order, n, w->rl.inv);
  }
  if (n)
    hipLaunchKernelGGL(relabel_rows_kernel, dim3(ew_grid((size_t)n * kWave)), dim3(256), 0, stream(), g->rl_order,
                       g->rl_pi, d_src_rp, in_edges ? g->in_col : g->col, in_edges ? g->in_eid : nullptr, sd.rp, n,
                       sd.col, sd.eid);
  OLV_HIP(hipGetLastError());
  OLV_TRY(build_seg_plan_rp(rp, g->tune.seg_bounds, sd.plan));
  OLV_HIP(hipStreamSynchronize(stream()));
  sd.ready = true;
  return OLV_OK;
}

// The source-segmented pull plan over the renumbered in-CSR (built once).
static int ensure_blocked(olv_graph *g, PrWork *w) {
  BlockedPlan &bp = w->rl.blk_in;
  if (bp.ready) return OLV_OK;
  OLV_TRY(ensure_side(g, w, true, false));
  const uint32_t n = g->n, m = g->m;
  std::vector<uint32_t> rp((size_t)n + 1), col(std::max<uint32_t>(m, 1));
  OLV_HIP(hipMemcpy(rp.data(), w->rl.in.rp, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
  if (m) OLV_HIP(hipMemcpy(col.data(), w->rl.in.col, (size_t)m * 4, hipMemcpyDeviceToHost));
  return build_blocked_plan(rp, col.data(), false, (uint32_t)cu_count(), kUnitEdges, bp);
}

// lslot[e] <- eid[lslot[e]] where lslot[e] is a slot
__global__ void slot_to_edge_kernel(uint32_t *__restrict__ lslot, const uint32_t *__restrict__ eid, uint32_t nent) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < nent; e += gridDim.x * blockDim.x) {
    const uint32_t s = lslot[e];
    if (s != kNoSlot) lslot[e] = eid[s];
  }
}

// The segmented pull plan of a weighted solve over the renumbered in-CSR
// (side 0) or out-CSR (side 1), with every entry's edge id and units small
// enough to stage their weights (built once per side).
static int ensure_wblocked(olv_graph *g, PrWork *w, int side) {
  BlockedPlan &bp = w->rl.blk_w[side];
  if (bp.ready) return OLV_OK;
  OLV_TRY(ensure_side(g, w, side == 0, true));
  const RelabelSide &sd = side == 0 ? w->rl.in : w->rl.out;
  const uint32_t n = g->n, m = g->m;
  std::vector<uint32_t> rp((size_t)n + 1), col(std::max<uint32_t>(m, 1));
  OLV_HIP(hipMemcpy(rp.data(), sd.rp, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
  if (m) OLV_HIP(hipMemcpy(col.data(), sd.col, (size_t)m * 4, hipMemcpyDeviceToHost));
  OLV_TRY(build_blocked_plan(rp, col.data(), true, (uint32_t)cu_count(), kUnitEdgesW, bp));
  hipLaunchKernelGGL(slot_to_edge_kernel, dim3(ew_grid(bp.nent)), dim3(256), 0, stream(), bp.lslot, sd.eid, bp.nent);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

// ----------------------------------------------------------------- ops
// testPageRank.cu:79-91: sum_{u->v} r_u/outdeg(u), r' = d*sum + (1-d)/n,
// err += |r' - r|.  contrib[u] = r_u * (1/outdeg(u)).
struct PrOp {
  const uint32_t *ci;
  const float *contrib_cur;
  float *contrib_next;
  float *r;
  const float *inv_outdeg;
  float d, base;
  __device__ __forceinline__ uint32_t idx(uint32_t k) const { return ld_stream(&ci[k]); }
  __device__ __forceinline__ float val(uint32_t, uint32_t, uint32_t u) const { return ld_gather(&contrib_cur[u]); }
  __device__ __forceinline__ void finish(uint32_t v, float s, float &err) const {
    const float rn = d * s + base;
    err += fabsf(rn - r[v]);
    r[v] = rn;
    contrib_next[v] = rn * inv_outdeg[v];
  }
};

// Summand of the segmented pull for plain PageRank: the staged contribution.
struct SliceVal {
  const float *sl = nullptr;
  __device__ __forceinline__ void bind(const float *s, const float *) { sl = s; }
  __device__ __forceinline__ float operator()(uint32_t, uint32_t i) const { return sl[i]; }
};

// Summand of a weighted segmented pull: the entry's staged weight times the
// staged source value (the iterate for the forward, lambda for the adjoint).
struct WSliceVal {
  const float *sl = nullptr, *w = nullptr;
  __device__ __forceinline__ void bind(const float *s, const float *ws) {
    sl = s;
    w = ws;
  }
  __device__ __forceinline__ float operator()(uint32_t k, uint32_t i) const { return w[k] * sl[i]; }
};

// The update of WprOp::finish for the combine of the segmented pull.
struct WprFin {
  const float *r_cur;
  float *r_next;
  const float *t;
  float d, omd;
  struct Pre {
    float r, t;
  };
  __device__ __forceinline__ Pre pre(uint32_t v) const { return Pre{r_cur[v], t[v]}; }
  __device__ __forceinline__ void finish(uint32_t v, float s, float &err, const Pre &p) const {
    const float rn = d * s + omd * p.t;
    err += fabsf(rn - p.r);
    r_next[v] = rn;
  }
};

// The update of AdjOp::finish for the combine of the segmented pull.
struct AdjFin {
  const float *l_cur;
  float *l_next;
  const float *gbar;
  float d;
  struct Pre {
    float l, g;
  };
  __device__ __forceinline__ Pre pre(uint32_t u) const { return Pre{l_cur[u], gbar[u]}; }
  __device__ __forceinline__ void finish(uint32_t u, float s, float &err, const Pre &p) const {
    const float ln = p.g + d * s;
    err += fabsf(ln - p.l);
    l_next[u] = ln;
  }
};

// The update of PrOp::finish for the combine of the segmented pull.
struct PrFin {
  float *contrib_next;
  float *r;
  const float *inv_outdeg;
  float d, base;
  struct Pre {
    float r, inv;
  };
  // the row's own inputs, loaded before its partials arrive
  __device__ __forceinline__ Pre pre(uint32_t v) const { return Pre{r[v], inv_outdeg[v]}; }
  __device__ __forceinline__ void finish(uint32_t v, float s, float &err, const Pre &p) const {
    const float rn = d * s + base;
    err += fabsf(rn - p.r);
    r[v] = rn;
    contrib_next[v] = rn * p.inv;
  }
};

// r' = d * sum_{u->v} w_uv r_u + (1-d) t_v over the in-CSR with weights
// permuted into in-slot order.
struct WprOp {
  const uint32_t *ci;
  const float *w;
  const float *r_cur;
  float *r_next;
  const float *t;
  float d, omd;
  __device__ __forceinline__ uint32_t idx(uint32_t k) const { return ld_stream(&ci[k]); }
  __device__ __forceinline__ float val(uint32_t, uint32_t k, uint32_t u) const { return ld_stream(&w[k]) * ld_gather(&r_cur[u]); }
  __device__ __forceinline__ void finish(uint32_t v, float s, float &err) const {
    const float rn = d * s + omd * t[v];
    err += fabsf(rn - r_cur[v]);
    r_next[v] = rn;
  }
};

// lambda'_u = gbar_u + d * sum_{u->v} w_uv lambda_v over the out-CSR.
struct AdjOp {
  const uint32_t *ci;
  const float *w;
  const float *l_cur;
  float *l_next;
  const float *gbar;
  float d;
  __device__ __forceinline__ uint32_t idx(uint32_t k) const { return ld_stream(&ci[k]); }
  __device__ __forceinline__ float val(uint32_t, uint32_t k, uint32_t v) const { return ld_stream(&w[k]) * ld_gather(&l_cur[v]); }
  __device__ __forceinline__ void finish(uint32_t u, float s, float &err) const {
    const float ln = gbar[u] + d * s;
    err += fabsf(ln - l_cur[u]);
    l_next[u] = ln;
  }
};

// grad_w[e=(u,v)] = d * r_u * lambda_v, written while walking the out-CSR
// (Mapped: a renumbered out-CSR, whose slot k holds edge eid[k]).
template <bool Mapped>
struct GradOp {
  const uint32_t *ci;
  const float *r;
  const float *lam;
  float *grad_w;
  const uint32_t *eid;
  float d;
  __device__ __forceinline__ uint32_t idx(uint32_t k) const { return ld_stream(&ci[k]); }
  __device__ __forceinline__ float val(uint32_t u, uint32_t k, uint32_t v) const {
    __builtin_nontemporal_store(d * r[u] * lam[v], &grad_w[Mapped ? ld_stream(&eid[k]) : k]);
    return 0.f;
  }
  __device__ __forceinline__ void finish(uint32_t, float, float &) const {}
};

// --------------------------------------------------------------- kernels
// Zero the control block and set eps (all threads of the calling workgroup).
__device__ __forceinline__ void ctl_reset(IterCtl *ctl, double eps) {
  uint32_t *w = reinterpret_cast<uint32_t *>(ctl);
  for (uint32_t i = threadIdx.x; i < sizeof(IterCtl) / 4; i += blockDim.x) w[i] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) ctl->eps = eps;
}

__global__ void pr_init_kernel(uint32_t n, float r0, const float *__restrict__ inv_outdeg, float *__restrict__ r,
                               float *__restrict__ contrib, IterCtl *__restrict__ ctl, double eps) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  for (uint32_t v = tid; v < n; v += gridDim.x * blockDim.x) {
    r[v] = r0;
    contrib[v] = r0 * inv_outdeg[v];
  }
  if (blockIdx.x == 0) ctl_reset(ctl, eps);
}

__global__ void ctl_init_kernel(IterCtl *ctl, double eps) { ctl_reset(ctl, eps); }

// part[b] = sum |g| over workgroup b's grid-stride share
__global__ void __launch_bounds__(256) l1_partial_kernel(const float *__restrict__ g, uint32_t n,
                                                         double *__restrict__ part) {
  __shared__ double s_red[256 / kWave];
  double s = 0.0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += fabs((double)g[i]);
  s = block_sum<256, double>(s, s_red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// eps = tol * sum |gbar| from k partials, added in a fixed order (one workgroup)
__global__ void __launch_bounds__(256) ctl_init_l1_kernel(IterCtl *ctl, const double *__restrict__ part, uint32_t k,
                                                          double tol) {
  __shared__ double s_red[256 / kWave];
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < k; i += 256) s += part[i];
  s = block_sum<256, double>(s, s_red);
  ctl_reset(ctl, 0.0);
  if (threadIdx.x == 0) ctl->eps = tol * s;
}

__global__ void permute_kernel(const float *__restrict__ w, const uint32_t *__restrict__ eid, uint32_t m,
                               float *__restrict__ out) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) out[k] = w[eid[k]];
}

// The weights w in slot order by eid for side 0 (in-CSR pulls) or 1 (the
// adjoint over the out-CSR), in the side's cached buffer.  With `reuse` the
// permutation is skipped when that buffer was last made from the same
// weights address and slot map (the caller vouches that the weights did not
// change since the call that made it); host-staged weights are never reused.
static int slot_weights(const olv_graph *g, PrWork *pw, int side, const float *w, const uint32_t *eid, bool reuse,
                        bool staged, const float **out) {
  if (!pw->wslot[side]) OLV_HIP(hipMalloc(reinterpret_cast<void **>(&pw->wslot[side]), std::max<size_t>(g->m, 1) * 4));
  if (!(reuse && !staged && pw->wslot_src[side] == w && pw->wslot_eid[side] == eid)) {
    if (g->m)
      hipLaunchKernelGGL(permute_kernel, dim3(ew_grid(g->m)), dim3(256), 0, stream(), w, eid, g->m, pw->wslot[side]);
    OLV_HIP(hipGetLastError());
    pw->wslot_src[side] = staged ? nullptr : w;
    pw->wslot_eid[side] = eid;
    // a call that did not vouch for these weights invalidates the other
    // side's copy of them too
    if (!reuse && pw->wslot_src[side ^ 1] == w) pw->wslot_src[side ^ 1] = nullptr;
  }
  *out = pw->wslot[side];
  return OLV_OK;
}

// out[e] = w[edge[e]], 0 where edge[e] is kNoSlot
__global__ void piece_weight_kernel(const float *__restrict__ w, const uint32_t *__restrict__ edge, uint32_t nent,
                                    float *__restrict__ out) {
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < nent; e += gridDim.x * blockDim.x) {
    const uint32_t x = edge[e];
    out[e] = x != kNoSlot ? w[x] : 0.f;
  }
}

// The weights w in the entry order of the weighted segmented pull of side 0
// (in-CSR) or 1 (out-CSR), cached like slot_weights' copies.
static int piece_weights(PrWork *pw, int side, const float *w, const uint32_t *eid, bool reuse, bool staged,
                         const float **out) {
  const BlockedPlan &bp = pw->rl.blk_w[side];
  if (!pw->wpiece[side]) OLV_HIP(hipMalloc(reinterpret_cast<void **>(&pw->wpiece[side]), (size_t)bp.nent * 4));
  if (!(reuse && !staged && pw->wpiece_src[side] == w && pw->wpiece_eid[side] == eid)) {
    hipLaunchKernelGGL(piece_weight_kernel, dim3(ew_grid(bp.nent)), dim3(256), 0, stream(), w, bp.lslot, bp.nent,
                       pw->wpiece[side]);
    OLV_HIP(hipGetLastError());
    pw->wpiece_src[side] = staged ? nullptr : w;
    pw->wpiece_eid[side] = eid;
    if (!reuse && pw->wpiece_src[side ^ 1] == w) pw->wpiece_src[side ^ 1] = nullptr;
  }
  *out = pw->wpiece[side];
  return OLV_OK;
}

// y[i] = a * x[idx[i]]
__global__ void gather_scale_kernel(const float *__restrict__ x, const uint32_t *__restrict__ idx, float a, uint32_t n,
                                    float *__restrict__ y) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = a * x[idx[i]];
}

__global__ void scale_kernel(const float *__restrict__ x, float a, uint32_t n, float *__restrict__ y) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = a * x[i];
}

static unsigned ew_grid(size_t n) {
  size_t b = (n + 255) / 256;
  return (unsigned)std::max<size_t>(1, std::min(b, (size_t)cu_count() * 8));
}

// hipEvent pairs bracketing each launch of the dominant kernel (OLV_PROFILE).
struct KernelProfile {
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  bool on = false;
  ~KernelProfile() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
  int reserve(size_t launches) {
    on = true;
    ev.resize(2 * launches, nullptr);
    for (auto &e : ev) OLV_HIP(hipEventCreate(&e));
    return OLV_OK;
  }
  hipEvent_t next() { return on && used < ev.size() ? ev[used++] : nullptr; }
  // after the stream has drained
  int total(double *ms, uint32_t *launches) {
    double s = 0;
    for (size_t i = 0; i + 1 < used; i += 2) {
      float t = 0.f;
      OLV_HIP(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
      s += t;
    }
    *ms = s;
    *launches = (uint32_t)(used / 2);
    return OLV_OK;
  }
};

template <class Op>
static int enqueue_iteration(const SegDev &d, const uint32_t *rp, const Op &op, PrWork *w, uint32_t k,
                             KernelProfile *prof = nullptr) {
  hipEvent_t a = prof ? prof->next() : nullptr, b = prof ? prof->next() : nullptr;
  if (a) OLV_HIP(hipEventRecord(a, stream()));
  launch_seg(d, rp, op, w->ctl, k);
  if (b) OLV_HIP(hipEventRecord(b, stream()));
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

// One iteration of a segmented pull: the segment pass over `src` with
// summand SV, then the combine applying `fin`.
template <class SV, bool kW, class Fin>
static int enqueue_blocked(const BlockedPlan &bp, const BlkDev &bd, const float *src, const Fin &fin, PrWork *w,
                           uint32_t k, KernelProfile *prof = nullptr) {
  // combine: a wave per listed big row, kCmbMidLanes lanes per listed mid
  // row, then a lane per row
  const unsigned cu = (unsigned)cu_count();
  const unsigned gb = std::min<unsigned>((bp.n_big + 3) / 4, cu * 4);
  const unsigned gm = std::min<unsigned>((bp.n_mid + 256 / kCmbMidLanes - 1) / (256 / kCmbMidLanes), cu * 16);
  const unsigned gc = gb + gm + ew_grid(bp.n);
  hipEvent_t a = prof ? prof->next() : nullptr, b = prof ? prof->next() : nullptr;
  if (a) OLV_HIP(hipEventRecord(a, stream()));
  hipLaunchKernelGGL((blk_segment_kernel<SV, kW>), dim3(bp.ngroups), dim3(kBlkThreads), 0, stream(), bd, SV{}, src,
                     w->ctl, k);
  hipLaunchKernelGGL(blk_combine_kernel<Fin>, dim3(gc), dim3(256), 0, stream(), bd, fin, w->ctl, k, gb, gm);
  if (b) OLV_HIP(hipEventRecord(b, stream()));
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

static int read_ctl(PrWork *w) {
  OLV_HIP(hipMemcpyAsync(w->h_ctl, w->ctl, sizeof(IterCtl), hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  return OLV_OK;
}

// Outcome after `enq` enqueued iterations (control block read back):
// updates performed, the last residual, and whether it met eps.
struct IterResult {
  uint32_t iters;
  double residual;
  bool converged;
};

static IterResult iter_result(const IterCtl &h, uint32_t enq) {
  IterResult r;
  if (h.done) {
    r.iters = h.iters;
    r.residual = h.residual;
    r.converged = true;
    return r;
  }
  double s = 0.0;
  if (enq)
    for (int i = 0; i < kErrSlots; ++i) s += h.errs[(enq - 1) % 3][i].v;
  r.iters = enq;
  r.residual = s;
  r.converged = enq > 0 && s < h.eps;
  return r;
}

// Host-pointer staging for OLV_PTR_HOST.
struct Stage {
  DevBuf buf;
  const float *in(const float *p, size_t n, bool host, int *st) {
    if (!host || !p) return p;
    if ((*st = buf.alloc(n * 4)) != OLV_OK) return nullptr;
    if (hipMemcpy(buf.p, p, n * 4, hipMemcpyHostToDevice) != hipSuccess) *st = fail(OLV_EHIP, "staging copy failed");
    return buf.as<float>();
  }
  float *out(float *p, size_t n, bool host, int *st) {
    if (!host || !p) return p;
    if ((*st = buf.alloc(n * 4)) != OLV_OK) return nullptr;
    return buf.as<float>();
  }
};

// Fixed point r = d * P^T r + (1-d) t over the in-slot weights ws (or,
// with bp, by the segmented pull of plan bp over the weights wp in its entry
// order), iterated from r0 = t until the L1 change drops below tol; the
// result is left in w->buf[res->iters & 1].  t must not be w->buf[0] or
// w->buf[1].
static int weighted_solve(const olv_graph *g, const SegDev &d, const uint32_t *rp, const uint32_t *ci, PrWork *w,
                          const float *ws, const float *t, float damping, float tol, uint32_t max_iters, Timer &timer,
                          IterResult *res, const BlockedPlan *bp = nullptr, const float *wp = nullptr) {
  OLV_HIP(hipMemcpyAsync(w->buf[0], t, (size_t)g->n * 4, hipMemcpyDeviceToDevice, stream()));
  hipLaunchKernelGGL(ctl_init_kernel, dim3(1), dim3(256), 0, stream(), w->ctl, (double)tol);
  const uint32_t group = 8;
  uint32_t it = 0;
  *res = IterResult{0, 0.0, false};
  for (; it < max_iters;) {
    const uint32_t stop = std::min(max_iters, it + group);
    for (; it < stop; ++it) {
      if (bp) {
        const WprFin fin{w->buf[it & 1], w->buf[(it + 1) & 1], t, damping, 1.0f - damping};
        OLV_TRY((enqueue_blocked<WSliceVal, true>(*bp, make_blk_dev(*bp, wp), w->buf[it & 1], fin, w, it)));
      } else {
        WprOp op{ci, ws, w->buf[it & 1], w->buf[(it + 1) & 1], t, damping, 1.0f - damping};
        OLV_TRY(enqueue_iteration(d, rp, op, w, it));
      }
    }
    timer.stop();
    OLV_TRY(read_ctl(w));
    *res = iter_result(*w->h_ctl, it);
    if (res->converged) break;
  }
  return OLV_OK;
}

// The in-CSR arrays a weighted pull runs over: the renumbered copy, or the
// graph's own with OLV_GRAPH_ORDER.  eid: each slot's edge id.
struct PullView {
  SegDev d;
  const uint32_t *rp, *ci, *eid;
  bool relabel;
};

static int pull_view(olv_graph *g, PrWork *w, unsigned flags, PullView *v) {
  v->relabel = (flags & OLV_GRAPH_ORDER) == 0;
  if (v->relabel) {
    OLV_TRY(ensure_side(g, w, true, true));
    v->d = make_seg_dev(w->rl.in.plan, g->tune);
    v->rp = w->rl.in.rp;
    v->ci = w->rl.in.col;
    v->eid = w->rl.in.eid;
  } else {
    OLV_TRY(ensure_plan(g, true));
    v->d = make_seg_dev(g->plan_in, g->tune);
    v->rp = g->in_row_ptr;
    v->ci = g->in_col;
    v->eid = g->in_eid;
  }
  return OLV_OK;
}

// out[i] = in[idx[i]] for i < n on the library stream.
static void gather_f32(const float *in, const uint32_t *idx, uint32_t n, float *out) {
  if (n) hipLaunchKernelGGL(gather_kernel<float>, dim3(ew_grid(n)), dim3(256), 0, stream(), in, idx, n, out);
}

}  // namespace olv

using namespace olv;

extern "C" {

int olv_pagerank_f32(const olv_graph *cg, float damping, uint32_t max_iters, double epsilon, float *ranks,
                     unsigned flags, olv_stats *st) {
  if (!cg || !ranks) return fail(OLV_EINVAL, "olv_pagerank_f32: NULL argument");
  olv_graph *g = const_cast<olv_graph *>(cg);
  const bool relabel = (flags & OLV_GRAPH_ORDER) == 0;
  if (!g->in_row_ptr) OLV_TRY(olv_graph_build_transpose(g));
  if (!relabel) OLV_TRY(ensure_plan(g, true));
  PrWork *w = nullptr;
  OLV_TRY(get_pr_work(g, &w));
  if (relabel) OLV_TRY(ensure_side(g, w, true, false));
  const bool blocked = relabel && !(flags & OLV_PR_ROW_PULL);
  if (blocked) OLV_TRY(ensure_blocked(g, w));
  const SegDev dr = make_seg_dev(relabel ? w->rl.in.plan : g->plan_in, g->tune);
  const uint32_t *rp = relabel ? w->rl.in.rp : g->in_row_ptr;
  const uint32_t *ci = relabel ? w->rl.in.col : g->in_col;
  const float *inv = relabel ? w->rl.inv : g->inv_outdeg;
  const bool host = (flags & OLV_PTR_HOST) != 0;
  int s = OLV_OK;
  Stage so;
  float *r = so.out(ranks, g->n, host, &s);
  OLV_TRY(s);
  float *ri = relabel ? w->buf[2] : r;  // iterate (new numbering when relabelled)
  KernelProfile prof;
  if (flags & OLV_PROFILE) OLV_TRY(prof.reserve(max_iters));
  Timer t;
  t.start();
  hipLaunchKernelGGL(pr_init_kernel, dim3(ew_grid(g->n)), dim3(256), 0, stream(), g->n, 1.0f / (float)g->n, inv, ri,
                     w->buf[0], w->ctl, epsilon);
  const float base = (float)((1.0 - (double)damping) / (double)g->n);
  if (blocked) {
    const BlockedPlan &bp = w->rl.blk_in;
    const BlkDev bd = make_blk_dev(bp);
    for (uint32_t it = 0; it < max_iters; ++it)
      OLV_TRY((enqueue_blocked<SliceVal, false>(bp, bd, w->buf[it & 1], PrFin{w->buf[(it + 1) & 1], ri, inv, damping, base},
                                                w, it, &prof)));
  } else {
    for (uint32_t it = 0; it < max_iters; ++it) {
      PrOp op{ci, w->buf[it & 1], w->buf[(it + 1) & 1], ri, inv, damping, base};
      OLV_TRY(enqueue_iteration(dr, rp, op, w, it, &prof));
    }
  }
  if (relabel && g->n)
    hipLaunchKernelGGL(gather_kernel<float>, dim3(ew_grid(g->n)), dim3(256), 0, stream(), (const float *)ri, g->rl_pi,
                       g->n, r);
  OLV_HIP(hipGetLastError());
  t.stop();
  const double ms = t.ms();
  OLV_TRY(read_ctl(w));
  const IterResult res = iter_result(*w->h_ctl, max_iters);
  double kms = 0;
  uint32_t kl = 0;
  if (prof.on) OLV_TRY(prof.total(&kms, &kl));
  if (host) OLV_HIP(hipMemcpy(ranks, r, (size_t)g->n * 4, hipMemcpyDeviceToHost));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = res.iters;
    st->residual = res.residual;
    st->edges_traversed = (uint64_t)g->m * res.iters;
    st->edges_inspected = st->edges_traversed;
    st->vertices_reached = g->n;
    st->kernel_ms = kms;
    st->kernel_launches = kl;
  }
  return OLV_OK;
}

int olv_pagerank_weighted_f32(const olv_graph *cg, const float *wts, const float *tele, float damping, float tol,
                              uint32_t max_iters, float *ranks, unsigned flags, olv_stats *st) {
  if (!cg || !wts || !tele || !ranks) return fail(OLV_EINVAL, "olv_pagerank_weighted_f32: NULL argument");
  olv_graph *g = const_cast<olv_graph *>(cg);
  PrWork *w = nullptr;
  OLV_TRY(get_pr_work(g, &w));
  PullView pv;
  OLV_TRY(pull_view(g, w, flags, &pv));
  const bool blocked = pv.relabel && !(flags & OLV_PR_ROW_PULL);
  if (blocked) OLV_TRY(ensure_wblocked(g, w, 0));
  const bool host = (flags & OLV_PTR_HOST) != 0;
  int s = OLV_OK;
  Stage sw, st_t, so;
  const float *dw = sw.in(wts, g->m, host, &s);
  OLV_TRY(s);
  const float *dt = st_t.in(tele, g->n, host, &s);
  OLV_TRY(s);
  float *r = so.out(ranks, g->n, host, &s);
  OLV_TRY(s);
  ScratchBuf tb;  // the teleport vector in the new numbering
  if (pv.relabel) OLV_TRY(tb.alloc(std::max<size_t>(g->n, 1) * 4));
  Timer t;
  t.start();
  const float *tt = dt;
  if (pv.relabel) {
    gather_f32(dt, g->rl_order, g->n, tb.as<float>());
    tt = tb.as<float>();
  }
  const float *ws = nullptr, *wp = nullptr;
  if (blocked) OLV_TRY(piece_weights(w, 0, dw, pv.eid, (flags & OLV_W_UNCHANGED) != 0, host, &wp));
  else OLV_TRY(slot_weights(g, w, 0, dw, pv.eid, (flags & OLV_W_UNCHANGED) != 0, host, &ws));
  IterResult res;
  OLV_TRY(weighted_solve(g, pv.d, pv.rp, pv.ci, w, ws, tt, damping, tol, max_iters, t, &res,
                         blocked ? &w->rl.blk_w[0] : nullptr, wp));
  const double ms = t.ms();
  const uint32_t iters = res.iters;
  if (pv.relabel) gather_f32(w->buf[iters & 1], g->rl_pi, g->n, r);
  else OLV_HIP(hipMemcpyAsync(r, w->buf[iters & 1], (size_t)g->n * 4, hipMemcpyDeviceToDevice, stream()));
  OLV_HIP(hipGetLastError());
  if (host) OLV_HIP(hipMemcpy(ranks, r, (size_t)g->n * 4, hipMemcpyDeviceToHost));
  else OLV_HIP(hipStreamSynchronize(stream()));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = iters;
    st->residual = res.residual;
    st->edges_traversed = (uint64_t)g->m * iters;
    st->edges_inspected = st->edges_traversed;
    st->vertices_reached = g->n;
  }
  if (!res.converged)
    return fail(OLV_ENOCONV, "olv_pagerank_weighted_f32: residual %.3e after %u iterations (tol %.3e)", res.residual,
                iters, (double)tol);
  return OLV_OK;
}

int olv_pagerank_jvp_f32(const olv_graph *cg, const float *wts, const float *ranks, const float *dwts,
                         const float *dtele, float damping, float tol, uint32_t max_iters, float *dranks, unsigned flags,
                         olv_stats *st) {
  if (!cg || !wts || !ranks || !dwts || !dtele || !dranks) return fail(OLV_EINVAL, "olv_pagerank_jvp_f32: NULL argument");
  if (!(damping >= 0.0f && damping < 1.0f)) return fail(OLV_EINVAL, "olv_pagerank_jvp_f32: damping must be in [0, 1)");
  olv_graph *g = const_cast<olv_graph *>(cg);
  PrWork *w = nullptr;
  OLV_TRY(get_pr_work(g, &w));
  PullView pv;
  OLV_TRY(pull_view(g, w, flags, &pv));
  const bool blocked = pv.relabel && !(flags & OLV_PR_ROW_PULL);
  if (blocked) OLV_TRY(ensure_wblocked(g, w, 0));
  const bool host = (flags & OLV_PTR_HOST) != 0;
  int s = OLV_OK;
  Stage sw, sr, sdw, sdt, so;
  const float *dw = sw.in(wts, g->m, host, &s);
  OLV_TRY(s);
  const float *dr_in = sr.in(ranks, g->n, host, &s);
  OLV_TRY(s);
  const float *ddw = sdw.in(dwts, g->m, host, &s);
  OLV_TRY(s);
  const float *ddt = sdt.in(dtele, g->n, host, &s);
  OLV_TRY(s);
  float *out = so.out(dranks, g->n, host, &s);
  OLV_TRY(s);
  ScratchBuf rb, tb, dwb;  // r and dt in the new numbering; dw in slot order
  if (pv.relabel) {
    OLV_TRY(rb.alloc(std::max<size_t>(g->n, 1) * 4));
    OLV_TRY(tb.alloc(std::max<size_t>(g->n, 1) * 4));
  }
  OLV_TRY(dwb.alloc(std::max<size_t>(g->m, 1) * 4));
  Timer t;
  t.start();
  if (pv.relabel) {
    gather_f32(dr_in, g->rl_order, g->n, rb.as<float>());
    gather_f32(ddt, g->rl_order, g->n, tb.as<float>());
    dr_in = rb.as<float>();
    ddt = tb.as<float>();
  }
  // source term s = d * dP^T r + (1-d) dt: one forward iteration with the
  // weights dw, applied to r, and the teleport dt
  if (g->m)
    hipLaunchKernelGGL(permute_kernel, dim3(ew_grid(g->m)), dim3(256), 0, stream(), ddw, pv.eid, g->m, dwb.as<float>());
  hipLaunchKernelGGL(ctl_init_kernel, dim3(1), dim3(256), 0, stream(), w->ctl, 0.0);
  WprOp src_op{pv.ci, dwb.as<float>(), dr_in, w->buf[2], ddt, damping, 1.0f - damping};
  OLV_TRY(enqueue_iteration(pv.d, pv.rp, src_op, w, 0));
  // dr = d * P^T dr + s, i.e. the forward fixed point with teleport s / (1-d)
  hipLaunchKernelGGL(scale_kernel, dim3(ew_grid(g->n)), dim3(256), 0, stream(), w->buf[2], 1.0f / (1.0f - damping),
                     g->n, w->buf[3]);
  const float *ws = nullptr, *wp = nullptr;
  if (blocked) OLV_TRY(piece_weights(w, 0, dw, pv.eid, (flags & OLV_W_UNCHANGED) != 0, host, &wp));
  else OLV_TRY(slot_weights(g, w, 0, dw, pv.eid, (flags & OLV_W_UNCHANGED) != 0, host, &ws));
  IterResult res;
  OLV_TRY(weighted_solve(g, pv.d, pv.rp, pv.ci, w, ws, w->buf[3], damping, tol, max_iters, t, &res,
                         blocked ? &w->rl.blk_w[0] : nullptr, wp));
  const double ms = t.ms();
  if (pv.relabel) gather_f32(w->buf[res.iters & 1], g->rl_pi, g->n, out);
  else OLV_HIP(hipMemcpyAsync(out, w->buf[res.iters & 1], (size_t)g->n * 4, hipMemcpyDeviceToDevice, stream()));
  OLV_HIP(hipGetLastError());
  if (host) OLV_HIP(hipMemcpy(dranks, out, (size_t)g->n * 4, hipMemcpyDeviceToHost));
  else OLV_HIP(hipStreamSynchronize(stream()));
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = res.iters;
    st->residual = res.residual;
    st->edges_traversed = (uint64_t)g->m * (res.iters + 1);
    st->edges_inspected = st->edges_traversed;
    st->vertices_reached = g->n;
  }
  if (!res.converged)
    return fail(OLV_ENOCONV, "olv_pagerank_jvp_f32: residual %.3e after %u iterations (tol %.3e)", res.residual,
                res.iters, (double)tol);
  return OLV_OK;
}

int olv_pagerank_vjp_f32(const olv_graph *cg, const float *wts, const float *ranks, const float *gbar, float damping,
                         float tol, uint32_t max_iters, float *grad_w, float *grad_t, float *lambda_out,
                         unsigned flags, olv_stats *st) {
  if (!cg || !wts || !ranks || !gbar) return fail(OLV_EINVAL, "olv_pagerank_vjp_f32: NULL argument");
  olv_graph *g = const_cast<olv_graph *>(cg);
  const bool relabel = (flags & OLV_GRAPH_ORDER) == 0;
  PrWork *w = nullptr;
  OLV_TRY(get_pr_work(g, &w));
  // the out-CSR the adjoint runs over: the renumbered copy (by the segmented
  // pull unless OLV_PR_ROW_PULL) or the graph's own
  const bool blocked = relabel && !(flags & OLV_PR_ROW_PULL);
  if (relabel) OLV_TRY(ensure_side(g, w, false, true));
  else OLV_TRY(ensure_plan(g, false));
  if (blocked) OLV_TRY(ensure_wblocked(g, w, 1));
  const SegDev d = make_seg_dev(relabel ? w->rl.out.plan : g->plan_out, g->tune);
  const uint32_t *rp = relabel ? w->rl.out.rp : g->row_ptr;
  const uint32_t *ci = relabel ? w->rl.out.col : g->col;
  const bool host = (flags & OLV_PTR_HOST) != 0;
  int s = OLV_OK;
  Stage sw, sr, sg, sgw, sgt, sl;
  const float *dw = sw.in(wts, g->m, host, &s);
  OLV_TRY(s);
  const float *dr = sr.in(ranks, g->n, host, &s);
  OLV_TRY(s);
  const float *dg = sg.in(gbar, g->n, host, &s);
  OLV_TRY(s);
  float *dgw = sgw.out(grad_w, g->m, host, &s);
  OLV_TRY(s);
  float *dgt = sgt.out(grad_t, g->n, host, &s);
  OLV_TRY(s);
  float *dl = sl.out(lambda_out, g->n, host, &s);
  OLV_TRY(s);
  ScratchBuf rb, gb;  // r and gbar in the new numbering
  if (relabel) {
    OLV_TRY(rb.alloc(std::max<size_t>(g->n, 1) * 4));
    OLV_TRY(gb.alloc(std::max<size_t>(g->n, 1) * 4));
  }
  Timer t;
  t.start();
  const float *wv = dw, *rv = dr, *gv = dg, *wp = nullptr;
  if (relabel) {
    if (blocked) OLV_TRY(piece_weights(w, 1, dw, w->rl.out.eid, (flags & OLV_W_UNCHANGED) != 0, host, &wp));
    else OLV_TRY(slot_weights(g, w, 1, dw, w->rl.out.eid, (flags & OLV_W_UNCHANGED) != 0, host, &wv));
    gather_f32(dr, g->rl_order, g->n, rb.as<float>());
    gather_f32(dg, g->rl_order, g->n, gb.as<float>());
    rv = rb.as<float>();
    gv = gb.as<float>();
  }
  // buffers 2/3 hold lambda (0/1 left free for a concurrent forward's state)
  float *lam[2] = {w->buf[2], w->buf[3]};
  OLV_HIP(hipMemcpyAsync(lam[0], gv, (size_t)g->n * 4, hipMemcpyDeviceToDevice, stream()));
  {
    // the stop threshold from gbar in the graph's numbering (the same sum in
    // both numberings)
    const uint32_t k = (uint32_t)std::min<size_t>(kPartials, ew_grid(g->n));
    hipLaunchKernelGGL(l1_partial_kernel, dim3(k), dim3(256), 0, stream(), dg, g->n, w->part);
    hipLaunchKernelGGL(ctl_init_l1_kernel, dim3(1), dim3(256), 0, stream(), w->ctl, w->part, k, (double)tol);
  }
  const uint32_t group = 8;
  uint32_t it = 0;
  IterResult res{0, 0.0, false};
  for (; it < max_iters;) {
    const uint32_t stop = std::min(max_iters, it + group);
    for (; it < stop; ++it) {
      if (blocked) {
        const BlockedPlan &bp = w->rl.blk_w[1];
        const AdjFin fin{lam[it & 1], lam[(it + 1) & 1], gv, damping};
        OLV_TRY((enqueue_blocked<WSliceVal, true>(bp, make_blk_dev(bp, wp), lam[it & 1], fin, w, it)));
      } else {
        AdjOp op{ci, wv, lam[it & 1], lam[(it + 1) & 1], gv, damping};
        OLV_TRY(enqueue_iteration(d, rp, op, w, it));
      }
    }
    OLV_TRY(read_ctl(w));
    res = iter_result(*w->h_ctl, it);
    if (res.converged) break;
  }
  const uint32_t iters = res.iters;
  const float *lf = lam[iters & 1];
  if (dgw && g->m) {
    // chunked hub rows write their edges inside the chunk blocks
    if (relabel) {
      GradOp<true> gop{ci, rv, lf, dgw, w->rl.out.eid, damping};
      launch_seg(d, rp, gop, (IterCtl *)nullptr, 0u);
    } else {
      GradOp<false> gop{ci, rv, lf, dgw, nullptr, damping};
      launch_seg(d, rp, gop, (IterCtl *)nullptr, 0u);
    }
  }
  if (dgt) {
    if (relabel) {
      hipLaunchKernelGGL(gather_scale_kernel, dim3(ew_grid(g->n)), dim3(256), 0, stream(), lf,
                         (const uint32_t *)g->rl_pi, 1.0f - damping, g->n, dgt);
    } else {
      hipLaunchKernelGGL(scale_kernel, dim3(ew_grid(g->n)), dim3(256), 0, stream(), lf, 1.0f - damping, g->n, dgt);
    }
  }
  if (dl) {
    if (relabel) gather_f32(lf, g->rl_pi, g->n, dl);
    else OLV_HIP(hipMemcpyAsync(dl, lf, (size_t)g->n * 4, hipMemcpyDeviceToDevice, stream()));
  }
  OLV_HIP(hipGetLastError());
  t.stop();
  const double ms = t.ms();
  if (host) {
    if (grad_w && g->m) OLV_HIP(hipMemcpy(grad_w, dgw, (size_t)g->m * 4, hipMemcpyDeviceToHost));
    if (grad_t) OLV_HIP(hipMemcpy(grad_t, dgt, (size_t)g->n * 4, hipMemcpyDeviceToHost));
    if (lambda_out) OLV_HIP(hipMemcpy(lambda_out, dl, (size_t)g->n * 4, hipMemcpyDeviceToHost));
  }
  if (st) {
    std::memset(st, 0, sizeof(*st));
    st->elapsed_ms = ms;
    st->iterations = iters;
    st->residual = res.residual;
    st->edges_traversed = (uint64_t)g->m * (iters + 1);
    st->edges_inspected = st->edges_traversed;
    st->vertices_reached = g->n;
  }
  if (!res.converged)
    return fail(OLV_ENOCONV, "olv_pagerank_vjp_f32: residual %.3e after %u iterations", res.residual, iters);
  return OLV_OK;
}

}  // extern "C"
