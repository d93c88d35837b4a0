// Runtime half of libolive_hip.so: errors, streams, device memory, fills,
// frontier conversions, graph upload and the in-edge (transposed) CSR.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bfs_core.hpp"
#include "device.hpp"
#include "internal.hpp"

namespace olv {

static thread_local char g_err[1024] = "";
char *error_buffer() { return g_err; }

static thread_local hipStream_t g_stream = nullptr;
hipStream_t stream() { return g_stream; }

static int g_device = -1;
static int g_cus = 0;

int current_device() {
  if (g_device < 0) {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) d = 0;
    g_device = d;
  }
  return g_device;
}

int cu_count() {
  if (g_cus == 0) {
    int d = current_device();
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, d) != hipSuccess || c <= 0) c = 256;
    g_cus = c;
  }
  return g_cus;
}

// ------------------------------------------------------------------ fills
template <class T>
__global__ void fill_kernel(T *__restrict__ p, T v, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

// 16-byte stores for the aligned bulk of a 4-byte fill.
__global__ void fill_u32x4_kernel(uint4 *__restrict__ p, uint32_t v, size_t count4) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  const uint4 q = make_uint4(v, v, v, v);
  for (; i < count4; i += stride) p[i] = q;
}

static unsigned fill_grid(size_t work) {
  size_t blocks = (work + 255) / 256;
  size_t cap = (size_t)cu_count() * 8;
  return (unsigned)std::max<size_t>(1, std::min(blocks, cap));
}

int fill_u32(uint32_t *p, uint32_t v, size_t count) {
  if (count == 0) return OLV_OK;
  uintptr_t addr = reinterpret_cast<uintptr_t>(p);
  size_t head = 0;
  if (addr % 16) head = std::min(count, (16 - addr % 16) / 4);
  if (head) hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, stream(), p, v, head);
  size_t rest = count - head;
  size_t n4 = rest / 4;
  if (n4) hipLaunchKernelGGL(fill_u32x4_kernel, dim3(fill_grid(n4)), dim3(256), 0, stream(),
                             reinterpret_cast<uint4 *>(p + head), v, n4);
  size_t tail = rest - n4 * 4;
  if (tail) hipLaunchKernelGGL(fill_kernel<uint32_t>, dim3(1), dim3(64), 0, stream(), p + head + n4 * 4, v, tail);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

// ------------------------------------------------ bitmap <-> queue
// One lane per bitmap word; block scan of popcounts; one global atomic per
// block for the output base.
template <int BLOCK>
__global__ void __launch_bounds__(BLOCK) bitmap_to_queue_kernel(const uint32_t *__restrict__ bm, uint32_t nwords,
                                                                uint32_t *__restrict__ queue,
                                                                uint32_t *__restrict__ count) {
  __shared__ uint32_t scan_s[BLOCK / kWave + 1];
  __shared__ uint32_t base_s;
  for (uint32_t w0 = blockIdx.x * BLOCK; w0 < nwords; w0 += gridDim.x * BLOCK) {
    uint32_t w = w0 + threadIdx.x;
    uint32_t word = w < nwords ? bm[w] : 0u;
    uint32_t total;
    uint32_t off = block_exclusive_scan<BLOCK>((uint32_t)__popc(word), scan_s, &total);
    if (threadIdx.x == 0) base_s = total ? atomicAdd(count, total) : 0u;
    __syncthreads();
    uint32_t pos = base_s + off;
    while (word) {
      uint32_t b = __ffs(word) - 1;
      queue[pos++] = w * 32u + b;
      word &= word - 1;
    }
    __syncthreads();
  }
}

__global__ void queue_to_bitmap_kernel(const uint32_t *__restrict__ queue, const uint32_t *__restrict__ count,
                                       uint32_t *__restrict__ bm, uint32_t n) {
  uint32_t c = *count;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c; i += gridDim.x * blockDim.x) {
    uint32_t v = queue[i];
    if (v < n) atomicOr(&bm[v >> 5], 1u << (v & 31));
  }
}

// Degree binning of a queue into three bucket queues (thread / wave /
// workgroup per vertex) with wave-aggregated atomics.
__global__ void frontier_bin_kernel(const uint32_t *__restrict__ queue, const uint32_t *__restrict__ count,
                                    const uint32_t *__restrict__ row_ptr, uint32_t thr_wave, uint32_t thr_block,
                                    uint32_t *__restrict__ q0, uint32_t *__restrict__ q1,
                                    uint32_t *__restrict__ q2, uint32_t *__restrict__ counts) {
  uint32_t c = *count;
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t base = blockIdx.x * blockDim.x; base < c; base += stride) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < c;
    uint32_t v = valid ? queue[i] : 0u;
    uint32_t deg = valid ? row_ptr[v + 1] - row_ptr[v] : 0u;
    int b = !valid ? -1 : (deg >= thr_block ? 2 : (deg >= thr_wave ? 1 : 0));
    uint32_t *qs[3] = {q0, q1, q2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      uint64_t mask = ballot(b == k);
      if (!mask) continue;
      uint32_t leader = __ffsll((unsigned long long)mask) - 1;
      uint32_t basek = 0;
      if (lane_id() == leader) basek = atomicAdd(&counts[k], (uint32_t)__popcll(mask));
      basek = __shfl(basek, leader, kWave);
      if (b == k) qs[k][basek + lane_rank(mask)] = v;
    }
  }
}

// Queue entries below hub_thr out-edges to out_q; every other entry becomes
// ceil(deg / hub_thr) hub items (vertex, chunk).  One atomic per workgroup
// and output per pass (a block-wide scan places the entries).
__global__ void __launch_bounds__(256)
    frontier_split_kernel(const uint32_t *__restrict__ queue, const uint32_t *__restrict__ count,
                          const uint32_t *__restrict__ row_ptr, uint32_t hub_thr, uint32_t *__restrict__ out_q,
                          uint32_t *__restrict__ out_count, uint2 *__restrict__ hubs, uint32_t *__restrict__ hub_count) {
  __shared__ uint32_t s_scan[256 / kWave + 1];
  __shared__ uint32_t s_base[2];
  const uint32_t c = *count;
  for (uint32_t base = blockIdx.x * 256; base < c; base += gridDim.x * 256) {
    const uint32_t i = base + threadIdx.x;
    uint32_t v = 0, deg = 0;
    if (i < c) {
      v = queue[i];
      deg = row_ptr[v + 1] - row_ptr[v];
    }
    const bool light = i < c && deg < hub_thr;
    const uint32_t nh = (i < c && !light) ? (deg + hub_thr - 1) / hub_thr : 0u;
    uint32_t tl, th;
    const uint32_t ol = block_exclusive_scan<256>(light ? 1u : 0u, s_scan, &tl);
    const uint32_t oh = block_exclusive_scan<256>(nh, s_scan, &th);
    if (threadIdx.x == 0) {
      s_base[0] = tl ? atomicAdd(out_count, tl) : 0u;
      s_base[1] = th ? atomicAdd(hub_count, th) : 0u;
    }
    __syncthreads();
    if (light) out_q[s_base[0] + ol] = v;
    for (uint32_t k = 0; k < nh; ++k) hubs[s_base[1] + oh + k] = make_uint2(v, k);
    __syncthreads();
  }
}

// --------------------------------------------------------- host transpose
// Stable counting transpose: in-lists come out sorted by source id, and
// in_eid[k] is the out-edge id of in-slot k.
static int host_transpose(uint32_t n, uint32_t m, const uint32_t *row_ptr, const uint32_t *col,
                          std::vector<uint32_t> &in_rp, std::vector<uint32_t> &in_col,
                          std::vector<uint32_t> &in_eid) {
  in_rp.assign((size_t)n + 1, 0);
  in_col.resize(std::max<uint32_t>(m, 1));
  in_eid.resize(std::max<uint32_t>(m, 1));
  int nt = host_threads();
  // bound the per-thread count arrays to ~1 GiB in total
  while (nt > 1 && (size_t)nt * n * 4 > (1ull << 30)) --nt;
  if (m < (1u << 16)) nt = 1;
  // split sources into nt ranges of about m/nt edges each
  std::vector<uint32_t> vb(nt + 1, n);
  vb[0] = 0;
  for (int t = 1; t < nt; ++t) {
    uint64_t target = (uint64_t)m * t / nt;
    vb[t] = (uint32_t)(std::upper_bound(row_ptr, row_ptr + n + 1, (uint32_t)target) - row_ptr - 1);
    if (vb[t] < vb[t - 1]) vb[t] = vb[t - 1];
  }
  std::vector<std::vector<uint32_t>> cnt(nt, std::vector<uint32_t>(n, 0));
  auto run = [&](auto fn) {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) th.emplace_back([&, t] { fn(t); });
    for (auto &x : th) x.join();
  };
  run([&](int t) {
    uint32_t *c = cnt[t].data();
    for (uint32_t e = row_ptr[vb[t]]; e < row_ptr[vb[t + 1]]; ++e) c[col[e]]++;
  });
  // in_rp and per-thread cursors, parallel over vertex ranges
  {
    uint64_t acc = 0;
    for (uint32_t v = 0; v < n; ++v) {
      in_rp[v] = (uint32_t)acc;
      for (int t = 0; t < nt; ++t) acc += cnt[t][v];
    }
    in_rp[n] = (uint32_t)acc;
  }
  run([&](int t) {
    uint32_t lo = (uint32_t)((uint64_t)n * t / nt), hi = (uint32_t)((uint64_t)n * (t + 1) / nt);
    for (uint32_t v = lo; v < hi; ++v) {
      uint32_t cur = in_rp[v];
      for (int s = 0; s < nt; ++s) {
        uint32_t c = cnt[s][v];
        cnt[s][v] = cur;
        cur += c;
      }
    }
  });
  run([&](int t) {
    uint32_t *cur = cnt[t].data();
    for (uint32_t u = vb[t]; u < vb[t + 1]; ++u)
      for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) {
        uint32_t k = cur[col[e]]++;
        in_col[k] = u;
        in_eid[k] = e;
      }
  });
  return OLV_OK;
}

static int upload(uint32_t **dst, const uint32_t *src, size_t count) {
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(dst), std::max<size_t>(count, 1) * 4));
  if (count) OLV_HIP(hipMemcpy(*dst, src, count * 4, hipMemcpyHostToDevice));
  return OLV_OK;
}

// First two in-neighbours of every vertex (0xffffffff where absent).
__global__ void in_head_kernel(const uint32_t *__restrict__ irp, const uint32_t *__restrict__ icol, uint32_t n,
                               uint2 *__restrict__ head) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    const uint32_t b = irp[v], d = irp[v + 1] - b;
    head[v] = make_uint2(d > 0 ? icol[b] : 0xffffffffu, d > 1 ? icol[b + 1] : 0xffffffffu);
  }
}

// Build g->in_head from the in-CSR on the device if it is missing (graphs
// whose in-CSR was assembled on the device, e.g. graph unions).
int ensure_in_head(olv_graph *g) {
  if (g->in_head || !g->in_row_ptr) return OLV_OK;
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&g->in_head), std::max<size_t>(g->n, 1) * sizeof(uint2)));
  if (g->n) {
    hipLaunchKernelGGL(in_head_kernel, dim3(fill_grid(g->n)), dim3(256), 0, g_stream, g->in_row_ptr, g->in_col, g->n,
                       g->in_head);
    OLV_HIP(hipGetLastError());
    OLV_HIP(hipStreamSynchronize(g_stream));
  }
  return OLV_OK;
}

static int attach_transpose(olv_graph *g, const uint32_t *h_rp, const uint32_t *h_col) {
  std::vector<uint32_t> irp, icol, ieid;
  OLV_TRY(host_transpose(g->n, g->m, h_rp, h_col, irp, icol, ieid));
  uint32_t mx = 0;
  for (uint32_t v = 0; v < g->n; ++v) mx = std::max(mx, irp[v + 1] - irp[v]);
  g->max_indeg = mx;
  OLV_TRY(upload(&g->in_row_ptr, irp.data(), (size_t)g->n + 1));
  OLV_TRY(upload(&g->in_col, icol.data(), g->m));
  OLV_TRY(upload(&g->in_eid, ieid.data(), g->m));
  // First two in-neighbours of every vertex (0xffffffff where absent): a
  // bottom-up step finds most parents there with one coalesced load.
  std::vector<uint2> head(std::max<uint32_t>(g->n, 1), make_uint2(0xffffffffu, 0xffffffffu));
  for (uint32_t v = 0; v < g->n; ++v) {
    const uint32_t d = irp[v + 1] - irp[v];
    if (d > 0) head[v].x = icol[irp[v]];
    if (d > 1) head[v].y = icol[irp[v] + 1];
  }
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&g->in_head), head.size() * sizeof(uint2)));
  OLV_HIP(hipMemcpy(g->in_head, head.data(), head.size() * sizeof(uint2), hipMemcpyHostToDevice));
  // Vertices without in-edges: only a seed can reach them.
  const size_t nw = ((size_t)g->n + 31) / 32;
  std::vector<uint32_t> no_in(std::max<size_t>(nw, 1), 0u);
  for (uint32_t v = 0; v < g->n; ++v)
    if (irp[v + 1] == irp[v]) no_in[v >> 5] |= 1u << (v & 31);
  OLV_TRY(upload(&g->no_in_bm, no_in.data(), nw));
  return OLV_OK;
}

__global__ void inv_outdeg_kernel(const uint32_t *__restrict__ rp, uint32_t n, float *__restrict__ inv) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x) {
    uint32_t d = rp[v + 1] - rp[v];
    inv[v] = d ? 1.0f / (float)d : 0.0f;
  }
}

int compute_inv_outdeg(const uint32_t *rp, uint32_t n, float *inv) {
  hipLaunchKernelGGL(inv_outdeg_kernel, dim3(fill_grid(n)), dim3(256), 0, g_stream, rp, n, inv);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

void SegPlan::release() {
  void *ps[] = {list, hub_vertex, hub_first, chunk_hub, chunk_partial, hub_done};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  list = hub_vertex = hub_first = chunk_hub = hub_done = nullptr;
  chunk_partial = nullptr;
  nhub = nchunks = 0;
}

// Group vertices by degree for the pull/push segment kernels.
int build_seg_plan(const olv_graph *g, bool in_edges, SegPlan &plan) {
  const uint32_t n = g->n;
  const uint32_t *d_rp = in_edges ? g->in_row_ptr : g->row_ptr;
  if (!d_rp) return fail(OLV_EINVAL, "segment plan needs the %s CSR", in_edges ? "in-edge" : "out-edge");
  std::vector<uint32_t> rp((size_t)n + 1);
  OLV_HIP(hipMemcpy(rp.data(), d_rp, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
  return build_seg_plan_rp(rp, g->tune.seg_bounds, plan);
}

bool seg_plan_stale(const SegPlan &plan, const Tuning &t) {
  return plan.bounds[0] != t.seg_bounds[0] || plan.bounds[1] != t.seg_bounds[1] || plan.bounds[2] != t.seg_bounds[2];
}

int build_seg_plan_rp(const std::vector<uint32_t> &rp, const uint32_t *bounds, SegPlan &plan) {
  plan.release();
  const uint32_t n = (uint32_t)(rp.size() - 1);
  // lanes per vertex: 1 (deg<=b0), 8 (<=b1), 64 (<=b2), workgroup (<=chunk), chunked hubs
  plan.bounds[0] = bounds[0];
  plan.bounds[1] = bounds[1];
  plan.bounds[2] = bounds[2];
  plan.chunk = 8192;
  plan.bounds[3] = plan.chunk;
  std::vector<uint32_t> lists[SegPlan::kBuckets];
  std::vector<uint32_t> hubs;
  std::vector<uint32_t> empty_rows;  // listed last in bucket 0: iterations may skip them
  for (uint32_t v = 0; v < n; ++v) {
    uint32_t d = rp[v + 1] - rp[v];
    if (d == 0) empty_rows.push_back(v);
    else if (d <= plan.bounds[0]) lists[0].push_back(v);
    else if (d <= plan.bounds[1]) lists[1].push_back(v);
    else if (d <= plan.bounds[2]) lists[2].push_back(v);
    else if (d <= plan.bounds[3]) lists[3].push_back(v);
    else hubs.push_back(v);
  }
  plan.nempty = (uint32_t)empty_rows.size();
  lists[0].insert(lists[0].end(), empty_rows.begin(), empty_rows.end());
  std::vector<uint32_t> all;
  all.reserve(n);
  plan.offsets[0] = 0;
  for (uint32_t b = 0; b < SegPlan::kBuckets; ++b) {
    all.insert(all.end(), lists[b].begin(), lists[b].end());
    plan.offsets[b + 1] = (uint32_t)all.size();
  }
  plan.contig = true;
  for (uint32_t b = 0; b < SegPlan::kBuckets && plan.contig; ++b) {
    const uint32_t o = plan.offsets[b];
    plan.shift[b] = o < plan.offsets[b + 1] ? all[o] - o : 0u;  // mod 2^32
    for (uint32_t i = o; i < plan.offsets[b + 1]; ++i)
      if (all[i] != i + plan.shift[b]) {
        plan.contig = false;
        break;
      }
  }
  OLV_TRY(upload(&plan.list, all.data(), all.size()));
  std::vector<uint32_t> first(hubs.size() + 1, 0), chub;
  for (size_t h = 0; h < hubs.size(); ++h) {
    uint32_t d = rp[hubs[h] + 1] - rp[hubs[h]];
    uint32_t k = (d + plan.chunk - 1) / plan.chunk;
    first[h + 1] = first[h] + k;
    for (uint32_t i = 0; i < k; ++i) chub.push_back((uint32_t)h);
  }
  plan.nhub = (uint32_t)hubs.size();
  plan.nchunks = first.back();
  OLV_TRY(upload(&plan.hub_vertex, hubs.data(), hubs.size()));
  OLV_TRY(upload(&plan.hub_first, first.data(), first.size()));
  OLV_TRY(upload(&plan.chunk_hub, chub.data(), chub.size()));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&plan.chunk_partial), std::max<size_t>(plan.nchunks, 1) * 4));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&plan.hub_done), std::max<size_t>(plan.nhub, 1) * 4));
  OLV_HIP(hipMemset(plan.hub_done, 0, std::max<size_t>(plan.nhub, 1) * 4));
  return OLV_OK;
}

static void graph_free(olv_graph *g) {
  if (!g) return;
  free_bfs_work(g);
  free_pr_work(g);
  free_sssp_bin(g);
  g->plan_in.release();
  g->plan_out.release();
  graph_free(g->twin);
  void *ps[] = {g->row_ptr, g->col, g->weights, g->in_row_ptr, g->in_col, g->in_eid, g->no_in_bm, g->in_head,
                g->inv_outdeg, g->union_map, g->rl_pi, g->rl_order};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  delete g;
}

int indeg_order(olv_graph *g, std::vector<uint32_t> *order_out) {
  if (g->rl_pi) {
    if (order_out) *order_out = g->h_rl_order;
    return OLV_OK;
  }
  if (!g->in_row_ptr) OLV_TRY(olv_graph_build_transpose(g));
  const uint32_t n = g->n;
  std::vector<uint32_t> in_rp((size_t)n + 1);
  OLV_HIP(hipMemcpy(in_rp.data(), g->in_row_ptr, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
  // counting sort by in-degree, descending, stable in the vertex id
  uint32_t maxd = 0;
  for (uint32_t v = 0; v < n; ++v) maxd = std::max(maxd, in_rp[v + 1] - in_rp[v]);
  std::vector<uint32_t> start((size_t)maxd + 2, 0);
  for (uint32_t v = 0; v < n; ++v) ++start[maxd - (in_rp[v + 1] - in_rp[v]) + 1];
  for (size_t k = 1; k < start.size(); ++k) start[k] += start[k - 1];
  std::vector<uint32_t> order(n), pi(n);
  for (uint32_t v = 0; v < n; ++v) {
    const uint32_t i = start[maxd - (in_rp[v + 1] - in_rp[v])]++;
    order[i] = v;
    pi[v] = i;
  }
  uint32_t *d_pi = nullptr, *d_order = nullptr;
  int st = upload(&d_pi, pi.data(), n);
  if (st == OLV_OK) st = upload(&d_order, order.data(), n);
  if (st != OLV_OK) {  // leave the graph without a half-built numbering
    if (d_pi) (void)hipFree(d_pi);
    if (d_order) (void)hipFree(d_order);
    return st;
  }
  g->rl_pi = d_pi;
  g->rl_order = d_order;
  g->h_rl_pi = std::move(pi);
  g->h_rl_order = std::move(order);
  if (order_out) *order_out = g->h_rl_order;
  return OLV_OK;
}

// s[e] = pi[u], d[e] = pi[col[e]] for every edge e = (u, col[e]); one wave
// per row.
__global__ void coo_relabel_kernel(const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col,
                                   const uint32_t *__restrict__ pi, uint32_t n, uint32_t *__restrict__ s,
                                   uint32_t *__restrict__ d) {
  const uint32_t waves = gridDim.x * (blockDim.x / kWave);
  const uint32_t lane = threadIdx.x % kWave;
  for (uint32_t u = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; u < n; u += waves) {
    const uint32_t b = rp[u], e = rp[u + 1], pu = pi[u];
    for (uint32_t k0 = b + lane; k0 < e; k0 += 4 * kWave) {  // four loads in flight per lane
      uint32_t c[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) c[i] = k0 + i * kWave < e ? col[k0 + i * kWave] : 0u;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (k0 + i * kWave < e) {
          s[k0 + i * kWave] = pu;
          d[k0 + i * kWave] = pi[c[i]];
        }
    }
  }
}

int ensure_twin(olv_graph *g) {
  if (g->twin) return OLV_OK;
  OLV_TRY(indeg_order(g, nullptr));
  ScratchBuf s, d;
  OLV_TRY(s.alloc(std::max<size_t>(g->m, 1) * 4));
  OLV_TRY(d.alloc(std::max<size_t>(g->m, 1) * 4));
  if (g->m) {
    hipLaunchKernelGGL(coo_relabel_kernel, dim3(fill_grid((size_t)g->n * kWave)), dim3(256), 0, stream(), g->row_ptr,
                       g->col, g->rl_pi, g->n, s.as<uint32_t>(), d.as<uint32_t>());
    OLV_HIP(hipGetLastError());
  }
  olv_graph *t = nullptr;
  OLV_TRY(olv_graph_from_coo(&t, g->n, g->m, s.as<uint32_t>(), d.as<uint32_t>(), g->weights, g->wdtype,
                             OLV_BUILD_TRANSPOSE, nullptr));
  OLV_HIP(hipStreamSynchronize(stream()));
  t->device = g->device;
  t->out_map = g->rl_order;
  g->twin = t;
  return OLV_OK;
}

}  // namespace olv

using namespace olv;

// ------------------------------------------------------------ scratch pool
namespace olv {
namespace {
struct ScratchPool {
  std::mutex mu;
  std::multimap<size_t, std::pair<int, void *>> free_blocks;  // size -> (device, block)
  std::unordered_map<void *, std::pair<int, size_t>> owned;    // block -> (device, size)
};
ScratchPool &pool() {
  static ScratchPool *p = new ScratchPool();  // never destroyed: blocks outlive static teardown
  return *p;
}
void release_cached(int device) {  // caller holds the lock
  ScratchPool &P = pool();
  for (auto it = P.free_blocks.begin(); it != P.free_blocks.end();) {
    if (device >= 0 && it->second.first != device) {
      ++it;
      continue;
    }
    (void)hipFree(it->second.second);
    P.owned.erase(it->second.second);
    it = P.free_blocks.erase(it);
  }
}
}  // namespace

void *scratch_get(size_t bytes) {
  const int dev = current_device();
  bytes = (bytes + 255) / 256 * 256;
  ScratchPool &P = pool();
  std::lock_guard<std::mutex> lk(P.mu);
  const size_t limit = std::max(2 * bytes, bytes + (size_t(1) << 20));
  for (auto it = P.free_blocks.lower_bound(bytes); it != P.free_blocks.end() && it->first <= limit; ++it) {
    if (it->second.first != dev) continue;
    void *p = it->second.second;
    P.free_blocks.erase(it);
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    release_cached(dev);  // give the cached blocks back and try once more
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  P.owned[p] = {dev, bytes};
  return p;
}

void scratch_put(void *p) {
  if (!p) return;
  (void)hipStreamSynchronize(stream());  // no queued work of this thread still uses the block
  ScratchPool &P = pool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.owned.find(p);
  if (it == P.owned.end()) return;
  P.free_blocks.insert({it->second.second, {it->second.first, p}});
}

}  // namespace olv

extern "C" {

const char *olv_last_error(void) { return g_err; }

int olv_version(void) { return 100; }

int olv_init(int device) {
  int count = 0;
  OLV_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(OLV_EINVAL, "olv_init: device %d not present (%d devices)", device, count);
  OLV_HIP(hipSetDevice(device));
  g_device = device;
  g_cus = 0;
  (void)cu_count();
  unsigned char *scratch_base = nullptr;  // set-up of the allocation-free entry point on this device
  OLV_TRY(noalloc_scratch(&scratch_base));
  return OLV_OK;
}

int olv_set_stream(void *s) {
  g_stream = reinterpret_cast<hipStream_t>(s);
  return OLV_OK;
}

void *olv_get_stream(void) { return reinterpret_cast<void *>(g_stream); }

int olv_free_cached_memory(void) {
  olv::ScratchPool &P = olv::pool();
  std::lock_guard<std::mutex> lk(P.mu);
  olv::release_cached(-1);
  return OLV_OK;
}

int olv_synchronize(void) {
  OLV_HIP(hipStreamSynchronize(g_stream));
  return OLV_OK;
}

int olv_device_info(int *cus, int *warp, size_t *total) {
  int d = current_device();
  hipDeviceProp_t p;
  OLV_HIP(hipGetDeviceProperties(&p, d));
  if (cus) *cus = p.multiProcessorCount;
  if (warp) *warp = p.warpSize;
  if (total) *total = p.totalGlobalMem;
  return OLV_OK;
}

int olv_malloc(void **dptr, size_t bytes) {
  if (!dptr) return fail(OLV_EINVAL, "olv_malloc: dptr is NULL");
  if (hipMalloc(dptr, bytes ? bytes : 4) != hipSuccess) {
    *dptr = nullptr;
    return fail(OLV_ENOMEM, "olv_malloc: %zu bytes failed", bytes);
  }
  return OLV_OK;
}

int olv_free(void *p) {
  if (p) OLV_HIP(hipFree(p));
  return OLV_OK;
}

int olv_host_malloc(void **hptr, size_t bytes) {
  if (!hptr) return fail(OLV_EINVAL, "olv_host_malloc: hptr is NULL");
  if (hipHostMalloc(hptr, bytes ? bytes : 4, hipHostMallocDefault) != hipSuccess) {
    *hptr = nullptr;
    return fail(OLV_ENOMEM, "olv_host_malloc: %zu bytes failed", bytes);
  }
  return OLV_OK;
}

int olv_host_free(void *p) {
  if (p) OLV_HIP(hipHostFree(p));
  return OLV_OK;
}

int olv_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  if (!bytes) return OLV_OK;
  OLV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
  OLV_HIP(hipStreamSynchronize(g_stream));
  return OLV_OK;
}

int olv_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  if (!bytes) return OLV_OK;
  OLV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
  OLV_HIP(hipStreamSynchronize(g_stream));
  return OLV_OK;
}

int olv_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  if (!bytes) return OLV_OK;
  OLV_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, g_stream));
  return OLV_OK;
}

int olv_fill(void *dptr, uint64_t value, size_t count, int elem_bytes) {
  if (!count) return OLV_OK;
  if (!dptr) return fail(OLV_EINVAL, "olv_fill: NULL pointer");
  switch (elem_bytes) {
    case 1:
      hipLaunchKernelGGL(fill_kernel<uint8_t>, dim3(fill_grid(count)), dim3(256), 0, g_stream,
                         static_cast<uint8_t *>(dptr), (uint8_t)value, count);
      break;
    case 4:
      return fill_u32(static_cast<uint32_t *>(dptr), (uint32_t)value, count);
    case 8:
      hipLaunchKernelGGL(fill_kernel<uint64_t>, dim3(fill_grid(count)), dim3(256), 0, g_stream,
                         static_cast<uint64_t *>(dptr), (uint64_t)value, count);
      break;
    default:
      return fail(OLV_EINVAL, "olv_fill: element size %d not supported (1, 4, 8)", elem_bytes);
  }
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

int olv_bitmap_words(uint32_t n, uint32_t *words) {
  if (!words) return fail(OLV_EINVAL, "olv_bitmap_words: NULL");
  *words = (n + 31) / 32;
  return OLV_OK;
}

int olv_bitmap_to_queue(const uint32_t *bitmap, uint32_t n, uint32_t *queue, uint32_t *count_dev) {
  if (!bitmap || !queue || !count_dev) return fail(OLV_EINVAL, "olv_bitmap_to_queue: NULL pointer");
  uint32_t nw = (n + 31) / 32;
  OLV_TRY(fill_u32(count_dev, 0u, 1));
  if (nw == 0) return OLV_OK;
  unsigned grid = std::max(1u, std::min((nw + 255) / 256, (unsigned)cu_count() * 4));
  hipLaunchKernelGGL(bitmap_to_queue_kernel<256>, dim3(grid), dim3(256), 0, g_stream, bitmap, nw, queue, count_dev);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

int olv_queue_to_bitmap(const uint32_t *queue, const uint32_t *count_dev, uint32_t *bitmap, uint32_t n, int clear) {
  if (!queue || !count_dev || !bitmap) return fail(OLV_EINVAL, "olv_queue_to_bitmap: NULL pointer");
  if (clear) OLV_TRY(fill_u32(bitmap, 0u, (n + 31) / 32));
  unsigned grid = std::max(1u, std::min((n + 255) / 256, (unsigned)cu_count() * 4));
  hipLaunchKernelGGL(queue_to_bitmap_kernel, dim3(grid), dim3(256), 0, g_stream, queue, count_dev, bitmap, n);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

int olv_frontier_bin(const olv_graph *g, const uint32_t *queue, const uint32_t *count_dev,
                     uint32_t *const bucket_queues[3], uint32_t *bucket_counts_dev) {
  if (!g || !queue || !count_dev || !bucket_queues || !bucket_counts_dev)
    return fail(OLV_EINVAL, "olv_frontier_bin: NULL pointer");
  OLV_TRY(fill_u32(bucket_counts_dev, 0u, 3));
  unsigned grid = std::max(1u, std::min((g->n + 255) / 256, (unsigned)cu_count() * 4));
  hipLaunchKernelGGL(frontier_bin_kernel, dim3(grid), dim3(256), 0, g_stream, queue, count_dev, g->row_ptr,
                     g->tune.td_wave_thr, g->tune.td_block_thr, bucket_queues[0], bucket_queues[1],
                     bucket_queues[2], bucket_counts_dev);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

int olv_frontier_split(const olv_graph *g, const uint32_t *queue, const uint32_t *count_dev, uint32_t *out_queue,
                       uint32_t *counts_dev, uint32_t *hub_pairs) {
  uint2 *hubs = reinterpret_cast<uint2 *>(hub_pairs);
  if (!g || !queue || !count_dev || !out_queue || !counts_dev || !hubs)
    return fail(OLV_EINVAL, "olv_frontier_split: NULL pointer");
  OLV_TRY(fill_u32(counts_dev, 0u, 2));
  const unsigned grid = std::max(1u, std::min((g->n + 255) / 256, (unsigned)cu_count() * 4));
  hipLaunchKernelGGL(frontier_split_kernel, dim3(grid), dim3(256), 0, g_stream, queue, count_dev, g->row_ptr,
                     g->tune.hub_thr, out_queue, counts_dev, hubs, counts_dev + 1);
  OLV_HIP(hipGetLastError());
  return OLV_OK;
}

size_t olv_hub_capacity(uint64_t m) { return hub_capacity(m); }

int olv_graph_buckets(const olv_graph *g, uint32_t *thread_max, uint32_t *wave_max) {
  if (!g) return fail(OLV_EINVAL, "olv_graph_buckets: NULL graph");
  if (thread_max) *thread_max = g->tune.td_wave_thr - 1;
  if (wave_max) *wave_max = g->tune.td_block_thr - 1;
  return OLV_OK;
}

int olv_graph_create(olv_graph **out, uint32_t n, uint32_t m, const uint32_t *row_ptr, const uint32_t *col,
                     const void *weights, int weight_dtype, unsigned flags) {
  if (!out) return fail(OLV_EINVAL, "olv_graph_create: out is NULL");
  *out = nullptr;
  if (n == 0) return fail(OLV_EINVAL, "olv_graph_create: graph has no vertices");
  if (!row_ptr || (m > 0 && !col)) return fail(OLV_EINVAL, "olv_graph_create: NULL CSR array");
  if (weight_dtype != OLV_W_NONE && weight_dtype != OLV_W_I32 && weight_dtype != OLV_W_F32)
    return fail(OLV_EINVAL, "olv_graph_create: unknown weight dtype %d", weight_dtype);
  if (weight_dtype != OLV_W_NONE && m > 0 && !weights) return fail(OLV_EINVAL, "olv_graph_create: NULL weights");
  (void)current_device();
  const bool host = (flags & OLV_PTR_HOST) != 0;
  std::vector<uint32_t> h_rp, h_col;
  const uint32_t *rp = row_ptr, *cl = col;
  if (!host) {
    h_rp.resize((size_t)n + 1);
    h_col.resize(std::max<uint32_t>(m, 1));
    OLV_HIP(hipMemcpy(h_rp.data(), row_ptr, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost));
    if (m) OLV_HIP(hipMemcpy(h_col.data(), col, (size_t)m * 4, hipMemcpyDeviceToHost));
    rp = h_rp.data();
    cl = h_col.data();
  }
  // validate the CSR on the host: monotone offsets, ids in range
  if (rp[0] != 0 || rp[n] != m) return fail(OLV_EINVAL, "olv_graph_create: row_ptr[0]=%u row_ptr[n]=%u, m=%u", rp[0], rp[n], m);
  uint32_t maxdeg = 0;
  for (uint32_t v = 0; v < n; ++v) {
    if (rp[v + 1] < rp[v]) return fail(OLV_EINVAL, "olv_graph_create: row_ptr decreases at %u", v);
    maxdeg = std::max(maxdeg, rp[v + 1] - rp[v]);
  }
  for (uint32_t e = 0; e < m; ++e)
    if (cl[e] >= n) return fail(OLV_EINVAL, "olv_graph_create: col[%u]=%u out of range (n=%u)", e, cl[e], n);
  olv_graph *g = new olv_graph();
  g->device = current_device();
  g->n = n;
  g->m = m;
  g->max_outdeg = maxdeg;
  int st = OLV_OK;
  do {
    if ((st = upload(&g->row_ptr, rp, (size_t)n + 1)) != OLV_OK) break;
    if ((st = upload(&g->col, cl, m)) != OLV_OK) break;
    if (weight_dtype != OLV_W_NONE) {
      if (hipMalloc(&g->weights, std::max<size_t>(m, 1) * 4) != hipSuccess) {
        st = fail(OLV_ENOMEM, "olv_graph_create: weights allocation failed");
        break;
      }
      if (m && hipMemcpy(g->weights, weights, (size_t)m * 4, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice) != hipSuccess) {
        st = fail(OLV_EHIP, "olv_graph_create: weights copy failed");
        break;
      }
      g->wdtype = weight_dtype;
    }
    if (hipMalloc(reinterpret_cast<void **>(&g->inv_outdeg), (size_t)n * 4) != hipSuccess) {
      st = fail(OLV_ENOMEM, "olv_graph_create: allocation failed");
      break;
    }
    hipLaunchKernelGGL(inv_outdeg_kernel, dim3(fill_grid(n)), dim3(256), 0, g_stream, g->row_ptr, n, g->inv_outdeg);
    if (flags & OLV_BUILD_TRANSPOSE)
      if ((st = attach_transpose(g, rp, cl)) != OLV_OK) break;
    if (hipStreamSynchronize(g_stream) != hipSuccess) {
      st = fail(OLV_EHIP, "olv_graph_create: upload failed");
      break;
    }
  } while (0);
  if (st != OLV_OK) {
    graph_free(g);
    return st;
  }
  *out = g;
  return OLV_OK;
}

int olv_graph_destroy(olv_graph *g) {
  if (g) {
    (void)hipStreamSynchronize(g_stream);
    graph_free(g);
  }
  return OLV_OK;
}

int olv_graph_build_transpose(olv_graph *g) {
  if (!g) return fail(OLV_EINVAL, "olv_graph_build_transpose: NULL graph");
  if (g->in_row_ptr) return OLV_OK;
  std::vector<uint32_t> rp((size_t)g->n + 1), cl(std::max<uint32_t>(g->m, 1));
  OLV_HIP(hipMemcpy(rp.data(), g->row_ptr, rp.size() * 4, hipMemcpyDeviceToHost));
  if (g->m) OLV_HIP(hipMemcpy(cl.data(), g->col, (size_t)g->m * 4, hipMemcpyDeviceToHost));
  return attach_transpose(g, rp.data(), cl.data());
}

int olv_graph_counts(const olv_graph *g, uint32_t *n, uint32_t *m) {
  if (!g) return fail(OLV_EINVAL, "olv_graph_counts: NULL graph");
  if (n) *n = g->n;
  if (m) *m = g->m;
  return OLV_OK;
}

int olv_graph_device_arrays(const olv_graph *g, const uint32_t **row_ptr, const uint32_t **col,
                            const void **weights, const uint32_t **in_row_ptr, const uint32_t **in_col,
                            const uint32_t **in_eid) {
  if (!g) return fail(OLV_EINVAL, "olv_graph_device_arrays: NULL graph");
  if (row_ptr) *row_ptr = g->row_ptr;
  if (col) *col = g->col;
  if (weights) *weights = g->weights;
  if (in_row_ptr) *in_row_ptr = g->in_row_ptr;
  if (in_col) *in_col = g->in_col;
  if (in_eid) *in_eid = g->in_eid;
  return OLV_OK;
}

int olv_graph_transpose_to_host(const olv_graph *g, uint32_t *in_row_ptr, uint32_t *in_col, uint32_t *in_eid) {
  if (!g) return fail(OLV_EINVAL, "olv_graph_transpose_to_host: NULL graph");
  if (!g->in_row_ptr) return fail(OLV_EINVAL, "olv_graph_transpose_to_host: transpose not built");
  if (in_row_ptr) OLV_HIP(hipMemcpy(in_row_ptr, g->in_row_ptr, ((size_t)g->n + 1) * 4, hipMemcpyDeviceToHost));
  if (in_col && g->m) OLV_HIP(hipMemcpy(in_col, g->in_col, (size_t)g->m * 4, hipMemcpyDeviceToHost));
  if (in_eid && g->m) OLV_HIP(hipMemcpy(in_eid, g->in_eid, (size_t)g->m * 4, hipMemcpyDeviceToHost));
  return OLV_OK;
}

int olv_set_buckets(olv_graph *g, uint32_t thread_max, uint32_t wave_max, uint32_t block_threads) {
  if (!g) return fail(OLV_EINVAL, "olv_set_buckets: NULL graph");
  if (block_threads != 0 && block_threads != 256 && block_threads != 512)
    return fail(OLV_EINVAL, "olv_set_buckets: block_threads must be 256 or 512");
  if (thread_max + 1 < 2 || wave_max <= thread_max) return fail(OLV_EINVAL, "olv_set_buckets: need 1 <= thread_max < wave_max");
  g->tune.td_wave_thr = thread_max + 1;
  g->tune.td_block_thr = wave_max + 1;
  if (block_threads) g->tune.block_threads = block_threads;
  return OLV_OK;
}

}  // extern "C"
