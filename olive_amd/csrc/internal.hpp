// Internal types shared by the translation units of libolive_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "error.hpp"
#include "olive_hip.h"

#ifndef OLV_HUB_CHUNK
#define OLV_HUB_CHUNK 1024
#endif

namespace olv {

hipStream_t stream();
int cu_count();
int current_device();

// Pull/push schedule of a CSR: vertices split into degree buckets, each run
// with a different number of lanes per vertex; rows longer than kChunk are
// split into chunks whose partial sums are combined in a fixed order.
struct SegPlan {
  static constexpr uint32_t kBuckets = 4;  // lanes per vertex: 1, 8, 64, block
  uint32_t *list = nullptr;        // vertex ids grouped by bucket
  uint32_t offsets[kBuckets + 1] = {0, 0, 0, 0, 0};
  uint32_t bounds[kBuckets] = {0, 0, 0, 0};  // max degree of each bucket
  // chunked rows (degree > chunk): (vertex, first chunk slot) and chunk table
  uint32_t nhub = 0;
  uint32_t nchunks = 0;
  uint32_t chunk = 0;
  uint32_t nempty = 0;              // rows without entries, at the end of bucket 0
  // every bucket's list is a run of consecutive ids (a renumbered CSR):
  // list[i] = i + shift[bucket]
  bool contig = false;
  uint32_t shift[kBuckets] = {0, 0, 0, 0};
  uint32_t *hub_vertex = nullptr;   // nhub
  uint32_t *hub_first = nullptr;    // nhub + 1 (chunk slot ranges)
  uint32_t *chunk_hub = nullptr;    // nchunks: owning hub index
  float *chunk_partial = nullptr;   // nchunks
  uint32_t *hub_done = nullptr;     // nhub arrival counters
  void release();
};

// Launch shapes and thresholds of one graph (olv_tuning in the C ABI; the
// defaults are the settings measured best on RMAT graphs).
struct Tuning {
  uint32_t bfs_expand_per_cu = 8;       // BFS expansion workgroups per CU (graph numbering)
  uint32_t bfs_expand_per_cu_twin = 6;  // the same on the renumbered twin
  uint32_t bfs_compact_per_cu = 8;      // BFS compaction workgroups per CU
  uint32_t td_block_thr = 256;  // degree >= this: whole workgroup per vertex
  uint32_t td_wave_thr = 64;    // degree >= this: one wave per vertex
  uint32_t hub_thr = OLV_HUB_CHUNK;  // degree >= this: split into hub items of hub_thr edges
  uint32_t block_threads = 256;
  float alpha = 14.0f;          // top-down -> bottom-up when m_f > m_u / alpha
  float beta = 24.0f;           // bottom-up -> top-down when n_f < n / beta
  uint32_t mbfs_pull_per_cu = 48;    // batched BFS bottom-up workgroups per CU
  uint32_t seg_grid_per_cu = 32;     // row pulls: workgroups per CU
  uint32_t seg_unroll = 8;           // row pulls: gathers in flight per lane (4, 8, 16)
  uint32_t seg_bounds[3] = {8, 96, 1024};  // row pulls: 1 lane / 8 lanes / a wave up to these degrees
  uint32_t sssp_relax_per_cu = 6;    // SSSP relaxation workgroups per CU (direct path)
  uint32_t sssp_delta_num = 32;      // SSSP near-far step: mean weight x this / mean out-degree
  uint32_t source = 0;               // 0 defaults, 1 measured, 2 tuning cache, 3 set by the caller
  double tune_ms = 0.0;
  bool tuned = false;
};
constexpr uint32_t kHubThrMin = 256;  // smallest hub item (sizes the hub item buffers)

struct BfsWork;   // bfs.hip
struct PrWork;    // pagerank.hip
struct SsspBinWork;  // sssp_bin.hip

}  // namespace olv

struct olv_graph {
  int device = 0;
  uint32_t n = 0, m = 0;
  uint32_t *row_ptr = nullptr, *col = nullptr;
  void *weights = nullptr;
  int wdtype = OLV_W_NONE;
  uint32_t *in_row_ptr = nullptr, *in_col = nullptr, *in_eid = nullptr;
  uint32_t *no_in_bm = nullptr;    // bitmap: in-degree 0 (built with the transpose)
  uint2 *in_head = nullptr;        // first two in-neighbours per vertex (built with the transpose)
  bool wstat_ready = false;        // int32 weight statistics (SSSP near-far step)
  int32_t wmin = 0;
  double wmean = 0.0;
  uint32_t max_outdeg = 0, max_indeg = 0;
  uint32_t sssp_rounds[4] = {0, 0, 0, 0};  // rounds of the last SSSP runs (first chunk = their maximum)
  float *inv_outdeg = nullptr;     // 1/outdeg (0 for sinks), fp32
  olv::Tuning tune;
  olv::SegPlan plan_in;   // buckets over the in-CSR (pull PageRank)
  olv::SegPlan plan_out;  // buckets over the out-CSR (adjoint, gradients)
  bool plan_in_ready = false, plan_out_ready = false;
  olv::BfsWork *bfs = nullptr;
  olv::PrWork *pr = nullptr;
  olv::SsspBinWork *sssp_bin = nullptr;  // binned SSSP work (for the last k sources)
  // In-degree numbering: vertices by descending in-degree, ties by id
  // (built on first use; shared by the PageRank pull and the BFS twin).
  uint32_t *rl_pi = nullptr;       // vertex -> new number
  uint32_t *rl_order = nullptr;    // new number -> vertex
  std::vector<uint32_t> h_rl_pi;   // host copies of rl_pi (seed mapping) and rl_order
  std::vector<uint32_t> h_rl_order;
  olv_graph *twin = nullptr;       // this graph renumbered by rl_pi (BFS runs on it)
  const uint32_t *out_map = nullptr;  // a twin: its vertex -> the original graph's (not owned)
  uint32_t *union_map = nullptr;   // multigraph: unused here
};

namespace olv {

// Device helpers implemented in runtime.hip
int fill_u32(uint32_t *p, uint32_t v, size_t count);
int build_seg_plan(const olv_graph *g, bool in_edges, SegPlan &plan);
// rp: host row offsets (n + 1); bounds: the bucket degree bounds (Tuning::seg_bounds)
int build_seg_plan_rp(const std::vector<uint32_t> &rp, const uint32_t *bounds, SegPlan &plan);
// plan was built with bounds other than the graph's current ones
bool seg_plan_stale(const SegPlan &plan, const Tuning &t);
// A graph never tuned takes its signature's entry in the tuning cache named
// by the OLV_TUNE_CACHE environment variable, if any (autotune.hip).
int apply_cached_tuning(olv_graph *g);
// In-degree numbering of g (g->rl_pi, g->rl_order, g->h_rl_pi), built
// once; needs the in-CSR.  order (optional) receives new number -> vertex.
int indeg_order(olv_graph *g, std::vector<uint32_t> *order);
// g->twin: g renumbered by its in-degree numbering, with its transpose.
int ensure_twin(olv_graph *g);
int ensure_in_head(olv_graph *g);
int compute_inv_outdeg(const uint32_t *rp, uint32_t n, float *inv);  // 1/outdeg (0 for sinks)
// Device exclusive prefix sum (ingest.hip): out[0..n), out[n] = total < 2^32.
int exclusive_scan(const uint32_t *in, uint32_t *out, size_t n);
void free_bfs_work(olv_graph *g);
// This device's address of the allocation-free BFS's static work array
// (looked up once per device).
int noalloc_scratch(unsigned char **out);
void free_pr_work(olv_graph *g);
void free_sssp_bin(olv_graph *g);
// Binned SSSP (sssp_bin.hip) for k instances from device sources; *used =
// false when the graph does not fit it (the caller then relaxes directly).
int sssp_bin_run(const olv_graph *g, const uint32_t *d_srcs, uint32_t k, const int32_t *weights, int32_t *dist,
                 int32_t delta, unsigned flags, olv_stats *st, bool *used);

// RAII device buffer for temporaries inside an entry point.
struct DevBuf {
  void *p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  int alloc(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess) {
      p = nullptr;
      return fail(OLV_ENOMEM, "device allocation of %zu bytes failed", bytes);
    }
    return OLV_OK;
  }
  template <class T>
  T *as() const { return static_cast<T *>(p); }
};

// Device scratch blocks kept for reuse by later calls (runtime.hip): a
// process-wide pool per device.  scratch_get returns a cached block of at
// least `bytes` (at most twice as large, or 1 MiB larger) or a new one;
// scratch_put waits for the calling thread's stream, then keeps the block.
// olv_free_cached_memory() returns the cached blocks to the device.
void *scratch_get(size_t bytes);
void scratch_put(void *p);

// RAII scratch buffer from the pool (for temporaries inside entry points).
struct ScratchBuf {
  void *p = nullptr;
  ScratchBuf() = default;
  ScratchBuf(const ScratchBuf &) = delete;
  ScratchBuf &operator=(const ScratchBuf &) = delete;
  ~ScratchBuf() { reset(); }
  void reset() {
    if (p) scratch_put(p);
    p = nullptr;
  }
  int alloc(size_t bytes) {
    reset();
    p = scratch_get(bytes ? bytes : 4);
    if (!p) return fail(OLV_ENOMEM, "device scratch of %zu bytes failed", bytes);
    return OLV_OK;
  }
  template <class T>
  T *as() const { return static_cast<T *>(p); }
};

// Event-timed region on the current stream.
struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  Timer() {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
  }
  ~Timer() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  void start() { (void)hipEventRecord(a, stream()); }
  void stop() { (void)hipEventRecord(b, stream()); }
  double ms() {
    float t = 0.f;
    (void)hipEventSynchronize(b);
    (void)hipEventElapsedTime(&t, a, b);
    return t;
  }
};

}  // namespace olv
