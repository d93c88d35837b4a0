/*
 * olive_hip.h -- C ABI of libolive_hip.so, the MI355X (gfx950) frontier engine.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * integer status (OLV_OK == 0) and never throws.  On failure the message is
 * available from olv_last_error() (thread-local).  Handles are not
 * thread-safe: one olv_graph is driven from one host thread at a time.
 *
 * Device pointers are the default for algorithm inputs and outputs (the
 * caller owns them, e.g. torch tensors or olv_malloc'd buffers).  Passing
 * OLV_PTR_HOST in `flags` makes the library stage host buffers itself.
 *
 * The reference engine (Oliver<V,E,A>, a header-only CUDA template API) has
 * no FFI; each declaration below names the reference code it replaces.
 * The functor-generic part of that API (user gather/reduce/cond/update
 * functors) lives in the C++ headers under olive_amd/include and calls the
 * type-independent entry points declared here.
 */
#ifndef OLIVE_HIP_H
#define OLIVE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  OLV_OK = 0,
  OLV_EINVAL = 1,      /* bad argument (null pointer, out-of-range id, size) */
  OLV_ENOMEM = 2,      /* device or host allocation failed                   */
  OLV_EHIP = 3,        /* a HIP runtime call failed                          */
  OLV_ERANGE = 4,      /* size beyond what the ABI supports (e.g. E >= 2^32) */
  OLV_EUNSUPPORTED = 5,/* feature not available for this graph / device      */
  OLV_ENOCONV = 6,     /* iterative solve hit its iteration cap              */
  OLV_EINTERNAL = 7    /* internal consistency check failed                  */
};

/* ---- flags ------------------------------------------------------------- */
#define OLV_PTR_HOST        0x0001u  /* inputs/outputs are host pointers            */
#define OLV_NO_DIRECTION_OPT 0x0002u /* BFS: top-down (push) only                    */
#define OLV_FORCE_BOTTOM_UP  0x0004u /* BFS: every level after the first pulls
                                        (bottom-up); builds the transpose         */
#define OLV_HOST_LOOP        0x0008u /* wait for every level on the host instead
                                        of enqueueing levels in chunks whose
                                        control lives on the device             */
#define OLV_PERSISTENT       0x0020u /* BFS/SSSP: all levels in one cooperative
                                        persistent kernel with a grid barrier
                                        (slower on MI355X; kept to compare)      */
#define OLV_PROFILE          0x0010u /* time each launch of the dominant kernel
                                        with hipEvents (olv_stats.kernel_ms)     */
#define OLV_GRAPH_ORDER      0x0040u /* olv_bfs, olv_bfs_multisource, the
                                        batched BFS calls and plans and the
                                        PageRank solves (plain, weighted, VJP,
                                        JVP): run in the graph's own vertex
                                        numbering instead of the in-degree
                                        renumbering (same results, bit for bit,
                                        except that the PageRank solves'
                                        default segmented pulls add their terms
                                        in another fixed order)                  */
#define OLV_BUILD_TRANSPOSE  0x0100u /* graph_create: also build the in-edge CSR    */
#define OLV_PR_ROW_PULL      0x0400u /* olv_pagerank_f32 and the weighted
                                        forward, JVP and VJP: one pull per row
                                        over the renumbered CSR instead of the
                                        source-segmented pull (kept to compare) */
#define OLV_W_UNCHANGED      0x0200u /* weighted PageRank, VJP, JVP: the weights
                                        at this address are the ones an earlier
                                        call on this graph received, unchanged;
                                        reuse the copy of them that call put in
                                        the solver's slot order (ignored for
                                        host pointers or when no such copy is
                                        kept)                                    */
#define OLV_SSSP_DIRECT      0x0800u /* olv_sssp, olv_sssp_batched: relax with a
                                        dist probe and device-scope atomics per
                                        edge instead of binning the candidates
                                        by destination partition (kept to
                                        compare; also taken when the binned
                                        scheme does not fit: more than 2^25
                                        virtual vertices or 2^32 bin entries) */

/* weight dtypes for olv_graph_create */
enum { OLV_W_NONE = 0, OLV_W_I32 = 1, OLV_W_F32 = 2 };

/* text formats for olv_graph_load_file */
enum { OLV_FMT_EDGE_LIST = 0, OLV_FMT_DIMACS = 1 };

/* value formats for olv_write_text: "%d\n" of int32, "%f\n" of float or
 * double (printed as C's printf prints the double) */
enum { OLV_TXT_I32 = 0, OLV_TXT_F32 = 1, OLV_TXT_F64 = 2 };

typedef struct olv_graph olv_graph;

/* Per-call statistics (all optional: pass NULL). */
typedef struct olv_stats {
  double   elapsed_ms;        /* device time of the algorithm, hipEvents      */
  uint64_t edges_traversed;   /* E_r: sum of out-degrees of reached vertices  */
  uint64_t vertices_reached;  /* V_r                                          */
  uint64_t edges_inspected;   /* edges actually read (bottom-up may skip)     */
  uint32_t iterations;        /* BFS/SSSP levels, PageRank iterations         */
  uint32_t bottom_up_levels;  /* BFS: levels run bottom-up                    */
  double   residual;          /* PageRank: last L1 delta                      */
  double   kernel_ms;         /* OLV_PROFILE: summed hipEvent time of the
                                 dominant kernel's launches                   */
  uint32_t kernel_launches;   /* OLV_PROFILE: number of those launches        */
  uint32_t reserved;
} olv_stats;

/* ---- runtime ----------------------------------------------------------- */
const char *olv_last_error(void);
int olv_version(void);                    /* 100*major + minor               */
int olv_init(int device);                 /* select device, query CU count   */
int olv_set_stream(void *hip_stream);     /* NULL = the null stream          */
void *olv_get_stream(void);
int olv_synchronize(void);                /* wait for the current stream     */
/* Return the device scratch blocks the library keeps between calls (ingest,
 * egress and scan temporaries) to the device.  No reference counterpart. */
int olv_free_cached_memory(void);
int olv_device_info(int *cu_count, int *warp_size, size_t *total_mem);

/* ---- device memory (replaces GRD<T>: grd.h:76-164) --------------------- */
int olv_malloc(void **dptr, size_t bytes);                 /* GRD::reserve  */
int olv_free(void *dptr);                                  /* GRD::del      */
int olv_host_malloc(void **hptr, size_t bytes);            /* pinned staging */
int olv_host_free(void *hptr);
int olv_memcpy_h2d(void *dst, const void *src, size_t bytes); /* GRD::cache  */
int olv_memcpy_d2h(void *dst, const void *src, size_t bytes); /* GRD::persist*/
int olv_memcpy_d2d(void *dst, const void *src, size_t bytes);
/* Fill `count` elements of 1/4/8 bytes with `value` on the device: replaces
 * the host loop + full H2D copy of GRD::allTo (grd.h:90-97). */
int olv_fill(void *dptr, uint64_t value, size_t count, int elem_bytes);

/* ---- frontier (replaces VertexSubset: vertexSubset.h:38-156, bitmap.h) -- *
 * A flag frontier is a bitmap of ceil(n/32) uint32 words plus an optional
 * "touched" queue; a queue frontier is uint32 ids plus a device count.    */
int olv_bitmap_words(uint32_t n, uint32_t *words);
/* Compact a bitmap into a queue: wave64 ballot/popcount + LDS scan.
 * count_dev (device, 1 uint32) receives the number of ids. */
int olv_bitmap_to_queue(const uint32_t *bitmap, uint32_t n, uint32_t *queue,
                        uint32_t *count_dev);
/* Set bits of the ids in queue[0..*count_dev) (bitmap cleared first if
 * clear != 0). */
int olv_queue_to_bitmap(const uint32_t *queue, const uint32_t *count_dev,
                        uint32_t *bitmap, uint32_t n, int clear);
/* Split queue[0..*count_dev) by out-degree into the graph's degree buckets
 * (thread / wave / workgroup per vertex).  bucket_queues holds 3 arrays of
 * capacity n; bucket_counts_dev holds 3 uint32 (zeroed here).
 * Replaces the fixed GroupSize of edgeFilterKernel (oliverKernel.h:63-66). */
int olv_frontier_bin(const olv_graph *g, const uint32_t *queue,
                     const uint32_t *count_dev, uint32_t *const bucket_queues[3],
                     uint32_t *bucket_counts_dev);
/* Split queue[0..*count_dev) for the push kernels: entries with fewer than
 * the graph's hub_thr out-edges go to out_queue (capacity n), every other
 * entry becomes ceil(deg / hub_thr) hub items, (vertex, chunk) pairs of
 * uint32 in hub_pairs (capacity olv_hub_capacity(m) pairs); counts_dev
 * holds 2 uint32 (queue length, hub items; zeroed here).  Replaces the
 * GroupSize lanes per vertex of edgeFilterKernel (oliverKernel.h:63-66). */
int olv_frontier_split(const olv_graph *g, const uint32_t *queue, const uint32_t *count_dev,
                       uint32_t *out_queue, uint32_t *counts_dev, uint32_t *hub_pairs);
/* Hub items the split can produce for a graph of m edges. */
size_t olv_hub_capacity(uint64_t m);
/* Degree-bucket thresholds used by olv_frontier_bin for this graph. */
int olv_graph_buckets(const olv_graph *g, uint32_t *thread_max,
                      uint32_t *wave_max);

/* ---- graph (replaces Oliver::readGraph oliver.h:190-202 and the CSR of
 *      csrGraph.h:54-84) ------------------------------------------------- */
/* row_ptr[n+1] and col[m] (uint32, CSR, out-edges); weights[m] of
 * weight_dtype or NULL.  Pointers are host pointers if OLV_PTR_HOST is set,
 * else device pointers (copied).  Zero-edge graphs are accepted (the
 * reference asserts on them, grd.h:77). */
int olv_graph_create(olv_graph **out, uint32_t n, uint32_t m,
                     const uint32_t *row_ptr, const uint32_t *col,
                     const void *weights, int weight_dtype, unsigned flags);
int olv_graph_destroy(olv_graph *g);
/* Build the in-edge CSR (needed by bottom-up BFS, batched BFS, PageRank). */
int olv_graph_build_transpose(olv_graph *g);
int olv_graph_counts(const olv_graph *g, uint32_t *n, uint32_t *m);
/* Device pointers of the resident arrays (NULL where absent).  in_eid maps an
 * in-edge slot to its out-edge id. */
int olv_graph_device_arrays(const olv_graph *g, const uint32_t **row_ptr,
                            const uint32_t **col, const void **weights,
                            const uint32_t **in_row_ptr,
                            const uint32_t **in_col, const uint32_t **in_eid);
/* Copy the in-edge CSR to the host (for tests). */
int olv_graph_transpose_to_host(const olv_graph *g, uint32_t *in_row_ptr,
                                uint32_t *in_col, uint32_t *in_eid);

/* ---- CSR ingest on the device (replaces CsrGraph::fromEdgeListFile
 *      csrGraph.h:106-171, CsrGraph::fromDimacsFile csrGraph.h:177-236 and
 *      the source-order clustering of edgeTuple.h:42-67) ------------------ */
/* Edge list in any order -> graph.  The edges are stable-sorted by source
 * on the device (edges of one source keep their input order); weights
 * (weight_dtype, or OLV_W_NONE) follow their edges.  Ids must be < n.
 * OLV_PTR_HOST: the arrays are host pointers; OLV_BUILD_TRANSPOSE: also
 * build the in-edge CSR on the device (in-lists ordered by source). */
int olv_graph_from_coo(olv_graph **out, uint32_t n, uint32_t m,
                       const uint32_t *src, const uint32_t *dst,
                       const void *weights, int weight_dtype, unsigned flags,
                       olv_stats *st);
/* Edge-list text -> graph, parsed on the device: '#' lines skipped, the
 * first other non-empty line is "n m", then exactly m "src dst [w]" lines.
 * Unlike csrGraph.h:106-171 the tuples need not be sorted (they are
 * stable-sorted by source) and a third column on every edge line is kept
 * as int32 weights (OLV_W_I32).  Errors name the 1-based line.
 * st (optional): elapsed_ms = device time of the whole ingest, iterations =
 * text lines; with OLV_PROFILE, kernel_ms = the line-parsing kernel. */
int olv_graph_from_edge_list_text(olv_graph **out, const char *text,
                                  size_t len, unsigned flags, olv_stats *st);
/* DIMACS adjacency text -> graph, parsed on the device: header "n m", then
 * line i after it lists the 1-based neighbours of vertex i-1; exactly 2m
 * neighbour entries (csrGraph.h:177-236). */
int olv_graph_from_dimacs_text(olv_graph **out, const char *text, size_t len,
                               unsigned flags, olv_stats *st);
/* Read a file (OLV_FMT_EDGE_LIST or OLV_FMT_DIMACS) into pinned host memory
 * and ingest it as above. */
int olv_graph_load_file(olv_graph **out, const char *path, int format,
                        unsigned flags, olv_stats *st);
/* Copy the out-CSR to the host (weights if present; weight_dtype gets the
 * graph's weight dtype).  Any pointer may be NULL. */
int olv_graph_csr_to_host(const olv_graph *g, uint32_t *row_ptr,
                          uint32_t *col, void *weights, int *weight_dtype);

/* ---- result egress (replaces Oliver::printVertices / printEdges
 *      oliver.h:204-212 over GRD::persist grd.h:118-124 and the per-element
 *      fprintf of GRD::print grd.h:130-134) ----------------------------- */
/* One line per value in the programs' text formats (BFS.cu:41-43 "%d",
 * PageRank.cu:42-44 "%f"), formatted on the device and streamed to `path`.
 * values[i] sits at byte i*stride (stride 0 = the value size), so one field
 * of a vertex struct can be written.  values is a device pointer unless
 * OLV_PTR_HOST.  Doubles of magnitude >= 2^63 give OLV_EUNSUPPORTED.
 * st (optional): edges_inspected = text bytes written; with OLV_PROFILE,
 * kernel_ms / kernel_launches = the formatting (write-pass) kernel. */
int olv_write_text(const char *path, const void *values, size_t count,
                   size_t stride, int fmt, unsigned flags, olv_stats *st);
/* Binary dump: 24-byte header ("OLVR", version 1, dtype tag, element size,
 * count; little-endian) then the values, copied through two pinned buffers
 * so device-to-host copies overlap the file writes. */
int olv_write_binary(const char *path, const void *values, size_t count,
                     size_t elem_bytes, int dtype, unsigned flags);
/* Read a dump written by olv_write_binary.  values == NULL reads only the
 * header; else the values go to `values` (device unless OLV_PTR_HOST),
 * OLV_ERANGE if they exceed capacity_bytes. */
int olv_read_binary(const char *path, void *values, size_t capacity_bytes,
                    uint64_t *count, uint32_t *elem_bytes, int *dtype,
                    unsigned flags);

/* ---- algorithms -------------------------------------------------------- */
/* BFS levels from src, INF = 0x7fffffff for unreached.  Replaces the BFS.cu
 * level loop (BFS.cu:124-148) over edgeFilter/vertexFilter
 * (oliver.h:56-113, oliverKernel.h:45-141) run to convergence.  The first
 * direction-optimising call on a graph also builds its in-degree renumbered
 * twin (a second copy of the graph and its transpose, kept until
 * olv_graph_destroy); later calls run on the twin. */
int olv_bfs(const olv_graph *g, uint32_t src, int32_t *levels, unsigned flags,
            olv_stats *st);

/* olv_bfs without any allocation in the call: no reference counterpart; for
 * callers that cannot allocate device memory (tight serving loops, code
 * that must not call hipMalloc).  Same levels as olv_bfs, computed in the
 * graph's own numbering (as OLV_GRAPH_ORDER).  Every work buffer lives in one
 * static 32 MiB device array per GPU (held until the process ends), which
 * fits graphs up to about 5.5 M vertices; larger graphs fail with
 * OLV_ERANGE.  `levels` must be a device pointer (OLV_PTR_HOST and
 * OLV_PERSISTENT are rejected).  Direction-optimising runs need the graph's
 * in-edge CSR with its in-list heads (OLV_BUILD_TRANSPOSE at creation, or one
 * olv_bfs call); otherwise pass OLV_NO_DIRECTION_OPT.  st->elapsed_ms is
 * host wall time.  All calls on one GPU share the array: calls that overlap
 * on one GPU (two threads, two streams) race; each call initialises what it
 * reads, so sequential calls need no clearing in between. */
int olv_bfs_noalloc(const olv_graph *g, uint32_t src, int32_t *levels,
                    unsigned flags, olv_stats *st);

/* SSSP distances over int32 weights (graph weights, or `weights` if not
 * NULL), INF = 0x7fffffff.  Frontier relaxation as SSSP.cu:146-163 with
 * cond(v, acc[v]) (SSSP.cu:72-75), run to the same fixed point in near-far
 * order; with non-negative weights each round's candidates are binned by
 * destination partition and min-reduced in LDS (OLV_SSSP_DIRECT: per-edge
 * probe and atomics instead).  The first binned call on a graph keeps its
 * bins (k x m pairs of 8 bytes) on the graph until olv_graph_destroy. */
int olv_sssp(const olv_graph *g, uint32_t src, const int32_t *weights,
             int32_t *dist, unsigned flags, olv_stats *st);

/* Unweighted PageRank in fp32 with fp32 accumulation, testPageRank.cu:57-104
 * semantics: r0 = 1/n, r <- d*sum_{u->v} r_u/outdeg(u) + (1-d)/n, no dangling
 * redistribution, stop after max_iters iterations or when the L1 delta drops
 * below epsilon.  Replaces PageRank.cu:122-133 (edgeMap/vertexMap/
 * vertexReduce, oliver.h:121-183). */
int olv_pagerank_f32(const olv_graph *g, float damping, uint32_t max_iters,
                     double epsilon, float *ranks, unsigned flags,
                     olv_stats *st);

/* Weighted, teleported PageRank fixed point (no reference counterpart;
 * reduces to olv_pagerank_f32 when w = 1/outdeg, t = 1/n):
 *   r = d * P^T r + (1-d) t,  P[u][v] = w_uv (w indexed like col),
 * iterated from r0 = t until ||r_k+1 - r_k||_1 < tol. */
int olv_pagerank_weighted_f32(const olv_graph *g, const float *w,
                              const float *t, float damping, float tol,
                              uint32_t max_iters, float *ranks, unsigned flags,
                              olv_stats *st);
/* Jacobian-vector product (forward mode) of the same fixed point: for a
 * tangent (dw, dt) at the solution r,
 *   dr = d * P^T dr + d * dP^T r + (1-d) dt,   (dP)[u][v] = dw_uv,
 * solved like the forward with teleport s/(1-d), s = d dP^T r + (1-d) dt.
 * <gbar, dr> equals <grad_w, dw> + <grad_t, dt> of olv_pagerank_vjp_f32
 * (the dot-product test).  No reference counterpart; damping in [0, 1). */
int olv_pagerank_jvp_f32(const olv_graph *g, const float *w, const float *r,
                         const float *dw, const float *dt, float damping,
                         float tol, uint32_t max_iters, float *dr,
                         unsigned flags, olv_stats *st);
/* Vector-Jacobian product of the fixed point above for L = <gbar, r>:
 *   lambda = gbar + d * P lambda          (adjoint solve on the out-CSR)
 *   grad_t = (1-d) lambda,   grad_w[e=(u,v)] = d * r_u * lambda_v.
 * Stops when ||lambda_k+1 - lambda_k||_1 < tol * ||gbar||_1.
 * grad_w / grad_t / lambda_out may be NULL. */
int olv_pagerank_vjp_f32(const olv_graph *g, const float *w,
                         const float *ranks, const float *gbar, float damping,
                         float tol, uint32_t max_iters, float *grad_w,
                         float *grad_t, float *lambda_out, unsigned flags,
                         olv_stats *st);

/* Source lists (`srcs`) of the multi-source entry points below are always
 * host arrays; OLV_PTR_HOST refers to the level / distance outputs. */

/* BFS from a set of sources: levels[v] = hops to the nearest source.  The
 * frontier of level 0 is the whole set (one persistent launch). */
int olv_bfs_multisource(const olv_graph *g, const uint32_t *srcs, uint32_t k,
                        int32_t *levels, unsigned flags, olv_stats *st);
/* k-source BFS in one bit-parallel sweep per group of 256 sources (no
 * reference counterpart; equal to k runs of olv_bfs).  levels is k x n, row s
 * = levels from srcs[s]. */
int olv_bfs_batched(const olv_graph *g, const uint32_t *srcs, uint32_t k,
                    int32_t *levels, unsigned flags, olv_stats *st);
/* k-source SSSP as k instances of the graph (virtual vertex s*n + v)
 * relaxed together in every round of the level pipeline (equal to k runs of
 * olv_sssp); dist is k x n. */
int olv_sssp_batched(const olv_graph *g, const uint32_t *srcs, uint32_t k,
                     const int32_t *weights, int32_t *dist, unsigned flags,
                     olv_stats *st);
/* Block-diagonal union of k graphs as a new handle; vertex_offsets (host,
 * k+1 entries, may be NULL) receives where each graph's ids start. */
int olv_graph_union(olv_graph *const *gs, uint32_t k, olv_graph **out,
                    uint32_t *vertex_offsets);
/* BFS on k independent graphs in one launch: graph i from srcs[i] into
 * levels[i] (n_i entries each; levels is a host array of k pointers). */
int olv_bfs_multigraph(olv_graph *const *gs, uint32_t k, const uint32_t *srcs,
                       int32_t *const *levels, unsigned flags, olv_stats *st);

/* ---- hipGraph plans ---------------------------------------------------- */
/* Capture a whole traversal (state reset, seeding, persistent kernel) once;
 * olv_plan_launch replays it on the current stream.  `levels` (device) is
 * bound at capture time.  A plan shares the graph's work buffers: do not
 * run two plans of one graph concurrently.  A plan runs with the graph's
 * tuning at capture time; a graph never tuned first takes its signature's
 * entry in the tuning cache named by the OLV_TUNE_CACHE environment
 * variable, when there is one (olv_autotune_ex). */
typedef struct olv_plan olv_plan;
int olv_plan_bfs(const olv_graph *g, const uint32_t *srcs, uint32_t k,
                 int32_t *levels, unsigned flags, olv_plan **out);
int olv_plan_bfs_batched(const olv_graph *g, const uint32_t *srcs, uint32_t k,
                         int32_t *levels, unsigned flags, olv_plan **out);
/* st (optional) waits for the replay and reports levels and time. */
int olv_plan_launch(olv_plan *p, olv_stats *st);
int olv_plan_destroy(olv_plan *p);

/* ---- autotuner --------------------------------------------------------- *
 * No reference counterpart: Olive's programs hand-pick GroupSize per run
 * (BFS.cu:127-135); here the launch shapes and thresholds of every kernel
 * family are run-time settings of the graph handle, searched on the graph
 * itself and kept in a cache keyed by the graph's signature. */
typedef struct olv_tuning {
  uint32_t bfs_expand_per_cu;       /* BFS expansion workgroups per CU (graph numbering) */
  uint32_t bfs_expand_per_cu_twin;  /* the same on the renumbered twin                   */
  uint32_t bfs_compact_per_cu;      /* BFS/SSSP compaction workgroups per CU             */
  uint32_t td_wave_thr;             /* push: out-degree >= this: a wave per vertex       */
  uint32_t td_block_thr;            /* push: out-degree >= this: a workgroup per vertex  */
  uint32_t hub_thr;                 /* push: out-degree >= this: items of hub_thr edges
                                       (>= 256)                                          */
  float alpha, beta;                /* BFS direction switch (Beamer)                     */
  uint32_t mbfs_pull_per_cu;        /* batched BFS bottom-up workgroups per CU           */
  uint32_t seg_grid_per_cu;         /* row pulls (weighted PageRank, adjoint, per-row
                                       PageRank): workgroups per CU                      */
  uint32_t seg_unroll;              /* row pulls: gathers in flight per lane (4, 8, 16)  */
  uint32_t seg_bounds[3];           /* row pulls: 1 lane per row up to seg_bounds[0]
                                       entries, 8 lanes up to [1], a wave up to [2],
                                       then a workgroup                                  */
  uint32_t sssp_relax_per_cu;       /* SSSP direct relaxation (OLV_SSSP_DIRECT)
                                       workgroups per CU                                 */
  uint32_t sssp_delta_num;          /* SSSP near-far step = mean weight x this / mean
                                       out-degree (1..1024; distances do not change)    */
  uint32_t source;                  /* 0 defaults, 1 measured by olv_autotune_ex, 2 read
                                       from a tuning cache, 3 set by the caller          */
  double tune_ms;                   /* host time olv_autotune_ex spent (0 on a hit)      */
} olv_tuning;

#define OLV_TUNE_BFS     0x1u  /* plain BFS: grids, bucket thresholds, hubs, alpha/beta */
#define OLV_TUNE_BATCHED 0x2u  /* 256-source BFS: bottom-up grid                         */
#define OLV_TUNE_PULL    0x4u  /* row pulls: grid, unroll, bucket bounds                 */
#define OLV_TUNE_SSSP    0x8u  /* SSSP near-far step, direct relaxation grid (int32 w)   */
#define OLV_TUNE_ALL     0xFu

int olv_get_tuning_ex(const olv_graph *g, olv_tuning *out);
/* Validates every field (EINVAL names the first bad one). */
int olv_set_tuning_ex(olv_graph *g, const olv_tuning *t);
/* Search the settings of the kernel families in `what` on this graph
 * (coordinate descent over candidate values, timing runs of the engine
 * itself) and keep the winner in the handle.  With cache_path (a JSON Lines
 * file), a graph whose signature is already in the cache takes the cached
 * settings without measuring, and a measured result is written back. */
int olv_autotune_ex(olv_graph *g, unsigned what, const char *cache_path, olv_tuning *out);
/* Signature of a graph: vertex and edge counts and a 64-bit hash of its
 * log2 out- and in-degree histograms, as "v<n>-e<m>-h<hex>" (key of the
 * tuning cache).  olv_csr_signature computes it from a host CSR. */
int olv_graph_signature(const olv_graph *g, char *key, size_t len);
int olv_csr_signature(uint32_t n, uint32_t m, const uint32_t *row_ptr, const uint32_t *col, char *key, size_t len);
/* The tuning cache (host only).  load: *found = 1 and *out filled when the
 * key is present; store: adds or replaces the key's line. */
int olv_tuning_cache_load(const char *path, const char *key, olv_tuning *out, int *found);
int olv_tuning_cache_store(const char *path, const char *key, const olv_tuning *t);
/* Plain-BFS tuning only, without a cache (olv_autotune_ex with OLV_TUNE_BFS);
 * returns the bucket thresholds in the form olv_set_buckets takes. */
int olv_autotune(olv_graph *g, uint32_t *thread_max, uint32_t *wave_max,
                 uint32_t *block_threads);
int olv_set_buckets(olv_graph *g, uint32_t thread_max, uint32_t wave_max,
                    uint32_t block_threads);
/* Direction switch (Beamer): top-down -> bottom-up when the next frontier's
 * out-edges exceed unexplored edges / alpha; back when the frontier has
 * fewer than n / beta vertices. */
int olv_set_direction(olv_graph *g, float alpha, float beta);
int olv_get_tuning(const olv_graph *g, uint32_t *wave_thr, uint32_t *block_thr,
                   float *alpha, float *beta);

/* ---- seeded graph generators (host memory, deterministic for a seed) ---- */
typedef struct olv_host_csr {
  uint32_t n;
  uint32_t m;
  uint32_t *row_ptr;  /* n+1 */
  uint32_t *col;      /* m, ascending within each row, duplicates kept */
  void *weights;      /* m or NULL */
  int weight_dtype;
} olv_host_csr;

/* R x C directed grid: id = C*row + col, edges to the 4 neighbours. */
int olv_gen_grid(uint32_t rows, uint32_t cols, olv_host_csr *out);
/* RMAT(scale, edge_factor, a, b, c) without noise or permutation; drops
 * self-loops when drop_self_loops != 0; duplicates kept. */
int olv_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b,
                 double c, uint64_t seed, int drop_self_loops,
                 olv_host_csr *out);
/* n vertices, m edges, endpoints uniform (self-loops kept). */
int olv_gen_uniform(uint32_t n, uint32_t m, uint64_t seed, olv_host_csr *out);
/* Attach int32 weights uniform in [lo, hi] (by CSR position). */
int olv_gen_weights_i32(olv_host_csr *g, int32_t lo, int32_t hi,
                        uint64_t seed);
/* Attach fp32 weights uniform in (0, 1], normalised to sum 1 per source. */
int olv_gen_weights_f32_normalized(olv_host_csr *g, uint64_t seed);
/* Flat-Dirichlet vector of length n (normalised Exp(1) draws). */
int olv_gen_dirichlet(uint32_t n, uint64_t seed, float *out);
/* Scale of graph i of a batch, uniform in [lo, hi]. */
uint32_t olv_gen_batch_scale(uint64_t seed, uint32_t i, uint32_t lo,
                             uint32_t hi);
uint64_t olv_gen_batch_seed(uint64_t seed, uint32_t i);
void olv_host_csr_free(olv_host_csr *g);

#ifdef __cplusplus
}
#endif
#endif /* OLIVE_HIP_H */
