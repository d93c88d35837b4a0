/*
 * oracle.c -- CPU restatement of the reference's algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (olive_amd/, include/)
 * links, imports or calls this file; only tests/, the smoke check and the
 * benchmark's CPU-baseline leg use it, as the checker.
 *
 * Each function names the reference code it restates.  The reference's own
 * checkers (testBFS.cu, testPageRank.cu) include cuda_runtime.h through
 * common.h and cannot be compiled here, so they are restated, not built.
 *
 * Pinning (what checks each function beyond re-typing it) is listed in the
 * header comment of each function and exercised by tests/test_oracle.py.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_INF 0x7fffffff

/*
 * BFS levels, restating testBFS.cu:56-79 (CLRS FIFO BFS; INF = 0x7fffffff,
 * testBFS.cu:56).  Pinned by: closed forms on the reference data files
 * (grid row+col, chain i, star/complete-DAG, isolated, singleton,
 * orderedGraph_9 = [0,1,2,1,2,3,4,3,4]), scipy.sparse.csgraph / networkx
 * shortest-path lengths, and the BFS edge invariants.
 * Returns the number of levels (max finite level + 1).
 */
int orc_bfs(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
            uint32_t src, int32_t *levels) {
  uint32_t *queue;
  uint32_t head = 0, tail = 0;
  int32_t maxlvl = 0;
  for (uint32_t i = 0; i < n; i++) levels[i] = ORC_INF;
  if (src >= n) return -1;
  queue = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  if (!queue) return -2;
  levels[src] = 0;
  queue[tail++] = src;
  while (head < tail) {
    uint32_t v = queue[head++];
    for (uint32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++) {
      uint32_t dst = col[e];
      if (levels[dst] == ORC_INF) {
        levels[dst] = levels[v] + 1;
        if (levels[dst] > maxlvl) maxlvl = levels[dst];
        queue[tail++] = dst;
      }
    }
  }
  free(queue);
  return maxlvl + 1;
}

/*
 * Level-synchronous frontier relaxation, restating the SSSP.cu:146-163 loop
 * over Oliver::edgeFilter (oliver.h:56-83: reset acc to INF, K1
 * oliverKernel.h:67-81 with gather = dist[u] + w, reduce = atomicMin,
 * flag dst) and Oliver::vertexFilter (oliver.h:92-113, K2
 * oliverKernel.h:134-139 with cond(dist, acc[v]) = dist > acc and
 * update dist = acc).  cond receives the accumulator, the only reading under
 * which SSSP.cu computes shortest paths (the reference kernel passes the id,
 * oliverKernel.h:112,135).  Run to convergence; a frontier still non-empty
 * after n rounds reports a negative cycle (SSSP.cu:153-157) as -1.
 * Pinned by: textbook Dijkstra (orc_dijkstra), scipy Dijkstra, the
 * unit-weight reduction to BFS, and the edge/tightness invariants.
 * Returns the number of rounds.
 */
int orc_sssp_frontier(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                      const int32_t *w, uint32_t src, int32_t *dist) {
  int32_t *acc;
  uint8_t *flag;
  uint32_t *q, *nq, qn = 0;
  int rounds = 0;
  for (uint32_t i = 0; i < n; i++) dist[i] = ORC_INF;
  if (src >= n) return -2;
  acc = (int32_t *)malloc(sizeof(int32_t) * n);
  flag = (uint8_t *)calloc(n, 1);
  q = (uint32_t *)malloc(sizeof(uint32_t) * n);
  nq = (uint32_t *)malloc(sizeof(uint32_t) * n);
  if (!acc || !flag || !q || !nq) {
    free(acc); free(flag); free(q); free(nq);
    return -3;
  }
  dist[src] = 0;
  q[qn++] = src;
  while (qn > 0) {
    uint32_t nqn = 0;
    if ((uint32_t)rounds == n) { rounds = -1; break; }
    for (uint32_t i = 0; i < n; i++) acc[i] = ORC_INF;
    for (uint32_t i = 0; i < qn; i++) {
      uint32_t u = q[i];
      for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
        uint32_t v = col[e];
        int32_t cand = (int32_t)((uint32_t)dist[u] + (uint32_t)w[e]);
        if (cand < acc[v]) acc[v] = cand;
        flag[v] = 1;
      }
    }
    for (uint32_t v = 0; v < n; v++) {
      if (!flag[v]) continue;
      flag[v] = 0;
      if (dist[v] > acc[v]) {
        dist[v] = acc[v];
        nq[nqn++] = v;
      }
    }
    { uint32_t *t = q; q = nq; nq = t; }
    qn = nqn;
    rounds++;
  }
  free(acc); free(flag); free(q); free(nq);
  return rounds;
}

/* ---- binary heap Dijkstra (textbook; independent check of the above) ---- */
typedef struct { int64_t d; uint32_t v; } orc_item;

static void heap_push(orc_item *h, uint64_t *hn, orc_item it) {
  uint64_t i = (*hn)++;
  h[i] = it;
  while (i > 0) {
    uint64_t p = (i - 1) / 2;
    if (h[p].d <= h[i].d) break;
    orc_item t = h[p]; h[p] = h[i]; h[i] = t;
    i = p;
  }
}

static orc_item heap_pop(orc_item *h, uint64_t *hn) {
  orc_item top = h[0];
  uint64_t i = 0, n = --(*hn);
  h[0] = h[n];
  for (;;) {
    uint64_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < n && h[l].d < h[m].d) m = l;
    if (r < n && h[r].d < h[m].d) m = r;
    if (m == i) break;
    orc_item t = h[m]; h[m] = h[i]; h[i] = t;
    i = m;
  }
  return top;
}

/* Dijkstra over non-negative int32 weights; INF = 0x7fffffff. */
int orc_dijkstra(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                 const int32_t *w, uint32_t src, int32_t *dist) {
  uint64_t m = row_ptr[n], hn = 0;
  orc_item *h;
  for (uint32_t i = 0; i < n; i++) dist[i] = ORC_INF;
  if (src >= n) return -2;
  h = (orc_item *)malloc(sizeof(orc_item) * (m + 1));
  if (!h) return -3;
  dist[src] = 0;
  heap_push(h, &hn, (orc_item){0, src});
  while (hn > 0) {
    orc_item it = heap_pop(h, &hn);
    if (it.d > dist[it.v]) continue;
    for (uint32_t e = row_ptr[it.v]; e < row_ptr[it.v + 1]; e++) {
      int64_t nd = it.d + w[e];
      uint32_t v = col[e];
      if (nd < dist[v]) {
        dist[v] = (int32_t)nd;
        heap_push(h, &hn, (orc_item){nd, v});
      }
    }
  }
  free(h);
  return 0;
}

/*
 * PageRank in fp64, restating testPageRank.cu:57-104: ranks = deltas = 1/N;
 * each iteration zeroes nghSums, scatters ranks[v]/outdeg(v) along out-edges
 * in source order, new = d*sum + (1-d)/N, err = sum |delta|; stop when
 * err < epsilon or after max_iters iterations (testPageRank `-max`
 * semantics).  No dangling-mass redistribution.
 * Pinned by: the star graph's two-state recurrence, isolated graph
 * (0.15/N after one step), scipy.sparse matrix power iteration, and the mass
 * identity sum r' = d * sum_{outdeg>0} r + (1-d).
 * Returns the number of rank updates performed.
 */
int orc_pagerank(uint32_t n, const uint32_t *row_ptr, const uint32_t *col,
                 int max_iters, double damping, double epsilon, double *ranks,
                 double *err_out) {
  double one_over_n = 1.0 / n, err = 0.0;
  double *sums = (double *)malloc(sizeof(double) * n);
  double *deltas = (double *)malloc(sizeof(double) * n);
  int iterations = 0, updates = 0;
  if (!sums || !deltas) { free(sums); free(deltas); return -3; }
  for (uint32_t i = 0; i < n; i++) { ranks[i] = one_over_n; deltas[i] = one_over_n; }
  while (iterations < max_iters) {
    for (uint32_t i = 0; i < n; i++) sums[i] = 0.0;
    for (uint32_t v = 0; v < n; v++) {
      uint32_t outdeg = row_ptr[v + 1] - row_ptr[v];
      for (uint32_t e = row_ptr[v]; e < row_ptr[v + 1]; e++)
        sums[col[e]] += ranks[v] / outdeg;
    }
    for (uint32_t v = 0; v < n; v++) {
      double nr = damping * sums[v] + (1 - damping) * one_over_n;
      deltas[v] = nr - ranks[v];
      ranks[v] = nr;
    }
    updates++;
    err = 0.0;
    for (uint32_t i = 0; i < n; i++) err += fabs(deltas[i]);
    if (err < epsilon) break;
    iterations++;
  }
  if (err_out) *err_out = err;
  free(sums); free(deltas);
  return updates;
}

/*
 * Weighted, teleported PageRank fixed point in fp64 (no reference
 * counterpart; equals orc_pagerank's update when w = 1/outdeg, t = 1/N):
 *   r_{k+1}[v] = d * sum_{u->v} w_uv r_k[u] + (1-d) t[v],  r_0 = t,
 * until ||r_{k+1} - r_k||_1 < tol.  Pinned by: the dense solve
 * r = (1-d)(I - d P^T)^{-1} t in tests (numpy LAPACK), and the reduction to
 * orc_pagerank.  Returns iterations, or -1 if max_iters was hit.
 */
int orc_pagerank_weighted(uint32_t n, const uint32_t *row_ptr,
                          const uint32_t *col, const double *w,
                          const double *t, double d, double tol, int max_iters,
                          double *r) {
  double *nr = (double *)malloc(sizeof(double) * n);
  int it = 0;
  if (!nr) return -3;
  for (uint32_t i = 0; i < n; i++) r[i] = t[i];
  for (;;) {
    double delta = 0.0;
    if (it == max_iters) { it = -1; break; }
    for (uint32_t i = 0; i < n; i++) nr[i] = 0.0;
    for (uint32_t u = 0; u < n; u++)
      for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++)
        nr[col[e]] += w[e] * r[u];
    for (uint32_t v = 0; v < n; v++) {
      double x = d * nr[v] + (1 - d) * t[v];
      delta += fabs(x - r[v]);
      r[v] = x;
    }
    it++;
    if (delta < tol) break;
  }
  free(nr);
  return it;
}

/*
 * Adjoint of the fixed point above for L = <gbar, r>:
 *   lambda_{k+1}[u] = gbar[u] + d * sum_{u->v} w_uv lambda_k[v],
 * lambda_0 = gbar, until ||lambda_{k+1} - lambda_k||_1 < tol * ||gbar||_1.
 * VJPs (computed by the caller): grad_t = (1-d) lambda,
 * grad_w[(u,v)] = d r_u lambda_v.  Pinned by: dense solve of
 * (I - d P) lambda = gbar, central finite differences of
 * orc_pagerank_weighted, and the dot-product identity.
 */
int orc_pagerank_adjoint(uint32_t n, const uint32_t *row_ptr,
                         const uint32_t *col, const double *w,
                         const double *gbar, double d, double tol,
                         int max_iters, double *lambda) {
  double *nl = (double *)malloc(sizeof(double) * n);
  double gnorm = 0.0;
  int it = 0;
  if (!nl) return -3;
  for (uint32_t i = 0; i < n; i++) { lambda[i] = gbar[i]; gnorm += fabs(gbar[i]); }
  for (;;) {
    double delta = 0.0;
    if (it == max_iters) { it = -1; break; }
    for (uint32_t u = 0; u < n; u++) {
      double s = 0.0;
      for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) s += w[e] * lambda[col[e]];
      nl[u] = gbar[u] + d * s;
    }
    for (uint32_t u = 0; u < n; u++) {
      delta += fabs(nl[u] - lambda[u]);
      lambda[u] = nl[u];
    }
    it++;
    if (delta < tol * gnorm) break;
  }
  free(nl);
  return it;
}

/*
 * Edge-list text -> CSR, restating CsrGraph::fromEdgeListFile
 * (csrGraph.h:106-171): lines whose first non-blank byte is '#' and blank
 * lines are skipped, the first other line is "n m" (csrGraph.h:127), then
 * one "src dst" line per edge read as two integers (csrGraph.h:134-136).
 * Two differences, the loader's documented contract: the tuples are
 * stable-sorted by source (counting sort; the sort csrGraph.h:149 leaves
 * commented out) and a third column present on every edge line is kept as
 * int32 weights (csrGraph.h:139 drops it).  Pinned by: equality with the
 * Python restatement oracle.load_edge_list on the reference's data files
 * and on generated text with comments, CRLF, tabs and signed weights.
 *
 * orc_edge_list_header: *n, *m from the header; returns 0, or -1 if none.
 * orc_edge_list_load: fills row_ptr[n+1], col[m], w[m] (w may be NULL) and
 * *has_w; returns 0, -(line number) for a malformed line, or -2^30 when the
 * number of edge lines differs from m.
 */
static int orc_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

static const char *orc_next_content(const char *p, const char *end, const char **eol, long long *line) {
  while (p < end) {
    const char *e = memchr(p, '\n', (size_t)(end - p));
    if (!e) e = end;
    const char *q = p;
    while (q < e && orc_blank(*q)) q++;
    ++*line;
    if (q < e && *q != '#') {
      *eol = e;
      return q;
    }
    p = e + 1;
  }
  return NULL;
}

/* Integers of [p, e): up to 3 values, returns the token count or -1. */
static int orc_tokens(const char *p, const char *e, long long *v) {
  int nt = 0;
  while (p < e) {
    while (p < e && orc_blank(*p)) p++;
    if (p >= e) break;
    const char *q = p;
    int neg = 0, nd = 0, ok = 1;
    long long x = 0;
    if (*q == '+' || *q == '-') neg = *q++ == '-';
    for (; q < e && !orc_blank(*q); q++, nd++) {
      if (*q < '0' || *q > '9' || nd >= 18) ok = 0;
      else x = x * 10 + (*q - '0');
    }
    if (nt < 3) {
      if (!ok || nd == 0) return -1;
      v[nt] = neg ? -x : x;
    }
    nt++;
    p = q;
  }
  return nt;
}

int orc_edge_list_header(const char *text, size_t len, long long *n, long long *m) {
  const char *eol;
  long long line = 0, v[3];
  const char *p = orc_next_content(text, text + len, &eol, &line);
  if (!p || orc_tokens(p, eol, v) < 2) return -1;
  *n = v[0];
  *m = v[1];
  return 0;
}

int orc_edge_list_load(const char *text, size_t len, uint32_t n, uint32_t m, uint32_t *row_ptr, uint32_t *col,
                       int32_t *w, int *has_w) {
  const char *end = text + len, *eol;
  long long line = 0, v[3];
  const char *p = orc_next_content(text, end, &eol, &line);
  uint32_t *src = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
  uint32_t *dst = (uint32_t *)malloc(sizeof(uint32_t) * (m ? m : 1));
  int32_t *wt = (int32_t *)malloc(sizeof(int32_t) * (m ? m : 1));
  uint64_t e = 0;
  int minw = 1, rc = 0;
  if (!src || !dst || !wt) rc = -3;
  while (rc == 0 && p && (p = orc_next_content(eol + (eol < end), end, &eol, &line))) {
    const int nt = orc_tokens(p, eol, v);
    if (nt < 2 || v[0] < 0 || v[0] >= n || v[1] < 0 || v[1] >= n) {
      rc = (int)-line;
      break;
    }
    if (e >= m) {
      rc = -(1 << 30);
      break;
    }
    src[e] = (uint32_t)v[0];
    dst[e] = (uint32_t)v[1];
    wt[e] = nt >= 3 ? (int32_t)v[2] : 0;
    if (nt < 3) minw = 0;
    e++;
  }
  if (rc == 0 && e != m) rc = -(1 << 30);
  if (rc == 0) {
    /* stable counting sort by source */
    memset(row_ptr, 0, sizeof(uint32_t) * ((size_t)n + 1));
    for (uint64_t i = 0; i < m; i++) row_ptr[src[i] + 1]++;
    for (uint32_t i = 0; i < n; i++) row_ptr[i + 1] += row_ptr[i];
    uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    if (!cur) rc = -3;
    else {
      memcpy(cur, row_ptr, sizeof(uint32_t) * ((size_t)n + 1));
      for (uint64_t i = 0; i < m; i++) {
        const uint32_t k = cur[src[i]]++;
        col[k] = dst[i];
        if (w) w[k] = wt[i];
      }
      free(cur);
    }
    *has_w = m > 0 && minw;
  }
  free(src);
  free(dst);
  free(wt);
  return rc;
}

/*
 * The reference programs' output loop: GRD::print (grd.h:130-134) calling
 * each vertex's print(), fprintf(outputFile, "%d\n", level) (BFS.cu:41-43)
 * or fprintf(outputFile, "%f\n", rank) (PageRank.cu:42-44).  is_double
 * selects "%f" over doubles, else "%d" over int32.  Pinned by: equality
 * with Python's "%d" formatting and with libc snprintf("%f").
 * Returns 0, or -1 if the file cannot be written.
 */
int orc_print_values(const char *path, const void *values, uint64_t count, int is_double) {
  FILE *f = fopen(path, "w");
  if (!f) return -1;
  for (uint64_t i = 0; i < count; i++) {
    if (is_double) fprintf(f, "%f\n", ((const double *)values)[i]);
    else fprintf(f, "%d\n", ((const int32_t *)values)[i]);
  }
  return fclose(f) == 0 ? 0 : -1;
}

/*
 * Checks rows of BFS levels against the two properties that fix them: the
 * levels testBFS.cu:56-79 computes are the unique array with
 *   (a) lvl[src] = 0, every other finite level >= 1, no negative level;
 *   (b) every edge (u, v) with lvl[u] finite has lvl[v] <= lvl[u] + 1;
 *   (c) every reached v != src has an in-edge (u, v) with lvl[u] = lvl[v] - 1.
 * (c) gives a path from src of length lvl[v], so lvl[v] >= dist(v); (b)
 * gives lvl[v] <= dist(v) along a shortest path and makes every vertex with
 * a reached in-neighbour reached; so lvl = dist, INF exactly where no path
 * exists.  Row r of `levels` (n entries) belongs to srcs[r].  Rows are split
 * over `threads` POSIX threads.  Returns -1 when every row passes, otherwise
 * the lowest failing row index; -2 on allocation failure.
 * Pinned by tests/test_oracle.py: rows from orc_bfs pass, and planted
 * mistakes (a level one too high, one too low, a reached vertex set to INF,
 * a second level-0 vertex) each fail.
 */
typedef struct {
  uint32_t n;
  const uint32_t *row_ptr, *col, *srcs;
  const int32_t *levels;
  uint32_t r0, r1;
  int64_t bad;  /* lowest failing row of this slice, or -1 */
  int err;
} orc_verify_job;

static int orc_verify_row(uint32_t n, const uint32_t *row_ptr, const uint32_t *col, uint32_t src,
                          const int32_t *lv, uint8_t *has_parent) {
  if (src >= n || lv[src] != 0) return 0;
  memset(has_parent, 0, n);
  for (uint32_t u = 0; u < n; u++) {
    const int32_t l = lv[u];
    if (l == ORC_INF) continue;
    if (l < 0 || (l == 0 && u != src)) return 0;
    for (uint32_t e = row_ptr[u]; e < row_ptr[u + 1]; e++) {
      const uint32_t v = col[e];
      const int32_t lvv = lv[v];
      if (lvv == ORC_INF || (int64_t)lvv > (int64_t)l + 1) return 0;
      if (lvv == l + 1) has_parent[v] = 1;
    }
  }
  for (uint32_t v = 0; v < n; v++)
    if (v != src && lv[v] != ORC_INF && !has_parent[v]) return 0;
  return 1;
}

static void *orc_verify_worker(void *arg) {
  orc_verify_job *j = (orc_verify_job *)arg;
  uint8_t *hp = (uint8_t *)malloc(j->n ? j->n : 1);
  j->bad = -1;
  j->err = hp == NULL;
  if (!hp) return NULL;
  for (uint32_t r = j->r0; r < j->r1; r++)
    if (!orc_verify_row(j->n, j->row_ptr, j->col, j->srcs[r], j->levels + (size_t)r * j->n, hp)) {
      j->bad = r;
      break;
    }
  free(hp);
  return NULL;
}

int64_t orc_bfs_verify(uint32_t n, const uint32_t *row_ptr, const uint32_t *col, uint32_t rows,
                       const uint32_t *srcs, const int32_t *levels, int threads) {
  if (threads < 1) threads = 1;
  if ((uint32_t)threads > rows) threads = rows ? (int)rows : 1;
  orc_verify_job *jobs = (orc_verify_job *)calloc((size_t)threads, sizeof(orc_verify_job));
  pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  if (!jobs || !tid) {
    free(jobs);
    free(tid);
    return -2;
  }
  for (int t = 0; t < threads; t++) {
    jobs[t] = (orc_verify_job){n, row_ptr, col, srcs, levels, (uint32_t)((uint64_t)rows * t / threads),
                               (uint32_t)((uint64_t)rows * (t + 1) / threads), -1, 0};
    if (pthread_create(&tid[t], NULL, orc_verify_worker, &jobs[t]) != 0) {
      orc_verify_worker(&jobs[t]);
      tid[t] = 0;
    }
  }
  int64_t bad = -1;
  int err = 0;
  for (int t = 0; t < threads; t++) {
    if (tid[t]) pthread_join(tid[t], NULL);
    err |= jobs[t].err;
    if (jobs[t].bad >= 0 && (bad < 0 || jobs[t].bad < bad)) bad = jobs[t].bad;
  }
  free(jobs);
  free(tid);
  return err ? -2 : bad;
}
