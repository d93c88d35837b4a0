// Frontier steps shared by the BFS, SSSP and batched drivers.  Each step is
// a __device__ function over a virtual block range [vb, vb + nvb) so the same
// code runs in the per-level kernels or inside one persistent kernel.
//
// Reference semantics (BFS.cu:38-79 over oliver.h:56-113): a vertex first
// reached while expanding level L gets level L+1; INF = 0x7fffffff.  Levels
// are unique, so push (top-down) and pull (bottom-up) steps may be mixed per
// level without changing any output.
//
// A frontier handed to a push step has two parts: a queue of vertices with
// out-degree below hub_thr, and a list of hub items (vertex, chunk) that cut
// each higher-degree out-list into hub_thr-edge pieces, one workgroup per
// piece -- so no single workgroup walks a 10^5-edge list alone.
#pragma once
#include "device.hpp"
#include "internal.hpp"

namespace olv {

// Level counters.  Every workgroup adds its totals with device-scope
// atomics, which on MI355X are performed at the memory side and serialise
// per address (≈12 ns each, measured: 4K same-address atomics cost a
// near-empty level ~48 us).  So a counter set is kCtrSlots copies, one per
// 128-byte line, and workgroup b adds to copy b % kCtrSlots.
constexpr int kCtrSlots = 32;

struct alignas(128) BfsCtr {
  unsigned long long mf;         // sum of out-degrees of vertices claimed
  unsigned long long inspected;  // edges read
  uint32_t nf;                   // vertices claimed
  uint32_t pad[27];
};

struct CtrSet {
  BfsCtr s[kCtrSlots];
};

struct CtrSum {
  unsigned long long mf, inspected;
  uint32_t nf;
};

// Zero a counter set (all threads of one workgroup).
__device__ __forceinline__ void ctr_clear(CtrSet *c) {
  uint32_t *w = reinterpret_cast<uint32_t *>(c);
  for (uint32_t i = threadIdx.x; i < sizeof(CtrSet) / 4; i += blockDim.x) w[i] = 0u;
}

// Totals of a counter set, for every thread of the calling workgroup (the
// first wave reads one slot per lane and reduces; all threads must call).
__device__ __forceinline__ CtrSum ctr_total(const CtrSet *c) {
  __shared__ CtrSum s_tot;
  if (threadIdx.x < kWave) {
    unsigned long long mf = 0, in = 0, nf = 0;
    if (threadIdx.x < kCtrSlots) {
      const BfsCtr &x = c->s[threadIdx.x];
      mf = x.mf;
      in = x.inspected;
      nf = x.nf;
    }
    mf = wave_sum(mf);
    in = wave_sum(in);
    nf = wave_sum(nf);
    if (threadIdx.x == 0) s_tot = CtrSum{mf, in, (uint32_t)nf};
  }
  __syncthreads();
  const CtrSum r = s_tot;
  __syncthreads();
  return r;
}

struct BfsArgs {
  const uint32_t *row_ptr;
  const uint32_t *col;
  const uint32_t *in_row_ptr;
  const uint32_t *in_col;
  const uint2 *in_head;  // first two in-neighbours (0xffffffff where absent)
  int32_t *levels;
  const uint32_t *omap;  // levels index of each vertex (renumbered twin), or nullptr: the vertex id
  uint8_t *visited;    // one byte per vertex: nonzero = claimed (padded to 32-vertex words)
  uint32_t *snap;      // bitmap of the claims as of the last frontier extraction
  uint32_t *queue;
  uint2 *hubs;
  uint32_t n;
  uint32_t wave_thr;   // degree >= wave_thr: one wave per vertex
  uint32_t block_thr;  // degree >= block_thr: whole workgroup per vertex
  uint32_t hub_thr;    // degree >= hub_thr: hub items of hub_thr edges
  // Mapped: stores go through omap (a template flag, not a run-time test:
  // the test in the claim path slowed small-level traversals by a quarter)
  template <bool Mapped>
  __device__ __forceinline__ void set_level(uint32_t v, int32_t l) const { levels[Mapped ? omap[v] : v] = l; }
};

// Hub items needed for the worst case of m edges (items of at least
// kHubThrMin edges).
inline size_t hub_capacity(uint64_t m) { return 2 * (m / kHubThrMin) + 64; }

// Visit edges e, e + stride, ... < end with U edges in flight per lane: the
// U target loads are issued together, then the U state loads, then the
// updates -- a lane's chain of dependent memory latencies is cut U-fold.
// Visitor: `Tgt tgt(e, payload)` (reads col[e] and other per-edge data),
// `S state(const Tgt &)` (reads the target's state), `void apply(const Tgt &,
// S, payload)` (claims / updates the target).
template <int U, class Visitor, class P>
__device__ __forceinline__ void visit_range(Visitor &vis, uint32_t e, uint32_t end, uint32_t stride, const P &pl) {
  using T = decltype(vis.tgt(0u, pl));
  using S = decltype(vis.state(vis.tgt(0u, pl)));
  for (uint64_t b = e; b < end; b += (uint64_t)U * stride) {
    T t[U];
    S s[U];
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (b + (uint64_t)k * stride < end) t[k] = vis.tgt((uint32_t)(b + (uint64_t)k * stride), pl);
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (b + (uint64_t)k * stride < end) s[k] = vis.state(t[k]);
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (b + (uint64_t)k * stride < end) vis.apply(t[k], s[k], pl);
  }
}

#ifndef OLV_VISIT_UNROLL
#define OLV_VISIT_UNROLL 4
#endif
constexpr int kVisitUnroll = OLV_VISIT_UNROLL;

// Workload-balanced push expansion of a frontier (queue[0, qsize) plus hub
// items hubs[0, nhubs)).  Each queued out-list is read by the whole
// workgroup (degree >= block_thr), by one wave (>= wave_thr), or -- for the
// remaining short lists -- by a block-wide scan that deals the edges of many
// vertices to consecutive lanes, so a wave always reads 64 consecutive
// column entries.
//
// Visitor: `uint32_t row(uint32_t q)` maps a frontier entry to its CSR row,
// `P load(uint32_t q)` gives the per-source payload read once per frontier
// entry; edges are handled through visit_range (tgt / state / apply).
template <int BLOCK, class Visitor>
__device__ void push_expand(const uint32_t *queue, uint32_t qsize, const uint2 *hubs, uint32_t nhubs,
                            const uint32_t *row_ptr, uint32_t wave_thr, uint32_t block_thr, uint32_t hub_chunk,
                            uint32_t vb, uint32_t nvb, Visitor &vis, unsigned long long &insp) {
  using P = decltype(vis.load(0u));
  __shared__ uint32_t s_owner, s_beg, s_end;
  __shared__ P s_own_payload;
  __shared__ uint32_t s_scan[BLOCK / kWave + 1];
  constexpr int U = kVisitUnroll;
  __shared__ uint32_t s_edges[BLOCK * U];
  __shared__ P s_payload[BLOCK * U];
  const uint32_t lane = lane_id();
  // hub pieces: one workgroup per item
  for (uint32_t h = vb; h < nhubs; h += nvb) {
    const uint2 it = hubs[h];
    const uint32_t u = vis.row(it.x);
    const uint32_t beg = row_ptr[u] + it.y * hub_chunk;
    const uint32_t e1 = row_ptr[u + 1];
    const uint32_t end = e1 < beg + hub_chunk ? e1 : beg + hub_chunk;
    const P pl = vis.load(it.x);
    if (threadIdx.x == 0) insp += end - beg;
    visit_range<U>(vis, beg + threadIdx.x, end, BLOCK, pl);
  }
  // Queue entries per workgroup and pass: a full workgroup's worth when the
  // queue fills the grid, else an even share, so that a short queue of
  // high-degree vertices is not left to the first workgroup.
  const uint32_t per = qsize >= nvb * BLOCK ? BLOCK : (qsize + nvb - 1) / nvb;
  for (uint32_t base = vb * per; base < qsize; base += nvb * per) {
    const uint32_t i = base + threadIdx.x;
    uint32_t beg = 0, end = 0;
    P pl{};
    if (threadIdx.x < per && i < qsize) {
      const uint32_t q = queue[i];
      const uint32_t u = vis.row(q);
      beg = row_ptr[u];
      end = row_ptr[u + 1];
      pl = vis.load(q);
    }
    insp += end - beg;
    // workgroup per vertex
    for (;;) {
      if (threadIdx.x == 0) s_owner = 0xffffffffu;
      __syncthreads();
      if (end - beg >= block_thr) s_owner = threadIdx.x;
      __syncthreads();
      const uint32_t owner = s_owner;
      if (owner == 0xffffffffu) break;
      if (threadIdx.x == owner) {
        s_beg = beg;
        s_end = end;
        s_own_payload = pl;
        beg = end;
      }
      __syncthreads();
      const uint32_t eb = s_beg, ee = s_end;
      const P opl = s_own_payload;
      visit_range<U>(vis, eb + threadIdx.x, ee, BLOCK, opl);
      __syncthreads();
    }
    // wave per vertex
    for (;;) {
      const uint64_t mask = ballot(end - beg >= wave_thr);
      if (!mask) break;
      const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
      const uint32_t eb = __shfl(beg, leader, kWave), ee = __shfl(end, leader, kWave);
      const P lpl = __shfl(pl, leader, kWave);
      if (lane == leader) beg = end;
      visit_range<U>(vis, eb + lane, ee, kWave, lpl);
    }
    // short lists: scan-based edge dealing
    const uint32_t rem = end - beg;
    uint32_t total;
    const uint32_t off = block_exclusive_scan<BLOCK>(rem, s_scan, &total);
    for (uint32_t done = 0; done < total; done += BLOCK * U) {
      const uint32_t lo = off > done ? off : done;
      const uint32_t hi = (off + rem) < (done + BLOCK * U) ? (off + rem) : (done + BLOCK * U);
      for (uint32_t k = lo; k < hi; ++k) {
        s_edges[k - done] = beg + (k - off);
        s_payload[k - done] = pl;
      }
      __syncthreads();
      const uint32_t cnt = total - done < BLOCK * U ? total - done : BLOCK * U;
      using T = decltype(vis.tgt(0u, pl));
      using S = decltype(vis.state(vis.tgt(0u, pl)));
      T t[U];
      S sv[U];
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (threadIdx.x + k * BLOCK < cnt) t[k] = vis.tgt(s_edges[threadIdx.x + k * BLOCK], s_payload[threadIdx.x + k * BLOCK]);
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (threadIdx.x + k * BLOCK < cnt) sv[k] = vis.state(t[k]);
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (threadIdx.x + k * BLOCK < cnt) vis.apply(t[k], sv[k], s_payload[threadIdx.x + k * BLOCK]);
      __syncthreads();
    }
  }
}

// BFS claim: a visitor that finds the vertex unvisited stores its visited
// byte and its level -- plain byte stores, no atomics (32 neighbouring
// vertices sharing a bitmap word made the claims serialise on the hot words
// of an RMAT graph).  Two visitors may both find the byte clear; both store
// the same values, and the level counters then count the vertex twice --
// they only steer the direction choice.  The next frontier is not recorded
// here: it is the set of claims new since the last extraction
// (frontier_words).
template <bool Mapped>
struct BfsVisitor {
  const BfsArgs &a;
  int32_t nlevel;
  uint32_t nf = 0;
  unsigned long long mf = 0;
  __device__ BfsVisitor(const BfsArgs &aa, int32_t l) : a(aa), nlevel(l) {}
  __device__ __forceinline__ uint32_t row(uint32_t q) const { return q; }
  __device__ __forceinline__ uint32_t load(uint32_t) const { return 0u; }
  __device__ __forceinline__ uint32_t tgt(uint32_t e, uint32_t) const { return a.col[e]; }
  __device__ __forceinline__ uint32_t state(uint32_t v) const { return a.visited[v]; }
  __device__ __forceinline__ void apply(uint32_t v, uint32_t seen, uint32_t) {
    if (seen) return;
    a.visited[v] = 1;
    a.template set_level<Mapped>(v, nlevel);
    nf += 1;
    mf += a.row_ptr[v + 1] - a.row_ptr[v];
  }
};

template <int BLOCK, bool Mapped>
__device__ void td_step(const BfsArgs &a, uint32_t qsize, uint32_t nhubs, int32_t nlevel, uint32_t vb, uint32_t nvb,
                        uint32_t &nf, unsigned long long &mf, unsigned long long &insp) {
  BfsVisitor<Mapped> vis(a, nlevel);
  push_expand<BLOCK>(a.queue, qsize, a.hubs, nhubs, a.row_ptr, a.wave_thr, a.block_thr, a.hub_thr, vb, nvb, vis,
                     insp);
  nf += vis.nf;
  mf += vis.mf;
}

// Bottom-up step: every unvisited vertex scans its in-list for a parent in
// the current frontier bitmap and stops at the first one.  Each lane carries kBuBatch vertices (grid-stride
// slots) at once: their state, row and in-list loads are issued together, so
// the chain of dependent memory latencies is paid once per kBuBatch slots.
// A lane scans at most kBuLaneSteps batches of kVisitUnroll in-edges per
// vertex on its own; in-lists left unfinished are then scanned by the whole
// wave, 64 x kVisitUnroll edges per step, one vertex at a time.
#ifndef OLV_BU_LANE_STEPS
#define OLV_BU_LANE_STEPS 4
#endif
#ifndef OLV_BU_BATCH
#define OLV_BU_BATCH 1
#endif
constexpr int kBuLaneSteps = OLV_BU_LANE_STEPS;
constexpr int kBuBatch = OLV_BU_BATCH;

template <int BLOCK, bool Mapped>
__device__ void bu_step(const BfsArgs &a, const uint32_t *front_bm, int32_t nlevel, uint32_t vb, uint32_t nvb,
                        uint32_t &nf, unsigned long long &mf, unsigned long long &insp) {
  constexpr int U = kVisitUnroll;
  constexpr int J = kBuBatch;
  const uint32_t lane = lane_id();
  const uint64_t stride = (uint64_t)nvb * (BLOCK / kWave) * kWave;
  const uint32_t gw = vb * (BLOCK / kWave) + threadIdx.x / kWave;
  for (uint64_t base = (uint64_t)gw * kWave; base < a.n; base += J * stride) {
    uint32_t e[J], ee[J];
    bool todo[J], found[J];
    bool any_todo = false;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const uint64_t v = base + j * stride + lane;
      const bool inb = v < a.n;
      todo[j] = inb && a.visited[v] == 0;
      any_todo |= todo[j];
      found[j] = false;
      e[j] = ee[j] = 0;
    }
    if (!ballot(any_todo)) continue;
    // the first two in-neighbours, one coalesced load per vertex (without
    // a head array every in-list is scanned from its first entry)
    uint2 hd[J];
    const bool heads = a.in_head != nullptr;
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const uint32_t v = (uint32_t)(base + j * stride) + lane;
      hd[j] = (todo[j] && heads) ? a.in_head[v] : make_uint2(0xffffffffu, 0xffffffffu);
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const bool okx = hd[j].x != 0xffffffffu, oky = hd[j].y != 0xffffffffu;
      const bool hit = (okx & bm_test(front_bm, okx ? hd[j].x : 0u)) | (oky & bm_test(front_bm, oky ? hd[j].y : 0u));
      found[j] = todo[j] && hit;
      if (todo[j]) insp += (uint32_t)okx + (uint32_t)oky;
    }
    // the rest of the in-list, from its third entry, for the vertices not done
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (todo[j] && !found[j] && (!heads || hd[j].y != 0xffffffffu)) {
        const uint32_t v = (uint32_t)(base + j * stride) + lane;
        e[j] = a.in_row_ptr[v] + (heads ? 2u : 0u);
        ee[j] = a.in_row_ptr[v + 1];
      }
    // own scans: U in-edges of each of the J vertices per step, all loads
    // and frontier tests issued together
    for (int st = 0; st < kBuLaneSteps; ++st) {
      uint32_t u[J][U];
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const bool act = todo[j] && !found[j] && e[j] < ee[j];
#pragma unroll
        for (int k = 0; k < U; ++k) u[j][k] = act && e[j] + k < ee[j] ? a.in_col[e[j] + k] : 0xffffffffu;
      }
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const bool act = todo[j] && !found[j] && e[j] < ee[j];
        bool hit = false;
#pragma unroll
        for (int k = 0; k < U; ++k) {
          const bool ok = u[j][k] != 0xffffffffu;
          hit |= ok & bm_test(front_bm, ok ? u[j][k] : 0u);  // word 0 stands in for a missing edge
        }
        if (act) {
          found[j] = hit;
          insp += ee[j] - e[j] < (uint32_t)U ? ee[j] - e[j] : (uint32_t)U;
          e[j] = ee[j] - e[j] < (uint32_t)U ? ee[j] : e[j] + U;
        }
      }
    }
    // unfinished in-lists: the whole wave scans them, one vertex at a time
#pragma unroll
    for (int j = 0; j < J; ++j) {
      uint64_t pend = ballot(todo[j] && !found[j] && e[j] < ee[j]);
      while (pend) {
        const uint32_t leader = __ffsll((unsigned long long)pend) - 1;
        pend &= pend - 1;
        const uint32_t lb = __shfl(e[j], leader, kWave), le = __shfl(ee[j], leader, kWave);
        bool any = false;
        for (uint32_t x = lb; x < le && !any; x += kWave * U) {
          bool hit = false;
          uint32_t u[U];
#pragma unroll
          for (int k = 0; k < U; ++k) {
            const uint32_t xe = x + k * kWave + lane;
            u[k] = xe < le ? a.in_col[xe] : 0xffffffffu;
          }
#pragma unroll
          for (int k = 0; k < U; ++k) {
            const bool ok = u[k] != 0xffffffffu;
            hit |= ok & bm_test(front_bm, ok ? u[k] : 0u);
          }
          any = ballot(hit) != 0;
          if (lane == 0) insp += le - x < (uint32_t)(kWave * U) ? le - x : (uint32_t)(kWave * U);
        }
        if (lane == leader) found[j] = any;
      }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
      const uint32_t v = (uint32_t)(base + j * stride) + lane;
      if (found[j]) {
        a.visited[v] = 1;
        a.template set_level<Mapped>(v, nlevel);
        nf += 1;
        mf += a.row_ptr[v + 1] - a.row_ptr[v];
      }
    }
  }
}

// Bitmap -> frontier.  One lane per word; bits whose row has out-degree >=
// hub_thr become hub items (vertex, chunk) appended with one atomic per hub,
// the rest are placed in the queue by a block scan of per-lane counts with
// one global atomic per block.  row(v) maps a bit to its CSR row;
// word_of(w) yields bitmap word w (called once per word, by its lane).
template <int BLOCK, class RowFn, class WordFn>
__device__ void compact_words(uint32_t nwords, WordFn word_of, const uint32_t *row_ptr, uint32_t hub_thr,
                              uint32_t *queue, uint32_t *qcount, uint2 *hubs, uint32_t *hcount, uint32_t vb,
                              uint32_t nvb, RowFn row) {
  __shared__ uint32_t s_scan[BLOCK / kWave + 1];
  __shared__ uint32_t s_base;
  for (uint32_t w0 = vb * BLOCK; w0 < nwords; w0 += nvb * BLOCK) {
    const uint32_t w = w0 + threadIdx.x;
    const uint32_t word = w < nwords ? word_of(w) : 0u;
    uint32_t small = 0;
    for (uint32_t x = word; x;) {
      const uint32_t b = __ffs(x) - 1;
      x &= x - 1;
      const uint32_t v = w * 32u + b;
      const uint32_t u = row(v);
      const uint32_t d = row_ptr[u + 1] - row_ptr[u];
      if (d >= hub_thr) {
        const uint32_t k = (d + hub_thr - 1) / hub_thr;
        const uint32_t hb = atomicAdd(hcount, k);
        for (uint32_t j = 0; j < k; ++j) hubs[hb + j] = make_uint2(v, j);
      } else {
        ++small;
      }
    }
    uint32_t total;
    const uint32_t off = block_exclusive_scan<BLOCK>(small, s_scan, &total);
    if (threadIdx.x == 0) s_base = total ? atomicAdd(qcount, total) : 0u;
    __syncthreads();
    uint32_t pos = s_base + off;
    for (uint32_t x = word; x;) {
      const uint32_t b = __ffs(x) - 1;
      x &= x - 1;
      const uint32_t v = w * 32u + b;
      const uint32_t u = row(v);
      if (row_ptr[u + 1] - row_ptr[u] < hub_thr) queue[pos++] = v;
    }
    __syncthreads();
  }
}

template <int BLOCK, class RowFn>
__device__ __forceinline__ void compact_step(const uint32_t *bm, uint32_t nwords, const uint32_t *row_ptr,
                                             uint32_t hub_thr, uint32_t *queue, uint32_t *qcount, uint2 *hubs,
                                             uint32_t *hcount, uint32_t vb, uint32_t nvb, RowFn row) {
  compact_words<BLOCK>(nwords, [bm](uint32_t w) { return bm[w]; }, row_ptr, hub_thr, queue, qcount, hubs, hcount,
                       vb, nvb, row);
}

struct IdRow {
  __device__ __forceinline__ uint32_t operator()(uint32_t v) const { return v; }
};

// The visited bytes and snapshot word of one 32-vertex word, loaded ahead
// of the control-state reads they do not depend on.
struct WordLoad {
  uint4 q0, q1;
  uint32_t snap;
};

__device__ __forceinline__ WordLoad load_word(const BfsArgs &a, uint32_t w) {
  const uint4 *p = reinterpret_cast<const uint4 *>(a.visited + (size_t)w * 32u);  // 32-byte aligned
  return WordLoad{p[0], p[1], a.snap[w]};
}

// Next frontier = visited bytes set since the last extraction: written to
// `front` (for a bottom-up step) and, if `to_queue`, compacted into the
// queue and hub items (for a top-down step); `snap` catches up.  `pre` is
// this lane's first word (index vb * BLOCK + threadIdx.x), already loaded.
template <int BLOCK>
__device__ void frontier_words(const BfsArgs &a, uint32_t *front, uint32_t nw, bool to_queue, uint32_t *qcount,
                               uint32_t *hcount, uint32_t vb, uint32_t nvb, const WordLoad &pre) {
  const uint32_t wpre = vb * BLOCK + threadIdx.x;
  auto word_of = [&](uint32_t w) -> uint32_t {
    const WordLoad x = w == wpre ? pre : load_word(a, w);
    const uint32_t parts[8] = {x.q0.x, x.q0.y, x.q0.z, x.q0.w, x.q1.x, x.q1.y, x.q1.z, x.q1.w};
    uint32_t v = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int b = 0; b < 4; ++b) v |= ((parts[i] >> (8 * b)) & 0xffu ? 1u : 0u) << (4 * i + b);
    const uint32_t nb = v & ~x.snap;
    if (nb) a.snap[w] = x.snap | nb;
    front[w] = nb;
    return nb;
  };
  if (to_queue)
    compact_words<BLOCK>(nw, word_of, a.row_ptr, a.hub_thr, a.queue, qcount, a.hubs, hcount, vb, nvb, IdRow());
  else
    for (uint32_t w = vb * BLOCK + threadIdx.x; w < nw; w += nvb * BLOCK) (void)word_of(w);
}

template <int BLOCK>
__device__ __forceinline__ void flush_counters(CtrSet *cs, uint32_t nf, unsigned long long mf,
                                               unsigned long long insp) {
  __shared__ unsigned long long s_red[BLOCK / kWave];
  const unsigned long long tnf = block_sum<BLOCK, unsigned long long>((unsigned long long)nf, s_red);
  const unsigned long long tmf = block_sum<BLOCK, unsigned long long>(mf, s_red);
  const unsigned long long tin = block_sum<BLOCK, unsigned long long>(insp, s_red);
  if (threadIdx.x == 0) {
    BfsCtr *c = &cs->s[blockIdx.x % kCtrSlots];
    if (tnf) atomicAdd(&c->nf, (uint32_t)tnf);
    if (tmf) atomicAdd(&c->mf, tmf);
    if (tin) atomicAdd(&c->inspected, tin);
  }
}

// Direction choice for the next level (Beamer et al., SC'12): go bottom-up
// when the next frontier's out-edges exceed the unexplored edges / alpha;
// return to top-down when the frontier holds fewer than n / beta vertices.
__host__ __device__ __forceinline__ bool choose_bottom_up(bool cur_bu, bool can_bu, uint64_t nf, uint64_t mf,
                                                          uint64_t mu, uint32_t n, float alpha, float beta) {
  if (!can_bu) return false;
  if (!cur_bu) return (double)mf * alpha > (double)mu;
  return (double)nf * beta >= (double)n;
}

// Seeds -> level-0 frontier (one workgroup): distinct seeds with out-degree
// below hub_thr go to the queue, the others become hub items.  Returns the
// out-degree sum of the distinct seeds in *deg (thread 0).
template <class Claim, class RowFn>
__device__ void seed_frontier(const uint32_t *seeds, uint32_t k, const uint32_t *row_ptr, uint32_t hub_thr,
                              uint32_t *queue, uint32_t *qcount, uint2 *hubs, uint32_t *hcount, Claim claim,
                              RowFn row, unsigned long long *deg) {
  __shared__ unsigned long long s_deg;
  if (threadIdx.x == 0) s_deg = 0;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) {
    const uint32_t q = claim(i);  // frontier entry, or ~0u if not new
    if (q == 0xffffffffu) continue;
    const uint32_t u = row(q);
    const uint32_t d = row_ptr[u + 1] - row_ptr[u];
    atomicAdd(&s_deg, (unsigned long long)d);
    if (d >= hub_thr) {
      const uint32_t kk = (d + hub_thr - 1) / hub_thr;
      const uint32_t hb = atomicAdd(hcount, kk);
      for (uint32_t j = 0; j < kk; ++j) hubs[hb + j] = make_uint2(q, j);
    } else {
      queue[atomicAdd(qcount, 1u)] = q;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *deg = s_deg;
}

}  // namespace olv
