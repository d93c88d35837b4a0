// Host construction of the source-segmented pull plan (blocked.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "blocked.hpp"

namespace olv {

namespace {

// Host threads for a pass over `work` items: one per 2^20 items, at most 16
// and at most the machine's.
unsigned host_threads(size_t work) {
  const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
  return (unsigned)std::min<size_t>({(size_t)16, (size_t)hc, work / (1u << 20) + 1});
}

// f(t) for t < T, on T - 1 new host threads and the caller (t = 0); a thread
// that cannot be started runs its share on the caller.
template <class F>
void host_parallel(unsigned T, const F &f) {
  std::vector<std::thread> th;
  th.reserve(T);
  for (unsigned t = 1; t < T; ++t) {
    try {
      th.emplace_back(f, t);
    } catch (...) {
      f(t);
    }
  }
  f(0);
  for (auto &x : th) x.join();
}

template <class T>
int upload_vec(T **dst, const std::vector<T> &src) {
  const size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(dst), bytes));
  if (!src.empty()) OLV_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return OLV_OK;
}

// Lanes and reads per piece by length: one lane with 2, 4 or 8 reads up to
// 2, 4, 8 entries, then 8 reads on 2, 4, ..., 32 lanes up to 16, 32, ...,
// 256 entries, and a wave up to kPieceMax.
constexpr uint32_t kBucketLanes[9] = {1, 1, 1, 2, 4, 8, 16, 32, 64};
constexpr uint32_t kBucketReads[9] = {2, 4, 8, 8, 8, 8, 8, 8, 8};
constexpr int kBuckets = 9;
static_assert(kBuckets <= (1 << kBucketBits), "a unit's bucket fits its descriptor");
static_assert(kPieceMax <= kBucketReads[kBuckets - 1] * kBucketLanes[kBuckets - 1], "the widest bucket covers a piece");
// the first bucket whose lanes and reads cover the piece (the segment pass
// sums every piece in one step): the fewest reads for short pieces, then the
// fewest lanes
inline uint32_t piece_bucket(uint32_t len) {
  uint32_t b = 0;
  while (b + 1 < kBuckets && len > kBucketReads[b] * kBucketLanes[b]) ++b;
  return b;
}

}  // namespace

int build_blocked_host(const std::vector<uint32_t> &rp, const uint32_t *col, bool with_slots, uint32_t groups,
                       uint32_t unit_edges, BlockedHost &h) {
  if (unit_edges < kPieceMax || unit_edges > kUnitEdges || unit_edges % 8)
    return fail(OLV_EINTERNAL, "blocked plan: %u entries per unit", unit_edges);
  const uint32_t n = (uint32_t)(rp.size() - 1);
  const uint32_t m = rp[n];
  h.n = n;
  h.m = m;
  uint32_t n_live = n;
  while (n_live > 0 && rp[n_live] == rp[n_live - 1]) --n_live;
  h.n_live = n_live;
  const uint32_t nseg = n ? ((n - 1) >> kSegBits) + 1 : 0;
  h.nseg = nseg;

  // 1. entries stably sorted by the segment of their source: within a
  //    segment, rows ascending and each row's entries in slot order (a
  //    counting sort over row ranges on host threads; each entry's slot,
  //    row and 16-bit offset carried along)
  std::vector<uint32_t> soff((size_t)nseg + 1, 0);
  std::vector<uint32_t> perm(m), prow(m);
  std::vector<uint16_t> loff(m);
  {
    const unsigned T = host_threads(m);
    std::vector<uint32_t> rb(T + 1, n);  // row ranges of about m / T entries
    for (unsigned t = 0; t < T; ++t)
      rb[t] = (uint32_t)(std::lower_bound(rp.begin(), rp.end() - 1, (uint32_t)((uint64_t)m * t / T)) - rp.begin());
    std::vector<uint32_t> cur((size_t)T * nseg, 0);  // per range: counts, then cursors
    std::vector<uint8_t> bad(T, 0);
    host_parallel(T, [&](unsigned t) {
      uint32_t *c = cur.data() + (size_t)t * nseg;
      for (uint32_t k = rp[rb[t]]; k < rp[rb[t + 1]]; ++k) {
        if (col[k] >= n) {
          bad[t] = 1;
          return;
        }
        c[col[k] >> kSegBits]++;
      }
    });
    for (unsigned t = 0; t < T; ++t)
      if (bad[t]) return fail(OLV_EINTERNAL, "blocked plan: an entry is out of range (n = %u)", n);
    uint32_t run = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
      soff[s] = run;
      for (unsigned t = 0; t < T; ++t) {
        const uint32_t x = cur[(size_t)t * nseg + s];
        cur[(size_t)t * nseg + s] = run;
        run += x;
      }
    }
    soff[nseg] = run;
    host_parallel(T, [&](unsigned t) {
      uint32_t *c = cur.data() + (size_t)t * nseg;
      for (uint32_t v = rb[t]; v < rb[t + 1]; ++v)
        for (uint32_t k = rp[v]; k < rp[v + 1]; ++k) {
          const uint32_t pos = c[col[k] >> kSegBits]++;
          perm[pos] = k;
          prow[pos] = v;
          loff[pos] = (uint16_t)(col[k] & (kSegSize - 1));
        }
    });
  }

  // 2. pieces: runs of one row within a segment, cut at kPieceMax entries;
  //    slots grouped by row in creation order (segment, then chunk)
  struct Piece {
    uint32_t start, len, slot;
  };
  std::vector<uint32_t> npc(n, 0);
  size_t npieces = 0;
  for (uint32_t s = 0; s < nseg; ++s)
    for (uint32_t p = soff[s]; p < soff[s + 1];) {
      uint32_t e = p;
      while (e < soff[s + 1] && prow[e] == prow[p]) ++e;
      const uint32_t chunks = (e - p + kPieceMax - 1) / kPieceMax;
      npc[prow[p]] += chunks;
      npieces += chunks;
      p = e;
    }
  if (npieces >= 0xffffffffull) return fail(OLV_ERANGE, "blocked plan: %zu pieces", npieces);
  std::vector<uint32_t> &cmb = h.cmb_rp;
  cmb.assign((size_t)n + 1, 0);
  for (uint32_t v = 0; v < n; ++v) cmb[v + 1] = cmb[v] + npc[v];
  // rows with more partials than one lane combines, listed by tier
  h.cmb_mid.clear();
  h.cmb_big.clear();
  for (uint32_t v = 0; v < n; ++v) {
    if (npc[v] > kCmbMid) h.cmb_big.push_back(v);
    else if (npc[v] > kCmbLight) h.cmb_mid.push_back(v);
  }
  std::vector<uint32_t> slot_cur(cmb.begin(), cmb.end() - 1);
  std::vector<Piece> order;  // final piece order: segment, then bucket (longest first), then row
  order.reserve(npieces);
  std::vector<uint4> &units = h.units;
  units.clear();
  std::vector<double> unit_cost;
  std::vector<Piece> byb[kBuckets];
  uint32_t edge_cursor = 0;  // first entry of the next unit
  for (uint32_t s = 0; s < nseg; ++s) {
    for (auto &b : byb) b.clear();
    for (uint32_t p = soff[s]; p < soff[s + 1];) {
      uint32_t e = p;
      while (e < soff[s + 1] && prow[e] == prow[p]) ++e;
      const uint32_t v = prow[p];
      for (uint32_t c = p; c < e; c += kPieceMax) {
        const uint32_t len = std::min(kPieceMax, e - c);
        byb[piece_bucket(len)].push_back(Piece{c, len, slot_cur[v]++});
      }
      p = e;
    }
    for (int b = kBuckets - 1; b >= 0; --b) {
      // units of at most unit_edges entries and kUnitPieces pieces over this
      // (segment, bucket) group
      size_t i = 0;
      while (i < byb[b].size()) {
        const uint32_t first = (uint32_t)order.size();
        uint32_t edges = 0, pieces = 0;
        while (i < byb[b].size() &&
               (pieces == 0 || (pieces < kUnitPieces && edges + byb[b][i].len <= unit_edges))) {
          edges += byb[b][i].len;
          ++pieces;
          order.push_back(byb[b][i++]);
        }
        units.push_back(make_uint4(s, (uint32_t)b, first, edge_cursor));
        for (uint32_t q = first; q < (uint32_t)order.size(); ++q) edge_cursor += order[q].len;
        // every unit starts on an 8-entry (16-byte) boundary: staged with
        // 16-byte loads and LDS stores
        if ((uint64_t)edge_cursor + 7 >= 0xffffffffull) return fail(OLV_ERANGE, "blocked plan: entries overflow");
        edge_cursor = (edge_cursor + 7) & ~7u;
        // cost in us, fitted to per-unit timestamps of RMAT-20 iterations
        // on MI355X: staging and the barriers around it ~1.3, then
        // kStepUs[b] per step (one piece per group of G lanes) and 0.04 per
        // thousand entries
        constexpr double kStepUs[kBuckets] = {0.15, 0.25, 0.41, 0.20, 0.29, 0.37, 0.38, 0.41, 0.37};
        const uint32_t per_step = kBlkThreads / kBucketLanes[b];
        const uint32_t steps = (pieces + per_step - 1) / per_step;
        unit_cost.push_back(1.3 + kStepUs[b] * steps + 4e-5 * edges);
      }
    }
  }

  // 3. entry arrays in piece order, each unit from its (aligned) first
  //    entry.  The last piece of a unit runs up to the next unit's first
  //    entry: the gap holds kSegSize, the staged slice's zero slot, as do 8
  //    spare entries at the end for the staging's 16-byte reads.
  std::vector<uint32_t> &lslot = h.lslot;
  std::vector<uint16_t> &pbeg = h.pbeg, &lidx = h.lidx;
  const uint32_t o = edge_cursor;
  pbeg.assign(npieces, 0);
  lidx.assign((size_t)o + 8, (uint16_t)kSegSize);
  lslot.assign(with_slots ? (size_t)o + 8 : 0, kNoSlot);
  {
    const unsigned T = host_threads(o);
    const size_t nu = units.size();
    host_parallel(T, [&](unsigned t) {  // units' entries are disjoint
      for (size_t un = nu * t / T; un < nu * (t + 1) / T; ++un) {
        const uint32_t e0 = units[un].w;
        uint32_t e = e0;
        const uint32_t qe = un + 1 < nu ? units[un + 1].z : (uint32_t)npieces;
        for (uint32_t q = units[un].z; q < qe; ++q) {
          const Piece &pc = order[q];
          pbeg[q] = (uint16_t)(e - e0);
          std::memcpy(&lidx[e], &loff[pc.start], pc.len * sizeof(uint16_t));
          if (with_slots) std::memcpy(&lslot[e], &perm[pc.start], pc.len * sizeof(uint32_t));
          e += pc.len;
        }
      }
    });
  }
  // each row's partials in segment order: cmb_pos[cmb_rp[v] + j] is the
  // piece (and partial) index of the row's j-th piece
  h.cmb_pos.assign(npieces, 0);
  {
    const unsigned T = host_threads(npieces);
    host_parallel(T, [&](unsigned t) {  // slots are a permutation of the pieces
      for (size_t q = npieces * t / T; q < npieces * (t + 1) / T; ++q) h.cmb_pos[order[q].slot] = (uint32_t)q;
    });
  }

  // 4. units to workgroups: longest first, each to the workgroup where it
  //    ends earliest, counting a slice load (~1 us) when the workgroup
  //    has no unit of that segment yet; each workgroup then runs its units
  //    in segment order.  A sentinel unit closes the list (its first piece
  //    and entry are the totals).
  const uint32_t nunits = (uint32_t)units.size();
  units.push_back(make_uint4(nseg, 0, (uint32_t)npieces, o));
  const uint32_t ng = std::max<uint32_t>(1, std::min<uint32_t>(groups ? groups : 1, std::max<uint32_t>(nunits, 1)));
  {
    std::vector<uint32_t> byc(nunits);
    for (uint32_t u = 0; u < nunits; ++u) byc[u] = u;
    std::stable_sort(byc.begin(), byc.end(), [&](uint32_t a, uint32_t b) { return unit_cost[a] > unit_cost[b]; });
    std::vector<double> load(ng, 0.0);
    load[0] = 1.5;  // workgroup 0 also applies the stop rule (a residual sum and two barriers)
    std::vector<std::vector<uint32_t>> mine(ng);
    std::vector<uint8_t> has((size_t)ng * nseg, 0);  // workgroup w already loads segment s
    constexpr double kSliceUs = 1.0;
    for (uint32_t u : byc) {
      const uint32_t sg = units[u].x;
      uint32_t best = 0;
      double bv = 0;
      for (uint32_t w = 0; w < ng; ++w) {
        const double v = load[w] + unit_cost[u] + (has[(size_t)w * nseg + sg] ? 0.0 : kSliceUs);
        if (w == 0 || v < bv) {
          bv = v;
          best = w;
        }
      }
      load[best] = bv;
      has[(size_t)best * nseg + sg] = 1;
      mine[best].push_back(u);
    }
    h.wg_ptr.assign(ng + 1, 0);
    h.wg_list.clear();
    h.wg_list.reserve(nunits);
    for (uint32_t w = 0; w < ng; ++w) {
      std::sort(mine[w].begin(), mine[w].end());  // unit order = segment order
      h.wg_ptr[w] = (uint32_t)h.wg_list.size();
      h.wg_list.insert(h.wg_list.end(), mine[w].begin(), mine[w].end());
    }
    h.wg_ptr[ng] = (uint32_t)h.wg_list.size();
  }
  static_assert(kUnitPieces < 0x10000 && kUnitEdges < 0x10000, "a unit's counts pack into 16 bits each");
  h.wg_units.resize(h.wg_list.size());
  for (size_t i = 0; i < h.wg_list.size(); ++i) {
    const uint4 a = units[h.wg_list[i]], b = units[h.wg_list[i] + 1];
    h.wg_units[i] = make_uint4(a.x << kBucketBits | a.y, a.z, a.w, (b.z - a.z) | (b.w - a.w) << 16);
  }
  h.wg_first.assign(ng, make_uint4(0, 0, 0, 0));
  for (uint32_t w = 0; w < ng; ++w)
    if (h.wg_ptr[w] < h.wg_ptr[w + 1]) h.wg_first[w] = h.wg_units[h.wg_ptr[w]];

  h.npieces = (uint32_t)npieces;
  h.nent = o + 8;
  h.nunits = nunits;
  h.ngroups = ng;
  return OLV_OK;
}

int build_blocked_plan(const std::vector<uint32_t> &rp, const uint32_t *col, bool with_slots, uint32_t groups,
                       uint32_t unit_edges, BlockedPlan &bp) {
  bp.release();
  BlockedHost h;
  OLV_TRY(build_blocked_host(rp, col, with_slots, groups, unit_edges, h));
  bp.unit_edges = unit_edges;
  bp.nent = h.nent;
  bp.n = h.n;
  bp.m = h.m;
  bp.n_live = h.n_live;
  bp.n_mid = (uint32_t)h.cmb_mid.size();
  bp.n_big = (uint32_t)h.cmb_big.size();
  bp.nseg = h.nseg;
  bp.npieces = h.npieces;
  bp.nunits = h.nunits;
  bp.ngroups = h.ngroups;
  OLV_TRY(upload_vec(&bp.wg_ptr, h.wg_ptr));
  OLV_TRY(upload_vec(&bp.wg_units, h.wg_units));
  OLV_TRY(upload_vec(&bp.wg_first, h.wg_first));
  OLV_TRY(upload_vec(&bp.pbeg, h.pbeg));
  OLV_TRY(upload_vec(&bp.lidx, h.lidx));
  if (with_slots) OLV_TRY(upload_vec(&bp.lslot, h.lslot));
  OLV_TRY(upload_vec(&bp.cmb_rp, h.cmb_rp));
  OLV_TRY(upload_vec(&bp.cmb_pos, h.cmb_pos));
  OLV_TRY(upload_vec(&bp.cmb_mid, h.cmb_mid));
  OLV_TRY(upload_vec(&bp.cmb_big, h.cmb_big));
  OLV_HIP(hipMalloc(reinterpret_cast<void **>(&bp.partial), std::max<size_t>(h.npieces, 1) * sizeof(float)));
  bp.ready = true;
  return OLV_OK;
}

}  // namespace olv
