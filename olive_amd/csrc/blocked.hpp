// Source-segmented pull for the PageRank iteration (CSR segmenting: Zhang,
// Kiriansky, Mendis, Amarasinghe, Zaharia, Matei, "Making Caches Work for
// Graph Analytics", IEEE BigData 2017, restated for a GPU with a 160 KB LDS
// per CU).
//
// The iterate's source vector is cut into segments of kSegSize consecutive
// vertices (64 KB of fp32, one LDS slice).  The in-edges of every
// destination are grouped by the segment of their source into "pieces"
// (destination, segment), each piece's entries kept in the row's slot order
// and stored as 16-bit offsets into its segment; pieces longer than
// kPieceMax entries are cut into chunks.  One iteration is two launches:
//
//   segment pass  each workgroup owns a run of work units (a unit: at most
//                 kUnitPieces pieces and kUnitEdges entries of one segment
//                 and one length bucket).  Per unit it stages, with
//                 coalesced loads, the unit's entries and piece offsets in
//                 LDS (and the segment of the source vector, when the
//                 segment changes), then sums every piece from LDS alone --
//                 1 to 64 lanes per piece by length -- and stores one
//                 partial per piece, in piece order (coalesced stores);
//   combine       per destination, its partials (listed by cmb_pos in
//                 segment order) are added in a fixed order -- by one lane
//                 for up to kCmbLight partials, 8 lanes up to kCmbMid, a
//                 wave beyond -- and the vertex update runs.
//
// A piece's terms are summed in a fixed order and the partials in segment
// order, so an iteration is bit-reproducible from run to run.  With the
// vertices numbered by descending in-degree (the renumbered CSRs of the
// PageRank solver) a skewed graph's in-edges fall into few pieces: RMAT-20
// has 16.8 M in-edges in 2.3 M pieces, most of them in segment 0.  Per
// iteration the pass reads 2 B per edge and 2 B per piece and writes 4 B per
// piece, and every gather is an LDS read.
#pragma once
#include <vector>

#include "device.hpp"
#include "internal.hpp"
#include "seg_core.hpp"

namespace olv {

constexpr uint32_t kSegBits = 14;
constexpr uint32_t kSegSize = 1u << kSegBits;  // sources per LDS slice (64 KB of fp32), plus a zero slot
static_assert(kSegSize <= 0xffff, "entries hold a slice offset or the zero slot in 16 bits");
constexpr int kBlkThreads = 1024;              // segment-pass workgroup (one per CU: 130-144 KB of LDS)
constexpr uint32_t kPieceMax = 512;            // longest piece (one wave step); longer ones are cut into chunks
constexpr uint32_t kUnitEdges = 32768;         // entries per work unit, at most (64 KB staged)
constexpr uint32_t kUnitEdgesW = 8192;         // with per-entry weights (16 + 32 KB staged)
constexpr uint32_t kNoSlot = 0xffffffffu;      // lslot of the alignment gaps
constexpr uint32_t kUnitPieces = 8192;         // pieces per work unit, at most (16 KB staged)
constexpr uint32_t kCmbLight = 16;             // partials of a row combined by one lane, at most
constexpr uint32_t kCmbMid = 128;              // by kCmbMidLanes lanes, at most; a wave beyond
constexpr uint32_t kCmbMidLanes = 8;           // (16 partials each, all in flight)

// Device arrays of a segmented pull over one CSR.
struct BlockedPlan {
  uint32_t n = 0, m = 0, n_live = 0;  // vertices, entries, rows with entries (a prefix of the rows)
  uint32_t n_mid = 0, n_big = 0;      // rows with kCmbLight < partials <= kCmbMid, and more
  uint32_t nseg = 0, npieces = 0, nunits = 0, ngroups = 0;
  uint32_t unit_edges = 0, nent = 0;  // entries per unit at most; entries with gaps and spares
  uint32_t *wg_ptr = nullptr;    // ngroups + 1: each segment-pass workgroup's range in wg_units
  uint4 *wg_units = nullptr;     // nunits: the units of each workgroup in segment order, packed as
                                 // (segment << 4 | bucket, first piece, first entry,
                                 // pieces | entries << 16); bucket 0..8: lanes and reads per piece
  uint4 *wg_first = nullptr;     // ngroups: each workgroup's first unit (zeros if it has none)
  uint16_t *pbeg = nullptr;      // npieces: each piece's first entry, counted from its unit's first
                                 // entry (a unit's last piece ends at the unit's aligned entry count)
  uint16_t *lidx = nullptr;      // source offset within the segment, in piece order; each unit starts
                                 // on an 8-entry boundary; the gaps and 8 spare entries at the end
                                 // hold kSegSize (the slice's zero slot)
  uint32_t *lslot = nullptr;     // (weighted solves only) CSR slot each entry came from, aligned as lidx
                                 // (kNoSlot in the gaps and spares); its owner may map it in place
  uint32_t *cmb_rp = nullptr;    // n + 1: each row's range in cmb_pos
  uint32_t *cmb_pos = nullptr;   // npieces: partial index of each row's pieces, in segment order
  uint32_t *cmb_mid = nullptr;   // n_mid: the rows with kCmbLight < partials <= kCmbMid, ascending
  uint32_t *cmb_big = nullptr;   // n_big: the rows with more than kCmbMid partials, ascending
  float *partial = nullptr;      // npieces
  bool ready = false;
  void release() {
    void *ps[] = {wg_ptr, wg_units, wg_first, pbeg, lidx, lslot, cmb_rp, cmb_pos, cmb_mid, cmb_big, partial};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    wg_units = wg_first = nullptr;
    wg_ptr = lslot = cmb_rp = cmb_pos = cmb_mid = cmb_big = nullptr;
    pbeg = lidx = nullptr;
    partial = nullptr;
    ready = false;
  }
};

// Host form of the plan (the arrays before upload).  units: nunits + 1
// (segment, bucket, first piece, first entry), the last a sentinel (-, -,
// npieces, entries); wg_list: the units of each workgroup.
struct BlockedHost {
  uint32_t n = 0, m = 0, n_live = 0, nseg = 0, npieces = 0, nunits = 0, ngroups = 0, nent = 0;
  std::vector<uint4> units, wg_units, wg_first;
  std::vector<uint32_t> wg_ptr, wg_list, lslot, cmb_rp, cmb_pos, cmb_mid, cmb_big;
  std::vector<uint16_t> pbeg, lidx;
};
int build_blocked_host(const std::vector<uint32_t> &rp, const uint32_t *col, bool with_slots, uint32_t groups,
                       uint32_t unit_edges, BlockedHost &h);

// Builds the plan for a CSR whose rows are given by host offsets rp (n + 1)
// and host entries col (m, source ids); rows must be ordered so that the rows
// without entries come last (n_live = first such row).  with_slots keeps the
// CSR slot of every entry (to gather per-slot data such as edge weights into
// piece order).  groups: segment-pass workgroups to balance the units over;
// unit_edges: entries per unit at most (kUnitEdges, or kUnitEdgesW for a
// pass that stages per-entry weights).
int build_blocked_plan(const std::vector<uint32_t> &rp, const uint32_t *col, bool with_slots, uint32_t groups,
                       uint32_t unit_edges, BlockedPlan &bp);

struct BlkDev {
  const uint32_t *wg_ptr;
  const uint4 *wg_units;
  const uint4 *wg_first;
  const uint16_t *pbeg;
  const uint16_t *lidx;
  const uint32_t *cmb_rp;
  const uint32_t *cmb_pos;
  const uint32_t *cmb_mid, *cmb_big;
  float *partial;
  const float *wpiece;  // per-entry weights in the plan's entry order (weighted passes)
  uint32_t n, n_live, n_mid, n_big;
};

inline BlkDev make_blk_dev(const BlockedPlan &p, const float *wpiece = nullptr) {
  return BlkDev{p.wg_ptr,  p.wg_units, p.wg_first, p.pbeg,  p.lidx, p.cmb_rp, p.cmb_pos, p.cmb_mid,
                p.cmb_big, p.partial,  wpiece,     p.n,     p.n_live, p.n_mid, p.n_big};
}

// Lane j's share of a piece of a unit whose pieces take G lanes each: the
// sum of sv(lidx[k]) over k = beg + j + i G, i < U (lidx and the slice in
// LDS), all reads in flight at once and the terms added pairwise.  Such a
// piece has at most U G entries (the length buckets), so this covers it;
// the k past its end (up to `end`, the unit's alignment padding, or beyond)
// read the slice's zero slot (kSegSize) instead of branching.
constexpr int kStepReads = 8;  // reads per lane, at most
constexpr uint32_t kBucketBits = 4;
static_assert(kPieceMax <= kStepReads * kWave, "a piece fits one step of a wave");
template <int G, int U, class SV>
__device__ __forceinline__ float piece_step(const SV &sv, const uint16_t *lidx, uint32_t beg, uint32_t end,
                                            uint32_t j) {
  uint32_t ix[U];
#pragma unroll
  for (int i = 0; i < U; ++i) {
    const uint32_t k = beg + j + (uint32_t)i * G;
    ix[i] = k < end ? (uint32_t)lidx[k] : kSegSize;
  }
  float acc[U];
#pragma unroll
  for (int i = 0; i < U; ++i) acc[i] = sv(beg + j + (uint32_t)i * G, ix[i]);
#pragma unroll
  for (int w = U / 2; w > 0; w >>= 1)
#pragma unroll
    for (int i = 0; i < w; ++i) acc[i] += acc[i + w];
  return acc[0];
}

// DPP move of x (rows of 16 lanes outside RowMask receive 0)
template <int Ctrl, int RowMask = 0xf>
__device__ __forceinline__ float blk_dpp(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), Ctrl, RowMask, 0xf, false));
}

// Sum over aligned groups of G lanes with DPP moves (no LDS round trip):
// inside rows of 16 lanes every lane of the group ends with the sum; 32
// lanes add the lower row's sum into the upper row (row_bcast:15), so there
// the upper row holds it; for the whole wave the four row sums are read by
// lane and every lane holds the sum.  Lane G - 1 of a group always does.
template <int G>
__device__ __forceinline__ float blk_group_sum(float v) {
  static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16 || G == 32 || G == kWave, "group of 2^i lanes");
  if constexpr (G == 1) return v;
  v += blk_dpp<0xb1>(v);                          // quad_perm [1,0,3,2]: lane ^ 1
  if constexpr (G >= 4) v += blk_dpp<0x4e>(v);    // quad_perm [2,3,0,1]: lane ^ 2
  if constexpr (G >= 8) v += blk_dpp<0x141>(v);   // row_half_mirror: lane 7 - i of 8
  if constexpr (G >= 16) v += blk_dpp<0x140>(v);  // row_mirror: lane 15 - i of 16
  if constexpr (G == 32) v += blk_dpp<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  if constexpr (G == kWave) {
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 15));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 31));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 47));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
    v = (r0 + r1) + (r2 + r3);
  }
  return v;
}

// Pieces of a unit summed G lanes each, U reads a lane; lane G - 1 of the
// group stores the partial.
template <int G, int U, class SV>
__device__ __forceinline__ void blk_pieces(const BlkDev &d, const SV &sv, const uint16_t *s_lidx,
                                           const uint16_t *s_pbeg, uint32_t np, uint32_t q0) {
  const uint32_t j = threadIdx.x % G;
  for (uint32_t q = threadIdx.x / G; q < np; q += kBlkThreads / G) {
    const float s = blk_group_sum<G>(piece_step<G, U>(sv, s_lidx, s_pbeg[q], s_pbeg[q + 1], j));
    if (j == G - 1) d.partial[q0 + q] = s;
  }
}

// Segment pass of iteration k.  With a control block (ctl != nullptr) it
// first applies the stop rule of the iteration loop: returns at once when
// done; workgroup 0 alone sets done when iteration k - 1 left an L1 residual
// below eps (the other workgroups' partials of that pass are then unused:
// the combine launch returns on done) and clears the residual slots of
// iteration k + 1.
// SV: void bind(const float *slice, const float *wts) and
// float operator()(uint32_t k, uint32_t i) const -- the summand of the
// unit's entry k whose source is slice[i] (wts: the unit's staged weights,
// kW only); it is also called with i = kSegSize (the zero slot) and k up to
// kPieceMax past a piece's end, where slice[i] is 0 and wts[k] finite.
// kW: units of at most kUnitEdgesW entries (plans built with that limit)
// whose weights (d.wpiece) are staged in LDS beside their offsets.
template <class SV, bool kW = false>
__global__ void __launch_bounds__(kBlkThreads) blk_segment_kernel(BlkDev d, SV sv, const float *__restrict__ src,
                                                                   IterCtl *ctl, uint32_t k) {
  constexpr uint32_t kUE = kW ? kUnitEdgesW : kUnitEdges;
  constexpr uint32_t kWS = kW ? kUnitEdgesW + kPieceMax : 4;  // staged weights and a finite margin
  __shared__ float slice[kSegSize + 4];  // slice[kSegSize] is the zero slot of the padding entries
  __shared__ __align__(16) uint16_t s_lidx[kUE];
  __shared__ __align__(16) float s_w[kWS];
  __shared__ uint16_t s_pbeg[kUnitPieces + 2];  // piece starts and the unit's end
  const uint32_t tid = threadIdx.x;
  const uint32_t u0 = d.wg_ptr[blockIdx.x], u1 = d.wg_ptr[blockIdx.x + 1];
  const uint4 first = d.wg_first[blockIdx.x];  // in flight with wg_ptr and during the checks
  if (ctl) {
    if (__hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    if (blockIdx.x == 0) {
      if (k > 0) {
        const double res = err_total(ctl, (k - 1) % 3);
        if (res < ctl->eps) {
          if (threadIdx.x == 0) {
            ctl->iters = k;
            ctl->residual = res;
            __hip_atomic_store(&ctl->done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
          return;
        }
      }
      double *z = &ctl->errs[(k + 1) % 3][0].v;
      for (uint32_t i = threadIdx.x; i < kErrSlots * (sizeof(ErrSlot) / sizeof(double)); i += kBlkThreads) z[i] = 0.0;
    }
  }
  if (u0 >= u1) return;
  if (tid == 0) slice[kSegSize] = 0.f;  // read after the loop's first barrier
  if constexpr (kW)
    for (uint32_t i = tid; i < kWS; i += kBlkThreads) s_w[i] = 0.f;  // finite past any unit's weights
  constexpr uint32_t kLi = kUE / (8 * kBlkThreads), kPi = kUnitPieces / kBlkThreads;
  constexpr uint32_t kWv = kW ? kUE / (4 * kBlkThreads) : 1;
  constexpr uint32_t kSl = kSegSize / (4 * kBlkThreads);
  // The next unit's staged data -- and its segment of the source vector when
  // that changes -- are loaded into registers while the current unit is
  // summed, and the descriptor of the unit after it too: no load waits on
  // another.  Entries move 8 at a time (units start on 8-entry boundaries).
  uint4 li[kLi];
  float4 wv[kWv];
  uint32_t pb[kPi];
  float4 sl[kSl];
  auto load_unit = [&](const uint4 c, uint32_t have) {
    const uint32_t np = c.w & 0xffffu, ne = c.w >> 16, seg = c.x >> kBucketBits;
    const uint4 *lsrc = reinterpret_cast<const uint4 *>(d.lidx + c.z);
#pragma unroll
    for (uint32_t j = 0; j < kLi; ++j) {
      const uint32_t i = tid + j * kBlkThreads;
      li[j] = i * 8 < ne ? lsrc[i] : make_uint4(0, 0, 0, 0);
    }
    if constexpr (kW) {
      const float4 *wsrc = reinterpret_cast<const float4 *>(d.wpiece + c.z);
#pragma unroll
      for (uint32_t j = 0; j < kWv; ++j) {
        const uint32_t i = tid + j * kBlkThreads;
        wv[j] = i * 4 < ne ? wsrc[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kPi; ++j) {
      const uint32_t i = tid + j * kBlkThreads;
      pb[j] = i < np ? d.pbeg[c.y + i] : 0u;
    }
    if (seg != have) {
      const uint32_t base = seg << kSegBits, cnt = d.n - base;
      if (cnt >= kSegSize) {  // uniform: the descriptor is scalar
#pragma unroll
        for (uint32_t j = 0; j < kSl; ++j)
          sl[j] = *reinterpret_cast<const float4 *>(src + base + (tid + j * kBlkThreads) * 4);
      } else {  // the last, partial segment
#pragma unroll
        for (uint32_t j = 0; j < kSl; ++j) {
          const uint32_t i = (tid + j * kBlkThreads) * 4;
          sl[j].x = i < cnt ? src[base + i] : 0.f;
          sl[j].y = i + 1 < cnt ? src[base + i + 1] : 0.f;
          sl[j].z = i + 2 < cnt ? src[base + i + 2] : 0.f;
          sl[j].w = i + 3 < cnt ? src[base + i + 3] : 0.f;
        }
      }
    }
  };
  // descriptors are the same for the whole workgroup: held in scalar registers
  auto scalar = [](const uint4 c) {
    return make_uint4(__builtin_amdgcn_readfirstlane(c.x), __builtin_amdgcn_readfirstlane(c.y),
                      __builtin_amdgcn_readfirstlane(c.z), __builtin_amdgcn_readfirstlane(c.w));
  };
  uint4 cd = scalar(first);  // the unit whose data the registers hold
  load_unit(cd, 0xffffffffu);
  uint4 nd = u0 + 1 < u1 ? d.wg_units[u0 + 1] : cd;
  uint32_t cur = 0xffffffffu;
  for (uint32_t u = u0; u < u1; ++u) {
    const uint32_t seg = cd.x >> kBucketBits, bucket = cd.x & ((1u << kBucketBits) - 1), q0 = cd.y;
    const uint32_t np = cd.w & 0xffffu, ne = cd.w >> 16;
    __syncthreads();  // the previous unit is done with the staged arrays
    if (seg != cur) {
#pragma unroll
      for (uint32_t j = 0; j < kSl; ++j) *reinterpret_cast<float4 *>(&slice[(tid + j * kBlkThreads) * 4]) = sl[j];
      cur = seg;
    }
#pragma unroll
    for (uint32_t j = 0; j < kLi; ++j) {
      const uint32_t i = tid + j * kBlkThreads;
      if (i * 8 < ne) reinterpret_cast<uint4 *>(s_lidx)[i] = li[j];
    }
    if constexpr (kW) {
#pragma unroll
      for (uint32_t j = 0; j < kWv; ++j) {
        const uint32_t i = tid + j * kBlkThreads;
        if (i * 4 < ne) reinterpret_cast<float4 *>(s_w)[i] = wv[j];
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < kPi; ++j) {
      const uint32_t i = tid + j * kBlkThreads;
      if (i < np) s_pbeg[i] = (uint16_t)pb[j];
    }
    if (tid == 0) s_pbeg[np] = (uint16_t)ne;
    if (u + 1 < u1) {
      cd = scalar(nd);
      load_unit(cd, seg);
      if (u + 2 < u1) nd = d.wg_units[u + 2];
    }
    __syncthreads();
    sv.bind(slice, s_w);
    // G lanes per piece and U reads per lane by bucket (pieces of up to U G
    // entries)
    switch (bucket) {
      case 0:  // up to 2 entries
        blk_pieces<1, 2>(d, sv, s_lidx, s_pbeg, np, q0);
        break;
      case 1:  // up to 4
        blk_pieces<1, 4>(d, sv, s_lidx, s_pbeg, np, q0);
        break;
      case 2:  // up to 8
        blk_pieces<1, 8>(d, sv, s_lidx, s_pbeg, np, q0);
        break;
      case 3:
        blk_pieces<2, 8>(d, sv, s_lidx, s_pbeg, np, q0);
        break;
      case 4:
        blk_pieces<4, 8>(d, sv, s_lidx, s_pbeg, np, q0);
        break;
      case 5:
        blk_pieces<8, 8>(d, sv, s_lidx, s_pbeg, np, q0);
        break;
      case 6:
        blk_pieces<16, 8>(d, sv, s_lidx, s_pbeg, np, q0);
        break;
      case 7:
        blk_pieces<32, 8>(d, sv, s_lidx, s_pbeg, np, q0);
        break;
      default:  // up to kPieceMax
        blk_pieces<kWave, 8>(d, sv, s_lidx, s_pbeg, np, q0);
        break;
    }
  }
}

// Combine of iteration k: row v's partials added in a fixed order, then
// fin.finish(v, sum, err, fin.pre(v)) (pre: the row's other inputs, loaded
// ahead); the residual goes to errs[k % 3] (ctl != nullptr).
// Rows by their number of partials: more than kCmbMid (the list cmb_big)
// a wave each -- lane j adds partials j, j + 64, ... (four loads in flight),
// then a DPP sum -- in the first big_blocks workgroups; kCmbLight + 1 to
// kCmbMid (cmb_mid) kCmbMidLanes = 8 lanes each -- lane j adds partials j,
// j + 8, ..., j + 120, all in flight, then a DPP sum -- in the next mid_blocks
// workgroups; every other row a lane each, up to kCmbLight loads in flight
// added in segment order, 256 consecutive rows per workgroup in the rest.
// Rows without entries (v >= n_live) get a constant update: iterations 0
// and 1 write it into both ping-pong buffers, later iterations skip them.
template <class Fin>
__global__ void __launch_bounds__(256) blk_combine_kernel(BlkDev d, Fin fin, IterCtl *ctl, uint32_t k,
                                                           uint32_t big_blocks, uint32_t mid_blocks) {
  if (ctl && __hip_atomic_load(&ctl->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  __shared__ float s_red[256 / kWave];
  const uint32_t nv = (!ctl || k >= 2) ? d.n_live : d.n;
  float err = 0.f;
  const uint32_t lane = threadIdx.x % kWave;
  if (blockIdx.x < big_blocks) {
    for (uint32_t i = (blockIdx.x * 256 + threadIdx.x) / kWave; i < d.n_big; i += big_blocks * (256 / kWave)) {
      const uint32_t v = d.cmb_big[i];  // the same for the whole wave
      const uint32_t a = d.cmb_rp[v], b = d.cmb_rp[v + 1];
      float s = 0.f;
      for (uint32_t q0 = a + lane; q0 < b; q0 += 4 * kWave) {
        uint32_t ix[4];
#pragma unroll
        for (uint32_t t = 0; t < 4; ++t) ix[t] = q0 + t * kWave < b ? d.cmb_pos[q0 + t * kWave] : 0u;
        float p[4];
#pragma unroll
        for (uint32_t t = 0; t < 4; ++t) p[t] = q0 + t * kWave < b ? d.partial[ix[t]] : 0.f;
#pragma unroll
        for (uint32_t t = 0; t < 4; ++t) s += p[t];
      }
      s = blk_group_sum<kWave>(s);
      if (lane == 0) fin.finish(v, s, err, fin.pre(v));
    }
  } else if (blockIdx.x < big_blocks + mid_blocks) {
    constexpr uint32_t G = kCmbMidLanes, kPer = kCmbMid / G;
    const uint32_t j = threadIdx.x % G;
    for (uint32_t i = ((blockIdx.x - big_blocks) * 256 + threadIdx.x) / G; i < d.n_mid; i += mid_blocks * (256 / G)) {
      const uint32_t v = d.cmb_mid[i];  // the same for the group's lanes (inside one DPP row)
      const uint32_t a = d.cmb_rp[v], b = d.cmb_rp[v + 1];
      uint32_t ix[kPer];
#pragma unroll
      for (uint32_t t = 0; t < kPer; ++t) ix[t] = a + j + t * G < b ? d.cmb_pos[a + j + t * G] : 0u;
      float p[kPer];
#pragma unroll
      for (uint32_t t = 0; t < kPer; ++t) p[t] = a + j + t * G < b ? d.partial[ix[t]] : 0.f;
      float s = 0.f;
#pragma unroll
      for (uint32_t t = 0; t < kPer; ++t) s += p[t];
      s = blk_group_sum<G>(s);
      if (j == 0) fin.finish(v, s, err, fin.pre(v));
    }
  } else {
    const uint32_t lb = blockIdx.x - big_blocks - mid_blocks, nlb = gridDim.x - big_blocks - mid_blocks;
    for (uint32_t v = lb * 256 + threadIdx.x; v < nv; v += nlb * 256) {
      const uint32_t a = d.cmb_rp[v], b = d.cmb_rp[v + 1];
      if (b - a > kCmbLight) continue;  // a listed row
      const auto pre = fin.pre(v);
      uint32_t ix[kCmbLight];
#pragma unroll
      for (uint32_t t = 0; t < kCmbLight; ++t) ix[t] = a + t < b ? d.cmb_pos[a + t] : 0u;
      float p[kCmbLight];
#pragma unroll
      for (uint32_t t = 0; t < kCmbLight; ++t) p[t] = a + t < b ? d.partial[ix[t]] : 0.f;
      float s = 0.f;
#pragma unroll
      for (uint32_t t = 0; t < kCmbLight; ++t)
        if (a + t < b) s += p[t];
      fin.finish(v, s, err, pre);
    }
  }
  if (!ctl) return;
  err = block_sum<256, float>(err, s_red);
  if (threadIdx.x == 0 && err != 0.f) atomicAdd(&ctl->errs[k % 3][blockIdx.x % kErrSlots].v, (double)err);
}

}  // namespace olv
