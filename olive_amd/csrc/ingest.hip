// CSR ingest on the device.
//
// * An edge list in any order (COO) becomes the out-CSR by a stable LSD
//   radix sort of the edge ids by source (8 bits per pass, wave64 ballot
//   ranking inside a tile, digit-major tile offsets from a device scan), so
//   the edges of a source keep their input order.  The in-CSR is the same
//   sort of the CSR positions by target: in-lists come out ordered by source
//   id, as the host transpose orders them.
// * Edge-list text (csrGraph.h:106-171 fromEdgeListFile) and DIMACS
//   adjacency text (csrGraph.h:177-236 fromDimacsFile) are split into lines
//   and parsed on the device, one lane per line.
//
// Text semantics (the reference's loaders, with two differences kept on
// purpose): '#' lines are skipped; the first other non-empty line is
// "n m"; then exactly m "src dst [w]" lines (columns after the third are
// ignored).  The
// reference never sorts (its stable sort is commented out at
// csrGraph.h:149, so it needs a file sorted by source) and drops the third
// column (csrGraph.h:139); here the tuples are stable-sorted by source and a
// third column is kept as int32 weights when every edge line has one.
// DIMACS: header "n m", then line i after the header lists the 1-based
// neighbours of vertex i-1 (blank lines are vertices without edges); the
// neighbour count must be 2m (csrGraph.h:229-232).

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "device.hpp"
#include "internal.hpp"

namespace olv {

namespace {

constexpr int kIB = 256;                     // workgroup size of the ingest kernels
constexpr int kIItems = 16;                  // items per lane in a tile
constexpr uint32_t kITile = kIB * kIItems;   // 4096 items per tile
constexpr int kRBits = 8;                    // radix digit width
constexpr int kRDigits = 1 << kRBits;
constexpr int kIWaves = kIB / kWave;
constexpr uint32_t kTextChunk = 64;          // text bytes per lane
constexpr uint32_t kTextTile = kIB * kTextChunk;

// Workgroups per CU of the grid-stride ingest kernels, and of the line
// parser (more lines in flight for its dependent byte reads: RMAT-22 text
// 2.21 ms with 64, 2.27 with 32, 2.39 with 16).
#ifndef OLV_INGEST_GRID_PER_CU
#define OLV_INGEST_GRID_PER_CU 16
#endif
#ifndef OLV_PARSE_GRID_PER_CU
#define OLV_PARSE_GRID_PER_CU 64
#endif
unsigned grid_for(size_t items, size_t per, size_t per_cu = OLV_INGEST_GRID_PER_CU) {
  const size_t b = (items + per - 1) / per;
  return (unsigned)std::max<size_t>(1, std::min(b, (size_t)cu_count() * per_cu));
}

template <class T>
T *take(DevBuf &b) {
  T *p = b.as<T>();
  b.p = nullptr;
  return p;
}

// ------------------------------------------------------------------ scans
// Exclusive prefix of x over the workgroup; `total` gets the workgroup sum.
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t x, uint32_t &total) {
  __shared__ uint32_t s_w[kIWaves];
  const uint32_t incl = wave_inclusive_scan(x);
  const uint32_t w = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) s_w[w] = incl;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (uint32_t i = 0; i < (uint32_t)kIWaves; ++i) {
    const uint32_t v = s_w[i];
    off += i < w ? v : 0u;
    tot += v;
  }
  __syncthreads();
  total = tot;
  return off + incl - x;
}

// sums[t] = sum of in over tile t.
__global__ void __launch_bounds__(kIB) scan_reduce_kernel(const uint32_t *__restrict__ in, size_t n,
                                                          uint32_t *__restrict__ sums) {
  const size_t base = (size_t)blockIdx.x * kITile;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kIItems; ++i) {
    const size_t k = base + (size_t)i * kIB + threadIdx.x;
    if (k < n) s += in[k];
  }
  uint32_t tot;
  (void)block_excl_sum(s, tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// sums[0..nt) -> exclusive prefix sums in place, sums[nt] = total (one workgroup).
__global__ void __launch_bounds__(kIB) scan_top_kernel(uint32_t *sums, uint32_t nt) {
  uint32_t carry = 0;
  for (size_t base = 0; base < nt; base += kITile) {
    uint32_t v[kIItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kIItems; ++i) {
      const size_t k = base + (size_t)threadIdx.x * kIItems + i;
      v[i] = k < nt ? sums[k] : 0u;
      s += v[i];
    }
    uint32_t tot;
    uint32_t run = carry + block_excl_sum(s, tot);
#pragma unroll
    for (int i = 0; i < kIItems; ++i) {
      const size_t k = base + (size_t)threadIdx.x * kIItems + i;
      if (k < nt) sums[k] = run;
      run += v[i];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) sums[nt] = carry;
}

// out = tile offset + exclusive prefix inside the tile (in == out allowed);
// the last tile also writes out[n] = total.
__global__ void __launch_bounds__(kIB) scan_down_kernel(const uint32_t *in, uint32_t *out, size_t n,
                                                        const uint32_t *__restrict__ sums, uint32_t nt) {
  const size_t base = (size_t)blockIdx.x * kITile + (size_t)threadIdx.x * kIItems;
  uint32_t v[kIItems];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kIItems; ++i) {
    v[i] = base + i < n ? in[base + i] : 0u;
    s += v[i];
  }
  uint32_t tot;
  uint32_t run = sums[blockIdx.x] + block_excl_sum(s, tot);
#pragma unroll
  for (int i = 0; i < kIItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (blockIdx.x == nt - 1 && threadIdx.x == 0) out[n] = sums[nt];
}

}  // namespace

// out[0..n) = exclusive prefix sums of in[0..n), out[n] = total (sum < 2^32).
// out must hold n + 1 words; in == out is allowed.
int exclusive_scan(const uint32_t *in, uint32_t *out, size_t n) {
  const size_t nt = std::max<size_t>(1, (n + kITile - 1) / kITile);
  if (nt >= (1ull << 32)) return fail(OLV_ERANGE, "scan of %zu items", n);
  ScratchBuf sums;
  OLV_TRY(sums.alloc((nt + 1) * 4));
  hipStream_t s = stream();
  hipLaunchKernelGGL(scan_reduce_kernel, dim3((unsigned)nt), dim3(kIB), 0, s, in, n, sums.as<uint32_t>());
  hipLaunchKernelGGL(scan_top_kernel, dim3(1), dim3(kIB), 0, s, sums.as<uint32_t>(), (uint32_t)nt);
  hipLaunchKernelGGL(scan_down_kernel, dim3((unsigned)nt), dim3(kIB), 0, s, in, out, n, sums.as<uint32_t>(),
                     (uint32_t)nt);
  OLV_HIP(hipGetLastError());
  return OLV_OK;  // `sums` goes back to the pool after this thread's stream drains
}

namespace {

// ------------------------------------------------------------- radix sort
// counts[d * ntiles + t] = keys of tile t whose digit at `shift` is d.
__global__ void __launch_bounds__(kIB) radix_count_kernel(const uint32_t *__restrict__ keys, uint32_t m,
                                                          uint32_t shift, uint32_t ntiles,
                                                          uint32_t *__restrict__ counts) {
  __shared__ uint32_t h[kIWaves][kRDigits];
  for (uint32_t i = threadIdx.x; i < (uint32_t)(kIWaves * kRDigits); i += kIB) (&h[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t w = threadIdx.x / kWave;
  const size_t base = (size_t)blockIdx.x * kITile;
#pragma unroll
  for (int i = 0; i < kIItems; ++i) {
    const size_t k = base + (size_t)i * kIB + threadIdx.x;
    if (k < m) atomicAdd(&h[w][(keys[k] >> shift) & (kRDigits - 1)], 1u);
  }
  __syncthreads();
  for (uint32_t d = threadIdx.x; d < (uint32_t)kRDigits; d += kIB) {
    uint32_t c = 0;
#pragma unroll
    for (int x = 0; x < kIWaves; ++x) c += h[x][d];
    counts[(size_t)d * ntiles + blockIdx.x] = c;
  }
}

// Value sources of the first pass (later passes read the previous output).
struct LoadIndex {
  __device__ uint32_t operator()(size_t k) const { return (uint32_t)k; }
};
struct LoadArr {
  const uint32_t *a;
  __device__ uint32_t operator()(size_t k) const { return a[k]; }
};
struct LoadPair {
  const uint32_t *a, *b;
  __device__ uint2 operator()(size_t k) const { return make_uint2(a[k], b[k]); }
};
struct LoadIndexAnd {
  const uint32_t *b;
  __device__ uint2 operator()(size_t k) const { return make_uint2((uint32_t)k, b[k]); }
};
template <class V>
struct LoadPrev {
  const V *a;
  __device__ V operator()(size_t k) const { return a[k]; }
};

// Stable scatter of one pass.  Wave w of tile t owns items
// t*kITile + w*kITile/kIWaves + i*64 + lane (i = 0..kIItems-1): in that order
// each item's rank among equal digits comes from 8 ballots (lanes with the
// same digit) plus the wave's running digit count in LDS; waves of a tile
// follow each other, tiles follow the scanned digit-major offsets.  The tile
// is first reordered by digit in LDS, so the stores of each digit's run go
// out as consecutive addresses.
static_assert(kIB == kRDigits, "one lane per digit in the offset step");
template <class V, class Load>
__global__ void __launch_bounds__(kIB) radix_scatter_kernel(const uint32_t *__restrict__ kin, Load load,
                                                            uint32_t *__restrict__ kout, V *__restrict__ vout,
                                                            uint32_t m, uint32_t shift, uint32_t ntiles,
                                                            const uint32_t *__restrict__ offs) {
  __shared__ uint32_t s_cnt[kIWaves][kRDigits];
  __shared__ uint32_t s_delta[kRDigits];
  __shared__ uint32_t s_key[kITile];
  __shared__ V s_val[kITile];
  for (uint32_t i = threadIdx.x; i < (uint32_t)(kIWaves * kRDigits); i += kIB) (&s_cnt[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t w = threadIdx.x / kWave, l = lane_id();
  const size_t base = (size_t)blockIdx.x * kITile + (size_t)w * (kITile / kIWaves);
  const uint64_t below_mask = (1ull << l) - 1ull;
  uint32_t key[kIItems], rank[kIItems];
  V val[kIItems];
#pragma unroll
  for (int i = 0; i < kIItems; ++i) {
    const size_t k = base + (size_t)i * kWave + l;
    const bool ok = k < m;
    key[i] = ok ? kin[k] : 0u;
    val[i] = ok ? load(k) : V{};
    const uint32_t d = (key[i] >> shift) & (kRDigits - 1);
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < kRBits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t r = 0;
    if (ok) {
      const uint32_t below = (uint32_t)__popcll(peers & below_mask);
      const uint32_t c0 = s_cnt[w][d];
      r = c0 + below;
      if (below == 0) s_cnt[w][d] = c0 + (uint32_t)__popcll(peers);  // first lane of the digit
    }
    rank[i] = r;
  }
  __syncthreads();
  {
    // lane d: tile-local start of digit d, each wave's start inside it, and
    // the shift from tile position to output position
    const uint32_t d = threadIdx.x;
    uint32_t c[kIWaves], tot = 0;
#pragma unroll
    for (int x = 0; x < kIWaves; ++x) {
      c[x] = s_cnt[x][d];
      tot += c[x];
    }
    uint32_t all;
    const uint32_t start = block_excl_sum(tot, all);
    uint32_t run = start;
#pragma unroll
    for (int x = 0; x < kIWaves; ++x) {
      s_cnt[x][d] = run;
      run += c[x];
    }
    s_delta[d] = offs[(size_t)d * ntiles + blockIdx.x] - start;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kIItems; ++i) {
    const size_t k = base + (size_t)i * kWave + l;
    if (k < m) {
      const uint32_t lp = s_cnt[w][(key[i] >> shift) & (kRDigits - 1)] + rank[i];
      s_key[lp] = key[i];
      s_val[lp] = val[i];
    }
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)min((size_t)kITile, (size_t)m - (size_t)blockIdx.x * kITile);
  for (uint32_t j = threadIdx.x; j < cnt; j += kIB) {
    const uint32_t k = s_key[j];
    const uint32_t pos = s_delta[(k >> shift) & (kRDigits - 1)] + j;
    kout[pos] = k;
    vout[pos] = s_val[j];
  }
}

struct SortBufs {
  ScratchBuf k[2], v[2], counts;
  int cur = 0;
  uint32_t *keys() const { return k[cur].as<uint32_t>(); }
  template <class V>
  V *vals() const { return v[cur].as<V>(); }
};

// Stable sort of items 0..m-1 by keys_in (keys < 2^key_bits), carrying a
// value of type V read by load0(k) in the first pass.  keys_in is left
// unchanged; the result is b.keys() / b.vals<V>().
template <class V, class Load0>
int radix_sort(const uint32_t *keys_in, Load0 load0, uint32_t m, uint32_t key_bits, SortBufs &b) {
  const uint32_t ntiles = (uint32_t)(((size_t)m + kITile - 1) / kITile);
  for (int i = 0; i < 2; ++i) {
    OLV_TRY(b.k[i].alloc((size_t)m * 4));
    OLV_TRY(b.v[i].alloc((size_t)m * sizeof(V)));
  }
  OLV_TRY(b.counts.alloc(((size_t)kRDigits * ntiles + 1) * 4));
  const uint32_t passes = std::max(1u, (key_bits + kRBits - 1) / kRBits);
  hipStream_t s = stream();
  uint32_t *counts = b.counts.as<uint32_t>();
  for (uint32_t p = 0; p < passes && ntiles; ++p) {
    const int out = (int)(p & 1u);
    const uint32_t shift = p * kRBits;
    const uint32_t *kin = p == 0 ? keys_in : b.k[out ^ 1].as<uint32_t>();
    hipLaunchKernelGGL(radix_count_kernel, dim3(ntiles), dim3(kIB), 0, s, kin, m, shift, ntiles, counts);
    OLV_HIP(hipGetLastError());
    OLV_TRY(exclusive_scan(counts, counts, (size_t)kRDigits * ntiles));
    if (p == 0)
      hipLaunchKernelGGL((radix_scatter_kernel<V, Load0>), dim3(ntiles), dim3(kIB), 0, s, kin, load0,
                         b.k[out].as<uint32_t>(), b.v[out].as<V>(), m, shift, ntiles, counts);
    else
      hipLaunchKernelGGL((radix_scatter_kernel<V, LoadPrev<V>>), dim3(ntiles), dim3(kIB), 0, s, kin,
                         LoadPrev<V>{b.v[out ^ 1].as<V>()}, b.k[out].as<uint32_t>(), b.v[out].as<V>(), m, shift,
                         ntiles, counts);
    OLV_HIP(hipGetLastError());
    b.cur = out;
  }
  return OLV_OK;
}

// a[i] = in[i].x, b[i] = in[i].y
__global__ void split_kernel(const uint2 *__restrict__ in, uint32_t m, uint32_t *__restrict__ a,
                             uint32_t *__restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
    const uint2 x = in[i];
    a[i] = x.x;
    b[i] = x.y;
  }
}

uint32_t bits_for(uint32_t n) {  // bits of the largest id n - 1
  uint32_t b = 0;
  while (b < 32 && ((uint64_t)(n - 1) >> b) != 0) ++b;
  return b;
}

// ------------------------------------------------------------ CSR pieces
// rp[v] = first position whose key is >= v, for sorted keys < n (rp[n] = m):
// a binary search per vertex, so long runs of rows without entries (ids
// that no key holds) cost no more than any other rows.
__global__ void rows_from_sorted_kernel(const uint32_t *__restrict__ keys, uint32_t m, uint32_t n,
                                        uint32_t *__restrict__ rp) {
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v <= n; v += (size_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = m;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (keys[mid] < (uint32_t)v) lo = mid + 1;
      else hi = mid;
    }
    rp[v] = lo;
  }
}

// src[e] = v for every edge e of row v.
__global__ void expand_rows_kernel(const uint32_t *__restrict__ rp, uint32_t n, uint32_t *__restrict__ src) {
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x)
    for (uint32_t e = rp[v]; e < rp[v + 1]; ++e) src[e] = (uint32_t)v;
}

// Bitmap of rows without entries.
__global__ void empty_rows_kernel(const uint32_t *__restrict__ rp, uint32_t n, uint32_t *__restrict__ bm) {
  const uint32_t nw = (n + 31) / 32;
  for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += (size_t)gridDim.x * blockDim.x) {
    uint32_t bits = 0;
    for (uint32_t b = 0; b < 32; ++b) {
      const size_t v = w * 32 + b;
      if (v < n && rp[v + 1] == rp[v]) bits |= 1u << b;
    }
    bm[w] = bits;
  }
}

// *out = max row length (one atomic per workgroup).
__global__ void __launch_bounds__(kIB) max_row_kernel(const uint32_t *__restrict__ rp, uint32_t n, uint32_t *out) {
  uint32_t mx = 0;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x)
    mx = max(mx, rp[v + 1] - rp[v]);
  mx = wave_max(mx);
  __shared__ uint32_t s[kIWaves];
  if (lane_id() == 0) s[threadIdx.x / kWave] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kIWaves; ++i) mx = max(mx, s[i]);
    atomicMax(out, mx);
  }
}

// flag |= 1 if any id is >= n.
__global__ void check_ids_kernel(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b, uint32_t m,
                                 uint32_t n, uint32_t *flag) {
  bool bad = false;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
    bad = bad || a[i] >= n || b[i] >= n;
  if (__ballot(bad) && lane_id() == 0) atomicOr(flag, 1u);
}

__global__ void copy_word_kernel(const uint32_t *src, uint32_t *dst) { *dst = *src; }

int read_word(const uint32_t *d, uint32_t *h) {
  OLV_HIP(hipMemcpyAsync(h, d, 4, hipMemcpyDeviceToHost, stream()));
  OLV_HIP(hipStreamSynchronize(stream()));
  return OLV_OK;
}

int max_row(const uint32_t *rp, uint32_t n, uint32_t *out) {
  ScratchBuf d;
  OLV_TRY(d.alloc(4));
  OLV_TRY(fill_u32(d.as<uint32_t>(), 0u, 1));
  hipLaunchKernelGGL(max_row_kernel, dim3(grid_for(n, kIB)), dim3(kIB), 0, stream(), rp, n, d.as<uint32_t>());
  OLV_HIP(hipGetLastError());
  return read_word(d.as<uint32_t>(), out);
}

// Device out-CSR (row_ptr, col, weights) plus esrc (source of every CSR
// position) -> a graph; `esrc` may be empty (rebuilt from row_ptr).
int assemble(uint32_t n, uint32_t m, DevBuf &rp, DevBuf &col, DevBuf &w, int wdtype, const uint32_t *esrc,
             bool transpose, olv_graph **out) {
  hipStream_t s = stream();
  DevBuf inv, irp, icol, ieid, noin;
  ScratchBuf esrc_buf;
  uint32_t maxout = 0, maxin = 0;
  OLV_TRY(max_row(rp.as<uint32_t>(), n, &maxout));
  OLV_TRY(inv.alloc((size_t)n * 4));
  OLV_TRY(compute_inv_outdeg(rp.as<uint32_t>(), n, inv.as<float>()));
  if (transpose) {
    if (!esrc && m) {
      OLV_TRY(esrc_buf.alloc((size_t)m * 4));
      hipLaunchKernelGGL(expand_rows_kernel, dim3(grid_for(n, kIB)), dim3(kIB), 0, s, rp.as<uint32_t>(), n,
                         esrc_buf.as<uint32_t>());
      OLV_HIP(hipGetLastError());
      esrc = esrc_buf.as<uint32_t>();
    }
    OLV_TRY(irp.alloc(((size_t)n + 1) * 4));
    OLV_TRY(icol.alloc((size_t)m * 4));
    OLV_TRY(ieid.alloc((size_t)m * 4));
    if (m) {
      // CSR positions sorted by target, carrying (position, source)
      SortBufs b;
      OLV_TRY(radix_sort<uint2>(col.as<uint32_t>(), LoadIndexAnd{esrc}, m, bits_for(n), b));
      hipLaunchKernelGGL(rows_from_sorted_kernel, dim3(grid_for((size_t)n + 1, kIB)), dim3(kIB), 0, s, b.keys(), m,
                         n, irp.as<uint32_t>());
      hipLaunchKernelGGL(split_kernel, dim3(grid_for(m, kIB)), dim3(kIB), 0, s, b.vals<uint2>(), m,
                         ieid.as<uint32_t>(), icol.as<uint32_t>());
      OLV_HIP(hipGetLastError());
      OLV_HIP(hipStreamSynchronize(s));  // sort buffers are freed on return
    } else {
      OLV_TRY(fill_u32(irp.as<uint32_t>(), 0u, (size_t)n + 1));
    }
    const size_t nw = ((size_t)n + 31) / 32;
    OLV_TRY(noin.alloc(std::max<size_t>(nw, 1) * 4));
    hipLaunchKernelGGL(empty_rows_kernel, dim3(grid_for(nw, kIB)), dim3(kIB), 0, s, irp.as<uint32_t>(), n,
                       noin.as<uint32_t>());
    OLV_HIP(hipGetLastError());
    OLV_TRY(max_row(irp.as<uint32_t>(), n, &maxin));
  }
  OLV_HIP(hipStreamSynchronize(s));
  olv_graph *g = new olv_graph();
  g->device = current_device();
  g->n = n;
  g->m = m;
  g->max_outdeg = maxout;
  g->row_ptr = take<uint32_t>(rp);
  g->col = take<uint32_t>(col);
  if (wdtype != OLV_W_NONE) {
    g->weights = take<void>(w);
    g->wdtype = wdtype;
  }
  g->inv_outdeg = take<float>(inv);
  if (transpose) {
    g->max_indeg = maxin;
    g->in_row_ptr = take<uint32_t>(irp);
    g->in_col = take<uint32_t>(icol);
    g->in_eid = take<uint32_t>(ieid);
    g->no_in_bm = take<uint32_t>(noin);
    const int st = ensure_in_head(g);
    if (st != OLV_OK) {
      olv_graph_destroy(g);
      return st;
    }
  }
  *out = g;
  return OLV_OK;
}

// COO on the device (ids checked) -> graph.
int build_from_coo(uint32_t n, uint32_t m, const uint32_t *src, const uint32_t *dst, const uint32_t *w, int wdtype,
                   bool transpose, olv_graph **out) {
  hipStream_t s = stream();
  DevBuf rp, col, wout;
  OLV_TRY(rp.alloc(((size_t)n + 1) * 4));
  if (wdtype != OLV_W_NONE || m == 0) OLV_TRY(col.alloc((size_t)m * 4));
  if (wdtype != OLV_W_NONE) OLV_TRY(wout.alloc((size_t)m * 4));
  if (m == 0) {
    OLV_TRY(fill_u32(rp.as<uint32_t>(), 0u, (size_t)n + 1));
    return assemble(n, m, rp, col, wout, wdtype, nullptr, transpose, out);
  }
  // edges sorted by source, carrying the target (and the weight)
  SortBufs b;
  if (wdtype == OLV_W_NONE) {
    OLV_TRY(radix_sort<uint32_t>(src, LoadArr{dst}, m, bits_for(n), b));
    OLV_TRY(col.alloc((size_t)m * 4));  // the sorted targets are the column array
    OLV_HIP(hipMemcpyAsync(col.p, b.vals<uint32_t>(), (size_t)m * 4, hipMemcpyDeviceToDevice, s));
  } else {
    OLV_TRY(radix_sort<uint2>(src, LoadPair{dst, w}, m, bits_for(n), b));
    hipLaunchKernelGGL(split_kernel, dim3(grid_for(m, kIB)), dim3(kIB), 0, s, b.vals<uint2>(), m,
                       col.as<uint32_t>(), wout.as<uint32_t>());
  }
  hipLaunchKernelGGL(rows_from_sorted_kernel, dim3(grid_for((size_t)n + 1, kIB)), dim3(kIB), 0, s, b.keys(), m, n,
                     rp.as<uint32_t>());
  OLV_HIP(hipGetLastError());
  return assemble(n, m, rp, col, wout, wdtype, b.keys(), transpose, out);
}

// ------------------------------------------------------------------- text
__device__ __forceinline__ bool is_space(uint8_t c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Newlines per text tile (bytes at and beyond len are zero).
__global__ void __launch_bounds__(kIB) nl_count_kernel(const uint8_t *__restrict__ t, size_t len,
                                                       uint32_t *__restrict__ cnt) {
  const size_t c0 = (size_t)blockIdx.x * kTextTile + (size_t)threadIdx.x * kTextChunk;
  uint32_t c = 0;
  if (c0 < len) {
    const uint4 *q = reinterpret_cast<const uint4 *>(t + c0);
#pragma unroll
    for (int i = 0; i < (int)(kTextChunk / 16); ++i) {
      const uint4 x = q[i];
      const uint32_t wv[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int b = 0; b < 4; ++b) c += ((wv[j] >> (8 * b)) & 0xFFu) == '\n';
    }
  }
  uint32_t tot;
  (void)block_excl_sum(c, tot);
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

// starts[1 + r] = position after the r-th newline.
__global__ void __launch_bounds__(kIB) nl_write_kernel(const uint8_t *__restrict__ t, size_t len,
                                                       const uint32_t *__restrict__ ofs,
                                                       unsigned long long *__restrict__ starts) {
  const size_t c0 = (size_t)blockIdx.x * kTextTile + (size_t)threadIdx.x * kTextChunk;
  uint32_t wv[kTextChunk / 4];
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < (int)(kTextChunk / 4); ++i) wv[i] = 0u;
  if (c0 < len) {
    const uint4 *q = reinterpret_cast<const uint4 *>(t + c0);
#pragma unroll
    for (int i = 0; i < (int)(kTextChunk / 16); ++i) {
      const uint4 x = q[i];
      wv[4 * i] = x.x;
      wv[4 * i + 1] = x.y;
      wv[4 * i + 2] = x.z;
      wv[4 * i + 3] = x.w;
    }
  }
#pragma unroll
  for (int i = 0; i < (int)kTextChunk; ++i) c += ((wv[i / 4] >> (8 * (i % 4))) & 0xFFu) == '\n';
  uint32_t tot;
  uint32_t r = ofs[blockIdx.x] + block_excl_sum(c, tot);
#pragma unroll
  for (int i = 0; i < (int)kTextChunk; ++i)
    if (((wv[i / 4] >> (8 * (i % 4))) & 0xFFu) == '\n') starts[1 + (size_t)(r++)] = c0 + i + 1;
}

// Sequential reads of text bytes through aligned dword loads (the text
// buffer is padded past its end, so the dword of any byte < len exists).
struct ByteReader {
  const uint32_t *w;
  size_t cached = ~(size_t)0;
  uint32_t word = 0;
  __device__ explicit ByteReader(const uint8_t *t) : w(reinterpret_cast<const uint32_t *>(t)) {}
  __device__ __forceinline__ uint8_t operator[](size_t p) {
    if ((p >> 2) != cached) {
      cached = p >> 2;
      word = w[cached];
    }
    return (uint8_t)(word >> (8 * (p & 3)));
  }
};

// Bytes of one line read front to back: the 8 words (32 bytes) from the
// line's first word are requested together when the reader is made; each
// time the position passes a word, the window shifts by one and the word 8
// further on is requested, so no byte waits on the load of the one before.
// Positions must not decrease between reads.  Reads stay within the text
// buffer's zero padding of at least 64 bytes past its end.
struct LineReader {
  static constexpr int kW = 8;
  const uint32_t *w;
  size_t cw;  // word index of wd[0]
  uint32_t wd[kW];
  __device__ LineReader(const uint8_t *t, size_t b) : w(reinterpret_cast<const uint32_t *>(t)), cw(b >> 2) {
#pragma unroll
    for (int i = 0; i < kW; ++i) wd[i] = w[cw + i];
  }
  __device__ __forceinline__ uint8_t operator[](size_t p) {
    while ((p >> 2) > cw) {
#pragma unroll
      for (int i = 0; i < kW - 1; ++i) wd[i] = wd[i + 1];
      wd[kW - 1] = w[cw + kW];
      ++cw;
    }
    return (uint8_t)(wd[0] >> (8 * (p & 3)));
  }
};

// is_space for a byte inside a line (never '\n'): ' ' or '\t' '\v' '\f' '\r'
__device__ __forceinline__ bool is_space_in_line(uint8_t c) { return c == ' ' || (uint8_t)(c - '\t') <= 4u; }

// parse_int for a line read by a LineReader: the same tokens and values,
// with the first 9 digits accumulated in 32 bits.
__device__ __forceinline__ bool parse_int_line(LineReader &t, size_t &p, size_t e, long long &val) {
  bool neg = false;
  uint8_t ch = t[p];
  if (ch == '+' || ch == '-') {
    neg = ch == '-';
    ++p;
  }
  uint32_t lo = 0;
  long long v = 0;
  int nd = 0;
  bool ok = true;
  for (; p < e; ++p) {
    ch = t[p];
    if (is_space_in_line(ch)) break;
    const uint32_t dgt = (uint32_t)ch - '0';
    if (dgt > 9u || nd >= 18) ok = false;
    else if (nd < 9) lo = lo * 10u + dgt;
    else v = (nd == 9 ? (long long)lo : v) * 10 + dgt;
    ++nd;
  }
  const long long x = nd <= 9 ? (long long)lo : v;
  val = neg ? -x : x;
  return ok && nd > 0;
}

// Parse one token starting at t[p] (not whitespace, p < e) as an integer:
// [+-]digits, at most 18 digits.  Advances p past the token.
template <class Reader>
__device__ __forceinline__ bool parse_int(Reader &t, size_t &p, size_t e, long long &val) {
  bool neg = false;
  if (t[p] == '+' || t[p] == '-') {
    neg = t[p] == '-';
    ++p;
  }
  long long v = 0;
  int nd = 0;
  bool ok = true;
  for (; p < e && !is_space(t[p]); ++p) {
    const uint8_t ch = t[p];
    if (ch < '0' || ch > '9' || nd >= 18) ok = false;
    else v = v * 10 + (ch - '0');
    ++nd;
  }
  val = neg ? -v : v;
  return ok && nd > 0;
}

// Line k spans [starts[k], starts[k+1] - 1).  Code: 0 comment, 1 blank, 2 content.
__device__ __forceinline__ uint32_t line_code(const uint8_t *text, const unsigned long long *starts, size_t k,
                                              size_t &b, size_t &e) {
  ByteReader t(text);
  b = starts[k];
  e = starts[k + 1] - 1;
  while (b < e && is_space(t[b])) ++b;
  if (b == e) return 1u;
  return t[b] == '#' ? 0u : 2u;
}

// flag[k] = 1 for the lines that `want` selects (bit c set for code c).
__global__ void line_flags_kernel(const uint8_t *__restrict__ t, const unsigned long long *__restrict__ starts,
                                  uint32_t nl, uint32_t want, uint32_t *__restrict__ flag) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += (size_t)gridDim.x * blockDim.x) {
    size_t b, e;
    flag[k] = (want >> line_code(t, starts, k, b, e)) & 1u;
  }
}

// *first = smallest line index whose code is 2 (one atomic per workgroup).
__global__ void __launch_bounds__(kIB) first_content_kernel(const uint8_t *__restrict__ t,
                                                            const unsigned long long *__restrict__ starts,
                                                            uint32_t nl, uint32_t *first) {
  uint32_t best = 0xffffffffu;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nl; k += (size_t)gridDim.x * blockDim.x) {
    size_t b, e;
    if (line_code(t, starts, k, b, e) == 2u) {
      best = (uint32_t)k;
      break;  // later k of this lane are larger
    }
  }
  best = wave_min(best);
  __shared__ uint32_t s[kIWaves];
  if (lane_id() == 0) s[threadIdx.x / kWave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kIWaves; ++i) best = min(best, s[i]);
    if (best != 0xffffffffu) atomicMin(first, best);
  }
}

// Error record: (line << 4) | reason, the smallest line wins.
enum : uint32_t { kErrToken = 1, kErrFew = 2, kErrId = 3 };

__device__ __forceinline__ void report(unsigned long long *err, size_t line, uint32_t why) {
  atomicMin(err, ((unsigned long long)line << 4) | why);
}

struct EdgeParse {
  const uint8_t *t;
  const unsigned long long *starts;
  const uint32_t *didx;  // exclusive scan of the content flags
  uint32_t nl, n, m;     // lines, vertices, edges kept
  uint32_t *src, *dst;
  int32_t *w;
  uint32_t *min_tok;     // smallest token count over the kept edge lines
  uint32_t *w_range;     // set if a third column does not fit int32
  unsigned long long *err;
};

// One lane per content line: edge didx - 1 (the first content line is the header).
__global__ void __launch_bounds__(kIB) edge_parse_kernel(EdgeParse a) {
  uint32_t mtok = 0xffffffffu;
  bool wbad = false;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.nl; k += (size_t)gridDim.x * blockDim.x) {
    const uint32_t di = a.didx[k];
    if (a.didx[k + 1] == di || di == 0 || di > a.m) continue;  // not content, header, or beyond m
    size_t b = a.starts[k], e = a.starts[k + 1] - 1;
    LineReader t(a.t, b);
    long long v[3] = {0, 0, 0};
    uint32_t ntok = 0;
    bool ok = true;
    for (;;) {
      while (b < e && is_space_in_line(t[b])) ++b;
      if (b >= e) break;
      long long x;
      const bool okt = parse_int_line(t, b, e, x);
      if (ntok < 3) {  // columns after the third are not read
        ok = ok && okt;
        v[ntok] = x;
      }
      ++ntok;
    }
    if (!ok) {
      report(a.err, k, kErrToken);
      continue;
    }
    if (ntok < 2) {
      report(a.err, k, kErrFew);
      continue;
    }
    if (v[0] < 0 || v[0] >= a.n || v[1] < 0 || v[1] >= a.n) {
      report(a.err, k, kErrId);
      continue;
    }
    mtok = min(mtok, ntok);
    if (ntok >= 3 && (v[2] < INT32_MIN || v[2] > INT32_MAX)) wbad = true;
    a.src[di - 1] = (uint32_t)v[0];
    a.dst[di - 1] = (uint32_t)v[1];
    a.w[di - 1] = (int32_t)v[2];
  }
  mtok = wave_min(mtok);
  const bool anyw = __ballot(wbad) != 0;
  __shared__ uint32_t s[kIWaves];
  __shared__ uint32_t s_w;
  if (threadIdx.x == 0) s_w = 0;
  __syncthreads();
  if (lane_id() == 0) {
    s[threadIdx.x / kWave] = mtok;
    if (anyw) s_w = 1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kIWaves; ++i) mtok = min(mtok, s[i]);
    if (mtok != 0xffffffffu) atomicMin(a.min_tok, mtok);
    if (s_w) atomicOr(a.w_range, 1u);
  }
}

struct AdjParse {
  const uint8_t *t;
  const unsigned long long *starts;
  const uint32_t *bidx;  // exclusive scan of the non-comment flags
  uint32_t nl, n, hb;    // lines, vertices, body index of the header
  uint32_t *deg;         // pass 1: tokens per vertex
  const uint32_t *rp;    // pass 2: row offsets
  uint32_t *col;
  unsigned long long *err;
};

// One lane per non-comment line after the header: vertex bidx - hb - 1.
template <bool WRITE>
__global__ void __launch_bounds__(kIB) adj_parse_kernel(AdjParse a) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.nl; k += (size_t)gridDim.x * blockDim.x) {
    const uint32_t bi = a.bidx[k];
    if (a.bidx[k + 1] == bi || bi <= a.hb) continue;
    const uint32_t v = bi - a.hb - 1;
    if (v >= a.n) continue;
    size_t b = a.starts[k], e = a.starts[k + 1] - 1;
    ByteReader t(a.t);
    uint32_t ntok = 0;
    uint32_t base = WRITE ? a.rp[v] : 0u;
    for (;;) {
      while (b < e && is_space(t[b])) ++b;
      if (b >= e) break;
      long long x;
      const bool ok = parse_int(t, b, e, x);
      if (!WRITE) {
        if (!ok) report(a.err, k, kErrToken);
        else if (x < 1 || x > a.n) report(a.err, k, kErrId);
      } else {
        a.col[base + ntok] = (uint32_t)(x - 1);
      }
      ++ntok;
    }
    if (!WRITE) a.deg[v] = ntok;
  }
}

// Host side of the header line: two integers.
bool parse_header(const std::string &s, long long &a, long long &b) {
  long long v[2] = {0, 0};
  int nt = 0;
  size_t p = 0;
  while (nt < 2) {
    while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\r' || s[p] == '\v' || s[p] == '\f')) ++p;
    if (p >= s.size()) break;
    size_t q = p;
    bool neg = false;
    if (s[q] == '+' || s[q] == '-') neg = s[q++] == '-';
    long long x = 0;
    int nd = 0;
    bool ok = true;
    for (; q < s.size() && !(s[q] == ' ' || s[q] == '\t' || s[q] == '\r' || s[q] == '\v' || s[q] == '\f'); ++q) {
      if (s[q] < '0' || s[q] > '9' || nd >= 18) ok = false;
      else x = x * 10 + (s[q] - '0');
      ++nd;
    }
    if (!ok || nd == 0) return false;
    v[nt++] = neg ? -x : x;
    p = q;
  }
  a = v[0];
  b = v[1];
  return nt == 2;
}

// Device copy of the text (zero padded to whole lane chunks) and its line
// starts: starts[k] for k < nl, starts[nl] = len + 1.
struct Lines {
  ScratchBuf text, starts;
  size_t len = 0;
  uint32_t nl = 0;
  const uint8_t *t() const { return text.as<uint8_t>(); }
  const unsigned long long *s() const { return starts.as<unsigned long long>(); }
};

__global__ void set_u64_kernel(unsigned long long *p, unsigned long long v) { *p = v; }

int split_lines(const char *text, size_t len, bool host, Lines &L) {
  hipStream_t s = stream();
  const size_t cap = (len + kTextChunk - 1) / kTextChunk * kTextChunk + kTextChunk;
  OLV_TRY(L.text.alloc(cap));
  OLV_HIP(hipMemsetAsync(L.text.p, 0, cap, s));
  if (len)
    OLV_HIP(hipMemcpyAsync(L.text.p, text, len, host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
  L.len = len;
  const size_t nt = std::max<size_t>(1, (len + kTextTile - 1) / kTextTile);
  if (nt >= (1ull << 31)) return fail(OLV_ERANGE, "text of %zu bytes", len);
  ScratchBuf cnt;
  OLV_TRY(cnt.alloc((nt + 1) * 4));
  hipLaunchKernelGGL(nl_count_kernel, dim3((unsigned)nt), dim3(kIB), 0, s, L.t(), len, cnt.as<uint32_t>());
  OLV_HIP(hipGetLastError());
  OLV_TRY(exclusive_scan(cnt.as<uint32_t>(), cnt.as<uint32_t>(), nt));
  uint32_t nnl = 0;
  OLV_TRY(read_word(cnt.as<uint32_t>() + nt, &nnl));
  if ((uint64_t)nnl + 1 >= (1ull << 32) - 1) return fail(OLV_ERANGE, "more than 2^32 lines");
  L.nl = nnl + 1;
  OLV_TRY(L.starts.alloc(((size_t)L.nl + 1) * 8));
  unsigned long long *st = L.starts.as<unsigned long long>();
  hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s, st, 0ull);
  hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s, st + L.nl, (unsigned long long)len + 1);
  hipLaunchKernelGGL(nl_write_kernel, dim3((unsigned)nt), dim3(kIB), 0, s, L.t(), len, cnt.as<uint32_t>(), st);
  OLV_HIP(hipGetLastError());
  OLV_HIP(hipStreamSynchronize(s));
  return OLV_OK;
}

// Bytes of line k (without the newline) copied to the host.
int fetch_line(const Lines &L, uint32_t k, std::string &out) {
  unsigned long long se[2];
  OLV_HIP(hipMemcpy(se, L.s() + k, 16, hipMemcpyDeviceToHost));
  const size_t b = se[0], e = se[1] - 1, nb = std::min<size_t>(e - b, 4096);
  out.assign(nb, '\0');
  if (nb) OLV_HIP(hipMemcpy(&out[0], L.t() + b, nb, hipMemcpyDeviceToHost));
  return OLV_OK;
}

int first_content_line(const Lines &L, uint32_t *k) {
  ScratchBuf d;
  OLV_TRY(d.alloc(4));
  OLV_TRY(fill_u32(d.as<uint32_t>(), 0xffffffffu, 1));
  hipLaunchKernelGGL(first_content_kernel, dim3(grid_for(L.nl, kIB)), dim3(kIB), 0, stream(), L.t(), L.s(), L.nl,
                     d.as<uint32_t>());
  OLV_HIP(hipGetLastError());
  return read_word(d.as<uint32_t>(), k);
}

int parse_error(const char *who, unsigned long long e) {
  static const char *why[] = {"", "a token is not an integer", "fewer than two integers", "a vertex id is out of range"};
  return fail(OLV_EINVAL, "%s: line %llu: %s", who, (e >> 4) + 1, why[std::min<unsigned long long>(e & 15, 3)]);
}

int read_err(const ScratchBuf &err, unsigned long long *h) {
  OLV_HIP(hipMemcpy(h, err.p, 8, hipMemcpyDeviceToHost));
  return OLV_OK;
}

int init_err(ScratchBuf &err) {
  OLV_TRY(err.alloc(8));
  const unsigned long long none = ~0ull;
  OLV_HIP(hipMemcpy(err.p, &none, 8, hipMemcpyHostToDevice));
  return OLV_OK;
}

}  // namespace
}  // namespace olv

using namespace olv;

extern "C" {

int olv_graph_from_coo(olv_graph **out, uint32_t n, uint32_t m, const uint32_t *src, const uint32_t *dst,
                       const void *weights, int weight_dtype, unsigned flags, olv_stats *st) {
  if (!out) return fail(OLV_EINVAL, "olv_graph_from_coo: out is NULL");
  *out = nullptr;
  if (n == 0) return fail(OLV_EINVAL, "olv_graph_from_coo: graph has no vertices");
  if (m > 0 && (!src || !dst)) return fail(OLV_EINVAL, "olv_graph_from_coo: NULL edge array");
  if (weight_dtype != OLV_W_NONE && weight_dtype != OLV_W_I32 && weight_dtype != OLV_W_F32)
    return fail(OLV_EINVAL, "olv_graph_from_coo: unknown weight dtype %d", weight_dtype);
  if (weight_dtype != OLV_W_NONE && m > 0 && !weights) return fail(OLV_EINVAL, "olv_graph_from_coo: NULL weights");
  (void)current_device();
  if (st) std::memset(st, 0, sizeof(*st));
  Timer t;
  t.start();
  const bool host = (flags & OLV_PTR_HOST) != 0;
  ScratchBuf ds, dd, dw, flag;
  const uint32_t *s = src, *d = dst, *w = static_cast<const uint32_t *>(weights);
  if (host && m) {
    OLV_TRY(ds.alloc((size_t)m * 4));
    OLV_TRY(dd.alloc((size_t)m * 4));
    OLV_HIP(hipMemcpy(ds.p, src, (size_t)m * 4, hipMemcpyHostToDevice));
    OLV_HIP(hipMemcpy(dd.p, dst, (size_t)m * 4, hipMemcpyHostToDevice));
    s = ds.as<uint32_t>();
    d = dd.as<uint32_t>();
    if (weight_dtype != OLV_W_NONE) {
      OLV_TRY(dw.alloc((size_t)m * 4));
      OLV_HIP(hipMemcpy(dw.p, weights, (size_t)m * 4, hipMemcpyHostToDevice));
      w = dw.as<uint32_t>();
    }
  }
  if (m) {
    OLV_TRY(flag.alloc(4));
    OLV_TRY(fill_u32(flag.as<uint32_t>(), 0u, 1));
    hipLaunchKernelGGL(check_ids_kernel, dim3(grid_for(m, kIB)), dim3(kIB), 0, stream(), s, d, m, n,
                       flag.as<uint32_t>());
    OLV_HIP(hipGetLastError());
    uint32_t bad = 0;
    OLV_TRY(read_word(flag.as<uint32_t>(), &bad));
    if (bad) return fail(OLV_EINVAL, "olv_graph_from_coo: a vertex id is >= n (%u)", n);
  }
  OLV_TRY(build_from_coo(n, m, s, d, w, weight_dtype, (flags & OLV_BUILD_TRANSPOSE) != 0, out));
  t.stop();
  if (st) st->elapsed_ms = t.ms();
  return OLV_OK;
}

int olv_graph_from_edge_list_text(olv_graph **out, const char *text, size_t len, unsigned flags,
                                  olv_stats *st) {
  static const char *who = "olv_graph_from_edge_list_text";
  if (!out) return fail(OLV_EINVAL, "%s: out is NULL", who);
  *out = nullptr;
  if (!text && len) return fail(OLV_EINVAL, "%s: NULL text", who);
  (void)current_device();
  if (st) std::memset(st, 0, sizeof(*st));
  Timer t;
  t.start();
  hipStream_t s = stream();
  Lines L;
  OLV_TRY(split_lines(text, len, (flags & OLV_PTR_HOST) != 0, L));
  // content lines (code 2) numbered in order; the first is the header
  ScratchBuf didx;
  OLV_TRY(didx.alloc(((size_t)L.nl + 1) * 4));
  hipLaunchKernelGGL(line_flags_kernel, dim3(grid_for(L.nl, kIB)), dim3(kIB), 0, s, L.t(), L.s(), L.nl, 1u << 2,
                     didx.as<uint32_t>());
  OLV_HIP(hipGetLastError());
  OLV_TRY(exclusive_scan(didx.as<uint32_t>(), didx.as<uint32_t>(), L.nl));
  uint32_t ncontent = 0, hline = 0;
  OLV_TRY(read_word(didx.as<uint32_t>() + L.nl, &ncontent));
  if (ncontent == 0) return fail(OLV_EINVAL, "%s: no header line", who);
  OLV_TRY(first_content_line(L, &hline));
  std::string hdr;
  OLV_TRY(fetch_line(L, hline, hdr));
  long long hn = 0, hm = 0;
  if (!parse_header(hdr, hn, hm)) return fail(OLV_EINVAL, "%s: line %u: header is not \"n m\"", who, hline + 1);
  if (hn <= 0 || hm < 0) return fail(OLV_EINVAL, "%s: header n=%lld m=%lld", who, hn, hm);
  if (hn >= (1ll << 32) || hm >= (1ll << 32)) return fail(OLV_ERANGE, "%s: header n=%lld m=%lld", who, hn, hm);
  const uint32_t n = (uint32_t)hn;
  if ((long long)ncontent - 1 != hm)
    return fail(OLV_EINVAL, "%s: header says %lld edges, found %u edge lines", who, hm, ncontent - 1);
  const uint32_t m = (uint32_t)hm;
  ScratchBuf src, dst, w, mt, wr, err;
  OLV_TRY(src.alloc((size_t)m * 4));
  OLV_TRY(dst.alloc((size_t)m * 4));
  OLV_TRY(w.alloc((size_t)m * 4));
  OLV_TRY(mt.alloc(4));
  OLV_TRY(wr.alloc(4));
  OLV_TRY(fill_u32(mt.as<uint32_t>(), 0xffffffffu, 1));
  OLV_TRY(fill_u32(wr.as<uint32_t>(), 0u, 1));
  OLV_TRY(init_err(err));
  EdgeParse a;
  a.t = L.t();
  a.starts = L.s();
  a.didx = didx.as<uint32_t>();
  a.nl = L.nl;
  a.n = n;
  a.m = m;
  a.src = src.as<uint32_t>();
  a.dst = dst.as<uint32_t>();
  a.w = w.as<int32_t>();
  a.min_tok = mt.as<uint32_t>();
  a.w_range = wr.as<uint32_t>();
  a.err = err.as<unsigned long long>();
  Timer tk;
  tk.start();
  hipLaunchKernelGGL(edge_parse_kernel, dim3(grid_for(L.nl, kIB, OLV_PARSE_GRID_PER_CU)), dim3(kIB), 0, s, a);
  tk.stop();
  OLV_HIP(hipGetLastError());
  OLV_HIP(hipStreamSynchronize(s));
  if (st && (flags & OLV_PROFILE)) {
    st->kernel_ms = tk.ms();
    st->kernel_launches = 1;
  }
  unsigned long long e = 0;
  OLV_TRY(read_err(err, &e));
  if (e != ~0ull) return parse_error(who, e);
  uint32_t min_tok = 0, w_range = 0;
  OLV_TRY(read_word(mt.as<uint32_t>(), &min_tok));
  OLV_TRY(read_word(wr.as<uint32_t>(), &w_range));
  const bool has_w = m > 0 && min_tok != 0xffffffffu && min_tok >= 3;
  if (has_w && w_range) return fail(OLV_EINVAL, "%s: an edge weight does not fit int32", who);
  L.text.reset();  // release the text before the sort
  L.starts.reset();
  didx.reset();
  OLV_TRY(build_from_coo(n, m, src.as<uint32_t>(), dst.as<uint32_t>(), has_w ? w.as<uint32_t>() : nullptr,
                         has_w ? OLV_W_I32 : OLV_W_NONE, (flags & OLV_BUILD_TRANSPOSE) != 0, out));
  t.stop();
  if (st) {
    st->elapsed_ms = t.ms();
    st->iterations = L.nl;  // text lines
  }
  return OLV_OK;
}

int olv_graph_from_dimacs_text(olv_graph **out, const char *text, size_t len, unsigned flags, olv_stats *st) {
  static const char *who = "olv_graph_from_dimacs_text";
  if (!out) return fail(OLV_EINVAL, "%s: out is NULL", who);
  *out = nullptr;
  if (!text && len) return fail(OLV_EINVAL, "%s: NULL text", who);
  (void)current_device();
  if (st) std::memset(st, 0, sizeof(*st));
  Timer t;
  t.start();
  hipStream_t s = stream();
  Lines L;
  OLV_TRY(split_lines(text, len, (flags & OLV_PTR_HOST) != 0, L));
  // non-comment lines (blank or content) numbered in order
  ScratchBuf bidx;
  OLV_TRY(bidx.alloc(((size_t)L.nl + 1) * 4));
  hipLaunchKernelGGL(line_flags_kernel, dim3(grid_for(L.nl, kIB)), dim3(kIB), 0, s, L.t(), L.s(), L.nl,
                     (1u << 1) | (1u << 2), bidx.as<uint32_t>());
  OLV_HIP(hipGetLastError());
  OLV_TRY(exclusive_scan(bidx.as<uint32_t>(), bidx.as<uint32_t>(), L.nl));
  uint32_t hline = 0xffffffffu;
  OLV_TRY(first_content_line(L, &hline));
  if (hline == 0xffffffffu) return fail(OLV_EINVAL, "%s: no header line", who);
  std::string hdr;
  OLV_TRY(fetch_line(L, hline, hdr));
  long long hn = 0, hm = 0;
  if (!parse_header(hdr, hn, hm)) return fail(OLV_EINVAL, "%s: line %u: header is not \"n m\"", who, hline + 1);
  if (hn <= 0 || hm < 0) return fail(OLV_EINVAL, "%s: header n=%lld m=%lld", who, hn, hm);
  if (hn >= (1ll << 32) || 2 * hm >= (1ll << 32)) return fail(OLV_ERANGE, "%s: header n=%lld m=%lld", who, hn, hm);
  const uint32_t n = (uint32_t)hn, m = (uint32_t)(2 * hm);
  uint32_t hb = 0;
  OLV_TRY(read_word(bidx.as<uint32_t>() + hline, &hb));
  ScratchBuf deg, err;
  DevBuf rp, col, none;
  OLV_TRY(deg.alloc(((size_t)n + 1) * 4));
  OLV_TRY(fill_u32(deg.as<uint32_t>(), 0u, (size_t)n + 1));
  OLV_TRY(init_err(err));
  AdjParse a;
  a.t = L.t();
  a.starts = L.s();
  a.bidx = bidx.as<uint32_t>();
  a.nl = L.nl;
  a.n = n;
  a.hb = hb;
  a.deg = deg.as<uint32_t>();
  a.rp = nullptr;
  a.col = nullptr;
  a.err = err.as<unsigned long long>();
  hipLaunchKernelGGL(adj_parse_kernel<false>, dim3(grid_for(L.nl, kIB)), dim3(kIB), 0, s, a);
  OLV_HIP(hipGetLastError());
  OLV_HIP(hipStreamSynchronize(s));
  unsigned long long e = 0;
  OLV_TRY(read_err(err, &e));
  if (e != ~0ull) return parse_error(who, e);
  // row offsets; the neighbour count must be 2m (csrGraph.h:229-232)
  uint64_t total = 0;
  {
    OLV_TRY(rp.alloc(((size_t)n + 1) * 4));
    OLV_TRY(exclusive_scan(deg.as<uint32_t>(), rp.as<uint32_t>(), n));
    uint32_t t32 = 0;
    OLV_TRY(read_word(rp.as<uint32_t>() + n, &t32));
    total = t32;
  }
  if (total != m) return fail(OLV_EINVAL, "%s: %llu neighbour entries, header says 2*%lld", who,
                              (unsigned long long)total, hm);
  OLV_TRY(col.alloc((size_t)m * 4));
  a.rp = rp.as<uint32_t>();
  a.col = col.as<uint32_t>();
  hipLaunchKernelGGL(adj_parse_kernel<true>, dim3(grid_for(L.nl, kIB)), dim3(kIB), 0, s, a);
  OLV_HIP(hipGetLastError());
  OLV_HIP(hipStreamSynchronize(s));
  L.text.reset();
  L.starts.reset();
  bidx.reset();
  deg.reset();
  OLV_TRY(assemble(n, m, rp, col, none, OLV_W_NONE, nullptr, (flags & OLV_BUILD_TRANSPOSE) != 0, out));
  t.stop();
  if (st) {
    st->elapsed_ms = t.ms();
    st->iterations = L.nl;
  }
  return OLV_OK;
}

int olv_graph_load_file(olv_graph **out, const char *path, int format, unsigned flags, olv_stats *st) {
  if (!out || !path) return fail(OLV_EINVAL, "olv_graph_load_file: NULL argument");
  *out = nullptr;
  if (format != OLV_FMT_EDGE_LIST && format != OLV_FMT_DIMACS)
    return fail(OLV_EINVAL, "olv_graph_load_file: unknown format %d", format);
  FILE *f = std::fopen(path, "rb");
  if (!f) return fail(OLV_EINVAL, "olv_graph_load_file: cannot open %s", path);
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(f);
    return fail(OLV_EINVAL, "olv_graph_load_file: cannot size %s", path);
  }
  char *buf = nullptr;
  if (hipHostMalloc(reinterpret_cast<void **>(&buf), std::max<size_t>(sz, 1), hipHostMallocDefault) != hipSuccess) {
    std::fclose(f);
    return fail(OLV_ENOMEM, "olv_graph_load_file: %ld bytes of pinned memory", sz);
  }
  const size_t got = std::fread(buf, 1, (size_t)sz, f);
  std::fclose(f);
  int rc = OLV_OK;
  if (got != (size_t)sz) rc = fail(OLV_EINVAL, "olv_graph_load_file: short read of %s", path);
  const unsigned fl = flags | OLV_PTR_HOST;
  if (rc == OLV_OK)
    rc = format == OLV_FMT_EDGE_LIST ? olv_graph_from_edge_list_text(out, buf, got, fl, st)
                                     : olv_graph_from_dimacs_text(out, buf, got, fl, st);
  (void)hipHostFree(buf);
  return rc;
}

int olv_graph_csr_to_host(const olv_graph *g, uint32_t *row_ptr, uint32_t *col, void *weights, int *weight_dtype) {
  if (!g) return fail(OLV_EINVAL, "olv_graph_csr_to_host: NULL graph");
  if (row_ptr) OLV_HIP(hipMemcpy(row_ptr, g->row_ptr, ((size_t)g->n + 1) * 4, hipMemcpyDeviceToHost));
  if (col && g->m) OLV_HIP(hipMemcpy(col, g->col, (size_t)g->m * 4, hipMemcpyDeviceToHost));
  if (weights && g->weights && g->m) OLV_HIP(hipMemcpy(weights, g->weights, (size_t)g->m * 4, hipMemcpyDeviceToHost));
  if (weight_dtype) *weight_dtype = g->wdtype;
  return OLV_OK;
}

}  // extern "C"
