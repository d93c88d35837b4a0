// Result egress.
//
// * Text: one value per line in the reference programs' formats -- "%d\n"
//   (BFS.cu:41-43, SSSP.cu:41-43) and "%f\n" (PageRank.cu:42-44) -- formatted
//   on the device (a length pass, a scan, a write pass) and streamed to the
//   file.  This replaces
//   Oliver::printVertices (oliver.h:204-207: a blocking copy of every vertex
//   value, GRD::persist grd.h:118-124, then one fprintf per vertex,
//   grd.h:130-134).
// * Binary: a 24-byte header (magic "OLVR", version, dtype, element size,
//   count) and the raw values.  Text and binary bytes leave the device
//   through two 8 MB pinned buffers, so the copy of one piece overlaps the
//   file write of the previous one; dumps are read back the same way.
//
// "%f" is formatted exactly as C's printf prints a double: the exact binary
// value rounded to six decimals, ties to even, "-" for a set sign bit
// (also on zero), "nan" / "inf" with their sign.

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "device.hpp"
#include "internal.hpp"

namespace olv {
namespace {

constexpr int kEB = 256;
constexpr size_t kTextChunkValues = size_t(1) << 23;   // values formatted per round
constexpr size_t kStageBytes = size_t(8) << 20;        // bytes per pinned staging buffer
constexpr uint32_t kMaxFieldChars = 32;                 // longest line of any format

unsigned grid_for(size_t items) {
  const size_t b = (items + kEB - 1) / kEB;
  return (unsigned)std::max<size_t>(1, std::min(b, (size_t)cu_count() * 16));
}

// Writes (or, with out == nullptr, only counts) the decimal digits of v.
__device__ __forceinline__ uint32_t put_u64(unsigned long long v, char *out) {
  uint32_t n = 1;
  for (unsigned long long t = v; t >= 10; t /= 10) ++n;
  if (out)
    for (uint32_t i = n; i-- > 0;) {
      out[i] = (char)('0' + v % 10);
      v /= 10;
    }
  return n;
}

__device__ __forceinline__ uint32_t put_str(const char *s, char *out) {
  uint32_t n = 0;
  for (; s[n]; ++n)
    if (out) out[n] = s[n];
  return n;
}

// "%d\n"
__device__ __forceinline__ uint32_t fmt_d(int32_t x, char *out) {
  uint32_t n = 0;
  long long v = x;
  if (v < 0) {
    if (out) out[0] = '-';
    n = 1;
    v = -v;
  }
  n += put_u64((unsigned long long)v, out ? out + n : nullptr);
  if (out) out[n] = '\n';
  return n + 1;
}

// "%f\n" of a double; returns 0 if |x| >= 2^63 (not supported).
__device__ __forceinline__ uint32_t fmt_f(double x, char *out) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
  const bool neg = bits >> 63;
  const uint32_t ex = (uint32_t)(bits >> 52) & 0x7FFu;
  const unsigned long long frac_bits = bits & ((1ull << 52) - 1);
  uint32_t n = 0;
  if (neg) {
    if (out) out[0] = '-';
    n = 1;
  }
  if (ex == 0x7FFu) {
    n += put_str(frac_bits ? "nan" : "inf", out ? out + n : nullptr);
    if (out) out[n] = '\n';
    return n + 1;
  }
  // |x| = m * 2^e
  const unsigned long long m = ex ? (frac_bits | (1ull << 52)) : frac_bits;
  const int e = ex ? (int)ex - 1075 : -1074;
  unsigned long long ip = 0, fr = 0;  // integer part, six decimals
  if (m != 0) {
    if (e >= 0) {
      if (e > 10) return 0;  // |x| >= 2^63
      ip = m << e;
    } else {
      const int s = -e;
      unsigned long long f = m;  // fraction numerator over 2^s
      if (s < 64) {
        ip = m >> s;
        f = m & ((1ull << s) - 1);
      }
      // fr = round_half_even(f * 10^6 / 2^s); f < 2^53, so the product < 2^73
      if (s <= 74) {
        const unsigned __int128 p = (unsigned __int128)f * 1000000u;
        const unsigned __int128 q = p >> s;
        const unsigned __int128 r = p - (q << s);
        const unsigned __int128 half = (unsigned __int128)1 << (s - 1);
        fr = (unsigned long long)q;
        if (r > half || (r == half && (fr & 1))) ++fr;
      }  // s > 74: f * 10^6 < 2^73 < 2^(s-1), rounds to 0
      if (fr == 1000000) {
        fr = 0;
        ++ip;
      }
    }
  }
  n += put_u64(ip, out ? out + n : nullptr);
  if (out) {
    out[n] = '.';
    for (int i = 6; i >= 1; --i) {
      out[n + i] = (char)('0' + fr % 10);
      fr /= 10;
    }
    out[n + 7] = '\n';
  }
  return n + 8;
}

__device__ __forceinline__ uint32_t fmt_value(const uint8_t *p, int fmt, char *out) {
  switch (fmt) {
    case OLV_TXT_I32:
      return fmt_d(*reinterpret_cast<const int32_t *>(p), out);
    case OLV_TXT_F32:
      return fmt_f((double)*reinterpret_cast<const float *>(p), out);
    default:
      return fmt_f(*reinterpret_cast<const double *>(p), out);
  }
}

// len[i] = characters of value i (0 = cannot be formatted).
__global__ void text_len_kernel(const uint8_t *__restrict__ base, size_t count, size_t stride, int fmt,
                                uint32_t *__restrict__ len, uint32_t *bad) {
  bool b = false;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t l = fmt_value(base + i * stride, fmt, nullptr);
    len[i] = l;
    b = b || l == 0;
  }
  if (__ballot(b) && lane_id() == 0) atomicOr(bad, 1u);
}

__global__ void text_write_kernel(const uint8_t *__restrict__ base, size_t count, size_t stride, int fmt,
                                  const uint32_t *__restrict__ off, char *__restrict__ out) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    (void)fmt_value(base + i * stride, fmt, out + off[i]);
}

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

struct Pinned {
  void *p = nullptr;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
  int alloc(size_t bytes) {
    if (hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return fail(OLV_ENOMEM, "pinned host allocation of %zu bytes failed", bytes);
    }
    return OLV_OK;
  }
};

// Device bytes to a file through two pinned buffers: the copy of piece k+1
// runs while piece k is written.
struct Stager {
  Pinned buf[2];
  hipEvent_t ev[2] = {nullptr, nullptr};
  size_t cap = 0;
  ~Stager() {
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  int init(size_t bytes) {
    cap = std::max<size_t>(bytes, 1);
    OLV_TRY(buf[0].alloc(cap));
    OLV_TRY(buf[1].alloc(cap));
    for (auto &e : ev) OLV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return OLV_OK;
  }
  int write(const uint8_t *src, size_t total, FILE *f, const char *who) {
    hipStream_t s = stream();
    int st = OLV_OK;
    size_t off = 0, k = 0;
    if (total && (hipMemcpyAsync(buf[0].p, src, std::min(cap, total), hipMemcpyDeviceToHost, s) != hipSuccess ||
                  hipEventRecord(ev[0], s) != hipSuccess))
      st = fail(OLV_EHIP, "%s: copy failed", who);
    while (st == OLV_OK && off < total) {
      const size_t c = std::min(cap, total - off);
      const size_t next = off + c;
      if (next < total && (hipMemcpyAsync(buf[(k + 1) & 1].p, src + next, std::min(cap, total - next),
                                          hipMemcpyDeviceToHost, s) != hipSuccess ||
                           hipEventRecord(ev[(k + 1) & 1], s) != hipSuccess))
        st = fail(OLV_EHIP, "%s: copy failed", who);
      if (st == OLV_OK && hipEventSynchronize(ev[k & 1]) != hipSuccess) st = fail(OLV_EHIP, "%s: copy failed", who);
      if (st == OLV_OK && std::fwrite(buf[k & 1].p, 1, c, f) != c) st = fail(OLV_EINVAL, "%s: file write failed", who);
      off = next;
      ++k;
    }
    (void)hipStreamSynchronize(s);  // no copy into the buffers outlives a failure
    return st;
  }
};

struct BinHeader {
  char magic[4];
  uint32_t version, dtype, elem_bytes;
  uint64_t count;
};
static_assert(sizeof(BinHeader) == 24, "binary header layout");

}  // namespace
}  // namespace olv

using namespace olv;

extern "C" {

int olv_write_text(const char *path, const void *values, size_t count, size_t stride, int fmt, unsigned flags,
                   olv_stats *st) {
  if (!path || (!values && count)) return fail(OLV_EINVAL, "olv_write_text: NULL argument");
  if (fmt != OLV_TXT_I32 && fmt != OLV_TXT_F32 && fmt != OLV_TXT_F64)
    return fail(OLV_EINVAL, "olv_write_text: unknown format %d", fmt);
  const size_t esz = fmt == OLV_TXT_F64 ? 8 : 4;
  if (stride == 0) stride = esz;
  if (stride < esz || stride % esz || reinterpret_cast<uintptr_t>(values) % esz)
    return fail(OLV_EINVAL, "olv_write_text: values must be %zu-byte aligned with a stride that is a multiple of %zu",
                esz, esz);
  (void)current_device();
  if (st) std::memset(st, 0, sizeof(*st));
  File file;
  file.f = std::fopen(path, "wb");
  if (!file.f) return fail(OLV_EINVAL, "olv_write_text: cannot open %s", path);
  if (count == 0) return OLV_OK;
  uint64_t chars = 0;
  hipStream_t s = stream();
  const bool host = (flags & OLV_PTR_HOST) != 0;
  const size_t chunk = std::min(count, kTextChunkValues);
  ScratchBuf staged, len, text, bad;
  size_t text_cap = 0;
  OLV_TRY(len.alloc((chunk + 1) * 4));
  OLV_TRY(bad.alloc(4));
  if (host) OLV_TRY(staged.alloc(chunk * stride));
  Stager out;
  OLV_TRY(out.init(std::min(kStageBytes, chunk * kMaxFieldChars)));
  for (size_t i0 = 0; i0 < count; i0 += chunk) {
    const size_t c = std::min(chunk, count - i0);
    const uint8_t *src = static_cast<const uint8_t *>(values) + i0 * stride;
    if (host) {
      OLV_HIP(hipMemcpyAsync(staged.p, src, (c - 1) * stride + esz, hipMemcpyHostToDevice, s));
      src = staged.as<uint8_t>();
    }
    OLV_TRY(fill_u32(bad.as<uint32_t>(), 0u, 1));
    hipLaunchKernelGGL(text_len_kernel, dim3(grid_for(c)), dim3(kEB), 0, s, src, c, stride, fmt, len.as<uint32_t>(),
                       bad.as<uint32_t>());
    OLV_HIP(hipGetLastError());
    OLV_TRY(exclusive_scan(len.as<uint32_t>(), len.as<uint32_t>(), c));
    uint32_t total = 0, b = 0;
    OLV_HIP(hipMemcpy(&total, len.as<uint32_t>() + c, 4, hipMemcpyDeviceToHost));
    OLV_HIP(hipMemcpy(&b, bad.p, 4, hipMemcpyDeviceToHost));
    if (b) return fail(OLV_EUNSUPPORTED, "olv_write_text: a value of magnitude >= 2^63 cannot be formatted");
    if (total > text_cap) {  // the text buffer is sized by the scan
      text.reset();
      OLV_TRY(text.alloc(total));
      text_cap = total;
    }
    Timer tk;
    tk.start();
    hipLaunchKernelGGL(text_write_kernel, dim3(grid_for(c)), dim3(kEB), 0, s, src, c, stride, fmt,
                       len.as<uint32_t>(), text.as<char>());
    tk.stop();
    OLV_HIP(hipGetLastError());
    if (st && (flags & OLV_PROFILE)) {
      st->kernel_ms += tk.ms();
      st->kernel_launches += 1;
    }
    chars += total;
    OLV_TRY(out.write(text.as<uint8_t>(), total, file.f, "olv_write_text"));
  }
  if (std::fclose(file.f) != 0) {
    file.f = nullptr;
    return fail(OLV_EINVAL, "olv_write_text: closing %s failed", path);
  }
  file.f = nullptr;
  if (st) st->edges_inspected = chars;  // text bytes written
  return OLV_OK;
}

int olv_write_binary(const char *path, const void *values, size_t count, size_t elem_bytes, int dtype,
                     unsigned flags) {
  if (!path || (!values && count) || elem_bytes == 0) return fail(OLV_EINVAL, "olv_write_binary: bad argument");
  (void)current_device();
  File file;
  file.f = std::fopen(path, "wb");
  if (!file.f) return fail(OLV_EINVAL, "olv_write_binary: cannot open %s", path);
  BinHeader h;
  std::memcpy(h.magic, "OLVR", 4);
  h.version = 1;
  h.dtype = (uint32_t)dtype;
  h.elem_bytes = (uint32_t)elem_bytes;
  h.count = count;
  if (std::fwrite(&h, sizeof(h), 1, file.f) != 1) return fail(OLV_EINVAL, "olv_write_binary: write to %s failed", path);
  const size_t total = count * elem_bytes;
  const uint8_t *src = static_cast<const uint8_t *>(values);
  if (flags & OLV_PTR_HOST) {
    if (total && std::fwrite(src, 1, total, file.f) != total)
      return fail(OLV_EINVAL, "olv_write_binary: write to %s failed", path);
  } else if (total) {
    Stager out;
    OLV_TRY(out.init(std::min(total, kStageBytes)));
    OLV_TRY(out.write(src, total, file.f, "olv_write_binary"));
  }
  if (std::fclose(file.f) != 0) {
    file.f = nullptr;
    return fail(OLV_EINVAL, "olv_write_binary: closing %s failed", path);
  }
  file.f = nullptr;
  return OLV_OK;
}

int olv_read_binary(const char *path, void *values, size_t capacity_bytes, uint64_t *count, uint32_t *elem_bytes,
                    int *dtype, unsigned flags) {
  if (!path) return fail(OLV_EINVAL, "olv_read_binary: NULL path");
  File file;
  file.f = std::fopen(path, "rb");
  if (!file.f) return fail(OLV_EINVAL, "olv_read_binary: cannot open %s", path);
  BinHeader h;
  if (std::fread(&h, sizeof(h), 1, file.f) != 1 || std::memcmp(h.magic, "OLVR", 4) != 0 || h.version != 1 ||
      h.elem_bytes == 0)
    return fail(OLV_EINVAL, "olv_read_binary: %s is not a result dump", path);
  if (count) *count = h.count;
  if (elem_bytes) *elem_bytes = h.elem_bytes;
  if (dtype) *dtype = (int)h.dtype;
  if (!values) return OLV_OK;  // header only
  const size_t total = (size_t)h.count * h.elem_bytes;
  if (total > capacity_bytes) return fail(OLV_ERANGE, "olv_read_binary: %zu bytes do not fit in %zu", total, capacity_bytes);
  uint8_t *dst = static_cast<uint8_t *>(values);
  if (flags & OLV_PTR_HOST) {
    if (total && std::fread(dst, 1, total, file.f) != total) return fail(OLV_EINVAL, "olv_read_binary: short file %s", path);
    return OLV_OK;
  }
  (void)current_device();
  hipStream_t s = stream();
  const size_t chunk = std::min(std::max<size_t>(total, 1), kStageBytes);
  Pinned buf[2];
  OLV_TRY(buf[0].alloc(chunk));
  OLV_TRY(buf[1].alloc(chunk));
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto &e : ev) OLV_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  int st = OLV_OK;
  size_t k = 0;
  for (size_t off = 0; st == OLV_OK && off < total; off += chunk, ++k) {
    const size_t c = std::min(chunk, total - off);
    // buffer k & 1 was last used by the copy two chunks back
    if (k >= 2 && hipEventSynchronize(ev[k & 1]) != hipSuccess) st = fail(OLV_EHIP, "olv_read_binary: copy failed");
    if (st == OLV_OK && std::fread(buf[k & 1].p, 1, c, file.f) != c) st = fail(OLV_EINVAL, "olv_read_binary: short file %s", path);
    if (st == OLV_OK && (hipMemcpyAsync(dst + off, buf[k & 1].p, c, hipMemcpyHostToDevice, s) != hipSuccess ||
                         hipEventRecord(ev[k & 1], s) != hipSuccess))
      st = fail(OLV_EHIP, "olv_read_binary: copy failed");
  }
  (void)hipStreamSynchronize(s);
  for (auto &e : ev) (void)hipEventDestroy(e);
  return st;
}

}  // extern "C"
