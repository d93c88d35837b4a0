// GPU-resident dataset of the Olive C++ API (grd.h:43-170 of the
// reference): a host buffer mirrored by a device buffer.  Device memory,
// fills and copies go through libolive_hip.so; allTo() fills on the device
// instead of the reference's host loop + full H2D copy (grd.h:90-97).
#ifndef OLIVE_GRD_H
#define OLIVE_GRD_H

#include <cstring>
#include <iostream>

#include "common.h"

template <typename T>
class GRD {
 public:
  T *elemsHost;    // host buffer
  T *elemsDevice;  // device buffer
  size_t length;
  int deviceId;

  GRD() : elemsHost(NULL), elemsDevice(NULL), length(0), deviceId(-1) {}

  /** Host element (device element inside device code). */
  __host__ __device__ inline T &operator[](size_t index) const {
#if defined(__HIP_DEVICE_COMPILE__)
    return elemsDevice[index];
#else
    return elemsHost[index];
#endif
  }

  inline size_t capacity() const { return length; }

  /** Allocate `len` elements on host and device (len 0 is allowed). */
  inline void reserve(size_t len, int id = 0) {
    deviceId = id;
    length = len;
    const size_t bytes = (len ? len : 1) * sizeof(T);
    elemsHost = reinterpret_cast<T *>(malloc(bytes));
    void *d = NULL;
    OLV_CHECK(olv_malloc(&d, bytes));
    elemsDevice = reinterpret_cast<T *>(d);
  }

  /** Set every element to x on the device (host mirror untouched). */
  void allTo(T x) {
    if (!length) return;
    if (sizeof(T) == 1 || sizeof(T) == 4 || sizeof(T) == 8) {
      uint64_t bits = 0;
      memcpy(&bits, &x, sizeof(T));
      OLV_CHECK(olv_fill(elemsDevice, bits, length, (int)sizeof(T)));
    } else {
      for (size_t i = 0; i < length; i++) elemsHost[i] = x;
      cache();
    }
  }

  /** Zero every byte on the device. */
  void clear() {
    if (length) OLV_CHECK(olv_fill(elemsDevice, 0, length * sizeof(T), 1));
  }

  /** elements[i] = x on host and device. */
  void set(size_t i, T x) {
    elemsHost[i] = x;
    OLV_CHECK(olv_memcpy_h2d(elemsDevice + i, elemsHost + i, sizeof(T)));
  }

  /** Device -> host. */
  inline void persist() {
    if (length) OLV_CHECK(olv_memcpy_d2h(elemsHost, elemsDevice, length * sizeof(T)));
  }

  /** Host -> device. */
  inline void cache() {
    if (length) OLV_CHECK(olv_memcpy_h2d(elemsDevice, elemsHost, length * sizeof(T)));
  }

  /** Call T::print() on every host element. */
  inline void print() {
    for (size_t i = 0; i < length; i++) elemsHost[i].print();
  }

  inline void peek() {
    persist();
    for (size_t i = 0; i < length; i++) std::cout << elemsHost[i] << " ";
    printf("\n");
  }

  inline void del() {
    if (deviceId < 0) return;
    free(elemsHost);
    if (elemsDevice) olv_free(elemsDevice);
    elemsHost = NULL;
    elemsDevice = NULL;
    length = 0;
    deviceId = -1;
  }
};

#endif  // OLIVE_GRD_H
