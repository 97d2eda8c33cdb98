// trn_dropin.hpp -- helpers shared by the drop-in sources that re-implement
// trinity's Metrics::run and Smooth::run on top of the C ABI
// (include/trinity_amd.h).  Compiled against the reference headers.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "trinity_amd.h"

namespace trinity_amd_dropin {

// One context per process, on the device named by TRINITY_AMD_DEVICE
// (default 0), created on first use and kept until exit.
trn_ctx* context();

// The reference has no error channel (void methods, asserts compiled out),
// so a failing device call ends the process with the library's message.
void check(int rc, const char* what);

// Device array that grows on demand and is kept between passes.
template <typename T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;

  // At least n elements; the contents are not kept when it grows.
  T* reserve(size_t n) {
    if (n > cap_) {
      if (p_) (void)hipFree(p_);
      p_ = nullptr;
      cap_ = 0;
      if (hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)) != hipSuccess) {
        std::fprintf(stderr, "trinity_amd: hipMalloc of %zu bytes failed\n", n * sizeof(T));
        std::exit(EXIT_FAILURE);
      }
      cap_ = n;
    }
    return p_;
  }
  T* upload(const T* host, size_t n) {
    reserve(n);
    if (n) check(hipMemcpy(p_, host, n * sizeof(T), hipMemcpyHostToDevice) == hipSuccess ? 0 : -2,
                 "hipMemcpy host->device");
    return p_;
  }
  void download(T* host, size_t n) const {
    if (n) check(hipMemcpy(host, p_, n * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess ? 0 : -2,
                 "hipMemcpy device->host");
  }
  T* get() const { return p_; }

 private:
  T* p_ = nullptr;
  size_t cap_ = 0;
};

// Device state shared by Metrics::run and Smooth::run of one process: the
// topology built from the mesh's element array and tags, rebuilt only when
// those changed since the previous pass (in the reference's loop a
// remeshing round changes them between two smoothings; Metrics::run and the
// next Smooth::run on an unchanged mesh share one build), and the vertex and
// element arrays of the passes.  Created on first use and kept until exit
// (never destroyed: the HIP runtime may be gone when static objects die).
class DeviceMesh {
 public:
  static DeviceMesh& get();

  // Device topology of (elems [3nt], tags [nv]); *built (may be NULL) tells
  // whether this call rebuilt it.
  trn_mesh* topology(int32_t nv, int32_t nt, const int32_t* elems, const uint8_t* tags,
                     bool* built = nullptr);

  DeviceBuffer<double> xy, tensor, solut, qualit;
  DeviceBuffer<int32_t> color, order;

 private:
  DeviceMesh() = default;
  trn_mesh* mesh_ = nullptr;
  int32_t nv_ = -1, nt_ = -1;
  std::vector<int32_t> elems_;
  std::vector<uint8_t> tags_;
};

}  // namespace trinity_amd_dropin
