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

// Device buffer owned for the duration of one pass.
template <typename T>
class DeviceArray {
 public:
  explicit DeviceArray(size_t n) : n_(n) {
    if (n_ && hipMalloc(reinterpret_cast<void**>(&p_), n_ * sizeof(T)) != hipSuccess) {
      std::fprintf(stderr, "trinity_amd: hipMalloc of %zu bytes failed\n", n_ * sizeof(T));
      std::exit(EXIT_FAILURE);
    }
  }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  ~DeviceArray() {
    if (p_) (void)hipFree(p_);
  }
  T* get() const { return p_; }
  void upload(const T* host) {
    if (n_) check(hipMemcpy(p_, host, n_ * sizeof(T), hipMemcpyHostToDevice) == hipSuccess ? 0 : -2,
                  "hipMemcpy host->device");
  }
  void download(T* host) const {
    if (n_) check(hipMemcpy(host, p_, n_ * sizeof(T), hipMemcpyDeviceToHost) == hipSuccess ? 0 : -2,
                  "hipMemcpy device->host");
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// RAII handle of a device topology built from the mesh's element array.
class DeviceTopology {
 public:
  DeviceTopology(int32_t nv, int32_t nt, const int32_t* tri, const uint8_t* tags);
  ~DeviceTopology();
  trn_mesh* get() const { return m_; }

 private:
  trn_mesh* m_ = nullptr;
};

}  // namespace trinity_amd_dropin
