// trn_dropin.cpp -- device context and topology handles of the drop-in.
#include "trn_dropin.hpp"

#include <mutex>

namespace trinity_amd_dropin {

trn_ctx* context() {
  static std::once_flag once;
  static trn_ctx* ctx = nullptr;
  std::call_once(once, [] {
    const char* env = std::getenv("TRINITY_AMD_DEVICE");
    const int device = env ? std::atoi(env) : 0;
    check(trn_ctx_create(device, &ctx), "trn_ctx_create");
  });
  return ctx;
}

void check(int rc, const char* what) {
  if (rc != TRN_OK) {
    std::fprintf(stderr, "trinity_amd: %s failed (%d): %s\n", what, rc, trn_last_error());
    std::exit(EXIT_FAILURE);
  }
}

DeviceTopology::DeviceTopology(int32_t nv, int32_t nt, const int32_t* tri, const uint8_t* tags) {
  // tri and tags may be host pointers: the library copies them
  check(trn_mesh_create(context(), nv, nt, tri, tags, nullptr, &m_), "trn_mesh_create");
}

DeviceTopology::~DeviceTopology() {
  if (m_) trn_mesh_destroy(m_);
}

}  // namespace trinity_amd_dropin
