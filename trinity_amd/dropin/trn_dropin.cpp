// trn_dropin.cpp -- device context and the shared device mesh of the drop-in.
#include "trn_dropin.hpp"

#include <cstring>
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

DeviceMesh& DeviceMesh::get() {
  static DeviceMesh* self = new DeviceMesh();
  return *self;
}

trn_mesh* DeviceMesh::topology(int32_t nv, int32_t nt, const int32_t* elems,
                               const uint8_t* tags, bool* built) {
  const size_t ne = 3 * (size_t)nt;
  const bool same = mesh_ != nullptr && nv == nv_ && nt == nt_ &&
                    std::memcmp(elems_.data(), elems, ne * sizeof(int32_t)) == 0 &&
                    std::memcmp(tags_.data(), tags, (size_t)nv) == 0;
  if (built) *built = !same;
  if (same) return mesh_;
  if (mesh_) trn_mesh_destroy(mesh_);
  mesh_ = nullptr;
  // elems and tags are host pointers: the library copies them
  check(trn_mesh_create(context(), nv, nt, elems, tags, nullptr, &mesh_), "trn_mesh_create");
  nv_ = nv;
  nt_ = nt;
  elems_.assign(elems, elems + ne);
  tags_.assign(tags, tags + nv);
  return mesh_;
}

}  // namespace trinity_amd_dropin
