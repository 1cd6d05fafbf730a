// Library context: one device, one stream, one workspace pool, last error text.
#pragma once

#include <string>

#include "vrc_internal.hpp"

struct vrc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  vrc::Workspace ws;
  std::string last_error;
};

namespace vrc {

// Runs f, converting exceptions into the C-ABI status convention:
// 0 ok, -1 bad argument, -2 HIP runtime failure, -3 other failure.
template <class F>
int guarded(vrc_ctx* c, F&& f) {
  if (!c) return -1;
  try {
    VRC_HIP(hipSetDevice(c->device));
    f();
    return 0;
  } catch (const ArgFailure& e) {
    c->last_error = e.what;
    return -1;
  } catch (const HipFailure& e) {
    c->last_error = e.what;
    return -2;
  } catch (const std::exception& e) {
    c->last_error = e.what();
    return -3;
  }
}

}  // namespace vrc
