// Device memory pool: blocks are cached by size and reused; hipMalloc only on a miss.
#include "vrc_internal.hpp"

namespace vrc {

Workspace::~Workspace() {
  for (auto& kv : owned_) (void)hipFree(kv.first);
}

DevBuf Workspace::get(size_t bytes) {
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~(size_t)255;
  auto it = free_.lower_bound(bytes);
  if (it != free_.end() && it->first <= bytes + bytes / 2 + (1u << 20)) {
    void* p = it->second;
    size_t b = it->first;
    free_.erase(it);
    return DevBuf(p, b, this);
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // release cached blocks and retry once (callers share one stream; a device
    // synchronize makes the cached blocks safe to free)
    (void)hipGetLastError();
    (void)hipDeviceSynchronize();
    trim();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw HipFailure("hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
  }
  owned_[p] = bytes;
  reserved_ += bytes;
  return DevBuf(p, bytes, this);
}

void Workspace::put(void* p, size_t bytes) {
  if (p) free_.emplace(bytes, p);
}

void Workspace::trim() {
  for (auto& kv : free_) {
    (void)hipFree(kv.second);
    reserved_ -= kv.first;
    owned_.erase(kv.second);
  }
  free_.clear();
}

}  // namespace vrc

// ---------------------------------------------------------------- stage timers
#include "vrc_context.hpp"

namespace vrc {

vrc_ctx*& tl_ctx() {
  static thread_local vrc_ctx* c = nullptr;
  return c;
}

static hipEvent_t take_event(vrc_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  VRC_HIP(hipEventCreate(&e));
  return e;
}

StageScope::StageScope(int st) : stage(st) {
  vrc_ctx* cur = tl_ctx();
  if (!cur || !cur->timing) return;
  c = cur;
  a = take_event(c);
  VRC_HIP(hipEventRecord(a, c->stream));
}

StageScope::~StageScope() {
  if (!c) return;
  hipEvent_t b = nullptr;
  try {
    b = take_event(c);
  } catch (...) {
    return;
  }
  if (hipEventRecord(b, c->stream) == hipSuccess) c->recs.push_back({stage, a, b});
}

}  // namespace vrc
