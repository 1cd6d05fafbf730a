// Device memory pool: blocks are cached by size and reused; hipMalloc only on a miss.
#include "vrc_internal.hpp"

namespace vrc {

Workspace::~Workspace() {
  for (auto& kv : owned_) (void)hipFree(kv.first);
}

DevBuf Workspace::get(size_t bytes) {
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~(size_t)255;
  if (arena_) {
    if (bytes > arena_bytes_ - arena_top_)
      throw ArgFailure("scratch region exhausted: a request of " + std::to_string(bytes) + " bytes with " +
                       std::to_string(arena_top_) + " of " + std::to_string(arena_bytes_) + " bytes in use");
    void* p = arena_ + arena_top_;
    arena_blocks_.push_back({arena_top_, false});
    arena_top_ += bytes;
    if (arena_top_ > arena_peak_) arena_peak_ = arena_top_;
    return DevBuf(p, bytes, this);
  }
  auto it = free_.lower_bound(bytes);
  if (it != free_.end() && it->first <= bytes + bytes / 2 + (1u << 20)) {
    void* p = it->second;
    size_t b = it->first;
    free_.erase(it);
    if (tracking_) tracked_[p] = b;
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
  if (tracking_) tracked_[p] = bytes;
  return DevBuf(p, bytes, this);
}

void Workspace::track_begin() {
  tracking_ = true;
  tracked_.clear();
}

std::vector<std::pair<void*, size_t>> Workspace::track_end(size_t* leased) {
  tracking_ = false;
  std::vector<std::pair<void*, size_t>> out;
  size_t held = 0;
  for (auto& kv : tracked_) {
    bool found = false;
    auto range = free_.equal_range(kv.second);
    for (auto it = range.first; it != range.second; ++it)
      if (it->second == kv.first) {
        free_.erase(it);
        found = true;
        break;
      }
    if (!found) {
      ++held;
      continue;
    }
    owned_.erase(kv.first);
    reserved_ -= kv.second;
    out.push_back(kv);
  }
  tracked_.clear();
  if (leased) *leased = held;
  return out;
}

void Workspace::put(void* p, size_t bytes) {
  if (!p) return;
  if (arena_lo_ && (char*)p >= arena_lo_ && (char*)p < arena_lo_ + arena_bytes_) {
    if (!arena_) return;  // released after arena_end: the region is not pooled
    const size_t off = (size_t)((char*)p - arena_);
    for (auto it = arena_blocks_.rbegin(); it != arena_blocks_.rend(); ++it)
      if (it->first == off) {
        it->second = true;
        break;
      }
    while (!arena_blocks_.empty() && arena_blocks_.back().second) {
      arena_top_ = arena_blocks_.back().first;
      arena_blocks_.pop_back();
    }
    return;
  }
  free_.emplace(bytes, p);
}

void Workspace::arena_begin(void* base, size_t bytes) {
  arena_ = arena_lo_ = (char*)base;
  arena_bytes_ = bytes;
  arena_top_ = arena_peak_ = 0;
  arena_blocks_.clear();
}

size_t Workspace::arena_end() {
  arena_ = nullptr;
  arena_blocks_.clear();
  arena_top_ = 0;
  return arena_peak_;
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
#include <algorithm>
#include <cstring>

#include "vrc_context.hpp"

namespace vrc {

PinnedPool::~PinnedPool() {
  for (auto& b : blocks) {
    (void)hipEventSynchronize(b.ev);
    (void)hipEventDestroy(b.ev);
    (void)hipHostFree(b.p);
  }
}

void PinnedPool::upload(hipStream_t s, void* dst, const void* host, size_t bytes) {
  if (bytes == 0) return;
  void* p = nullptr;
  if (hold) {  // recording a graph: the block belongs to the recording
    VRC_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    hold->push_back(p);
    std::memcpy(p, host, bytes);
    VRC_HIP(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, s));
    return;
  }
  Block* blk = nullptr;
  for (auto& b : blocks)
    if (b.bytes >= bytes && hipEventQuery(b.ev) == hipSuccess && (!blk || b.bytes < blk->bytes)) blk = &b;
  if (!blk) {
    Block b{nullptr, std::max<size_t>(bytes, 4096), nullptr};
    VRC_HIP(hipHostMalloc(&b.p, b.bytes, hipHostMallocDefault));
    VRC_HIP(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
    blocks.push_back(b);
    blk = &blocks.back();
  }
  std::memcpy(blk->p, host, bytes);
  VRC_HIP(hipMemcpyAsync(dst, blk->p, bytes, hipMemcpyHostToDevice, s));
  VRC_HIP(hipEventRecord(blk->ev, s));
}

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
