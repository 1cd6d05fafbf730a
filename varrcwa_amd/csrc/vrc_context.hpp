// Library context: one device, one stream, one workspace pool, last error text.
#pragma once

#include <atomic>
#include <string>
#include <vector>

#include "vrc_internal.hpp"

namespace vrc {
// Pinned host staging for host -> device copies: a block is reused once the copy that
// read it has completed (an event recorded after the copy), so uploads need no host
// synchronization. While a graph is being recorded (hold != nullptr) fresh blocks are
// taken and handed to the recording, which keeps them as long as its graph lives.
struct PinnedPool {
  struct Block {
    void* p;
    size_t bytes;
    hipEvent_t ev;
  };
  std::vector<Block> blocks;
  std::vector<void*>* hold = nullptr;
  ~PinnedPool();
  // Copies host[0, bytes) into a pinned block and enqueues its copy to dst on stream s.
  void upload(hipStream_t s, void* dst, const void* host, size_t bytes);
};
}  // namespace vrc

// Reference counted: samplers and integrators hold a reference, so the pool and the
// context outlive every object that leases device memory from them.
struct vrc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  vrc::Workspace ws;
  vrc::PinnedPool pinned;
  std::string last_error;
  std::atomic<int> refs{1};
  // stage timers (HIP events on the context stream), see vrc::StageScope
  bool timing = false;
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> event_pool;
  // QR iteration statistics accumulated while timing: matrices, sweeps, windows,
  // chase ticks, apply ticks, small-block ticks
  long long qr_stats[6] = {0, 0, 0, 0, 0, 0};
  // numerical-failure record (device, HEALTH_INTS ints; see health.hip) and the host-side
  // failure noted by the integrators themselves (code 0 = none)
  int* health = nullptr;
  int host_code = 0;
  std::string host_msg;
  ~vrc_ctx() {
    if (health) (void)hipFree(health);
    for (auto& r : recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : event_pool) (void)hipEventDestroy(e);
  }
};

namespace vrc {
inline void ctx_retain(vrc_ctx* c) { c->refs.fetch_add(1); }
inline void ctx_release(vrc_ctx* c) {
  if (c->refs.fetch_sub(1) == 1) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    delete c;
  }
}
}  // namespace vrc

namespace vrc {

// Stage identifiers of the timers (also the index in vrc_timing_read).
enum Stage : int {
  ST_FOURIER = 0,    // (a) Fourier/Toeplitz assembly kernels
  ST_PQ = 1,         // (b) fe33 factorization + P/Q formation (includes ST_FOURIER)
  ST_HESS = 2,       // (c) Hessenberg reduction + Q formation
  ST_QR = 3,         // (c) multishift QR kernel
  ST_EIGVEC = 4,     // (c) triangular eigenvectors + back-transform + normalisation
  ST_EIGPOST = 5,    // (c) PQ product, gamma, V = Q W / gamma
  ST_LAYER_S = 6,    // (d) slab S-matrices
  ST_STAR = 7,       // (d) star products
  ST_PROJECT = 8,    // (d) end projections
  ST_ADJOINT = 9,    // reverse pass: slab, star-product and projection adjoints
  ST_COUNT = 10
};

vrc_ctx*& tl_ctx();

// Records a pair of events around a stage on the active context's stream when
// timing is enabled; no-op otherwise.
struct StageScope {
  vrc_ctx* c = nullptr;
  int stage;
  hipEvent_t a = nullptr;
  explicit StageScope(int st);
  ~StageScope();
};

// Runs f, converting exceptions into the C-ABI status convention:
// 0 ok, -1 bad argument, -2 HIP runtime failure, -3 other failure.
template <class F>
int guarded(vrc_ctx* c, F&& f) {
  if (!c) return -1;
  struct Bind {
    vrc_ctx* prev;
    explicit Bind(vrc_ctx* c) : prev(tl_ctx()) { tl_ctx() = c; }
    ~Bind() { tl_ctx() = prev; }
  } bind(c);
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
