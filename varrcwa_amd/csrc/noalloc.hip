// Allocation-free star product: the kernels of vrc_star_product with every temporary
// taken from one fixed device array per GPU instead of the context's memory pool, for
// callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The array is a file-scope __device__ symbol, so each GPU the library runs on holds its
// own copy (SCRATCH_BYTES) from the code object's load until the process ends. Host code
// never passes the symbol's name to a kernel: its device address is looked up once per
// device with hipGetSymbolAddress, with that device current, when a context is created on
// that device, and kept in g_scratch[device]. Every call on a GPU shares that GPU's array
// and leaves its contents behind; calls that overlap on one GPU (two streams, contexts or
// threads) race on it and must be serialized by the caller. Calls on one context's stream
// are ordered and safe.
#include <mutex>
#include <stdexcept>

#include "vrc_ops.hpp"
#include "varrcwa.h"

namespace {

constexpr size_t SCRATCH_BYTES = (size_t)VRC_NOALLOC_SCRATCH_BYTES;  // per GPU
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};

}  // namespace

namespace vrc {

// Looks up this device's copy of the scratch array with that device current (the caller's
// current device is restored). Called by vrc_ctx_create; later calls for the same device
// return at once.
void noalloc_init_device(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw ArgFailure("device index out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return;
  int prev = -1;
  hipError_t e = hipGetDevice(&prev);
  if (e != hipSuccess) throw HipFailure(std::string("hipGetDevice failed: ") + hipGetErrorString(e));
  if (prev != device) {
    e = hipSetDevice(device);
    if (e != hipSuccess) throw HipFailure(std::string("hipSetDevice failed: ") + hipGetErrorString(e));
  }
  void* p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess || !p)
    throw HipFailure(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
  g_scratch[device] = p;
}

static void* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  return (device >= 0 && device < MAX_DEVICES) ? g_scratch[device] : nullptr;
}

}  // namespace vrc

using namespace vrc;

int vrc_star_product_noalloc(vrc_ctx* c, int N, int batch, void* Tdd, void* Tuu, void* Rud, void* Rdu,
                             const void* tdd, const void* tuu, const void* rud, const void* rdu, long ss) {
  return guarded(c, [&] {
    if (!(N > 0 && batch >= 0 && Tdd && Tuu && Rud && Rdu && tdd && tuu && rud && rdu))
      throw ArgFailure("bad arguments");
    if (batch == 0) return;
    const size_t want = star_product_scratch_bytes(N, batch);
    if (want > SCRATCH_BYTES)
      throw ArgFailure("vrc_star_product_noalloc: N = " + std::to_string(N) + ", batch = " + std::to_string(batch) +
                       " needs " + std::to_string(want) + " bytes of scratch, above the fixed " +
                       std::to_string(SCRATCH_BYTES) + " bytes (use vrc_star_product or a smaller batch)");
    void* base = scratch_of(c->device);
    if (!base) throw ArgFailure("scratch array of this device was not set up (create the context first)");
    struct Arena {
      Workspace& ws;
      bool open = true;
      Arena(Workspace& w, void* b) : ws(w) { ws.arena_begin(b, SCRATCH_BYTES); }
      size_t end() {
        open = false;
        return ws.arena_end();
      }
      ~Arena() {
        if (open) ws.arena_end();
      }
    } arena(c->ws, base);
    SMat S;
    S.Tdd = (cd*)Tdd;
    S.Tuu = (cd*)Tuu;
    S.Rud = (cd*)Rud;
    S.Rdu = (cd*)Rdu;
    star_product(*c, N, batch, S, (const cd*)tdd, (const cd*)tuu, (const cd*)rud, (const cd*)rdu, ss);
    // the size check above rests on star_product_scratch_bytes; a call that went above it
    // (still inside the array, or it would have thrown) means that estimate is stale
    const size_t peak = arena.end();
    if (peak > want)
      throw std::runtime_error("internal: star product used " + std::to_string(peak) + " bytes of scratch, " +
                               "above its estimate of " + std::to_string(want));
  });
}
