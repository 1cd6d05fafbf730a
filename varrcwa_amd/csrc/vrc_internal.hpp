// Internal (C++) interfaces shared between the kernel files and the C ABI layer.
#pragma once

#include <map>
#include <mutex>
#include <utility>

#include "vrc_common.hpp"

namespace vrc {

class Workspace;

// A device buffer leased from a Workspace; returned to the pool when destroyed.
// All work of one context runs on one stream, so a buffer returned here can be
// reused by later launches on that stream without synchronization.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  Workspace* owner = nullptr;
  DevBuf() = default;
  DevBuf(void* p, size_t b, Workspace* o) : ptr(p), bytes(b), owner(o) {}
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes), owner(o.owner) { o.owner = nullptr; o.ptr = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept;
  ~DevBuf();
  template <class T>
  T* as() const { return (T*)ptr; }
};

class Workspace {
 public:
  Workspace() = default;
  ~Workspace();
  DevBuf get(size_t bytes);
  void put(void* p, size_t bytes);
  size_t reserved() const { return reserved_; }
  void trim();  // release cached blocks (caller must have synchronized the stream)

  // Arena mode: between arena_begin and arena_end every get() is carved from the
  // caller's region [base, base + bytes) as a stack (a block is reclaimed once it and
  // every block above it are released) and nothing is allocated; a request that does
  // not fit throws ArgFailure. Blocks of the region returned after arena_end are
  // ignored. arena_end returns the high-water mark in bytes.
  void arena_begin(void* base, size_t bytes);
  size_t arena_end();

  // Graph recording: every block handed out between track_begin and track_end is noted;
  // track_end takes the released ones out of the pool and returns them (pointer, bytes):
  // from then on the caller owns them and frees them with hipFree. *leased counts the
  // noted blocks that are still leased (they stay with their holders).
  void track_begin();
  std::vector<std::pair<void*, size_t>> track_end(size_t* leased);
  bool tracking() const { return tracking_; }

 private:
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> owned_;
  size_t reserved_ = 0;
  char* arena_ = nullptr;       // active region (nullptr: pool mode)
  char* arena_lo_ = nullptr;    // last region used, for put() after arena_end
  size_t arena_bytes_ = 0, arena_top_ = 0, arena_peak_ = 0;
  std::vector<std::pair<size_t, bool>> arena_blocks_;  // (offset, released), bottom to top
  bool tracking_ = false;
  std::map<void*, size_t> tracked_;
};

inline DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
  if (this != &o) {
    if (owner) owner->put(ptr, bytes);
    ptr = o.ptr; bytes = o.bytes; owner = o.owner;
    o.owner = nullptr; o.ptr = nullptr;
  }
  return *this;
}
inline DevBuf::~DevBuf() {
  if (owner) owner->put(ptr, bytes);
}

// ---------------------------------------------------------------- numerical failures (health.hip)
// Positive status codes of the top-level calls (and the per-instance info of batched
// stacks): 1 eigensolver not converged, 2 zero pivot in an LU, 3 zero propagation
// constant, 4 non-finite results, 5 varRCWA depth limit with the error above pe.
enum HealthCode : int { H_OK = 0, H_EIG = 1, H_LU = 2, H_GAMMA0 = 3, H_NONFINITE = 4, H_DEPTH = 5 };
constexpr int HEALTH_INTS = 4;
// Record the first non-zero info / zero gamma of a launch in the active context's record
// (no-op without an active context).
void health_note_info(hipStream_t s, const int* info, int batch, int code);
void health_note_gamma(hipStream_t s, const cd* gam, long count);

// ---------------------------------------------------------------- blas.hip
void zgemm_strided(hipStream_t s, int opA, int opB, int m, int n, int k, cd alpha, const cd* A, long lda, long sA,
                   const cd* B, long ldb, long sB, cd beta, cd* C, long ldc, long sC, int batch);
void zgemm_set_big(bool on);  // large-tile kernel on/off (default on)
void zgemm_set_big_min_k(int k);  // smallest k for the large-tile kernel
void zgemm_set_big_waves(int w);  // wavefronts of the large-tile kernel (4 or 8)
void zgemm_set_small_waves(int w);  // wavefronts of the 64x64-tile kernel (4 or 8)
void zgemm_set_big_n(int n);        // columns of the large tile (64 or 128)
void zgemm_desc(hipStream_t s, int opA, int opB, const GemmDesc* d_desc, int count, int max_m, int max_n);
void zgemm_get_knobs(int* big_n, int* big_waves, int* small_waves, int* big_min_k);
void zgeam(hipStream_t s, int m, int n, cd a, const cd* A, long lda, long sA, cd b, const cd* B, long ldb, long sB,
           cd* C, long ldc, long sC, int batch);
void zset(hipStream_t s, int m, int n, cd diag, cd off, cd* C, long ldc, long sC, int batch);
void zadd_diag(hipStream_t s, int n, cd v, const cd* vec, long svec, cd* C, long ldc, long sC, int batch);
void zdscale(hipStream_t s, int m, int n, const cd* left, long sl, const cd* right, long sr, int inv_right,
             const cd* A, long lda, long sA, cd* C, long ldc, long sC, int batch);
void zmaxabs(hipStream_t s, int m, int n, const cd* A, long lda, long sA, double* out, int batch);

// ---------------------------------------------------------------- lu.hip
void zgetrf(hipStream_t s, int n, cd* A, long lda, long sA, int* ipiv, long sP, int* info, int batch, Workspace& ws);
// X = A^{-1} [B1 | B2] (nrhs columns each; X holds 2 nrhs columns and aliases neither B).
void zgetrs_pair(hipStream_t s, int n, int nrhs, const cd* LU, long lda, long sA, const int* ipiv, long sP,
                 const cd* B1, long ldb1, long sB1, const cd* B2, long ldb2, long sB2, cd* X, long ldx, long sX,
                 int batch, Workspace& ws);
void zgetrs(hipStream_t s, int n, int nrhs, const cd* LU, long lda, long sA, const int* ipiv, long sP, const cd* B,
            long ldb, long sB, cd* X, long ldx, long sX, int batch, Workspace& ws);
// Y <- U^{-1} Y for upper-triangular U (non-unit diagonal), in place.
void ztrsm_upper(hipStream_t s, int n, int nrhs, const cd* U, long ldu, long sU, cd* Y, long ldy, long sY,
                 int batch, Workspace& ws);
void zgetri(hipStream_t s, int n, const cd* LU, long lda, long sA, const int* ipiv, long sP, cd* Ainv, long ldi,
            long sI, int batch, Workspace& ws);
// Peak workspace bytes of zgetrf and zgetrs_pair for these sizes.
size_t zgetrf_scratch_bytes(int n, int batch);
size_t zgetrs_pair_scratch_bytes(int n, int nrhs, int batch);

// ---------------------------------------------------------------- eig.hip
// Right eigen-decomposition of a strided batch of general n x n matrices
// (A is destroyed). w: n eigenvalues per entry; VR: unit-2-norm eigenvectors with
// the largest component real (LAPACK zgeev convention); info per entry
// (0 ok, k>0: QR iteration failed with eigenvalues k..n unconverged).
void zgeev_batched(hipStream_t s, int n, cd* A, long lda, long sA, cd* w, long sw, cd* VR, long ldv, long sV,
                   int* info, int batch, Workspace& ws);

// Schur factors (W = diag(dbal) Z X, X unit upper triangular, not normalized) and the
// normalized eigenvectors from them.
void zgeev_factors(hipStream_t s, int n, cd* A, long lda, long sA, cd* w, long sw, cd* Z, cd* X, cd* dbal,
                   int* info, int batch, Workspace& ws);
void zgeev_vectors(hipStream_t s, int n, const cd* Z, const cd* X, const cd* dbal, cd* VR, long ldv, long sV,
                   int batch);
void zgeev_set_aed(int window);
void zgeev_set_max_sweeps(int n);  // QR sweep budget per matrix (0 = automatic)
void zgeev_set_bulges(int nb);  // bulges per QR sweep (1..16)
int zgeev_get_aed();
int zgeev_get_bulges();
void zgeev_set_qr_variant(int v);   // 0: W = 64, register products, 1: W = 64, LDS-staged, 2: W = 48,
                                    // 4: W = 32; -1 (default): by N and batch
int zgeev_get_qr_variant();
// matrices of one eigensolve launch that fill every QR workgroup slot (256 CUs; two slots
// per CU for variant 2, four for variant 4)
int zgeev_full_launch(int n);
// Sizes (in units) of the launches over which `units` units of `per_unit` n x n eigensolves
// each are solved: full launches of zgeev_full_launch(n) matrices, then one launch of the
// rest, whose QR kernel is chosen for its size. A QR launch takes as long as its busiest
// CU: the multi-matrix kernels only pay off in full launches (427 matrices of the two-per-CU
// kernel take as long as 512), while the one-per-CU kernel takes one round for any launch
// of up to 256.
std::vector<int> zgeev_launch_plan(int n, int units, int per_unit);
void zgeev_set_bulges_w48(int nb);  // bulges per sweep of the W = 48 kernel (1..12)
void zgeev_set_bulges_w32(int nb);  // bulges per sweep of the W = 32 kernel (1..7)
int zgeev_get_bulges_w48();
int zgeev_get_bulges_w32();
// Bit v set for each QR kernel variant v launched since the previous call (then cleared).
unsigned zgeev_take_qr_kernels();

// ---------------------------------------------------------------- autotune.cpp
// The kernel switches an autotuned harmonic count overrides (process-wide values).
struct TuneKnobs {
  int aed, bulges, big_n, big_waves, small_waves, big_min_k;
  int qr_variant;              // -1 (automatic), 0, 2 or 4 (zgeev_set_qr_variant)
  int bulges_w48, bulges_w32;  // bulges per sweep of QR variants 2 and 4 (bulges: 0 and 1)
};
TuneKnobs tune_get();
void tune_put(const TuneKnobs& k);
// Puts the autotuned switches of N in place before a call that solves N x N systems (or
// the caller's own values back when N has no autotuned entry).
void tune_apply(int N);
// A manual setting of a switch becomes the caller's own value.
void tune_manual_begin();
// gamma = sqrt(lambda) with the reference's branch rule.
void gamma_sqrt(hipStream_t s, int n, const cd* lam, long sl, cd* gam, long sg, int batch);

// ---------------------------------------------------------------- fourier.hip / operators
struct LayerGeom;  // defined in vrc_geom.hpp

}  // namespace vrc
