// Shared definitions for the varRCWA MI355X (gfx950) library.
//
// Storage convention everywhere: complex128 interleaved (re, im) as HIP double2,
// column-major with an explicit leading dimension; harmonic index m*ny + n
// (x order outer, y order inner), field rows [x-harmonics; y-harmonics].
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

typedef double2 cd;
typedef double d4 __attribute__((ext_vector_type(4)));

namespace vrc {

// ----------------------------------------------------------------- complex ops
__host__ __device__ inline cd mk(double r, double i) { cd z; z.x = r; z.y = i; return z; }
__host__ __device__ inline cd cadd(cd a, cd b) { return mk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cd csub(cd a, cd b) { return mk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cd cmul(cd a, cd b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline cd cconj(cd a) { return mk(a.x, -a.y); }
__host__ __device__ inline cd cneg(cd a) { return mk(-a.x, -a.y); }
__host__ __device__ inline cd cscale(cd a, double s) { return mk(a.x * s, a.y * s); }
// a*conj(b)
__host__ __device__ inline cd cmulc(cd a, cd b) { return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
__host__ __device__ inline double cabs1(cd a) { return fabs(a.x) + fabs(a.y); }
__host__ __device__ inline double cnorm2(cd a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline double cabsv(cd a) { return hypot(a.x, a.y); }
// Smith's algorithm for complex division (robust against overflow).
__host__ __device__ inline cd cdiv(cd a, cd b) {
  if (fabs(b.y) <= fabs(b.x)) {
    double r = b.y / b.x, den = b.x + b.y * r;
    return mk((a.x + a.y * r) / den, (a.y - a.x * r) / den);
  }
  double r = b.x / b.y, den = b.y + b.x * r;
  return mk((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}
__host__ __device__ inline cd crecip(cd b) { return cdiv(mk(1.0, 0.0), b); }
// Principal complex square root without cancellation (Kahan's formulation):
// branch cut on the negative real axis, result has Re >= 0.
__host__ __device__ inline cd csqrt_principal(cd z) {
  double x = z.x, y = z.y;
  if (x == 0.0 && y == 0.0) return mk(0.0, y);
  double t = sqrt(0.5 * (fabs(x) + hypot(x, y)));
  if (x >= 0.0) return mk(t, y / (2.0 * t));
  return mk(fabs(y) / (2.0 * t), copysign(t, y));
}
__host__ __device__ inline cd cexp(cd z) {
  double e = exp(z.x);
  double s, c;
  sincos(z.y, &s, &c);
  return mk(e * c, e * s);
}

// ------------------------------------------------------------------ errors
struct Error {
  int code;
  std::string msg;
};

#define VRC_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      throw vrc::HipFailure(std::string(#call) + ": " + hipGetErrorString(e_));   \
    }                                                                              \
  } while (0)

struct HipFailure {
  std::string what;
  explicit HipFailure(std::string w) : what(std::move(w)) {}
};
struct ArgFailure {
  std::string what;
  explicit ArgFailure(std::string w) : what(std::move(w)) {}
};

inline void launch_check(const char* name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw HipFailure(std::string("launch ") + name + ": " + hipGetErrorString(e));
}

inline int cdiv_up(long a, long b) { return (int)((a + b - 1) / b); }

// Operation on an operand of a matrix product.
enum Op : int { OP_N = 0, OP_T = 1, OP_C = 2 };

// One entry of a descriptor-driven batched GEMM: C = alpha*op(A)*op(B) + beta*C.
// Entries with m<=0, n<=0 are skipped; k<=0 scales C by beta.
struct GemmDesc {
  int m, n, k, pad_;
  const cd* A;
  const cd* B;
  cd* C;
  long lda, ldb, ldc;
  cd alpha, beta;
};

// gamma = sqrt(lambda) with the branch rule of WaveEquationSolver's GammaSqrtOp
// (src/rcwa/WaveEquationSolver.cpp:8-22): Im >= 0, flipped if |Im| < |Re| and Re < 0.
__device__ inline cd gamma_branch(cd lam) {
  cd g = csqrt_principal(lam);
  if (g.y < 0.0) g = cneg(g);
  if (fabs(g.y) < fabs(g.x) && g.x < 0.0) g = cneg(g);
  return g;
}

}  // namespace vrc
