// Batched complex128 non-Hermitian eigensolver for gfx950 (replaces the reference's
// LAPACKE_zgeev / magma_zgeev call, src/rcwa/MKLEigenSolver.hpp:45-55,
// src/gpu/RedhefferIntegratorGPU.cpp:216-229).
//
// Pipeline for a strided batch of n x n matrices:
//  1. Hessenberg reduction H = Q^H A Q: blocked Householder, panel of HB columns
//     reduced by one workgroup per matrix (the trailing-matrix product A*v is the
//     only memory-bound part), trailing update as three MFMA ZGEMMs.
//  2. Q formed explicitly from the stored reflectors (backward, blocked, ZGEMM).
//  3. Schur form T = Z^H H Z by a multishift QR sweep with tightly packed 3x3 bulges:
//     one workgroup per matrix; the bulge chain is chased inside a <=64x64 diagonal
//     window held in LDS, the window transform U is accumulated in LDS and applied
//     to the off-window rows/columns of H and to Z with in-kernel MFMA products.
//     Shifts are eigenvalues of the trailing block; blocks of size <= QR_NMIN are
//     finished by single-shift QR inside LDS.
//  4. Eigenvectors of T by a tiled back-substitution (64x64 tiles; one workgroup per
//     block column), then VR = Z * X (ZGEMM) and LAPACK zgeev normalisation.
#include <atomic>

#include "vrc_common.hpp"
#include "vrc_internal.hpp"
#include "vrc_context.hpp"

namespace vrc {

constexpr double ULP = 2.220446049250313e-16;      // 2^-52
constexpr double SAFMIN = 2.2250738585072014e-308;
constexpr int HB = 64;         // Hessenberg panel width
constexpr int PT = 1024;       // threads of the Hessenberg panel workgroup (16 waves)
constexpr int GV_MAXG = 8;     // most column groups of the panel's matrix-vector product
constexpr int QT = 512;        // threads of the QR workgroup (8 waves; the window MFMAs use 4)
constexpr int QW = 64;         // maximal QR window
constexpr int NBUL = 16;       // most bulges per sweep (2 shifts each)
constexpr int QR_NMIN = 48;    // active blocks this small are finished in LDS
constexpr int EXC_SWEEPS = 6;  // sweeps without deflation before exceptional shifts
constexpr int AED_NW = 24;     // default deflation window (measured best of 16..64 at N = 882)
constexpr int AED_NIBBLE = 14; // skip the sweep when the window deflated more than this percentage

__device__ inline double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Householder reflector G = I - tau v v^H, v = (1, x*scale), with G^H (alpha; x) = (beta; 0).
__device__ inline void house(cd alpha, double xnorm, double& beta, cd& tau, cd& scale) {
  if (xnorm == 0.0 && alpha.y == 0.0) {
    beta = alpha.x;
    tau = mk(0.0, 0.0);
    scale = mk(0.0, 0.0);
    return;
  }
  double nrm = hypot(hypot(alpha.x, alpha.y), xnorm);
  beta = alpha.x >= 0.0 ? -nrm : nrm;
  tau = mk((beta - alpha.x) / beta, -alpha.y / beta);
  scale = crecip(mk(alpha.x - beta, alpha.y));
}

// The same reflector for alpha and x = (y, z) (z = 0 for a 2-vector) with one square root
// and two divisions: sqrt(|alpha|^2 + |y|^2 + |z|^2) directly and 1/(alpha - beta) as
// conj(d) / |d|^2. Used when every component lies in [2^-400, 2^400] (no overflow or
// underflow in the squares); otherwise the scaled form above.
__device__ inline void house_small(cd alpha, cd y, cd z, double& beta, cd& tau, cd& scale) {
  const double m = fmax(fmax(fmax(fabs(alpha.x), fabs(alpha.y)), fmax(fabs(y.x), fabs(y.y))),
                        fmax(fabs(z.x), fabs(z.y)));
  const double xn2 = cnorm2(y) + cnorm2(z);
  if (!(m <= 0x1p400 && (m >= 0x1p-400 || m == 0.0))) {
    house(alpha, hypot(cabsv(y), cabsv(z)), beta, tau, scale);
    return;
  }
  if (xn2 == 0.0 && alpha.y == 0.0) {
    beta = alpha.x;
    tau = mk(0.0, 0.0);
    scale = mk(0.0, 0.0);
    return;
  }
  const double nrm = sqrt(cnorm2(alpha) + xn2);
  beta = alpha.x >= 0.0 ? -nrm : nrm;
  const double rb = 1.0 / beta;
  tau = mk((beta - alpha.x) * rb, -alpha.y * rb);
  const double dx = alpha.x - beta, dy = alpha.y;
  const double id = 1.0 / (dx * dx + dy * dy);
  scale = mk(dx * id, -dy * id);
}

// ===================================================================== Hessenberg
// out[i] = sum_{r>=r0} conj(Vf[r, p+i]) * x[r] for i < j, x in LDS.
constexpr int GV_W = 8;  // trailing-product loads in flight per thread

// One complex value read with the non-temporal hint (streamed once per column, not reused
// from cache).
typedef double vrc_d2 __attribute__((ext_vector_type(2)));
__device__ inline cd load_stream(const cd* p) {
  const vrc_d2 v = __builtin_nontemporal_load(reinterpret_cast<const vrc_d2*>(p));
  return mk(v.x, v.y);
}

// out[i] = sum_{r >= r0} conj(Vf[r, p+i]) x[r] for i < j (x in LDS): wave w reduces the
// columns i = w, w + PTT/64, ... (the columns proceed in parallel, four loads in flight per
// lane, rows summed in increasing order within a lane).
template <int PTT>
__device__ void panel_multi_dot(int N, int r0, int p, int j, const cd* Vf, const cd* x, cd* out) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int i = wave; i < j; i += PTT / 64) {
    const cd* vc = Vf + (long)(p + i) * N;
    double sr = 0.0, si = 0.0;
    int r = r0 + lane;
    for (; r + 192 < N; r += 256) {
      cd a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = vc[r + 64 * u];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const cd xb = x[r + 64 * u];
        sr += a[u].x * xb.x + a[u].y * xb.y;
        si += a[u].x * xb.y - a[u].y * xb.x;
      }
    }
    for (; r < N; r += 64) {
      const cd a = vc[r], xb = x[r];
      sr += a.x * xb.x + a.y * xb.y;
      si += a.x * xb.y - a.y * xb.x;
    }
    sr = wsum(sr);
    si = wsum(si);
    if (lane == 0) out[i] = mk(sr, si);
  }
  __syncthreads();
}

// acc - sum_{i < j} M[r + i ld] v[i], summed in increasing i with eight loads in flight.
__device__ inline cd sub_row_dot(cd acc, const cd* M, long ld, int r, int j, const cd* v) {
  int i = 0;
  for (; i + 8 <= j; i += 8) {
    cd m[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = M[r + (long)(i + u) * ld];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = csub(acc, cmul(m[u], v[i + u]));
  }
  for (; i < j; ++i) acc = csub(acc, cmul(M[r + (long)i * ld], v[i]));
  return acc;
}

// PTT threads: 1024 (16 waves) in general, 256 or 512 when N - 1 fits, so that several
// matrices share a CU (at most 128 VGPRs; the panel's T factor packed upper-triangular).
template <int PTT>
__global__ __launch_bounds__(PTT) __attribute__((amdgpu_waves_per_eu(4))) void gehrd_panel_kernel(
    int N, int p, int nb, cd* A, long lda, long sA, cd* Vf,
                                                          long sV, cd* Y, long sY, cd* Tp, long sT, cd* tau,
                                                          long sTau) {
  const long b = blockIdx.x;
  A += b * sA;
  Vf += b * sV;
  Y += b * sY;
  Tp += b * sT;
  tau += b * sTau;
  extern __shared__ cd dyn[];
  cd* bvec = dyn;
  cd* vvec = dyn + N;
  __shared__ cd Tl[HB * (HB + 1) / 2];  // T(i, j), i <= j, at j (j + 1) / 2 + i
  __shared__ cd vec1[HB], vec2[HB];
  __shared__ double rn[PTT / 64];
  __shared__ cd gpart[PTT];  // per-thread partial sums of the column-split product
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // rows below the panel; when they fit in the workgroup, the columns of the trailing
  // product are split over G groups of R threads (partial sums reduced through LDS)
  const int R = N - p - 1;
  const int G = (R > 0 && R <= PTT) ? (PTT / R < GV_MAXG ? PTT / R : GV_MAXG) : 1;

  for (int j = 0; j < nb; ++j) {
    const int c = p + j;
    if (t < j) vec1[t] = cconj(Vf[c + (long)(p + t) * N]);
    __syncthreads();
    // column c of A0*Q_j, rows below p (rows <= p are updated after the panel)
    for (int r = p + 1 + t; r < N; r += PTT) bvec[r] = sub_row_dot(A[r + (long)c * lda], Y, N, r, j, vec1);
    __syncthreads();
    if (j > 0) {
      // left transform by Q_j^H on rows >= p+1
      panel_multi_dot<PTT>(N, p + 1, p, j, Vf, bvec, vec2);
      if (t < j) {
        cd s = mk(0.0, 0.0);
        for (int l = 0; l <= t; ++l) s = cadd(s, cmul(cconj(Tl[t * (t + 1) / 2 + l]), vec2[l]));
        vec1[t] = s;
      }
      __syncthreads();
      for (int r = p + 1 + t; r < N; r += PTT) bvec[r] = sub_row_dot(bvec[r], Vf + (long)p * N, N, r, j, vec1);
      __syncthreads();
    }
    // final entries of column c and the reflector from rows c+1..N-1
    double part = 0.0;
    for (int r = p + 1 + t; r < N; r += PTT) {
      if (r <= c) A[r + (long)c * lda] = bvec[r];
      if (r >= c + 2) part += cnorm2(bvec[r]);
    }
    part = wsum(part);
    if (lane == 0) rn[wave] = part;
    __syncthreads();
    double xn2 = 0.0;
    for (int w = 0; w < PTT / 64; ++w) xn2 += rn[w];
    const double xnorm = sqrt(xn2);
    double beta;
    cd tv, scale;
    house(bvec[c + 1], xnorm, beta, tv, scale);
    const bool ident = (tv.x == 0.0 && tv.y == 0.0);
    for (int r = t; r < N; r += PTT) {
      cd v;
      if (r <= c) v = mk(0.0, 0.0);
      else if (r == c + 1) v = mk(1.0, 0.0);
      else v = ident ? mk(0.0, 0.0) : cmul(bvec[r], scale);
      vvec[r] = v;
      Vf[r + (long)c * N] = v;
      if (r == c + 1) A[r + (long)c * lda] = ident ? bvec[r] : mk(beta, 0.0);
      else if (r > c + 1) A[r + (long)c * lda] = mk(0.0, 0.0);
    }
    if (t == 0) tau[c] = tv;
    __syncthreads();
    // z = V_j^H v  (rows >= c+1)
    if (j > 0) panel_multi_dot<PTT>(N, c + 1, p, j, Vf, vvec, vec2);
    // y_j = tau*(A0[:, c+1:] v - Y_j z) for the rows below p (the rows above are one
    // ZGEMM after the panel). The trailing-matrix product is the memory-bound part.
    if (R > 0 && R <= PTT) {
      // thread (g, i): row p+1+i, columns c+1+GV_W g+GV_W Gk .. +GV_W (GV_W loads in flight)
      const int g = t / R, i = t - g * R;
      double sr = 0.0, si = 0.0;
      if (g < G) {
        const int r = p + 1 + i;
        for (int col0 = c + 1 + GV_W * g; col0 < N; col0 += GV_W * G) {
          cd a[GV_W];
#pragma unroll
          for (int u = 0; u < GV_W; ++u) {
            const bool in = col0 + u < N;
            a[u] = in ? load_stream(A + r + (long)(col0 + u) * lda) : mk(0.0, 0.0);
          }
#pragma unroll
          for (int u = 0; u < GV_W; ++u) {
            const cd v = col0 + u < N ? vvec[col0 + u] : mk(0.0, 0.0);
            sr += a[u].x * v.x - a[u].y * v.y;
            si += a[u].x * v.y + a[u].y * v.x;
          }
        }
      }
      gpart[t] = mk(sr, si);
      __syncthreads();
      if (t < R) {
        const int r = p + 1 + t;
        cd acc = gpart[t];
        for (int q = 1; q < G; ++q) acc = cadd(acc, gpart[q * R + t]);
        acc = sub_row_dot(acc, Y, N, r, j, vec2);
        Y[r + (long)j * N] = cmul(tv, acc);
      }
    } else
    for (int rb = p + 1 + t; rb < N; rb += 4 * PTT) {
      double gr[4] = {0.0, 0.0, 0.0, 0.0}, gi[4] = {0.0, 0.0, 0.0, 0.0};
      int col = c + 1;
      for (; col + 3 < N; col += 4) {
        cd v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = vvec[col + u];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = rb + PTT * q;
          if (r < N) {
            const cd* ar = A + r + (long)col * lda;
            cd a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = ar[(long)u * lda];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              gr[q] += a[u].x * v[u].x - a[u].y * v[u].y;
              gi[q] += a[u].x * v[u].y + a[u].y * v[u].x;
            }
          }
        }
      }
      for (; col < N; ++col) {
        const cd v = vvec[col];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int r = rb + PTT * q;
          if (r < N) {
            const cd a = A[r + (long)col * lda];
            gr[q] += a.x * v.x - a.y * v.y;
            gi[q] += a.x * v.y + a.y * v.x;
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = rb + PTT * q;
        if (r < N) {
          cd acc = mk(gr[q], gi[q]);
          for (int i = 0; i < j; ++i) acc = csub(acc, cmul(Y[r + (long)i * N], vec2[i]));
          Y[r + (long)j * N] = cmul(tv, acc);
        }
      }
    }
    if (t < j) {
      cd s = mk(0.0, 0.0);
      for (int l = t; l < j; ++l) s = cadd(s, cmul(Tl[l * (l + 1) / 2 + t], vec2[l]));
      Tl[j * (j + 1) / 2 + t] = cneg(cmul(tv, s));
    }
    if (t == 0) Tl[j * (j + 1) / 2 + j] = tv;
    __syncthreads();
  }
  for (int e = t; e < nb * nb; e += PTT) {
    int i = e % nb, jj = e / nb;
    Tp[i + (long)jj * HB] = (i <= jj) ? Tl[jj * (jj + 1) / 2 + i] : mk(0.0, 0.0);
  }
}

static int n_reflectors(int n) { return n >= 3 ? n - 2 : 0; }

// A <- H (Hessenberg, zeros below the subdiagonal). Vf (n x n) receives the
// reflector vectors by column, Tall (npanel x HB x HB) the panel T factors.
static void gehrd(hipStream_t s, int n, cd* A, long lda, long sA, cd* Vf, cd* Tall, int batch, Workspace& ws) {
  const int nref = n_reflectors(n);
  if (nref == 0) return;
  const int npanel = cdiv_up(nref, HB);
  DevBuf Yb = ws.get(sizeof(cd) * (size_t)n * HB * batch);
  DevBuf taub = ws.get(sizeof(cd) * (size_t)n * batch);
  DevBuf Wb = ws.get(sizeof(cd) * (size_t)HB * n * batch);
  DevBuf W2b = ws.get(sizeof(cd) * (size_t)HB * n * batch);
  cd* Y = Yb.as<cd>();
  cd* W = Wb.as<cd>();
  cd* W2 = W2b.as<cd>();
  const long sV = (long)n * n, sY = (long)n * HB, sT = (long)npanel * HB * HB, sW = (long)HB * n;
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  for (int p = 0; p < nref; p += HB) {
    const int nb = nref - p < HB ? nref - p : HB;
    cd* Tp = Tall + (long)(p / HB) * HB * HB;
    if (n - 1 <= 256)
      gehrd_panel_kernel<256><<<batch, 256, 2 * sizeof(cd) * n, s>>>(n, p, nb, A, lda, sA, Vf, sV, Y, sY, Tp, sT,
                                                                     taub.as<cd>(), n);
    else if (n - 1 <= 512)
      gehrd_panel_kernel<512><<<batch, 512, 2 * sizeof(cd) * n, s>>>(n, p, nb, A, lda, sA, Vf, sV, Y, sY, Tp, sT,
                                                                     taub.as<cd>(), n);
    else
      gehrd_panel_kernel<PT><<<batch, PT, 2 * sizeof(cd) * n, s>>>(n, p, nb, A, lda, sA, Vf, sV, Y, sY, Tp, sT,
                                                                   taub.as<cd>(), n);
    launch_check("gehrd_panel");
    // rows 0..p: Y = A0 V T, then the panel columns' A0 Q = A0 - Y V^H
    zgemm_strided(s, OP_N, OP_N, p + 1, nb, n - p - 1, one, A + (long)(p + 1) * lda, lda, sA,
                  Vf + (p + 1) + (long)p * n, n, sV, zero, W, n, sW, batch);
    zgemm_strided(s, OP_N, OP_N, p + 1, nb, nb, one, W, n, sW, Tp, HB, sT, zero, Y, n, sY, batch);
    if (nb > 1)
      zgemm_strided(s, OP_N, OP_C, p + 1, nb - 1, nb, mone, Y, n, sY, Vf + (p + 1) + (long)p * n, n, sV, one,
                    A + (long)(p + 1) * lda, lda, sA, batch);
    const int c0 = p + nb, ncols = n - c0;
    if (ncols <= 0) continue;
    zgemm_strided(s, OP_N, OP_C, n, ncols, nb, mone, Y, n, sY, Vf + c0 + (long)p * n, n, sV, one,
                  A + (long)c0 * lda, lda, sA, batch);
    zgemm_strided(s, OP_C, OP_N, nb, ncols, n - p - 1, one, Vf + (p + 1) + (long)p * n, n, sV,
                  A + (p + 1) + (long)c0 * lda, lda, sA, zero, W, HB, sW, batch);
    zgemm_strided(s, OP_C, OP_N, nb, ncols, nb, one, Tp, HB, sT, W, HB, sW, zero, W2, HB, sW, batch);
    zgemm_strided(s, OP_N, OP_N, n - p - 1, ncols, nb, mone, Vf + (p + 1) + (long)p * n, n, sV, W2, HB, sW, one,
                  A + (p + 1) + (long)c0 * lda, lda, sA, batch);
  }
}

// Q = H_0 H_1 ... H_{nref-1} into Z (n x n).
static void orghr(hipStream_t s, int n, const cd* Vf, const cd* Tall, cd* Z, long ldz, long sZ, int batch,
                  Workspace& ws) {
  zset(s, n, n, mk(1.0, 0.0), mk(0.0, 0.0), Z, ldz, sZ, batch);
  const int nref = n_reflectors(n);
  if (nref == 0) return;
  const int npanel = cdiv_up(nref, HB);
  DevBuf Wb = ws.get(sizeof(cd) * (size_t)HB * n * batch);
  DevBuf W2b = ws.get(sizeof(cd) * (size_t)HB * n * batch);
  const long sV = (long)n * n, sT = (long)npanel * HB * HB, sW = (long)HB * n;
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  for (int pi = npanel - 1; pi >= 0; --pi) {
    const int p = pi * HB;
    const int nb = nref - p < HB ? nref - p : HB;
    const int m = n - p - 1;
    const cd* V = Vf + (p + 1) + (long)p * n;
    cd* Zs = Z + (p + 1) + (long)(p + 1) * ldz;
    zgemm_strided(s, OP_C, OP_N, nb, m, m, one, V, n, sV, Zs, ldz, sZ, zero, Wb.as<cd>(), HB, sW, batch);
    zgemm_strided(s, OP_N, OP_N, nb, m, nb, one, Tall + (long)pi * HB * HB, HB, sT, Wb.as<cd>(), HB, sW, zero,
                  W2b.as<cd>(), HB, sW, batch);
    zgemm_strided(s, OP_N, OP_N, m, m, nb, mone, V, n, sV, W2b.as<cd>(), HB, sW, one, Zs, ldz, sZ, batch);
  }
}

// ===================================================================== LDS matrix products
typedef cd LdsMat[QW][QW + 1];

// acc = op(A) * B over K = 4 * ks (<= 64) for the 64x64 output tile; CT selects
// op(A)(i,k) = conj(A[k][i]). Windows narrower than 64 (zero padded) take fewer k steps.
template <bool CT, bool FULL = true>
__device__ __forceinline__ void mm64(const LdsMat& As, const LdsMat& Bs, int ks, d4 (&accR)[2][2],
                                     d4 (&accI)[2][2]) {
  if (FULL) ks = QW / 4;  // compile-time trip count for the full window
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }
  for (int s = 0; s < ks; ++s) {
    const int kr = 4 * s + (lane >> 4);
    double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int i = wr * 32 + q * 16 + (lane & 15);
      cd a = CT ? cconj(As[kr][i]) : As[i][kr];
      ar[q] = a.x;
      ai[q] = a.y;
      nai[q] = -a.y;
      const int j = wc * 32 + q * 16 + (lane & 15);
      cd bb = Bs[kr][j];
      br[q] = bb.x;
      bi[q] = bb.y;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accR[i][j], 0, 0, 0);
        accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi[j], accR[i][j], 0, 0, 0);
        accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], accI[i][j], 0, 0, 0);
        accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], accI[i][j], 0, 0, 0);
      }
  }
}

// Calls f(i, j, value) for every output element owned by this thread.
template <class F>
__device__ __forceinline__ void mm64_for_each(const d4 (&accR)[2][2], const d4 (&accI)[2][2], F f) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ii = wr * 32 + i * 16 + (lane >> 4) + 4 * r;
        const int jj = wc * 32 + j * 16 + (lane & 15);
        f(ii, jj, mk(accR[i][j][r], accI[i][j][r]));
      }
}

// The same product on eight waves (QR workgroup): wave w owns rows 32*(w>>2) .. +32 and
// columns 16*(w&3) .. +16 of the 64x64 output (two 16x16 blocks).
static_assert(QT == 512, "the QR window multiply and its staging assume eight waves");
template <bool CT, bool FULL>
__device__ __forceinline__ void mm64w8(const LdsMat& As, const LdsMat& Bs, int ks, d4 (&accR)[2], d4 (&accI)[2]) {
  if (FULL) ks = QW / 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 2, wc = wave & 3;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    accR[i] = d4{0.0, 0.0, 0.0, 0.0};
    accI[i] = d4{0.0, 0.0, 0.0, 0.0};
  }
  for (int s = 0; s < ks; ++s) {
    const int kr = 4 * s + (lane >> 4);
    double ar[2], ai[2], nai[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int i = wr * 32 + q * 16 + (lane & 15);
      cd a = CT ? cconj(As[kr][i]) : As[i][kr];
      ar[q] = a.x;
      ai[q] = a.y;
      nai[q] = -a.y;
    }
    const cd bb = Bs[kr][wc * 16 + (lane & 15)];
    const double br = bb.x, bi = bb.y;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      accR[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br, accR[i], 0, 0, 0);
      accR[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi, accR[i], 0, 0, 0);
      accI[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi, accI[i], 0, 0, 0);
      accI[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br, accI[i], 0, 0, 0);
    }
  }
}

template <class F>
__device__ __forceinline__ void mm64w8_for_each(const d4 (&accR)[2], const d4 (&accI)[2], F f) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 2, wc = wave & 3;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ii = wr * 32 + i * 16 + (lane >> 4) + 4 * r;
      const int jj = wc * 16 + (lane & 15);
      f(ii, jj, mk(accR[i][r], accI[i][r]));
    }
}

// M[r0:r0+w, c0:c1] <- U^H * M[...]  (U: w x w in LDS, zero padded), chunk buffer C;
// QT threads. The next 64-column chunk is fetched into registers while the current one
// is multiplied.
__device__ void apply_left_u(const LdsMat& U, LdsMat& C, int w, cd* M, long ld, int r0, int c0, int c1) {
  const int t = threadIdx.x;
  const int k = t & 63, j0 = t >> 6;  // element (k, j0 + 8q) of a chunk, q < 8
  cd pre[8];
  if (c0 < c1) {
    const int cw = c1 - c0 < QW ? c1 - c0 : QW;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      pre[q] = (k < w && j0 + 8 * q < cw) ? M[(r0 + k) + (long)(c0 + j0 + 8 * q) * ld] : mk(0.0, 0.0);
  }
  for (int cc = c0; cc < c1; cc += QW) {
    const int cw = c1 - cc < QW ? c1 - cc : QW;
#pragma unroll
    for (int q = 0; q < 8; ++q) C[k][j0 + 8 * q] = pre[q];
    __syncthreads();
    const int cn = cc + QW;
    if (cn < c1) {
      const int cwn = c1 - cn < QW ? c1 - cn : QW;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        pre[q] = (k < w && j0 + 8 * q < cwn) ? M[(r0 + k) + (long)(cn + j0 + 8 * q) * ld] : mk(0.0, 0.0);
    }
    d4 aR[2], aI[2];
    if (w == QW)
      mm64w8<true, true>(U, C, QW / 4, aR, aI);
    else
      mm64w8<true, false>(U, C, (w + 3) >> 2, aR, aI);
    mm64w8_for_each(aR, aI, [&](int i, int j, cd v) {
      if (i < w && j < cw) M[(r0 + i) + (long)(cc + j) * ld] = v;
    });
    __syncthreads();
  }
}
// M[r0:r1, c0:c0+w] <- M[...] * U, 64-row chunks, next chunk prefetched into registers.
__device__ void apply_right_u(const LdsMat& U, LdsMat& C, int w, cd* M, long ld, int r0, int r1, int c0) {
  const int t = threadIdx.x;
  const int i = t & 63, k0 = t >> 6;  // element (i, k0 + 8q) of a chunk, q < 8
  cd pre[8];
  if (r0 < r1) {
    const int rh = r1 - r0 < QW ? r1 - r0 : QW;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      pre[q] = (i < rh && k0 + 8 * q < w) ? M[(r0 + i) + (long)(c0 + k0 + 8 * q) * ld] : mk(0.0, 0.0);
  }
  for (int rr = r0; rr < r1; rr += QW) {
    const int rh = r1 - rr < QW ? r1 - rr : QW;
#pragma unroll
    for (int q = 0; q < 8; ++q) C[i][k0 + 8 * q] = pre[q];
    __syncthreads();
    const int rn = rr + QW;
    if (rn < r1) {
      const int rhn = r1 - rn < QW ? r1 - rn : QW;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        pre[q] = (i < rhn && k0 + 8 * q < w) ? M[(rn + i) + (long)(c0 + k0 + 8 * q) * ld] : mk(0.0, 0.0);
    }
    d4 aR[2], aI[2];
    if (w == QW)
      mm64w8<false, true>(C, U, QW / 4, aR, aI);
    else
      mm64w8<false, false>(C, U, (w + 3) >> 2, aR, aI);
    mm64w8_for_each(aR, aI, [&](int ii, int j, cd v) {
      if (ii < rh && j < w) M[(rr + ii) + (long)(c0 + j) * ld] = v;
    });
    __syncthreads();
  }
}
// ===================================================================== QR iteration
// The QR workgroup's LDS: the W x W window, its accumulated transform U and the bulge
// and reduction scratch. W = 64 with T = 512 threads fills most of a CU's LDS (one
// workgroup per CU); W = 48 with T = 256 stays under half of it, so two workgroups (two
// matrices) share a CU and one's serial bulge chase overlaps the other's window products.
template <int W>
using QMat = cd[W][W + 1];
template <int W, int T>
struct QrLds {
  QMat<W> Hw;
  QMat<W> U;
  cd shifts[2 * NBUL];
  cd rv1[NBUL], rv2[NBUL], rtau[NBUL];
  double rbeta[NBUL];
  int rsize[NBUL], rrow[NBUL], rcol[NBUL];
  int redi[T];
  cd aedv[64];  // written by every lane of a wavefront
  int ctl[4];
};
static_assert(sizeof(QrLds<48, 256>) <= 80 * 1024, "two W = 48 QR workgroups must fit one CU's LDS");
static_assert(sizeof(QrLds<32, 128>) <= 40 * 1024, "four W = 32 QR workgroups must fit one CU's LDS");

// Window products from registers (no LDS chunk buffer, no barrier): with the 16x16x4 FP64
// MFMA, lane l supplies A[l & 15][l >> 4] and B[l >> 4][l & 15] and holds
// D[(l >> 4) + 4 q][l & 15], q < 4. Both products are formed transposed so that a lane's
// global loads and stores run along a column (16 consecutive rows per instruction).
//
// M[r0:r0+w, c0:c1] <- U^H M[...]: as D = M^T conj(U) per 16 columns; wave v owns columns
// c0 + 16 v + 16 NW k, reads their w rows into registers and writes them back.
template <int W, int NW>
__device__ void apply_left_reg(const QMat<W>& U, int w, cd* M, long ld, int r0, int c0, int c1) {
  constexpr int KS = W / 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kk = lane >> 4, jl = lane & 15;
  const int ks = (w + 3) >> 2, rbn = (w + 15) >> 4;
  for (int cc = c0 + 16 * wave; cc < c1; cc += 16 * NW) {
    cd a[KS];
    {
      const int col = cc + jl;
      const cd* src = M + r0 + (long)col * ld;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        const int r = 4 * s + kk;
        a[s] = (col < c1 && r < w) ? src[r] : mk(0.0, 0.0);
      }
    }
    // one 16-row block of the result at a time: a[] and one accumulator pair live
    for (int rb = 0; rb < rbn; ++rb) {
      d4 accR = d4{0.0, 0.0, 0.0, 0.0}, accI = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        if (s < ks) {
          const cd u = U[4 * s + kk][16 * rb + jl];  // B[k][i] = conj(U[k][i])
          accR = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s].x, u.x, accR, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s].x, -u.y, accI, 0, 0, 0);
          accR = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s].y, u.y, accR, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s].y, u.x, accI, 0, 0, 0);
        }
      }
      // D[j][i] = (U^H M)[i][cc + j], j = kk + 4 q, i = 16 rb + jl
      const int i = 16 * rb + jl;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = cc + kk + 4 * q;
        if (i < w && col < c1) M[(r0 + i) + (long)col * ld] = mk(accR[q], accI[q]);
      }
    }
  }
}

// M[r0:r1, c0:c0+w] <- M[...] U: as D = U^T M^T per 16 rows; wave v owns rows
// r0 + 16 v + 16 NW k, reads their w columns into registers and writes them back.
template <int W, int NW>
__device__ void apply_right_reg(const QMat<W>& U, int w, cd* M, long ld, int r0, int r1, int c0) {
  constexpr int KS = W / 4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kk = lane >> 4, jl = lane & 15;
  const int ks = (w + 3) >> 2, cbn = (w + 15) >> 4;
  for (int rr = r0 + 16 * wave; rr < r1; rr += 16 * NW) {
    const int row = rr + jl;
    cd b[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int k = 4 * s + kk;
      b[s] = (row < r1 && k < w) ? M[row + (long)(c0 + k) * ld] : mk(0.0, 0.0);
    }
    for (int cb = 0; cb < cbn; ++cb) {
      d4 accR = d4{0.0, 0.0, 0.0, 0.0}, accI = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        if (s < ks) {
          const cd u = U[4 * s + kk][16 * cb + jl];  // A[j][k] = U[k][j]
          accR = __builtin_amdgcn_mfma_f64_16x16x4f64(u.x, b[s].x, accR, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f64_16x16x4f64(u.x, b[s].y, accI, 0, 0, 0);
          accR = __builtin_amdgcn_mfma_f64_16x16x4f64(-u.y, b[s].y, accR, 0, 0, 0);
          accI = __builtin_amdgcn_mfma_f64_16x16x4f64(u.y, b[s].x, accI, 0, 0, 0);
        }
      }
      // D[j][r] = (M U)[rr + r][16 cb + j], j = kk + 4 q, r = jl
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int j = 16 * cb + kk + 4 * q;
        if (row < r1 && j < w) M[row + (long)(c0 + j) * ld] = mk(accR[q], accI[q]);
      }
    }
  }
}

// Negligible-subdiagonal test (Ahues-Tisseur form) on an LDS block, row k of [l..i].
template <class Mat>
__device__ inline bool negl_lds(const Mat& Hw, int k, int l, int i, double smlnum) {
  const cd hs = Hw[k][k - 1];
  const double ahs = cabs1(hs);
  if (ahs <= smlnum) return true;
  double tst = cabs1(Hw[k - 1][k - 1]) + cabs1(Hw[k][k]);
  if (tst == 0.0) {
    if (k - 2 >= l) tst += cabs1(Hw[k - 1][k - 2]);
    if (k + 1 <= i) tst += cabs1(Hw[k + 1][k]);
  }
  if (ahs <= ULP * tst) {
    const double a1 = ahs, a2 = cabs1(Hw[k - 1][k]);
    const double ab = fmax(a1, a2), ba = fmin(a1, a2);
    const double b1 = cabs1(Hw[k][k]), b2 = cabs1(csub(Hw[k - 1][k - 1], Hw[k][k]));
    const double aa = fmax(b1, b2), bb = fmin(b1, b2);
    const double s = aa + ab;
    if (ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)))) return true;
  }
  return false;
}

// Eigenvalue of the 2x2 block [[a, b], [c, d]] closest to d.
__device__ inline cd wilkinson(cd a, cd b, cd c, cd d) {
  cd u = cmul(csqrt_principal(b), csqrt_principal(c));
  double su = cabs1(u);
  if (su == 0.0) return d;
  cd x = cscale(csub(a, d), 0.5);
  double sx = cabs1(x);
  double sc = fmax(su, sx);
  cd xs = cscale(x, 1.0 / sc), us = cscale(u, 1.0 / sc);
  cd y = cscale(csqrt_principal(cadd(cmul(xs, xs), cmul(us, us))), sc);
  if (sx > 0.0) {
    cd xn = cscale(x, 1.0 / sx);
    if (xn.x * y.x + xn.y * y.y < 0.0) y = cneg(y);
  }
  return csub(d, cmul(u, cdiv(u, cadd(x, y))));
}

// Orders the LDS accesses of one wavefront: every lane's earlier LDS writes are seen by
// every lane's later reads (LDS executes a wavefront's instructions in order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Complete Schur decomposition of the s x s Hessenberg block in Hw (s <= QW) by
// single-shift QR, run by ONE wavefront (lane = column / row index), accumulating the
// transform into U when wantU (U must hold the identity or an earlier transform).
// Subdiagonal entries are set to exact zeros on deflation, so Hw ends upper
// triangular. Returns 0 on success, else (index of the unconverged eigenvalue + 1).
template <class Mat>
__device__ int wave_schur(Mat& Hw, Mat& U, int s, bool wantU, int lane) {
  const double smlnum = SAFMIN * ((double)s / ULP);
  const int itmax = 30 * (s > 10 ? s : 10);
  int i = s - 1;
  while (i >= 0) {
    int l = 0;
    bool conv = false;
    for (int its = 0; its <= itmax; ++its) {
      const bool ng = (lane >= 1 && lane <= i) ? negl_lds(Hw, lane, 0, i, smlnum) : false;
      const unsigned long long mask = __ballot(ng);
      const int k = mask ? 63 - __clzll(mask) : 0;
      wave_sync();
      if (k > 0 && lane == 0) Hw[k][k - 1] = mk(0.0, 0.0);
      wave_sync();
      l = k;
      if (l >= i) {
        conv = true;
        break;
      }
      cd sh;
      if (its == 10) {
        sh = cadd(Hw[l][l], mk(0.75 * cabs1(Hw[l + 1][l]), 0.0));
      } else if (its == 20) {
        sh = cadd(Hw[i][i], mk(0.75 * cabs1(Hw[i][i - 1]), 0.0));
      } else {
        sh = wilkinson(Hw[i - 1][i - 1], Hw[i - 1][i], Hw[i][i - 1], Hw[i][i]);
      }
      for (int k2 = l; k2 < i; ++k2) {
        cd x, y;
        if (k2 == l) {
          x = csub(Hw[l][l], sh);
          y = Hw[l + 1][l];
        } else {
          x = Hw[k2][k2 - 1];
          y = Hw[k2 + 1][k2 - 1];
        }
        double beta;
        cd tau, scale;
        house_small(x, y, mk(0.0, 0.0), beta, tau, scale);
        const cd v1 = cmul(y, scale);
        wave_sync();
        if (tau.x == 0.0 && tau.y == 0.0) continue;
        if (lane < s) {
          const int j = lane;
          if (k2 > l && j == k2 - 1) {
            Hw[k2][j] = mk(beta, 0.0);
            Hw[k2 + 1][j] = mk(0.0, 0.0);
          } else {
            const cd h1 = Hw[k2][j], h2 = Hw[k2 + 1][j];
            const cd sm = cmul(cconj(tau), cadd(h1, cmul(cconj(v1), h2)));
            Hw[k2][j] = csub(h1, sm);
            Hw[k2 + 1][j] = csub(h2, cmul(sm, v1));
          }
        }
        wave_sync();
        if (lane < s) {
          const cd c1 = Hw[lane][k2], c2 = Hw[lane][k2 + 1];
          const cd sm = cmul(tau, cadd(c1, cmul(c2, v1)));
          Hw[lane][k2] = csub(c1, sm);
          Hw[lane][k2 + 1] = csub(c2, cmul(sm, cconj(v1)));
          if (wantU) {
            const cd u1 = U[lane][k2], u2 = U[lane][k2 + 1];
            const cd su = cmul(tau, cadd(u1, cmul(u2, v1)));
            U[lane][k2] = csub(u1, su);
            U[lane][k2 + 1] = csub(u2, cmul(su, cconj(v1)));
          }
        }
        wave_sync();
      }
    }
    if (!conv) return i + 1;
    i = l - 1;
  }
  return 0;
}

// Complex Givens rotation (LAPACK zlartg semantics): [cs sn; -conj(sn) cs] [f; g] = [r; 0].
__device__ inline void givens(cd f, cd g, double& cs, cd& sn) {
  const double f1 = cabsv(f), g1 = cabsv(g);
  if (g1 == 0.0) {
    cs = 1.0;
    sn = mk(0.0, 0.0);
  } else if (f1 == 0.0) {
    cs = 0.0;
    sn = cscale(cconj(g), 1.0 / g1);
  } else {
    const double d = sqrt(f1 * f1 + g1 * g1);
    cs = f1 / d;
    sn = cscale(cmul(cscale(f, 1.0 / f1), cconj(g)), 1.0 / d);
  }
}

// Swap the adjacent diagonal entries k, k+1 of the n x n upper-triangular T by a Givens
// rotation (LAPACK ztrexc step), updating the Schur vectors U (n rows). One wavefront.
template <class Mat>
__device__ void wave_swap(Mat& T, Mat& U, int n, int k, int lane) {
  const cd t11 = T[k][k], t22 = T[k + 1][k + 1];
  double cs;
  cd sn;
  givens(T[k][k + 1], csub(t22, t11), cs, sn);
  wave_sync();
  if (lane >= k + 2 && lane < n) {  // rows k, k+1 from column k+2 on
    const cd x = T[k][lane], y = T[k + 1][lane];
    T[k][lane] = cadd(cscale(x, cs), cmul(sn, y));
    T[k + 1][lane] = csub(cscale(y, cs), cmul(cconj(sn), x));
  }
  const cd snc = cconj(sn);
  if (lane < k) {  // columns k, k+1 above row k
    const cd x = T[lane][k], y = T[lane][k + 1];
    T[lane][k] = cadd(cscale(x, cs), cmul(snc, y));
    T[lane][k + 1] = csub(cscale(y, cs), cmul(sn, x));
  }
  if (lane < n) {
    const cd x = U[lane][k], y = U[lane][k + 1];
    U[lane][k] = cadd(cscale(x, cs), cmul(snc, y));
    U[lane][k + 1] = csub(cscale(y, cs), cmul(sn, x));
  }
  if (lane == 0) {
    T[k][k] = t22;
    T[k + 1][k + 1] = t11;
  }
  wave_sync();
}

// T <- G^H T on rows r0..r0+len-1 (columns c0..n-1), then T <- T G on those columns
// (rows 0..rows_t-1) and U <- U G (rows 0..n-1), G = I - tau v v^H with v in vv[0..len).
template <class Mat>
__device__ void wave_reflect(Mat& T, Mat& U, const cd* vv, cd tau, int r0, int len, int c0, int n, int rows_t,
                             int lane) {
  const cd ct = cconj(tau);
  if (lane >= c0 && lane < n) {
    cd d = mk(0.0, 0.0);
    for (int q = 0; q < len; ++q) d = cadd(d, cmul(cconj(vv[q]), T[r0 + q][lane]));
    d = cmul(ct, d);
    for (int q = 0; q < len; ++q) T[r0 + q][lane] = csub(T[r0 + q][lane], cmul(vv[q], d));
  }
  wave_sync();
  if (lane < n) {
    if (lane < rows_t) {
      cd d = mk(0.0, 0.0);
      for (int q = 0; q < len; ++q) d = cadd(d, cmul(T[lane][r0 + q], vv[q]));
      d = cmul(tau, d);
      for (int q = 0; q < len; ++q) T[lane][r0 + q] = csub(T[lane][r0 + q], cmul(d, cconj(vv[q])));
    }
    cd e = mk(0.0, 0.0);
    for (int q = 0; q < len; ++q) e = cadd(e, cmul(U[lane][r0 + q], vv[q]));
    e = cmul(tau, e);
    for (int q = 0; q < len; ++q) U[lane][r0 + q] = csub(U[lane][r0 + q], cmul(e, cconj(vv[q])));
  }
  wave_sync();
}

// QR kernel configuration: window W, T threads, RA: window products from registers
// (apply_left_reg / apply_right_reg) instead of through an LDS chunk buffer.
template <int W_, int T_, bool RA_>
struct QrCfg {
  static constexpr int W = W_, T = T_;
  static constexpr bool RA = RA_;
};

template <class C>
__device__ void lds_load_block(QrLds<C::W, C::T>& L, const cd* H, long ldh, int r0, int w) {
  for (int e = threadIdx.x; e < C::W * C::W; e += C::T) {
    const int i = e % C::W, j = e / C::W;
    L.Hw[i][j] = (i < w && j < w) ? H[(r0 + i) + (long)(r0 + j) * ldh] : mk(0.0, 0.0);
    L.U[i][j] = (i == j && i < w) ? mk(1.0, 0.0) : mk(0.0, 0.0);
  }
  __syncthreads();
}

template <class C>
__device__ void lds_store_block(QrLds<C::W, C::T>& L, cd* H, long ldh, int r0, int w) {
  __syncthreads();
  for (int e = threadIdx.x; e < w * w; e += C::T) {
    int i = e % w, j = e / w;
    H[(r0 + i) + (long)(r0 + j) * ldh] = L.Hw[i][j];
  }
  __syncthreads();
}

// Apply the window transform U (on indices ws..ws+w-1) to everything outside the window.
// The three regions (rows of the window right of it, columns of the window above it, Z)
// are disjoint; the register form needs one barrier at the end only.
template <class C>
__device__ void apply_window(QrLds<C::W, C::T>& L, int N, cd* H, long ldh, cd* Z, long ldz, int ws, int w) {
  const int we = ws + w - 1;
  if constexpr (C::RA) {
    constexpr int NW = C::T / 64;
    if (we + 1 < N) apply_left_reg<C::W, NW>(L.U, w, H, ldh, ws, we + 1, N);
    if (ws > 0) apply_right_reg<C::W, NW>(L.U, w, H, ldh, 0, ws, ws);
    apply_right_reg<C::W, NW>(L.U, w, Z, ldz, 0, N, ws);
    __syncthreads();
  } else {
    static_assert(C::W == QW && C::T == QT, "the LDS-staged window products are written for W = 64, 512 threads");
    if (we + 1 < N) apply_left_u(L.U, L.Hw, w, H, ldh, ws, we + 1, N);
    if (ws > 0) apply_right_u(L.U, L.Hw, w, H, ldh, 0, ws, ws);
    apply_right_u(L.U, L.Hw, w, Z, ldz, 0, N, ws);
  }
}

// Aggressive early deflation (the deflation-window step of LAPACK zlaqr3) on the
// trailing nw x nw window of the active block [lo, hi]: Schur form of the window,
// spike deflation test with reordering of the undeflatable eigenvalues to the top,
// re-reduction of the undeflated part to Hessenberg form, and the window transform
// applied to the rest of H and to Z. Returns the number of deflated eigenvalues (-1 if
// the window's QR did not converge; nothing is changed then). The undeflated
// eigenvalues are left in L.aedv[0..*nsa) for use as shifts.
template <class C>
__device__ int aed(QrLds<C::W, C::T>& L, int N, cd* H, long ldh, cd* Z, long ldz, int lo, int hi, int nw,
                    int* nsa) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int kwtop = hi - nw + 1;
  const cd spike = (kwtop > lo) ? H[kwtop + (long)(kwtop - 1) * ldh] : mk(0.0, 0.0);
  lds_load_block<C>(L, H, ldh, kwtop, nw);
  if (wave == 0) {
    const int info = wave_schur(L.Hw, L.U, nw, true, lane);
    if (lane == 0) L.ctl[1] = info;
    if (info == 0) {
      auto& T = L.Hw;
      auto& U = L.U;
      const double smlnum = SAFMIN * ((double)N / ULP);
      const double as = cabs1(spike);
      int ns = nw, ilst = 0;
      for (int knt = 0; knt < nw; ++knt) {
        const int j = ns - 1;
        double foo = cabs1(T[j][j]);
        if (foo == 0.0) foo = as;
        const bool defl = as * cabs1(U[0][j]) <= fmax(smlnum, ULP * foo);
        wave_sync();
        if (defl) {
          --ns;
        } else {
          for (int k = j - 1; k >= ilst; --k) wave_swap(T, U, nw, k, lane);
          ++ilst;
        }
      }
      if (lane < ns) L.aedv[lane] = T[lane][lane];
      wave_sync();
      cd newspike = mk(0.0, 0.0);
      if (ns > 0 && (spike.x != 0.0 || spike.y != 0.0)) {
        // shifts are kept in ctl-free storage: copy the undeflated eigenvalues first
        cd ev = lane < ns ? L.aedv[lane] : mk(0.0, 0.0);
        if (ns > 1) {
          // reflect the spike s * conj(U[0][0..ns)) onto e1 (zlaqr3 "reflect spike back")
          cd w = lane < ns ? cconj(U[0][lane]) : mk(0.0, 0.0);
          const double xn = sqrt(wsum(lane >= 1 && lane < ns ? cnorm2(w) : 0.0));
          const cd alpha = cconj(U[0][0]);
          double beta;
          cd tau, scale;
          house(alpha, xn, beta, tau, scale);
          wave_sync();
          L.aedv[lane] = lane == 0 ? mk(1.0, 0.0) : (lane < ns ? cmul(w, scale) : mk(0.0, 0.0));
          wave_sync();
          wave_reflect(T, U, L.aedv, tau, 0, ns, 0, nw, ns, lane);
          // Hessenberg reduction of T[0:ns, 0:ns] (columns 0..ns-3)
          for (int c = 0; c + 2 < ns; ++c) {
            const int len = ns - c - 1;
            const cd a0 = T[c + 1][c];
            const cd xv = (lane >= c + 2 && lane < ns) ? T[lane][c] : mk(0.0, 0.0);
            const double xnc = sqrt(wsum(cnorm2(xv)));
            double bc;
            cd tc, sc;
            house(a0, xnc, bc, tc, sc);
            wave_sync();
            if (lane < len) L.aedv[lane] = lane == 0 ? mk(1.0, 0.0) : cmul(T[c + 1 + lane][c], sc);
            wave_sync();
            if (lane == c) {
              T[c + 1][c] = mk(bc, 0.0);
              for (int r = c + 2; r < ns; ++r) T[r][c] = mk(0.0, 0.0);
            }
            wave_sync();
            if (tc.x != 0.0 || tc.y != 0.0) wave_reflect(T, U, L.aedv, tc, c + 1, len, c + 1, nw, ns, lane);
          }
        }
        newspike = cmul(spike, cconj(U[0][0]));
        wave_sync();
        if (lane < ns) L.aedv[lane] = ev;  // shifts for the next sweep
        wave_sync();
      }
      if (lane == 0) {
        L.ctl[2] = ns;
        L.shifts[0] = newspike;  // staging; read back before the next use of shifts
      }
    }
  }
  __syncthreads();
  if (L.ctl[1] != 0) {
    __syncthreads();
    *nsa = 0;
    return -1;
  }
  const int ns = L.ctl[2];
  const cd newspike = L.shifts[0];
  __syncthreads();
  lds_store_block<C>(L, H, ldh, kwtop, nw);
  if (t == 0 && kwtop > lo) H[kwtop + (long)(kwtop - 1) * ldh] = newspike;
  __syncthreads();
  apply_window<C>(L, N, H, ldh, Z, ldz, kwtop, nw);
  *nsa = ns;
  return nw - ns;
}

// Largest k in [1, hi] whose subdiagonal is negligible (sets it to zero); 0 if none.
template <class C>
__device__ int find_lo(QrLds<C::W, C::T>& L, cd* H, long ldh, int N, int hi) {
  const int t = threadIdx.x;
  const double smlnum = SAFMIN * ((double)N / ULP);
  int best = 0;
  for (int k = hi - t; k >= 1; k -= C::T) {
    const cd hs = H[k + (long)(k - 1) * ldh];
    const double ahs = cabs1(hs);
    bool ng = ahs <= smlnum;
    if (!ng) {
      const cd hkk = H[k + (long)k * ldh], hk1 = H[(k - 1) + (long)(k - 1) * ldh];
      double tst = cabs1(hk1) + cabs1(hkk);
      if (tst == 0.0) {
        if (k >= 2) tst += cabs1(H[(k - 1) + (long)(k - 2) * ldh]);
        if (k + 1 <= hi) tst += cabs1(H[(k + 1) + (long)k * ldh]);
      }
      if (ahs <= ULP * tst) {
        const double a2 = cabs1(H[(k - 1) + (long)k * ldh]);
        const double ab = fmax(ahs, a2), ba = fmin(ahs, a2);
        const double b1 = cabs1(hkk), b2 = cabs1(csub(hk1, hkk));
        const double aa = fmax(b1, b2), bb = fmin(b1, b2);
        const double s = aa + ab;
        ng = ba * (ab / s) <= fmax(smlnum, ULP * (bb * (aa / s)));
      }
    }
    if (ng) {
      best = k;
      break;
    }
  }
  L.redi[t] = best;
  __syncthreads();
  for (int w = C::T / 2; w > 0; w >>= 1) {
    if (t < w && L.redi[t + w] > L.redi[t]) L.redi[t] = L.redi[t + w];
    __syncthreads();
  }
  const int lo = L.redi[0];
  __syncthreads();
  if (t == 0 && lo > 0) H[lo + (long)(lo - 1) * ldh] = mk(0.0, 0.0);
  __syncthreads();
  return lo;
}

// One multishift sweep over the active block [lo, hi] with nbul bulges (shifts in L.shifts).
template <class C>
__device__ void qr_sweep(QrLds<C::W, C::T>& L, int N, cd* H, long ldh, cd* Z, long ldz, int lo, int hi, int nbul,
                         long long* st) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int adv = C::W - 3 * nbul - 1;
  const int T_end = (hi - 2) - (lo - 1) + 3 * (nbul - 1) + 1;
  for (int t0 = 0; t0 < T_end; t0 += adv) {
    const int t1 = t0 + adv < T_end ? t0 + adv : T_end;
    int pmin = lo - 1 + t0 - 3 * (nbul - 1);
    if (pmin < lo - 1) pmin = lo - 1;
    int pmax = lo - 1 + t1 - 1;
    if (pmax > hi - 2) pmax = hi - 2;
    const int ws = pmin > lo ? pmin : lo;
    const int we = pmax + 4 < hi ? pmax + 4 : hi;
    const int w = we - ws + 1;
    const long long c0 = wall_clock64();
    lds_load_block<C>(L, H, ldh, ws, w);
    for (int tt = t0; tt < t1; ++tt) {
      // reflectors of all bulges at this step, from the current window state
      if (t < nbul) {
        const int bb = t;
        const int p = lo - 1 + tt - 3 * bb;
        int size = 0;
        if (p >= lo - 1 && p <= hi - 2) {
          cd x, y, z = mk(0.0, 0.0);
          size = (p + 3 <= hi) ? 3 : 2;
          if (p == lo - 1) {
            const int a = lo - ws;
            const cd s1 = L.shifts[2 * bb], s2 = L.shifts[2 * bb + 1];
            const cd h11 = L.Hw[a][a], h21 = L.Hw[a + 1][a], h12 = L.Hw[a][a + 1];
            const cd h22 = L.Hw[a + 1][a + 1];
            const cd h32 = (size == 3) ? L.Hw[a + 2][a + 1] : mk(0.0, 0.0);
            const double sc = cabs1(csub(h11, s2)) + cabs1(h21);
            if (sc == 0.0) {
              x = y = mk(0.0, 0.0);
            } else {
              const cd h21s = cscale(h21, 1.0 / sc);
              x = cadd(cmul(h21s, h12), cmul(csub(h11, s1), cscale(csub(h11, s2), 1.0 / sc)));
              y = cmul(h21s, csub(cadd(h11, h22), cadd(s1, s2)));
              z = cmul(h21s, h32);
            }
            L.rcol[bb] = -1;
          } else {
            const int a = p - ws;
            x = L.Hw[a + 1][a];
            y = L.Hw[a + 2][a];
            if (size == 3) z = L.Hw[a + 3][a];
            L.rcol[bb] = a;
          }
          double beta;
          cd tau, scale;
          house_small(x, y, z, beta, tau, scale);
          L.rv1[bb] = cmul(y, scale);
          L.rv2[bb] = cmul(z, scale);
          L.rtau[bb] = tau;
          L.rbeta[bb] = beta;
          L.rrow[bb] = p + 1 - ws;
          if (tau.x == 0.0 && tau.y == 0.0) size = 0;
        }
        L.rsize[bb] = size;
      }
      __syncthreads();
      // row operations: wave v takes bulges v, v + T/64, ...; lane = column. Columns left
      // of the bulge's own column hold exact zeros in its rows and are skipped.
      for (int bb = wave; bb < nbul; bb += C::T / 64) {
        const int sz = L.rsize[bb];
        if (sz == 0) continue;
        const int r1 = L.rrow[bb], rc = L.rcol[bb];
        const int j = (r1 > 0 ? r1 - 1 : 0) + lane;
        if (j >= w) continue;
        if (j == rc) {
          L.Hw[r1][j] = mk(L.rbeta[bb], 0.0);
          L.Hw[r1 + 1][j] = mk(0.0, 0.0);
          if (sz == 3) L.Hw[r1 + 2][j] = mk(0.0, 0.0);
          continue;
        }
        const cd v1 = L.rv1[bb], v2 = L.rv2[bb], ct = cconj(L.rtau[bb]);
        cd h1 = L.Hw[r1][j], h2 = L.Hw[r1 + 1][j];
        cd h3 = sz == 3 ? L.Hw[r1 + 2][j] : mk(0.0, 0.0);
        cd sm = cadd(h1, cmul(cconj(v1), h2));
        if (sz == 3) sm = cadd(sm, cmul(cconj(v2), h3));
        sm = cmul(ct, sm);
        L.Hw[r1][j] = csub(h1, sm);
        L.Hw[r1 + 1][j] = csub(h2, cmul(sm, v1));
        if (sz == 3) L.Hw[r1 + 2][j] = csub(h3, cmul(sm, v2));
      }
      __syncthreads();
      // column operations on the window (rows down to the bulge's last row + 1; below it
      // the columns hold exact zeros) and on U (every row): pairs (matrix, bulge) over the
      // waves, lane = row
      for (int q = wave; q < 2 * nbul; q += C::T / 64) {
        const int mat = q >= nbul ? 1 : 0, bb = q - mat * nbul;
        const int sz = L.rsize[bb];
        if (sz == 0) continue;
        const int r1 = L.rrow[bb];
        const int i = lane;
        if (i >= w || (mat == 0 && i > r1 + 3)) continue;
        auto& M = mat == 0 ? L.Hw : L.U;
        const cd v1 = L.rv1[bb], v2 = L.rv2[bb], tu = L.rtau[bb];
        cd c1 = M[i][r1], c2 = M[i][r1 + 1];
        cd c3 = sz == 3 ? M[i][r1 + 2] : mk(0.0, 0.0);
        cd sm = cadd(c1, cmul(c2, v1));
        if (sz == 3) sm = cadd(sm, cmul(c3, v2));
        sm = cmul(tu, sm);
        M[i][r1] = csub(c1, sm);
        M[i][r1 + 1] = csub(c2, cmul(sm, cconj(v1)));
        if (sz == 3) M[i][r1 + 2] = csub(c3, cmul(sm, cconj(v2)));
      }
      __syncthreads();
    }
    lds_store_block<C>(L, H, ldh, ws, w);
    const long long c1 = wall_clock64();
    apply_window<C>(L, N, H, ldh, Z, ldz, ws, w);
    const long long c2 = wall_clock64();
    st[1] += 1;
    st[2] += c1 - c0;
    st[3] += c2 - c1;
  }
}

// stats (nullable): per matrix {sweeps, windows, window-chase ticks, off-window-apply ticks,
// small-block ticks} of the constant-rate device clock.
// use_aed: deflation-window size (0 = no aggressive early deflation); max_bul: bulges
// per sweep (<= NBUL).
template <class C>
__device__ void hqr_body(QrLds<C::W, C::T>& L, int N, cd* H, long ldh, long sH, cd* Z, long ldz, long sZ, cd* wout,
                         long sw, int* info, long long* stats, int use_aed, int max_bul, int max_sweeps) {
  long long st[5] = {0, 0, 0, 0, 0};
  const long b = blockIdx.x;
  H += b * sH;
  Z += b * sZ;
  wout += b * sw;
  const int t = threadIdx.x;
  // blocks this small are finished by the single-wave QR in the window; at least 8 chase
  // steps per window
  constexpr int NMIN = QR_NMIN < C::W ? QR_NMIN : C::W;
  constexpr int BMAX = (C::W - 9) / 3 < NBUL ? (C::W - 9) / 3 : NBUL;
  if (max_bul > BMAX) max_bul = BMAX;
  if (use_aed > C::W - 1) use_aed = C::W - 1;
  int ihi = N - 1;
  int sweeps = 0, last_progress = 0, prev_lo = -1, prev_hi = N;
  const int itmax = max_sweeps > 0 ? max_sweeps : 30 * (N > 10 ? N : 10);
  int fail = 0;
  while (ihi >= 0) {
    const int lo = find_lo<C>(L, H, ldh, N, ihi);
    const int m = ihi - lo + 1;
    if (lo != prev_lo || ihi != prev_hi) last_progress = sweeps;
    prev_lo = lo;
    prev_hi = ihi;
    if (m <= NMIN) {
      const long long c0 = wall_clock64();
      lds_load_block<C>(L, H, ldh, lo, m);
      if ((t >> 6) == 0) {
        const int r = wave_schur(L.Hw, L.U, m, true, t & 63);
        if (t == 0) L.ctl[1] = r;
      }
      __syncthreads();
      const int r = L.ctl[1];
      lds_store_block<C>(L, H, ldh, lo, m);
      if (r != 0) {
        fail = lo + r;
        break;
      }
      apply_window<C>(L, N, H, ldh, Z, ldz, lo, m);
      st[4] += wall_clock64() - c0;
      ihi = lo - 1;
      continue;
    }
    if (sweeps >= itmax) {
      fail = ihi + 1;
      break;
    }
    int nbul = (m - 2) / 6;
    if (nbul > max_bul) nbul = max_bul;
    if (nbul < 1) nbul = 1;
    // aggressive early deflation on the trailing window
    int nsa = 0;
    if (use_aed) {
      const long long c0 = wall_clock64();
      const int nw = m - 1 < use_aed ? m - 1 : use_aed;
      const int nd = aed<C>(L, N, H, ldh, Z, ldz, lo, ihi, nw, &nsa);
      st[4] += wall_clock64() - c0;
      if (nd > 0) {
        ihi -= nd;
        last_progress = sweeps;
        prev_hi = ihi;
        // a productive window: look again before sweeping (LAPACK's nibble rule)
        if (100 * nd > AED_NIBBLE * nw || ihi - lo + 1 <= NMIN) {
          ++sweeps;
          continue;
        }
      }
      if (nd < 0) nsa = 0;
      const int mm = ihi - lo + 1;
      nbul = (mm - 2) / 6;
      if (nbul > max_bul) nbul = max_bul;
      if (nbul < 1) nbul = 1;
    }
    int ns = 2 * nbul;
    const bool exceptional = (sweeps - last_progress) > 0 && ((sweeps - last_progress) % EXC_SWEEPS) == 0;
    if (exceptional) {
      if (t < ns) {
        const int k = ihi - t;
        cd hk = H[k + (long)k * ldh];
        double sub = cabs1(H[k + (long)(k - 1) * ldh]);
        L.shifts[t] = cadd(hk, mk(0.75 * sub, 0.0));
      }
      __syncthreads();
    } else if (nsa >= nbul) {
      // the undeflated window eigenvalues nearest the bottom serve as shifts
      if (nsa < ns) {
        nbul = nsa / 2 > 0 ? nsa / 2 : 1;
        ns = 2 * nbul;
      }
      if (t < ns) L.shifts[t] = L.aedv[nsa - ns + t];
      __syncthreads();
    } else {
      const int r0 = ihi - ns + 1;
      for (int e = t; e < C::W * C::W; e += C::T) {
        const int i = e % C::W, j = e / C::W;
        L.Hw[i][j] = (i < ns && j < ns) ? H[(r0 + i) + (long)(r0 + j) * ldh] : mk(0.0, 0.0);
      }
      __syncthreads();
      if ((t >> 6) == 0) wave_schur(L.Hw, L.U, ns, false, t & 63);
      __syncthreads();
      if (t < ns) L.shifts[t] = L.Hw[t][t];
      __syncthreads();
    }
    qr_sweep<C>(L, N, H, ldh, Z, ldz, lo, ihi, nbul, st);
    ++sweeps;
  }
  __syncthreads();
  for (int i = t; i < N; i += C::T) wout[i] = H[i + (long)i * ldh];
  if (t == 0) {
    info[b] = fail;
    if (stats) {
      st[0] = sweeps;
      for (int k = 0; k < 5; ++k) stats[b * 5 + k] = st[k];
    }
  }
}

// Kernels of the three configurations (one workgroup per matrix): W = 64 on 512 threads with
// the register (hqr_kernel) or the LDS-staged (hqr_lds_kernel) window products, and W = 48
// on 256 threads, two workgroups per CU (at most 256 VGPRs per lane).
#define HQR_ARGS                                                                                       \
  int N, cd *H, long ldh, long sH, cd *Z, long ldz, long sZ, cd *wout, long sw, int *info, long long *stats, \
      int use_aed, int max_bul, int max_sweeps
#define HQR_PASS N, H, ldh, sH, Z, ldz, sZ, wout, sw, info, stats, use_aed, max_bul, max_sweeps
__global__ __launch_bounds__(512) void hqr_kernel(HQR_ARGS) {
  __shared__ QrLds<64, 512> L;
  hqr_body<QrCfg<64, 512, true>>(L, HQR_PASS);
}
__global__ __launch_bounds__(512) void hqr_lds_kernel(HQR_ARGS) {
  __shared__ QrLds<64, 512> L;
  hqr_body<QrCfg<64, 512, false>>(L, HQR_PASS);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void hqr_w48_kernel(HQR_ARGS) {
  __shared__ QrLds<48, 256> L;
  hqr_body<QrCfg<48, 256, true>>(L, HQR_PASS);
}
// W = 32 on 128 threads: four workgroups (matrices) per CU, for small N.
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) void hqr_w32_kernel(HQR_ARGS) {
  __shared__ QrLds<32, 128> L;
  hqr_body<QrCfg<32, 128, true>>(L, HQR_PASS);
}
#undef HQR_ARGS
#undef HQR_PASS

// ===================================================================== eigenvectors
// X (n x n, pre-zeroed) <- upper-triangular eigenvector matrix of upper-triangular T:
// (T - T_kk I) x_k = 0, x_k(k) = 1, tiny pivots replaced by smin (LAPACK ztrevc rule).
__global__ __launch_bounds__(256) void trevc_kernel(int N, const cd* T, long ldt, long sT, cd* X, long ldx,
                                                    long sX) {
  __shared__ LdsMat Ts;
  __shared__ LdsMat Xs;
  __shared__ cd lam[QW];
  __shared__ double smin[QW];
  const long b = blockIdx.y;
  T += b * sT;
  X += b * sX;
  const int K = blockIdx.x;
  const int k0 = K * QW;
  const int bk = N - k0 < QW ? N - k0 : QW;
  const int t = threadIdx.x;
  for (int e = t; e < QW * QW; e += 256) {
    int i = e & 63, j = e >> 6;
    Ts[i][j] = (i < bk && j < bk) ? T[(k0 + i) + (long)(k0 + j) * ldt] : mk(0.0, 0.0);
  }
  __syncthreads();
  if (t < bk) {
    lam[t] = Ts[t][t];
    smin[t] = fmax(ULP * cabs1(lam[t]), SAFMIN * ((double)N / ULP));
  }
  __syncthreads();
  // diagonal tile: thread c solves its own column
  if (t < QW) {
    const int c = t;
    for (int i = 0; i < QW; ++i) Xs[i][c] = mk(0.0, 0.0);
    if (c < bk) {
      Xs[c][c] = mk(1.0, 0.0);
      for (int i = c - 1; i >= 0; --i) {
        cd s = mk(0.0, 0.0);
        for (int j = i + 1; j <= c; ++j) s = cadd(s, cmul(Ts[i][j], Xs[j][c]));
        cd d = csub(Ts[i][i], lam[c]);
        if (cabs1(d) < smin[c]) d = mk(smin[c], 0.0);
        Xs[i][c] = cneg(cdiv(s, d));
      }
    }
  }
  __syncthreads();
  for (int e = t; e < bk * bk; e += 256) {
    int i = e % bk, j = e / bk;
    X[(k0 + i) + (long)(k0 + j) * ldx] = Xs[i][j];
  }
  // right-hand sides for the rows above: X[0:k0, K] = -T[0:k0, K] * X_KK
  for (int rr = 0; rr < k0; rr += QW) {
    __syncthreads();
    for (int e = t; e < QW * QW; e += 256) {
      int i = e & 63, k = e >> 6;
      Ts[i][k] = (k < bk) ? T[(rr + i) + (long)(k0 + k) * ldt] : mk(0.0, 0.0);
    }
    __syncthreads();
    d4 aR[2][2], aI[2][2];
    mm64<false>(Ts, Xs, QW / 4, aR, aI);
    mm64_for_each(aR, aI, [&](int i, int j, cd v) {
      if (j < bk) X[(rr + i) + (long)(k0 + j) * ldx] = cneg(v);
    });
  }
  // block back-substitution upward
  for (int I = K - 1; I >= 0; --I) {
    const int i0 = I * QW;
    __syncthreads();
    for (int e = t; e < QW * QW; e += 256) {
      int i = e & 63, j = e >> 6;
      Ts[i][j] = T[(i0 + i) + (long)(i0 + j) * ldt];
      Xs[i][j] = (j < bk) ? X[(i0 + i) + (long)(k0 + j) * ldx] : mk(0.0, 0.0);
    }
    __syncthreads();
    if (t < bk) {
      const int c = t;
      for (int i = QW - 1; i >= 0; --i) {
        cd s = Xs[i][c];
        for (int j = i + 1; j < QW; ++j) s = csub(s, cmul(Ts[i][j], Xs[j][c]));
        cd d = csub(Ts[i][i], lam[c]);
        if (cabs1(d) < smin[c]) d = mk(smin[c], 0.0);
        Xs[i][c] = cdiv(s, d);
      }
    }
    __syncthreads();
    for (int e = t; e < QW * bk; e += 256) {
      int i = e & 63, j = e >> 6;
      X[(i0 + i) + (long)(k0 + j) * ldx] = Xs[i][j];
    }
    for (int rr = 0; rr < i0; rr += QW) {
      __syncthreads();
      for (int e = t; e < QW * QW; e += 256) {
        int i = e & 63, k = e >> 6;
        Ts[i][k] = T[(rr + i) + (long)(i0 + k) * ldt];
      }
      __syncthreads();
      d4 aR[2][2], aI[2][2];
      mm64<false>(Ts, Xs, QW / 4, aR, aI);
      mm64_for_each(aR, aI, [&](int i, int j, cd v) {
        if (j < bk) {
          cd* p = X + (rr + i) + (long)(k0 + j) * ldx;
          *p = csub(*p, v);
        }
      });
    }
  }
}

// Unit 2-norm, then rotate so that the largest-modulus component (first on ties) is real.
__global__ __launch_bounds__(256) void normalize_cols_kernel(int N, cd* V, long ldv, long sV) {
  const long b = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = blockIdx.x * 4 + wave;
  if (col >= N) return;
  cd* v = V + b * sV + (long)col * ldv;
  double ss = 0.0;
  for (int r = lane; r < N; r += 64) ss += cnorm2(v[r]);
  ss = wsum(ss);
  const double scl = 1.0 / sqrt(ss);
  double best = -1.0;
  int bi = 0x7fffffff;
  for (int r = lane; r < N; r += 64) {
    cd x = cscale(v[r], scl);
    double m = cnorm2(x);
    if (m > best) { best = m; bi = r; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    double ob = __shfl_xor(best, o, 64);
    int oi = __shfl_xor(bi, o, 64);
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
  }
  const cd vk = cscale(v[bi], scl);
  const cd rot = cscale(cconj(vk), 1.0 / sqrt(best));
  for (int r = lane; r < N; r += 64) {
    cd x = cmul(cscale(v[r], scl), rot);
    if (r == bi) x.y = 0.0;
    v[r] = x;
  }
}

// ===================================================================== balancing
// Diagonal similarity scaling by powers of two (the scaling half of LAPACK zgebal,
// job 'S', reference LAPACK 3.5+ norms): for each index i, scale row i by 1/f and
// column i by f when that reduces ||col i||_2 + ||row i||_2 below 0.95 of its value;
// sweeps repeat until no index changes (at most 64 sweeps). scale[i] receives the
// accumulated factor; eigenvectors of the original matrix are diag(scale) times those
// of the balanced one. One workgroup per matrix.
__global__ __launch_bounds__(256) void balance_kernel(int N, cd* A, long lda, long sA, cd* scale, long ss) {
  __shared__ double red[4][4];
  __shared__ int changed;
  const long b = blockIdx.x;
  A += b * sA;
  scale += b * ss;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const double sfmin1 = SAFMIN / (ULP * 2.0), sfmax1 = 1.0 / sfmin1;
  const double sfmin2 = sfmin1 * 2.0, sfmax2 = 1.0 / sfmin2;
  for (int i = t; i < N; i += 256) scale[i] = mk(1.0, 0.0);
  __syncthreads();
  for (int sweep = 0; sweep < 64; ++sweep) {
    if (t == 0) changed = 0;
    __syncthreads();
    for (int i = 0; i < N; ++i) {
      // column and row 2-norms (with the diagonal) and largest moduli
      double c2 = 0.0, r2 = 0.0, ca = 0.0, ra = 0.0;
      for (int k = t; k < N; k += 256) {
        const cd cv = A[k + (long)i * lda], rv = A[i + (long)k * lda];
        c2 += cnorm2(cv);
        r2 += cnorm2(rv);
        ca = fmax(ca, cnorm2(cv));
        ra = fmax(ra, cnorm2(rv));
      }
      c2 = wsum(c2);
      r2 = wsum(r2);
      for (int o = 32; o > 0; o >>= 1) {
        ca = fmax(ca, __shfl_xor(ca, o, 64));
        ra = fmax(ra, __shfl_xor(ra, o, 64));
      }
      if (lane == 0) {
        red[0][wave] = c2;
        red[1][wave] = r2;
        red[2][wave] = ca;
        red[3][wave] = ra;
      }
      __syncthreads();
      double c = sqrt(red[0][0] + red[0][1] + red[0][2] + red[0][3]);
      double r = sqrt(red[1][0] + red[1][1] + red[1][2] + red[1][3]);
      ca = sqrt(fmax(fmax(red[2][0], red[2][1]), fmax(red[2][2], red[2][3])));
      ra = sqrt(fmax(fmax(red[3][0], red[3][1]), fmax(red[3][2], red[3][3])));
      __syncthreads();
      if (c == 0.0 || r == 0.0) continue;
      double g = r / 2.0, f = 1.0;
      const double s = c + r;
      while (c < g && fmax(f, fmax(c, ca)) < sfmax2 && fmin(r, fmin(g, ra)) > sfmin2) {
        f *= 2.0;
        c *= 2.0;
        ca *= 2.0;
        r /= 2.0;
        g /= 2.0;
        ra /= 2.0;
      }
      g = c / 2.0;
      while (g >= r && fmax(r, ra) < sfmax2 && fmin(fmin(f, c), fmin(g, ca)) > sfmin2) {
        f /= 2.0;
        c /= 2.0;
        g /= 2.0;
        ca /= 2.0;
        r *= 2.0;
        ra *= 2.0;
      }
      if (c + r >= 0.95 * s) continue;
      const double sc = scale[i].x;
      if (f < 1.0 && sc < 1.0 && f * sc <= sfmin1) continue;
      if (f > 1.0 && sc > 1.0 && sc >= sfmax1 / f) continue;
      const double gi = 1.0 / f;
      for (int k = t; k < N; k += 256) {
        A[i + (long)k * lda] = cscale(A[i + (long)k * lda], gi);
      }
      __syncthreads();
      for (int k = t; k < N; k += 256) {
        A[k + (long)i * lda] = cscale(A[k + (long)i * lda], f);
      }
      if (t == 0) {
        scale[i] = mk(sc * f, 0.0);
        changed = 1;
      }
      __syncthreads();
    }
    __syncthreads();
    if (!changed) break;
    __syncthreads();
  }
}

static int g_aed_window = AED_NW;
static int g_bulges = 12;
static int g_max_sweeps = 0;  // 0: 30 * max(N, 10), LAPACK zlahqr's iteration budget
void zgeev_set_max_sweeps(int n) { g_max_sweeps = n < 0 ? 0 : n; }
void zgeev_set_aed(int window) { g_aed_window = window < 0 ? 0 : (window > QW ? QW : window); }
void zgeev_set_bulges(int nb) { g_bulges = nb < 1 ? 1 : (nb > NBUL ? NBUL : nb); }
// 0: hqr_kernel (W = 64, window products from registers), 1: hqr_lds_kernel (W = 64,
// window products staged through LDS), 2: hqr_w48_kernel (W = 48, two matrices per CU),
// 4: hqr_w32_kernel (W = 32, four matrices per CU); -1: automatic, 4 for N <= 256 in launches
// of at least 384 matrices and for N <= 512 in launches of at least 768, 2 for N <= 1024 in
// launches of at least 384, else 0 (measured: at N = 98 and 242 the four-per-CU kernel wins
// from about 1.5 matrices per CU, at N = 450 from about 3; at N = 882 the two-per-CU kernel in
// 512-matrix launches is 2.7 % faster than the one-per-CU kernel in 256-matrix launches)
static int g_qr_variant = -1;
static int qr_variant_for(int n, int batch) {
  if (g_qr_variant >= 0) return g_qr_variant;
  if ((n <= 256 && batch >= 384) || (n <= 512 && batch >= 768)) return 4;
  if (n > 512 && n <= 1024 && batch >= 384) return 2;
  return 0;
}
static int g_bulges_w48 = 8;  // bulges per sweep of the W = 48 kernel (<= 12: >= 11 chase steps per window)
static int g_bulges_w32 = 6;  // bulges per sweep of the W = 32 kernel (<= 7: >= 10 chase steps per window)
void zgeev_set_qr_variant(int v) { g_qr_variant = (v < -1 || v > 4 || v == 3) ? -1 : v; }
void zgeev_set_bulges_w48(int nb) { g_bulges_w48 = nb < 1 ? 1 : (nb > 12 ? 12 : nb); }
void zgeev_set_bulges_w32(int nb) { g_bulges_w32 = nb < 1 ? 1 : (nb > 7 ? 7 : nb); }
int zgeev_get_bulges_w48() { return g_bulges_w48; }
int zgeev_get_bulges_w32() { return g_bulges_w32; }
int zgeev_get_qr_variant() { return g_qr_variant; }
// bit v set: QR kernel variant v was launched since the last zgeev_take_qr_kernels()
static std::atomic<unsigned> g_qr_ran{0};
unsigned zgeev_take_qr_kernels() { return g_qr_ran.exchange(0u); }
// QR workgroup slots of the GPU (256 CUs) for a variant: matrices one round of its
// workgroups holds
static int qr_slots(int v) { return v == 4 ? 1024 : (v == 2 ? 512 : 256); }
int zgeev_full_launch(int n) {
  return qr_slots(g_qr_variant >= 0 ? g_qr_variant : (n <= 512 ? 4 : (n <= 1024 ? 2 : 0)));
}
std::vector<int> zgeev_launch_plan(int n, int units, int per_unit) {
  std::vector<int> plan;
  if (units <= 0) return plan;
  if (per_unit < 1) per_unit = 1;
  const int full = zgeev_full_launch(n);
  const int per = full / per_unit > 1 ? full / per_unit : 1;
  plan.assign(units / per, per);
  if (units % per) plan.push_back(units % per);
  return plan;
}
int zgeev_get_aed() { return g_aed_window; }
int zgeev_get_bulges() { return g_bulges; }

// Schur factors of a strided batch (A destroyed): A = D Z T Z^H D^{-1} with D = diag(dbal)
// (balancing), Z unitary (ld n), T upper triangular (left in A); X (ld n, unit upper
// triangular) holds the eigenvectors of T, so that W = D Z X are eigenvectors of A
// (not normalized).
void zgeev_factors(hipStream_t s, int n, cd* A, long lda, long sA, cd* w, long sw, cd* Z, cd* X, cd* dbal,
                   int* info, int batch, Workspace& ws) {
  if (n <= 0 || batch <= 0) return;
  const long nn = (long)n * n;
  const int nref = n_reflectors(n);
  const int npanel = nref > 0 ? cdiv_up(nref, HB) : 1;
  DevBuf Tb = ws.get(sizeof(cd) * (size_t)npanel * HB * HB * batch);
  cd* Vf = X;  // the reflectors live in X until the eigenvectors overwrite them
  {
    StageScope st(ST_HESS);
    balance_kernel<<<batch, 256, 0, s>>>(n, A, lda, sA, dbal, n);
    launch_check("balance");
    gehrd(s, n, A, lda, sA, Vf, Tb.as<cd>(), batch, ws);
    orghr(s, n, Vf, Tb.as<cd>(), Z, n, nn, batch, ws);
  }
  {
    StageScope st(ST_QR);
    vrc_ctx* c = tl_ctx();
    const bool want = c && c->timing;
    DevBuf sb;
    if (want) sb = ws.get(sizeof(long long) * 5 * batch);
    long long* stp = want ? sb.as<long long>() : nullptr;
    const int qv = qr_variant_for(n, batch);
    g_qr_ran.fetch_or(1u << qv, std::memory_order_relaxed);
    if (qv == 4) {
      hqr_w32_kernel<<<batch, 128, 0, s>>>(n, A, lda, sA, Z, n, nn, w, sw, info, stp, g_aed_window, g_bulges_w32,
                                           g_max_sweeps);
    } else if (qv == 2) {
      hqr_w48_kernel<<<batch, 256, 0, s>>>(n, A, lda, sA, Z, n, nn, w, sw, info, stp,
                                           g_aed_window < 48 ? g_aed_window : 47, g_bulges_w48, g_max_sweeps);
    } else if (qv == 1) {
      hqr_lds_kernel<<<batch, QT, 0, s>>>(n, A, lda, sA, Z, n, nn, w, sw, info, stp, g_aed_window, g_bulges,
                                          g_max_sweeps);
    } else {
      hqr_kernel<<<batch, QT, 0, s>>>(n, A, lda, sA, Z, n, nn, w, sw, info, stp, g_aed_window, g_bulges,
                                      g_max_sweeps);
    }
    launch_check("hqr");
    health_note_info(s, info, batch, H_EIG);
    if (want) {
      std::vector<long long> h(5 * (size_t)batch);
      VRC_HIP(hipMemcpyAsync(h.data(), sb.ptr, sizeof(long long) * h.size(), hipMemcpyDeviceToHost, s));
      VRC_HIP(hipStreamSynchronize(s));
      c->qr_stats[0] += batch;
      for (int b = 0; b < batch; ++b)
        for (int k = 0; k < 5; ++k) c->qr_stats[1 + k] += h[(size_t)b * 5 + k];
    }
  }
  StageScope st(ST_EIGVEC);
  zset(s, n, n, mk(0.0, 0.0), mk(0.0, 0.0), X, n, nn, batch);
  dim3 g(cdiv_up(n, QW), batch);
  trevc_kernel<<<g, 256, 0, s>>>(n, A, lda, sA, X, n, nn);
  launch_check("trevc");
}

// VR = D Z X with LAPACK zgeev's normalisation (unit 2-norm, largest component real).
void zgeev_vectors(hipStream_t s, int n, const cd* Z, const cd* X, const cd* dbal, cd* VR, long ldv, long sV,
                   int batch) {
  if (n <= 0 || batch <= 0) return;
  StageScope st(ST_EIGVEC);
  const long nn = (long)n * n;
  zgemm_strided(s, OP_N, OP_N, n, n, n, mk(1.0, 0.0), Z, n, nn, X, n, nn, mk(0.0, 0.0), VR, ldv, sV, batch);
  zdscale(s, n, n, dbal, n, nullptr, 0, 0, VR, ldv, sV, VR, ldv, sV, batch);  // undo the balancing
  dim3 gn(cdiv_up(n, 4), batch);
  normalize_cols_kernel<<<gn, 256, 0, s>>>(n, VR, ldv, sV);
  launch_check("normalize_cols");
}

void zgeev_batched(hipStream_t s, int n, cd* A, long lda, long sA, cd* w, long sw, cd* VR, long ldv, long sV,
                   int* info, int batch, Workspace& ws) {
  if (n <= 0 || batch <= 0) return;
  const long nn = (long)n * n;
  DevBuf Zb = ws.get(sizeof(cd) * (size_t)nn * batch), Xb = ws.get(sizeof(cd) * (size_t)nn * batch);
  DevBuf Db = ws.get(sizeof(cd) * (size_t)n * batch);
  zgeev_factors(s, n, A, lda, sA, w, sw, Zb.as<cd>(), Xb.as<cd>(), Db.as<cd>(), info, batch, ws);
  zgeev_vectors(s, n, Zb.as<cd>(), Xb.as<cd>(), Db.as<cd>(), VR, ldv, sV, batch);
}

__global__ void gamma_sqrt_kernel(int n, const cd* lam, long sl, cd* gam, long sg) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  gam[b * sg + i] = gamma_branch(lam[b * sl + i]);
}

void gamma_sqrt(hipStream_t s, int n, const cd* lam, long sl, cd* gam, long sg, int batch) {
  if (n <= 0 || batch <= 0) return;
  dim3 g(cdiv_up(n, 256), batch);
  gamma_sqrt_kernel<<<g, 256, 0, s>>>(n, lam, sl, gam, sg);
  launch_check("gamma_sqrt");
  if (sg == n) health_note_gamma(s, gam, (long)n * batch);
}

}  // namespace vrc
