// Part (a): Fourier convolution matrices of a piecewise-constant permittivity layer
// (replaces FourierSolver2D::solve, src/rcwa/FourierSolver2D.cpp:12-351, and the CUDA
// kernels of src/gpu/CudaFourierSolver2D.cu).
//
// For a batch of layers (same nx, ny) the three rules are produced as:
//   fe33 (Laurent, ContinuousXY):  F2[p][q] = sum_i FX_i[p] FY1_i[q]; fe33 = BlockToeplitz(F2)
//   fe11 (inverse rule X):         G11[(a,c)][q] = sum_i inv(Toep(FXinv_i))[a][c] FY1_i[q]
//   fe22 (inverse rule Y):         G22[(b,d)][m] = sum_i' inv(Toep(FYinv_i'))[b][d] FX1_i'[m]
// where every 1D coefficient is the exact integral of a piecewise constant:
//   E(x0,x1,m,d) = f/d * exp(-i pi m (x0+x1)/d) * sinc(m (x1-x0)/d) * (x1-x0)
// (src/rcwa/FourierSolver1D.cpp:11-41). The final expansion writes the three n x n
// matrices once (the HBM-bound step); everything before it is O(cells * harmonics).
#include "vrc_common.hpp"
#include "vrc_internal.hpp"
#include "vrc_geom.hpp"

namespace vrc {

constexpr double PI_REF = 3.1415926535897932;  // src/rcwa/defns.h:29

// One term of FourierSolver1D::solve, evaluated in the reference's operation order.
__device__ inline cd fterm(cd f, double x0, double x1, int m, double d) {
  const double dx = x1 - x0;
  if (m == 0) return cscale(cscale(f, 1.0 / d), dx);
  cd t = mk(f.x / d, f.y / d);
  const double ang = ((double)m * PI_REF) * (x1 + x0) / d;
  double sn, cs;
  sincos(ang, &sn, &cs);
  t = cmul(t, mk(cs, -sn));
  const double sx = (double)m * dx / d;
  const double sc = (sx == 0.0) ? 1.0 : sin(PI_REF * sx) / (PI_REF * sx);
  return cscale(cscale(t, sc), dx);
}

// Output slots per layer l of the 1D coefficient kernel:
//  FX [MY][M1], FXi[MY][M1], FY1[MY][M2], FYi[MX][M2], FX1[MX][M1]
__global__ void fourier_coef_kernel(FourierDev g) {
  const int l = blockIdx.y;
  const int ncx = g.ncx[l], ncy = g.ncy[l];
  const double Lx = g.Lx[l], Ly = g.Ly[l];
  const int M1 = 2 * g.nx - 1, M2 = 2 * g.ny - 1;
  const int hx = (M1 - 1) / 2, hy = (M2 - 1) / 2;
  const long MX = g.MX, MY = g.MY;
  const cd* val = g.val + (long)l * MY * MX;
  const double* xu = g.xu + (long)l * (MX + 1);
  const double* xs = g.xs + (long)l * (MX + 1);
  const double* ys = g.ys + (long)l * (MY + 1);
  const long nFX = MY * M1, nFY1 = MY * M2, nFYi = MX * M2, nFX1 = MX * M1;
  const long total = 2 * nFX + nFY1 + nFYi + nFX1;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    if (e < 2 * nFX) {
      if (g.fftx[l]) continue;  // fourier_fft_x_kernel
      const bool inv = e >= nFX;
      const long ee = inv ? e - nFX : e;
      const int i = (int)(ee / M1), p = (int)(ee % M1);
      cd acc = mk(0.0, 0.0);
      if (i < ncy) {
        const int m = p - hx;
        for (int j = 0; j < ncx; ++j) {
          cd f = val[(long)i * MX + j];
          if (inv) f = crecip(f);
          acc = cadd(acc, fterm(f, xu[j], xu[j + 1], m, Lx));
        }
      }
      (inv ? g.FXi : g.FX)[(long)l * nFX + ee] = acc;
    } else if (e < 2 * nFX + nFY1) {
      const long ee = e - 2 * nFX;
      const int i = (int)(ee / M2), q = (int)(ee % M2);
      cd r = mk(0.0, 0.0);
      if (i < ncy) r = fterm(mk(1.0, 0.0), ys[i], ys[i + 1], q - hy, Ly);
      g.FY1[(long)l * nFY1 + ee] = r;
    } else if (e < 2 * nFX + nFY1 + nFYi) {
      if (g.ffty[l]) continue;  // fourier_fft_y_kernel
      const long ee = e - 2 * nFX - nFY1;
      const int i = (int)(ee / M2), q = (int)(ee % M2);
      cd acc = mk(0.0, 0.0);
      if (i < ncx) {
        for (int s = 0; s < ncy; ++s)
          acc = cadd(acc, fterm(crecip(val[(long)s * MX + i]), ys[s], ys[s + 1], q - hy, Ly));
      }
      g.FYi[(long)l * nFYi + ee] = acc;
    } else {
      const long ee = e - 2 * nFX - nFY1 - nFYi;
      const int i = (int)(ee / M1), p = (int)(ee % M1);
      cd r = mk(0.0, 0.0);
      if (i < ncx) r = fterm(mk(1.0, 0.0), xs[i], xs[i + 1], p - hx, Lx);
      g.FX1[(long)l * nFX1 + ee] = r;
    }
  }
}

// ---------------------------------------------------------------- uniform pixel grids
// For G cells of width d/G starting at x0 the exact coefficient is a scaled DFT:
//   F_m = sum_j f_j (1/G) exp(-i pi m (2 x0/d + (2j+1)/G)) sinc(m/G)
//       = (1/G) sinc(m/G) exp(-i pi m (1/G + 2 x0/d)) DFT(f)_(m mod G),
// so each strip (row or column) is transformed once by a radix-2 FFT in LDS and the
// 2 nx - 1 (2 ny - 1) needed orders are read out with that correction.

// In-place radix-2 decimation-in-time FFT of G = 2^lg points already in bit-reversed
// order; tw[k] = exp(-2 pi i k / G), k < G / 2. Ends with a barrier.
__device__ void lds_fft(cd* a, int G, int lg, const cd* tw) {
  for (int st = 1; st <= lg; ++st) {
    const int half = 1 << (st - 1);
    const int step = G >> st;
    for (int b = threadIdx.x; b < G / 2; b += blockDim.x) {
      const int k = b & (half - 1);
      const int i0 = ((b >> (st - 1)) << st) + k, i1 = i0 + half;
      const cd v = cmul(a[i1], tw[k * step]);
      const cd u = a[i0];
      a[i0] = cadd(u, v);
      a[i1] = csub(u, v);
    }
    __syncthreads();
  }
}

__device__ inline void fft_twiddles(cd* tw, int G) {
  for (int k = threadIdx.x; k < G / 2; k += blockDim.x) {
    double sn, cs;
    sincospi(2.0 * k / G, &sn, &cs);
    tw[k] = mk(cs, -sn);
  }
}

// (1/G) sinc(m/G) exp(-i pi m (1/G + 2 x0/d)) for order m.
__device__ inline cd fft_correction(int m, int G, double x0, double d) {
  const double invG = 1.0 / G;
  if (m == 0) return mk(invG, 0.0);
  const double sx = (double)m * invG;
  const double sc = sin(PI_REF * sx) / (PI_REF * sx);
  double sn, cs;
  sincos(PI_REF * (double)m * (invG + 2.0 * x0 / d), &sn, &cs);
  return mk(invG * sc * cs, -invG * sc * sn);
}

// FX (Laurent rows) and FXi (rows of 1/eps) of every strip of the layers whose x cells
// form a uniform power-of-two grid; one workgroup per (strip, layer). Dynamic LDS:
// 1.5 G complex values.
__global__ __launch_bounds__(256) void fourier_fft_x_kernel(FourierDev g) {
  const int l = blockIdx.y, i = blockIdx.x;
  const int lg = g.fftx[l];
  if (!lg) return;
  if (i >= g.ncy[l]) {  // strips beyond the layer's count: zero tables (the contraction may read them)
    const int M1 = 2 * g.nx - 1;
    const long o = (long)l * g.MY * M1 + (long)i * M1;
    for (int p = threadIdx.x; p < M1; p += blockDim.x) g.FX[o + p] = g.FXi[o + p] = mk(0.0, 0.0);
    return;
  }
  extern __shared__ cd fsm[];
  const int G = 1 << lg;
  cd* a = fsm;
  cd* tw = fsm + G;
  fft_twiddles(tw, G);
  const long MX = g.MX, MY = g.MY;
  const int M1 = 2 * g.nx - 1, hx = g.nx - 1;
  const double d = g.Lx[l], x0 = g.xu[(long)l * (MX + 1)];
  const cd* row = g.val + ((long)l * MY + i) * MX;
  const long o = (long)l * MY * M1 + (long)i * M1;
  for (int pass = 0; pass < 2; ++pass) {
    for (int j = threadIdx.x; j < G; j += blockDim.x) {
      const cd f = row[j];
      a[__brev((unsigned)j) >> (32 - lg)] = pass ? crecip(f) : f;
    }
    __syncthreads();
    lds_fft(a, G, lg, tw);
    for (int p = threadIdx.x; p < M1; p += blockDim.x) {
      const int m = p - hx;
      const int k = ((m % G) + G) % G;
      (pass ? g.FXi : g.FX)[o + p] = cmul(a[k], fft_correction(m, G, x0, d));
    }
    __syncthreads();
  }
}

// FYi (columns of 1/eps along y) of the layers whose y strips form a uniform
// power-of-two grid; one workgroup per (x cell, layer).
__global__ __launch_bounds__(256) void fourier_fft_y_kernel(FourierDev g) {
  const int l = blockIdx.y, i = blockIdx.x;
  const int lg = g.ffty[l];
  if (!lg) return;
  if (i >= g.ncx[l]) {
    const int M2 = 2 * g.ny - 1;
    const long o = (long)l * g.MX * M2 + (long)i * M2;
    for (int q = threadIdx.x; q < M2; q += blockDim.x) g.FYi[o + q] = mk(0.0, 0.0);
    return;
  }
  extern __shared__ cd fsm[];
  const int G = 1 << lg;
  cd* a = fsm;
  cd* tw = fsm + G;
  fft_twiddles(tw, G);
  const long MX = g.MX, MY = g.MY;
  const int M2 = 2 * g.ny - 1, hy = g.ny - 1;
  const double d = g.Ly[l], y0 = g.ys[(long)l * (MY + 1)];
  const cd* col = g.val + (long)l * MY * MX + i;
  for (int s = threadIdx.x; s < G; s += blockDim.x) a[__brev((unsigned)s) >> (32 - lg)] = crecip(col[(long)s * MX]);
  __syncthreads();
  lds_fft(a, G, lg, tw);
  const long o = (long)l * MX * M2 + (long)i * M2;
  for (int q = threadIdx.x; q < M2; q += blockDim.x) {
    const int m = q - hy;
    const int k = ((m % G) + G) % G;
    g.FYi[o + q] = cmul(a[k], fft_correction(m, G, y0, d));
  }
}

// The same inverse for k <= TINV_SMALL_K with one wavefront per matrix and LDS sized to
// the matrix (k x (2k + 1) complex): many matrices share a CU (the fixed 64 x 129 array of
// the kernel below allows one). Pivot: first maximum of |re| + |im| in the column, found
// by a wave reduction (the same row the serial search picks).
constexpr int TINV_SMALL_K = 44;  // k (2k + 1) * 16 B <= 64 KB
__global__ __launch_bounds__(64) void toeplitz_inv_small_kernel(int k, const cd* v, long sv, cd* out, long so,
                                                                const int* count, int per_layer) {
  extern __shared__ cd tsm[];
  const int ld = 2 * k + 1;
  const long id = blockIdx.x;
  const int layer = (int)(id / per_layer), slot = (int)(id % per_layer);
  v += id * sv;
  out += id * so;
  const int t = threadIdx.x;
  if (slot >= count[layer]) {  // beyond the layer's strips: zero (the contraction may read it)
    for (int e = t; e < k * k; e += 64) out[e] = mk(0.0, 0.0);
    return;
  }
  for (int e = t; e < k * 2 * k; e += 64) {
    const int r = e % k, c = e / k;
    tsm[r * ld + c] = (c < k) ? v[k - 1 + r - c] : ((c - k == r) ? mk(1.0, 0.0) : mk(0.0, 0.0));
  }
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    double best = -1.0;
    int p = 0x7fffffff;
    for (int r = j + t; r < k; r += 64) {
      const double a = cabs1(tsm[r * ld + j]);
      if (a > best) {
        best = a;
        p = r;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double ob = __shfl_xor(best, o, 64);
      const int op = __shfl_xor(p, o, 64);
      if (ob > best || (ob == best && op < p)) {
        best = ob;
        p = op;
      }
    }
    if (!(p >= j && p < k)) p = j;  // no comparable candidate (NaN column): keep the diagonal
    if (p != j) {
      for (int c = t; c < 2 * k; c += 64) {
        const cd tmp = tsm[j * ld + c];
        tsm[j * ld + c] = tsm[p * ld + c];
        tsm[p * ld + c] = tmp;
      }
    }
    __syncthreads();
    const cd inv = crecip(tsm[j * ld + j]);
    __syncthreads();
    for (int c = t; c < 2 * k; c += 64) tsm[j * ld + c] = cmul(tsm[j * ld + c], inv);
    __syncthreads();
    for (int e = t; e < k * 2 * k; e += 64) {
      const int r = e % k, c = e / k;
      if (r != j && c != j) tsm[r * ld + c] = csub(tsm[r * ld + c], cmul(tsm[r * ld + j], tsm[j * ld + c]));
    }
    __syncthreads();
    for (int r = t; r < k; r += 64)
      if (r != j) tsm[r * ld + j] = mk(0.0, 0.0);
    __syncthreads();
  }
  for (int e = t; e < k * k; e += 64) {
    const int r = e % k, c = e / k;
    out[r + (long)c * k] = tsm[r * ld + k + c];
  }
}

// Inverse of the k x k Toeplitz matrix T(r,c) = v[k-1 + r - c] (ToeplitzMatrixXcs,
// src/rcwa/ToeplitzMatrixXcs.cpp:8-44) by Gauss-Jordan elimination with partial
// pivoting in LDS; one workgroup per matrix. Inactive slots (beyond the layer's
// strip/column count) are zeroed.
__global__ __launch_bounds__(256) void toeplitz_inv_kernel(int k, const cd* v, long sv, cd* out, long so,
                                                           const int* count, int per_layer) {
  __shared__ cd M[64][129];
  __shared__ int piv;
  const long id = blockIdx.x;
  const int layer = (int)(id / per_layer), slot = (int)(id % per_layer);
  v += id * sv;
  out += id * so;
  const int t = threadIdx.x;
  if (slot >= count[layer]) {  // beyond the layer's strips: zero (the contraction may read it)
    for (int e = t; e < k * k; e += 256) out[e] = mk(0.0, 0.0);
    return;
  }
  for (int e = t; e < k * 2 * k; e += 256) {
    int r = e % k, c = e / k;
    M[r][c] = (c < k) ? v[k - 1 + r - c] : ((c - k == r) ? mk(1.0, 0.0) : mk(0.0, 0.0));
  }
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    if (t == 0) {
      int p = j;
      double best = cabs1(M[j][j]);
      for (int r = j + 1; r < k; ++r) {
        double a = cabs1(M[r][j]);
        if (a > best) { best = a; p = r; }
      }
      piv = p;
    }
    __syncthreads();
    const int p = piv;
    if (p != j) {
      for (int c = t; c < 2 * k; c += 256) {
        cd tmp = M[j][c];
        M[j][c] = M[p][c];
        M[p][c] = tmp;
      }
    }
    __syncthreads();
    const cd inv = crecip(M[j][j]);
    __syncthreads();
    for (int c = t; c < 2 * k; c += 256) M[j][c] = cmul(M[j][c], inv);
    __syncthreads();
    for (int e = t; e < k * 2 * k; e += 256) {
      int r = e % k, c = e / k;
      if (r != j && c != j) M[r][c] = csub(M[r][c], cmul(M[r][j], M[j][c]));
    }
    __syncthreads();
    for (int r = t; r < k; r += 256)
      if (r != j) M[r][j] = mk(0.0, 0.0);
    __syncthreads();
  }
  for (int e = t; e < k * k; e += 256) {
    int r = e % k, c = e / k;
    out[r + (long)c * k] = M[r][k + c];
  }
}

static void toeplitz_inverses(hipStream_t s, int k, const cd* v, long sv, cd* out, long so, const int* count,
                              int per_layer, int nmat) {
  if (nmat <= 0) return;
  if (k <= TINV_SMALL_K) {
    const size_t lds = sizeof(cd) * (size_t)k * (2 * k + 1);
    toeplitz_inv_small_kernel<<<nmat, 64, lds, s>>>(k, v, sv, out, so, count, per_layer);
    launch_check("toeplitz_inv_small");
  } else {
    toeplitz_inv_kernel<<<nmat, 256, 0, s>>>(k, v, sv, out, so, count, per_layer);
    launch_check("toeplitz_inv");
  }
}

// F2[p][q], G11[(a,c)][q], G22[(b,d)][m] (sums over strips/columns in ascending order).
__global__ void fourier_contract_kernel(FourierDev g) {
  const int l = blockIdx.y;
  const int nx = g.nx, ny = g.ny, M1 = 2 * nx - 1, M2 = 2 * ny - 1;
  const int ncx = g.ncx[l], ncy = g.ncy[l];
  const long MX = g.MX, MY = g.MY;
  const long nF2 = (long)M1 * M2, nG11 = (long)nx * nx * M2, nG22 = (long)ny * ny * M1;
  const cd* FX = g.FX + (long)l * MY * M1;
  const cd* FY1 = g.FY1 + (long)l * MY * M2;
  const cd* FX1 = g.FX1 + (long)l * MX * M1;
  const cd* tX = g.tFFx + (long)l * MY * nx * nx;
  const cd* tY = g.tFFy + (long)l * MX * ny * ny;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nF2 + nG11 + nG22;
       e += (long)gridDim.x * blockDim.x) {
    cd acc = mk(0.0, 0.0);
    if (e < nF2) {
      const int p = (int)(e % M1), q = (int)(e / M1);
      for (int i = 0; i < ncy; ++i) acc = cadd(acc, cmul(FX[(long)i * M1 + p], FY1[(long)i * M2 + q]));
      g.F2[(long)l * nF2 + e] = acc;
    } else if (e < nF2 + nG11) {
      const long ee = e - nF2;
      const int ac = (int)(ee % (nx * nx)), q = (int)(ee / (nx * nx));
      for (int i = 0; i < ncy; ++i) acc = cadd(acc, cmul(FY1[(long)i * M2 + q], tX[(long)i * nx * nx + ac]));
      g.G11[(long)l * nG11 + ee] = acc;
    } else {
      const long ee = e - nF2 - nG11;
      const int bd = (int)(ee % (ny * ny)), m = (int)(ee / (ny * ny));
      for (int i = 0; i < ncx; ++i) acc = cadd(acc, cmul(FX1[(long)i * M1 + m], tY[(long)i * ny * ny + bd]));
      g.G22[(long)l * nG22 + ee] = acc;
    }
  }
}

// Expansion into the three n x n column-major matrices (row index a*ny+b).
__global__ void fourier_expand_kernel(FourierDev g, cd* fe11, cd* fe22, cd* fe33, long sfe) {
  const int l = blockIdx.z;
  const int nx = g.nx, ny = g.ny, n = nx * ny, M1 = 2 * nx - 1, M2 = 2 * ny - 1;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= n) return;
  const int a = r / ny, bb = r % ny, cc = c / ny, d = c % ny;
  const long o = (long)l * sfe + r + (long)c * n;
  const cd* F2 = g.F2 + (long)l * M1 * M2;
  const cd* G11 = g.G11 + (long)l * nx * nx * M2;
  const cd* G22 = g.G22 + (long)l * ny * ny * M1;
  fe33[o] = F2[(a - cc + nx - 1) + (long)(bb - d + ny - 1) * M1];
  fe11[o] = G11[(a + cc * nx) + (long)(bb - d + ny - 1) * nx * nx];
  fe22[o] = G22[(bb + d * ny) + (long)(a - cc + nx - 1) * ny * ny];
}

void fourier_tables(hipStream_t s, const FourierDev& g, int nlayers) {
  const int M1 = 2 * g.nx - 1, M2 = 2 * g.ny - 1;
  {
    dim3 grid(64, nlayers);
    fourier_coef_kernel<<<grid, 256, 0, s>>>(g);
    launch_check("fourier_coef");
  }
  if (g.fft_max > 0) {
    const size_t lds = sizeof(cd) * (size_t)g.fft_max * 3 / 2;
    fourier_fft_x_kernel<<<dim3((unsigned)g.MY, nlayers), 256, lds, s>>>(g);
    launch_check("fourier_fft_x");
    fourier_fft_y_kernel<<<dim3((unsigned)g.MX, nlayers), 256, lds, s>>>(g);
    launch_check("fourier_fft_y");
  }
  if (g.nx > 64 || g.ny > 64) throw ArgFailure("harmonic count per axis above 64 is not supported");
  toeplitz_inverses(s, g.nx, g.FXi, M1, g.tFFx, (long)g.nx * g.nx, g.ncy, (int)g.MY, nlayers * (int)g.MY);
  toeplitz_inverses(s, g.ny, g.FYi, M2, g.tFFy, (long)g.ny * g.ny, g.ncx, (int)g.MX, nlayers * (int)g.MX);
}

// The contraction as three batched products over the strip (column) index, for layers with
// many strips (pixel grids): F2 = FX^T FY1, G11 = tFFx FY1 (over strips), G22 = tFFy FX1
// (over x cells); slots beyond a layer's count hold zeros.
static void fourier_contract_gemm(hipStream_t s, const FourierDev& g, int nlayers) {
  const int nx = g.nx, ny = g.ny, M1 = 2 * nx - 1, M2 = 2 * ny - 1;
  const long MX = g.MX, MY = g.MY, nxx = (long)nx * nx, nyy = (long)ny * ny;
  const cd one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  zgemm_strided(s, OP_N, OP_T, M1, M2, (int)MY, one, g.FX, M1, MY * M1, g.FY1, M2, MY * M2, zero, g.F2, M1,
                (long)M1 * M2, nlayers);
  zgemm_strided(s, OP_N, OP_T, (int)nxx, M2, (int)MY, one, g.tFFx, nxx, MY * nxx, g.FY1, M2, MY * M2, zero, g.G11,
                nxx, nxx * M2, nlayers);
  zgemm_strided(s, OP_N, OP_T, (int)nyy, M1, (int)MX, one, g.tFFy, nyy, MX * nyy, g.FX1, M1, MX * M1, zero, g.G22,
                nyy, nyy * M1, nlayers);
}

void fourier_assemble(hipStream_t s, const FourierDev& g, int nlayers, cd* fe11, cd* fe22, cd* fe33, long sfe) {
  fourier_tables(s, g, nlayers);
  if (g.MX >= 64 || g.MY >= 64) {
    fourier_contract_gemm(s, g, nlayers);
  } else {
    dim3 grid(64, nlayers);
    fourier_contract_kernel<<<grid, 256, 0, s>>>(g);
    launch_check("fourier_contract");
  }
  const int n = g.nx * g.ny;
  dim3 grid(cdiv_up(n, 256), n, nlayers);
  fourier_expand_kernel<<<grid, 256, 0, s>>>(g, fe11, fe22, fe33, sfe);
  launch_check("fourier_expand");
}

// ===================================================================== shape tangents
// Derivative of a 1D coefficient at the coordinate x moving (FourierSolver1D::dSolve,
// src/rcwa/FourierSolver1D.cpp:93-143): one side's value / d * exp(-i 2 m pi x / d).
__device__ inline cd dterm(cd f, double x, int m, double d) {
  const cd t = mk(f.x / d, f.y / d);
  if (m == 0) return t;
  double sn, cs;
  sincos((2.0 * m * PI_REF) * x / d, &sn, &cs);
  return cmul(t, mk(cs, -sn));
}

// Per parameter p (layer 0 of g; cfs[r * npara + p]: coefficient of the coordinate at
// sorted position r, 0 when it is not varied):
//  dFX [p][i][m] = sum_r cfs[r][p] (val[i][r-1] - val[i][r]) term at xs[r]  (dSolve_v2 of f,
//                  src/rcwa/FourierSolver1D.cpp:195-249, for the Laurent rule)
//  dFXi[p][i][m]   likewise for 1 / val (the inverse rule along x)
//  dFX1[p][c][m] = cfs[c][p] (-term at xs[c]) + cfs[c+1][p] (term at xs[c+1])  (dSolve of the
//                  x-cell indicator at its two ends, src/rcwa/FourierSolver2D.cpp:188-212)
__global__ void shape_coef_kernel(FourierDev g, int npara, const double* cfs, cd* dFX, cd* dFXi, cd* dFX1) {
  const int ncx = g.ncx[0], ncy = g.ncy[0];
  const double Lx = g.Lx[0];
  const int M1 = 2 * g.nx - 1, hx = g.nx - 1;
  const long MX = g.MX, MY = g.MY;
  const long nFX = MY * M1, nFX1 = MX * M1, per = 2 * nFX + nFX1;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < per * npara; e += (long)gridDim.x * blockDim.x) {
    const int p = (int)(e / per);
    const long ee = e % per;
    if (ee < 2 * nFX) {
      const bool inv = ee >= nFX;
      const long k = inv ? ee - nFX : ee;
      const int i = (int)(k / M1), m = (int)(k % M1) - hx;
      cd acc = mk(0.0, 0.0);
      if (i < ncy) {
        for (int r = 1; r < ncx; ++r) {
          const double a = cfs[(long)r * npara + p];
          if (a == 0.0) continue;
          cd fl = g.val[(long)i * MX + r - 1], fr = g.val[(long)i * MX + r];
          if (inv) {
            fl = crecip(fl);
            fr = crecip(fr);
          }
          const cd d = cadd(dterm(fl, g.xs[r], m, Lx), dterm(mk(-fr.x, -fr.y), g.xs[r], m, Lx));
          acc = cadd(acc, cscale(d, a));
        }
      }
      (inv ? dFXi : dFX)[(long)p * nFX + k] = acc;
    } else {
      const long k = ee - 2 * nFX;
      const int c = (int)(k / M1), m = (int)(k % M1) - hx;
      cd acc = mk(0.0, 0.0);
      if (c < ncx) {
        const double a0 = cfs[(long)c * npara + p], a1 = cfs[(long)(c + 1) * npara + p];
        if (a0 != 0.0) acc = cadd(acc, cscale(dterm(mk(-1.0, 0.0), g.xs[c], m, Lx), a0));
        if (a1 != 0.0) acc = cadd(acc, cscale(dterm(mk(1.0, 0.0), g.xs[c + 1], m, Lx), a1));
      }
      dFX1[(long)p * nFX1 + k] = acc;
    }
  }
}

// D[b] = Toeplitz(v[b]) (k x k column-major, D(r,c) = v[k-1 + r - c]).
__global__ void toeplitz_dense_kernel(int k, const cd* v, long sv, cd* D, long sD) {
  const long b = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= k * k) return;
  const int r = e % k, c = e / k;
  D[b * sD + e] = v[b * sv + k - 1 + r - c];
}

void fourier_tangents(hipStream_t s, const FourierDev& g, int npara, const double* cfs, const ShapeWork& w,
                      cd* dfe11, cd* dfe22, cd* dfe33, long sfe) {
  const int nx = g.nx, ny = g.ny, n = nx * ny, M1 = 2 * nx - 1;
  const long MX = g.MX, MY = g.MY, nxx = (long)nx * nx;
  shape_coef_kernel<<<64, 256, 0, s>>>(g, npara, cfs, w.dFX, w.dFXi, w.dFX1);
  launch_check("shape_coef");
  // inverse rule along x: d inv(T) = -inv(T) dT inv(T) per strip (src/rcwa/FourierSolver2D.cpp:76-79)
  dim3 tg(cdiv_up(nxx, 256), (unsigned)(npara * MY));
  toeplitz_dense_kernel<<<tg, 256, 0, s>>>(nx, w.dFXi, M1, w.D, nxx);
  launch_check("toeplitz_dense");
  for (int p = 0; p < npara; ++p) {
    cd* Dp = w.D + (long)p * MY * nxx;
    cd* Yp = w.Y + (long)p * MY * nxx;
    cd* Tp = w.dtFF + (long)p * MY * nxx;
    zgemm_strided(s, OP_N, OP_N, nx, nx, nx, mk(1.0, 0.0), Dp, nx, nxx, g.tFFx, nx, nxx, mk(0.0, 0.0), Yp, nx, nxx,
                  (int)MY);
    zgemm_strided(s, OP_N, OP_N, nx, nx, nx, mk(-1.0, 0.0), g.tFFx, nx, nxx, Yp, nx, nxx, mk(0.0, 0.0), Tp, nx, nxx,
                  (int)MY);
  }
  // the forward contraction and expansion over the tangent tables
  for (int p = 0; p < npara; ++p) {
    FourierDev gp = g;
    gp.FX = w.dFX + (long)p * MY * M1;
    gp.tFFx = w.dtFF + (long)p * MY * nxx;
    gp.FX1 = w.dFX1 + (long)p * MX * M1;
    gp.F2 = w.F2;
    gp.G11 = w.G11;
    gp.G22 = w.G22;
    fourier_contract_kernel<<<dim3(64, 1), 256, 0, s>>>(gp);
    launch_check("fourier_contract");
    dim3 grid(cdiv_up(n, 256), n, 1);
    fourier_expand_kernel<<<grid, 256, 0, s>>>(gp, dfe11 + p * sfe, dfe22 + p * sfe, dfe33 + p * sfe, 0);
    launch_check("fourier_expand");
  }
}

}  // namespace vrc
