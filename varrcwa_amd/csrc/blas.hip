// Complex128 dense kernels for gfx950: FP64-MFMA ZGEMM (strided and descriptor
// batched) and the elementwise matrix kernels used between the factorizations.
//
// ZGEMM tiles: 64x64 and 128x64 output per workgroup (8 wavefronts by default; 4-wave
// variants selectable) from v_mfma_f64_16x16x4f64 blocks, K staged through LDS in
// chunks of 16 complex values, real and imaginary planes kept separate so that one
// complex multiply-add is four real MFMAs (standard 4M product, no 3M/Gauss trick).
#include "vrc_common.hpp"
#include "vrc_internal.hpp"

namespace vrc {

constexpr int GT_M = 64, GT_N = 64, GT_K = 16, GT_THREADS = 256;

template <int OPA, int OPB>
__device__ __forceinline__ void zgemm_tile(int m, int n, int k, cd alpha, const cd* __restrict__ A, long lda,
                                           const cd* __restrict__ B, long ldb, cd beta, cd* __restrict__ C,
                                           long ldc, int tm, int tn) {
  __shared__ double As_re[GT_K][GT_M + 1];
  __shared__ double As_im[GT_K][GT_M + 1];
  __shared__ double Bs_re[GT_K][GT_N + 1];
  __shared__ double Bs_im[GT_K][GT_N + 1];

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int row0 = tm * GT_M, col0 = tn * GT_N;

  d4 accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }

  // op(A) tile (GT_M x GT_K) and op(B) tile (GT_K x GT_N), 4 elements each per thread;
  // the next K chunk is fetched into registers while the current one is multiplied
  cd pa[4], pb[4];
  auto fetch = [&](int kk) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cd v = mk(0.0, 0.0);
      if (OPA == OP_N) {
        const int gr = row0 + (t & 63), gc = kk + (t >> 6) + 4 * i;
        if (gr < m && gc < k) v = A[gr + (long)gc * lda];
      } else {
        const int gr = row0 + (t >> 4) + 16 * i, gc = kk + (t & 15);
        if (gr < m && gc < k) v = A[gc + (long)gr * lda];
        if (OPA == OP_C) v.y = -v.y;
      }
      pa[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      cd v = mk(0.0, 0.0);
      if (OPB == OP_N) {
        const int gc = kk + (t & 15), gj = col0 + (t >> 4) + 16 * i;
        if (gc < k && gj < n) v = B[gc + (long)gj * ldb];
      } else {
        const int gc = kk + (t >> 6) + 4 * i, gj = col0 + (t & 63);
        if (gc < k && gj < n) v = B[gj + (long)gc * ldb];
        if (OPB == OP_C) v.y = -v.y;
      }
      pb[i] = v;
    }
  };
  fetch(0);
  for (int kk = 0; kk < k; kk += GT_K) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (OPA == OP_N) {
        As_re[(t >> 6) + 4 * i][t & 63] = pa[i].x;
        As_im[(t >> 6) + 4 * i][t & 63] = pa[i].y;
      } else {
        As_re[t & 15][(t >> 4) + 16 * i] = pa[i].x;
        As_im[t & 15][(t >> 4) + 16 * i] = pa[i].y;
      }
      if (OPB == OP_N) {
        Bs_re[t & 15][(t >> 4) + 16 * i] = pb[i].x;
        Bs_im[t & 15][(t >> 4) + 16 * i] = pb[i].y;
      } else {
        Bs_re[(t >> 6) + 4 * i][t & 63] = pb[i].x;
        Bs_im[(t >> 6) + 4 * i][t & 63] = pb[i].y;
      }
    }
    __syncthreads();
    if (kk + GT_K < k) fetch(kk + GT_K);
#pragma unroll
    for (int s = 0; s < GT_K / 4; ++s) {
      const int kr = s * 4 + (lane >> 4);
      double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        ar[b] = As_re[kr][wr * 32 + b * 16 + (lane & 15)];
        ai[b] = As_im[kr][wr * 32 + b * 16 + (lane & 15)];
        nai[b] = -ai[b];
        br[b] = Bs_re[kr][wc * 32 + b * 16 + (lane & 15)];
        bi[b] = Bs_im[kr][wc * 32 + b * 16 + (lane & 15)];
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
    __syncthreads();
  }

  // ---- epilogue: D layout of v_mfma_f64_16x16x4f64: register r of a lane holds row lane/16 + 4*r,
  // column lane%16 (established on gfx950 against a reference product)
  const bool beta0 = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gc = col0 + wc * 32 + j * 16 + (lane & 15);
      if (gc >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = row0 + wr * 32 + i * 16 + (lane >> 4) + 4 * r;
        if (gr >= m) continue;
        cd acc = mk(accR[i][j][r], accI[i][j][r]);
        cd out = cmul(alpha, acc);
        cd* cp = C + gr + (long)gc * ldc;
        if (!beta0) out = cadd(out, cmul(beta, *cp));
        *cp = out;
      }
    }
}

// The 64x64 tile on eight waves (512 threads): wave w owns rows 32*(w>>2) .. +32 and
// columns 16*(w&3) .. +16 (2x1 blocks); staging as above, 2 elements per thread.
template <int OPA, int OPB>
__device__ __forceinline__ void zgemm_tile8(int m, int n, int k, cd alpha, const cd* __restrict__ A, long lda,
                                            const cd* __restrict__ B, long ldb, cd beta, cd* __restrict__ C,
                                            long ldc, int tm, int tn) {
  __shared__ double As_re[GT_K][GT_M + 1];
  __shared__ double As_im[GT_K][GT_M + 1];
  __shared__ double Bs_re[GT_K][GT_N + 1];
  __shared__ double Bs_im[GT_K][GT_N + 1];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wr = wave >> 2, wc = wave & 3;
  const int row0 = tm * GT_M, col0 = tn * GT_N;
  d4 accR[2], accI[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    accR[i] = d4{0.0, 0.0, 0.0, 0.0};
    accI[i] = d4{0.0, 0.0, 0.0, 0.0};
  }
  // element q of this thread: A (row t&63, k (t>>6) + 8q) or (row (t>>4) + 32q, k t&15);
  // B (k t&15, col (t>>4) + 32q) or (k (t>>6) + 8q, col t&63)
  cd pa[2], pb[2];
  auto fetch = [&](int kk) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      cd v = mk(0.0, 0.0);
      if (OPA == OP_N) {
        const int gr = row0 + (t & 63), gc = kk + (t >> 6) + 8 * i;
        if (gr < m && gc < k) v = A[gr + (long)gc * lda];
      } else {
        const int gr = row0 + (t >> 4) + 32 * i, gc = kk + (t & 15);
        if (gr < m && gc < k) v = A[gc + (long)gr * lda];
        if (OPA == OP_C) v.y = -v.y;
      }
      pa[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      cd v = mk(0.0, 0.0);
      if (OPB == OP_N) {
        const int gc = kk + (t & 15), gj = col0 + (t >> 4) + 32 * i;
        if (gc < k && gj < n) v = B[gc + (long)gj * ldb];
      } else {
        const int gc = kk + (t >> 6) + 8 * i, gj = col0 + (t & 63);
        if (gc < k && gj < n) v = B[gj + (long)gc * ldb];
        if (OPB == OP_C) v.y = -v.y;
      }
      pb[i] = v;
    }
  };
  fetch(0);
  for (int kk = 0; kk < k; kk += GT_K) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      if (OPA == OP_N) {
        As_re[(t >> 6) + 8 * i][t & 63] = pa[i].x;
        As_im[(t >> 6) + 8 * i][t & 63] = pa[i].y;
      } else {
        As_re[t & 15][(t >> 4) + 32 * i] = pa[i].x;
        As_im[t & 15][(t >> 4) + 32 * i] = pa[i].y;
      }
      if (OPB == OP_N) {
        Bs_re[t & 15][(t >> 4) + 32 * i] = pb[i].x;
        Bs_im[t & 15][(t >> 4) + 32 * i] = pb[i].y;
      } else {
        Bs_re[(t >> 6) + 8 * i][t & 63] = pb[i].x;
        Bs_im[(t >> 6) + 8 * i][t & 63] = pb[i].y;
      }
    }
    __syncthreads();
    if (kk + GT_K < k) fetch(kk + GT_K);
#pragma unroll
    for (int s = 0; s < GT_K / 4; ++s) {
      const int kr = s * 4 + (lane >> 4);
      double ar[2], ai[2], nai[2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        ar[b] = As_re[kr][wr * 32 + b * 16 + (lane & 15)];
        ai[b] = As_im[kr][wr * 32 + b * 16 + (lane & 15)];
        nai[b] = -ai[b];
      }
      const double br = Bs_re[kr][wc * 16 + (lane & 15)], bi = Bs_im[kr][wc * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        accR[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br, accR[i], 0, 0, 0);
        accR[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi, accR[i], 0, 0, 0);
        accI[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi, accI[i], 0, 0, 0);
        accI[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br, accI[i], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const bool beta0 = (beta.x == 0.0 && beta.y == 0.0);
  const int gc = col0 + wc * 16 + (lane & 15);
  if (gc >= n) return;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int gr = row0 + wr * 32 + i * 16 + (lane >> 4) + 4 * r;
      if (gr >= m) continue;
      cd out = cmul(alpha, mk(accR[i][r], accI[i][r]));
      cd* cp = C + gr + (long)gc * ldc;
      if (!beta0) out = cadd(out, cmul(beta, *cp));
      *cp = out;
    }
}

// ---------------------------------------------------------------- large-tile ZGEMM
// 128x64 output per 256-thread workgroup: 2x2 wavefronts, each 64x32 = 4x2 blocks of
// v_mfma_f64_16x16x4f64 (re and im accumulators: 128 registers). K in chunks of 16;
// the next chunk is fetched from memory into registers while the current one is
// multiplied out of LDS. LDS rows are padded by one double so that both the staging
// stores (along m or along k) and the MFMA operand reads are free of bank conflicts.
constexpr int GB_M = 128, GB_N = 64, GB_K = 16;

template <int OPA, int OPB>
__device__ __forceinline__ void zgemm_tile_big(int m, int n, int k, cd alpha, const cd* __restrict__ A, long lda,
                                               const cd* __restrict__ B, long ldb, cd beta, cd* __restrict__ C,
                                               long ldc, int tm, int tn) {
  __shared__ double As_re[GB_K][GB_M + 1];
  __shared__ double As_im[GB_K][GB_M + 1];
  __shared__ double Bs_re[GB_K][GB_N + 1];
  __shared__ double Bs_im[GB_K][GB_N + 1];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int row0 = tm * GB_M, col0 = tn * GB_N;

  d4 accR[4][2], accI[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }
  // element e = t + 256*i of the A tile (8 per thread) and of the B tile (4 per thread)
  auto a_coord = [&](int e, int& r, int& c) {
    if (OPA == OP_N) { r = e & 127; c = e >> 7; } else { c = e & 15; r = e >> 4; }
  };
  auto b_coord = [&](int e, int& c, int& j) {
    if (OPB == OP_N) { c = e & 15; j = e >> 4; } else { j = e & 63; c = e >> 6; }
  };
  cd ra[8], rb[4];
  auto fetch = [&](int kk) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      int r, c;
      a_coord(t + 256 * i, r, c);
      const int gr = row0 + r, gc = kk + c;
      cd v = mk(0.0, 0.0);
      if (gr < m && gc < k) v = (OPA == OP_N) ? A[gr + (long)gc * lda] : A[gc + (long)gr * lda];
      if (OPA == OP_C) v.y = -v.y;
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int c, j;
      b_coord(t + 256 * i, c, j);
      const int gc = kk + c, gj = col0 + j;
      cd v = mk(0.0, 0.0);
      if (gc < k && gj < n) v = (OPB == OP_N) ? B[gc + (long)gj * ldb] : B[gj + (long)gc * ldb];
      if (OPB == OP_C) v.y = -v.y;
      rb[i] = v;
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      int r, c;
      a_coord(t + 256 * i, r, c);
      As_re[c][r] = ra[i].x;
      As_im[c][r] = ra[i].y;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int c, j;
      b_coord(t + 256 * i, c, j);
      Bs_re[c][j] = rb[i].x;
      Bs_im[c][j] = rb[i].y;
    }
  };
  fetch(0);
  for (int kk = 0; kk < k; kk += GB_K) {
    stash();
    __syncthreads();
    if (kk + GB_K < k) fetch(kk + GB_K);
#pragma unroll
    for (int s = 0; s < GB_K / 4; ++s) {
      const int kr = s * 4 + (lane >> 4);
      double ar[4], ai[4], nai[4], br[2], bi[2];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        ar[b] = As_re[kr][wr * 64 + b * 16 + (lane & 15)];
        ai[b] = As_im[kr][wr * 64 + b * 16 + (lane & 15)];
        nai[b] = -ai[b];
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        br[b] = Bs_re[kr][wc * 32 + b * 16 + (lane & 15)];
        bi[b] = Bs_im[kr][wc * 32 + b * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accR[i][j], 0, 0, 0);
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi[j], accR[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], accI[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], accI[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  const bool beta0 = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gc = col0 + wc * 32 + j * 16 + (lane & 15);
      if (gc >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = row0 + wr * 64 + i * 16 + (lane >> 4) + 4 * r;
        if (gr >= m) continue;
        cd out = cmul(alpha, mk(accR[i][j][r], accI[i][j][r]));
        cd* cp = C + gr + (long)gc * ldc;
        if (!beta0) out = cadd(out, cmul(beta, *cp));
        *cp = out;
      }
    }
}

// The 128x64 tile on eight waves (512 threads): wave w owns rows 32*(w>>1) .. +32 and
// columns 32*(w&1) .. +32 (2x2 blocks of 16x16); the K chunk is staged as above.
constexpr int GB8_THREADS = 512;

template <int OPA, int OPB>
__device__ __forceinline__ void zgemm_tile_big8(int m, int n, int k, cd alpha, const cd* __restrict__ A, long lda,
                                                const cd* __restrict__ B, long ldb, cd beta, cd* __restrict__ C,
                                                long ldc, int tm, int tn) {
  __shared__ double As_re[GB_K][GB_M + 1];
  __shared__ double As_im[GB_K][GB_M + 1];
  __shared__ double Bs_re[GB_K][GB_N + 1];
  __shared__ double Bs_im[GB_K][GB_N + 1];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int row0 = tm * GB_M, col0 = tn * GB_N;

  d4 accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }
  // element e = t + 512*i of the A tile (4 per thread) and of the B tile (2 per thread)
  auto a_coord = [&](int e, int& r, int& c) {
    if (OPA == OP_N) { r = e & 127; c = e >> 7; } else { c = e & 15; r = e >> 4; }
  };
  auto b_coord = [&](int e, int& c, int& j) {
    if (OPB == OP_N) { c = e & 15; j = e >> 4; } else { j = e & 63; c = e >> 6; }
  };
  cd ra[4], rb[2];
  auto fetch = [&](int kk) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int r, c;
      a_coord(t + GB8_THREADS * i, r, c);
      const int gr = row0 + r, gc = kk + c;
      cd v = mk(0.0, 0.0);
      if (gr < m && gc < k) v = (OPA == OP_N) ? A[gr + (long)gc * lda] : A[gc + (long)gr * lda];
      if (OPA == OP_C) v.y = -v.y;
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      int c, j;
      b_coord(t + GB8_THREADS * i, c, j);
      const int gc = kk + c, gj = col0 + j;
      cd v = mk(0.0, 0.0);
      if (gc < k && gj < n) v = (OPB == OP_N) ? B[gc + (long)gj * ldb] : B[gj + (long)gc * ldb];
      if (OPB == OP_C) v.y = -v.y;
      rb[i] = v;
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int r, c;
      a_coord(t + GB8_THREADS * i, r, c);
      As_re[c][r] = ra[i].x;
      As_im[c][r] = ra[i].y;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      int c, j;
      b_coord(t + GB8_THREADS * i, c, j);
      Bs_re[c][j] = rb[i].x;
      Bs_im[c][j] = rb[i].y;
    }
  };
  fetch(0);
  for (int kk = 0; kk < k; kk += GB_K) {
    stash();
    __syncthreads();
    if (kk + GB_K < k) fetch(kk + GB_K);
#pragma unroll
    for (int s = 0; s < GB_K / 4; ++s) {
      const int kr = s * 4 + (lane >> 4);
      double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        ar[b] = As_re[kr][wr * 32 + b * 16 + (lane & 15)];
        ai[b] = As_im[kr][wr * 32 + b * 16 + (lane & 15)];
        nai[b] = -ai[b];
        br[b] = Bs_re[kr][wc * 32 + b * 16 + (lane & 15)];
        bi[b] = Bs_im[kr][wc * 32 + b * 16 + (lane & 15)];
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
    __syncthreads();
  }
  const bool beta0 = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gc = col0 + wc * 32 + j * 16 + (lane & 15);
      if (gc >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = row0 + wr * 32 + i * 16 + (lane >> 4) + 4 * r;
        if (gr >= m) continue;
        cd out = cmul(alpha, mk(accR[i][j][r], accI[i][j][r]));
        cd* cp = C + gr + (long)gc * ldc;
        if (!beta0) out = cadd(out, cmul(beta, *cp));
        *cp = out;
      }
    }
}

// 128x128 tile on eight waves: wave w owns rows 64*(w>>2) .. +64 and columns 32*(w&3) .. +32
// (4x2 blocks of 16x16); K chunk of 16 staged through LDS as above.
constexpr int GQ_N = 128;

template <int OPA, int OPB>
__device__ __forceinline__ void zgemm_tile_sq(int m, int n, int k, cd alpha, const cd* __restrict__ A, long lda,
                                              const cd* __restrict__ B, long ldb, cd beta, cd* __restrict__ C,
                                              long ldc, int tm, int tn) {
  __shared__ double As_re[GB_K][GB_M + 1];
  __shared__ double As_im[GB_K][GB_M + 1];
  __shared__ double Bs_re[GB_K][GQ_N + 1];
  __shared__ double Bs_im[GB_K][GQ_N + 1];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wr = wave >> 2, wc = wave & 3;
  const int row0 = tm * GB_M, col0 = tn * GQ_N;
  d4 accR[4][2], accI[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = d4{0.0, 0.0, 0.0, 0.0};
      accI[i][j] = d4{0.0, 0.0, 0.0, 0.0};
    }
  // element e = t + 512*i of the A tile (4 per thread) and of the B tile (4 per thread)
  auto a_coord = [&](int e, int& r, int& c) {
    if (OPA == OP_N) { r = e & 127; c = e >> 7; } else { c = e & 15; r = e >> 4; }
  };
  auto b_coord = [&](int e, int& c, int& j) {
    if (OPB == OP_N) { c = e & 15; j = e >> 4; } else { j = e & 127; c = e >> 7; }
  };
  cd ra[4], rb[4];
  auto fetch = [&](int kk) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int r, c;
      a_coord(t + GB8_THREADS * i, r, c);
      const int gr = row0 + r, gc = kk + c;
      cd v = mk(0.0, 0.0);
      if (gr < m && gc < k) v = (OPA == OP_N) ? A[gr + (long)gc * lda] : A[gc + (long)gr * lda];
      if (OPA == OP_C) v.y = -v.y;
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int c, j;
      b_coord(t + GB8_THREADS * i, c, j);
      const int gc = kk + c, gj = col0 + j;
      cd v = mk(0.0, 0.0);
      if (gc < k && gj < n) v = (OPB == OP_N) ? B[gc + (long)gj * ldb] : B[gj + (long)gc * ldb];
      if (OPB == OP_C) v.y = -v.y;
      rb[i] = v;
    }
  };
  auto stash = [&]() {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int r, c;
      a_coord(t + GB8_THREADS * i, r, c);
      As_re[c][r] = ra[i].x;
      As_im[c][r] = ra[i].y;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int c, j;
      b_coord(t + GB8_THREADS * i, c, j);
      Bs_re[c][j] = rb[i].x;
      Bs_im[c][j] = rb[i].y;
    }
  };
  fetch(0);
  for (int kk = 0; kk < k; kk += GB_K) {
    stash();
    __syncthreads();
    if (kk + GB_K < k) fetch(kk + GB_K);
#pragma unroll
    for (int s = 0; s < GB_K / 4; ++s) {
      const int kr = s * 4 + (lane >> 4);
      double ar[4], ai[4], nai[4], br[2], bi[2];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        ar[b] = As_re[kr][wr * 64 + b * 16 + (lane & 15)];
        ai[b] = As_im[kr][wr * 64 + b * 16 + (lane & 15)];
        nai[b] = -ai[b];
      }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        br[b] = Bs_re[kr][wc * 32 + b * 16 + (lane & 15)];
        bi[b] = Bs_im[kr][wc * 32 + b * 16 + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], accR[i][j], 0, 0, 0);
          accR[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi[j], accR[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], accI[i][j], 0, 0, 0);
          accI[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], accI[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  const bool beta0 = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gc = col0 + wc * 32 + j * 16 + (lane & 15);
      if (gc >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = row0 + wr * 64 + i * 16 + (lane >> 4) + 4 * r;
        if (gr >= m) continue;
        cd out = cmul(alpha, mk(accR[i][j][r], accI[i][j][r]));
        cd* cp = C + gr + (long)gc * ldc;
        if (!beta0) out = cadd(out, cmul(beta, *cp));
        *cp = out;
      }
    }
}

// 1-D grid over (batch, n tile, m tile) with m tiles fastest. Workgroup ids are dealt
// to the 8 XCDs round-robin; the id is remapped so that every XCD works on one
// contiguous range of tiles (tiles of the same matrices share that XCD's L2).
template <int OPA, int OPB>
__global__ __launch_bounds__(GT_THREADS) void zgemm_big_kernel(int m, int n, int k, cd alpha, const cd* A, long lda,
                                                               long sA, const cd* B, long ldb, long sB, cd beta,
                                                               cd* C, long ldc, long sC, int gm, int gn, long total) {
  const long id = blockIdx.x;
  const long q = total / 8, rr = total % 8;
  const long x = id % 8, slot = id / 8;
  const long logical = x * q + (x < rr ? x : rr) + slot;
  const long per = (long)gm * gn;
  const long b = logical / per;
  const int rem = (int)(logical - b * per);
  const int tn = rem / gm, tm = rem - tn * gm;
  zgemm_tile_big<OPA, OPB>(m, n, k, alpha, A + b * sA, lda, B + b * sB, ldb, beta, C + b * sC, ldc, tm, tn);
}

template <int OPA, int OPB>
__global__ __launch_bounds__(GB8_THREADS) void zgemm_big8_kernel(int m, int n, int k, cd alpha, const cd* A, long lda,
                                                                 long sA, const cd* B, long ldb, long sB, cd beta,
                                                                 cd* C, long ldc, long sC, int gm, int gn,
                                                                 long total) {
  const long id = blockIdx.x;
  const long q = total / 8, rr = total % 8;
  const long x = id % 8, slot = id / 8;
  const long logical = x * q + (x < rr ? x : rr) + slot;
  const long per = (long)gm * gn;
  const long b = logical / per;
  const int rem = (int)(logical - b * per);
  const int tn = rem / gm, tm = rem - tn * gm;
  zgemm_tile_big8<OPA, OPB>(m, n, k, alpha, A + b * sA, lda, B + b * sB, ldb, beta, C + b * sC, ldc, tm, tn);
}

template <int OPA, int OPB>
__global__ __launch_bounds__(GB8_THREADS) void zgemm_sq_kernel(int m, int n, int k, cd alpha, const cd* A, long lda,
                                                               long sA, const cd* B, long ldb, long sB, cd beta,
                                                               cd* C, long ldc, long sC, int gm, int gn, long total) {
  const long id = blockIdx.x;
  const long q = total / 8, rr = total % 8;
  const long x = id % 8, slot = id / 8;
  const long logical = x * q + (x < rr ? x : rr) + slot;
  const long per = (long)gm * gn;
  const long b = logical / per;
  const int rem = (int)(logical - b * per);
  const int tn = rem / gm, tm = rem - tn * gm;
  zgemm_tile_sq<OPA, OPB>(m, n, k, alpha, A + b * sA, lda, B + b * sB, ldb, beta, C + b * sC, ldc, tm, tn);
}

// Scale-only path for k == 0 (C = beta*C), shared by both batch modes.
__device__ __forceinline__ void scale_tile(int m, int n, cd beta, cd* C, long ldc, int tm, int tn) {
  const bool beta0 = (beta.x == 0.0 && beta.y == 0.0);
  for (int e = threadIdx.x; e < GT_M * GT_N; e += GT_THREADS) {
    int r = e % GT_M, c = e / GT_M;
    int gr = tm * GT_M + r, gc = tn * GT_N + c;
    if (gr < m && gc < n) {
      cd* p = C + gr + (long)gc * ldc;
      *p = beta0 ? mk(0.0, 0.0) : cmul(beta, *p);
    }
  }
}

template <int OPA, int OPB>
__global__ __launch_bounds__(GT_THREADS) void zgemm_strided_kernel(int m, int n, int k, cd alpha, const cd* A,
                                                                   long lda, long sA, const cd* B, long ldb,
                                                                   long sB, cd beta, cd* C, long ldc, long sC) {
  const int tm = blockIdx.x, tn = blockIdx.y;
  const long b = blockIdx.z;
  if (k <= 0) {
    scale_tile(m, n, beta, C + b * sC, ldc, tm, tn);
    return;
  }
  zgemm_tile<OPA, OPB>(m, n, k, alpha, A + b * sA, lda, B + b * sB, ldb, beta, C + b * sC, ldc, tm, tn);
}

template <int OPA, int OPB>
__global__ __launch_bounds__(512) void zgemm_strided8_kernel(int m, int n, int k, cd alpha, const cd* A, long lda,
                                                            long sA, const cd* B, long ldb, long sB, cd beta, cd* C,
                                                            long ldc, long sC) {
  const int tm = blockIdx.x, tn = blockIdx.y;
  const long b = blockIdx.z;
  zgemm_tile8<OPA, OPB>(m, n, k, alpha, A + b * sA, lda, B + b * sB, ldb, beta, C + b * sC, ldc, tm, tn);
}

template <int OPA, int OPB>
__global__ __launch_bounds__(GT_THREADS) void zgemm_desc_kernel(const GemmDesc* __restrict__ desc) {
  const GemmDesc d = desc[blockIdx.z];
  const int tm = blockIdx.x, tn = blockIdx.y;
  if (d.m <= 0 || d.n <= 0) return;
  if (tm * GT_M >= d.m || tn * GT_N >= d.n) return;
  if (d.k <= 0) {
    scale_tile(d.m, d.n, d.beta, d.C, d.ldc, tm, tn);
    return;
  }
  zgemm_tile<OPA, OPB>(d.m, d.n, d.k, d.alpha, d.A, d.lda, d.B, d.ldb, d.beta, d.C, d.ldc, tm, tn);
}

static bool g_big_gemm = true;  // large-tile kernel for m >= 128, n >= 64 (switchable for comparison)
static int g_big_waves = 8;     // wavefronts of the large-tile kernel (4 or 8)
void zgemm_set_big_waves(int w) { g_big_waves = w == 8 ? 8 : 4; }
static int g_small_waves = 8;   // wavefronts of the 64x64-tile kernel (4 or 8)
static int g_big_n = 128;       // columns of the large tile (128 for n >= 128, else 64)
void zgemm_set_big_n(int nn) { g_big_n = nn == 128 ? 128 : 64; }
void zgemm_set_small_waves(int w) { g_small_waves = w == 8 ? 8 : 4; }
static int g_big_min_k = 129;   // ... and k at least this (short-k updates run better on the 64x64 tile)
void zgemm_set_big(bool on) { g_big_gemm = on; }
void zgemm_set_big_min_k(int k) { g_big_min_k = k; }
void zgemm_get_knobs(int* big_n, int* big_waves, int* small_waves, int* big_min_k) {
  *big_n = g_big_n;
  *big_waves = g_big_waves;
  *small_waves = g_small_waves;
  *big_min_k = g_big_min_k;
}

void zgemm_strided(hipStream_t s, int opA, int opB, int m, int n, int k, cd alpha, const cd* A, long lda,
                   long sA, const cd* B, long ldb, long sB, cd beta, cd* C, long ldc, long sC, int batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  if (g_big_gemm && g_big_n == 128 && k > 0 && k >= g_big_min_k && m >= GB_M && n >= GQ_N) {
    const int gm = cdiv_up(m, GB_M), gn = cdiv_up(n, GQ_N);
    const long total = (long)gm * gn * batch;
    dim3 grid((unsigned)total);
#define VRC_SQ(OA, OB) \
  zgemm_sq_kernel<OA, OB><<<grid, GB8_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, gm, gn, total)
    switch (opA * 3 + opB) {
      case 0: VRC_SQ(OP_N, OP_N); break;
      case 1: VRC_SQ(OP_N, OP_T); break;
      case 2: VRC_SQ(OP_N, OP_C); break;
      case 3: VRC_SQ(OP_T, OP_N); break;
      case 4: VRC_SQ(OP_T, OP_T); break;
      case 5: VRC_SQ(OP_T, OP_C); break;
      case 6: VRC_SQ(OP_C, OP_N); break;
      case 7: VRC_SQ(OP_C, OP_T); break;
      default: VRC_SQ(OP_C, OP_C); break;
    }
#undef VRC_SQ
    launch_check("zgemm_sq");
    return;
  }
  if (g_big_gemm && k > 0 && k >= g_big_min_k && m >= GB_M && n >= GB_N) {
    const int gm = cdiv_up(m, GB_M), gn = cdiv_up(n, GB_N);
    const long total = (long)gm * gn * batch;
    dim3 grid((unsigned)total);
#define VRC_BIG(OA, OB)                                                                                              \
  if (g_big_waves == 8)                                                                                              \
    zgemm_big8_kernel<OA, OB><<<grid, GB8_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, \
                                                           gm, gn, total);                                           \
  else                                                                                                               \
    zgemm_big_kernel<OA, OB><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC,   \
                                                         gm, gn, total)
    switch (opA * 3 + opB) {
      case 0: VRC_BIG(OP_N, OP_N); break;
      case 1: VRC_BIG(OP_N, OP_T); break;
      case 2: VRC_BIG(OP_N, OP_C); break;
      case 3: VRC_BIG(OP_T, OP_N); break;
      case 4: VRC_BIG(OP_T, OP_T); break;
      case 5: VRC_BIG(OP_T, OP_C); break;
      case 6: VRC_BIG(OP_C, OP_N); break;
      case 7: VRC_BIG(OP_C, OP_T); break;
      default: VRC_BIG(OP_C, OP_C); break;
    }
#undef VRC_BIG
    launch_check("zgemm_big");
    return;
  }
  dim3 grid(cdiv_up(m, GT_M), cdiv_up(n, GT_N), batch);
  if (g_small_waves == 8 && k > 0) {
#define VRC_S8(OA, OB) \
  zgemm_strided8_kernel<OA, OB><<<grid, 512, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC)
    switch (opA * 3 + opB) {
      case 0: VRC_S8(OP_N, OP_N); break;
      case 1: VRC_S8(OP_N, OP_T); break;
      case 2: VRC_S8(OP_N, OP_C); break;
      case 3: VRC_S8(OP_T, OP_N); break;
      case 4: VRC_S8(OP_T, OP_T); break;
      case 5: VRC_S8(OP_T, OP_C); break;
      case 6: VRC_S8(OP_C, OP_N); break;
      case 7: VRC_S8(OP_C, OP_T); break;
      default: VRC_S8(OP_C, OP_C); break;
    }
#undef VRC_S8
    launch_check("zgemm_strided8");
    return;
  }
  switch (opA * 3 + opB) {
    case 0: zgemm_strided_kernel<OP_N, OP_N><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 1: zgemm_strided_kernel<OP_N, OP_T><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 2: zgemm_strided_kernel<OP_N, OP_C><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 3: zgemm_strided_kernel<OP_T, OP_N><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 4: zgemm_strided_kernel<OP_T, OP_T><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 5: zgemm_strided_kernel<OP_T, OP_C><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 6: zgemm_strided_kernel<OP_C, OP_N><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    case 7: zgemm_strided_kernel<OP_C, OP_T><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
    default: zgemm_strided_kernel<OP_C, OP_C><<<grid, GT_THREADS, 0, s>>>(m, n, k, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC); break;
  }
  launch_check("zgemm_strided");
}

void zgemm_desc(hipStream_t s, int opA, int opB, const GemmDesc* d_desc, int count, int max_m, int max_n) {
  if (count <= 0 || max_m <= 0 || max_n <= 0) return;
  dim3 grid(cdiv_up(max_m, GT_M), cdiv_up(max_n, GT_N), count);
  switch (opA * 3 + opB) {
    case 0: zgemm_desc_kernel<OP_N, OP_N><<<grid, GT_THREADS, 0, s>>>(d_desc); break;
    case 1: zgemm_desc_kernel<OP_N, OP_T><<<grid, GT_THREADS, 0, s>>>(d_desc); break;
    case 2: zgemm_desc_kernel<OP_N, OP_C><<<grid, GT_THREADS, 0, s>>>(d_desc); break;
    case 3: zgemm_desc_kernel<OP_T, OP_N><<<grid, GT_THREADS, 0, s>>>(d_desc); break;
    case 4: zgemm_desc_kernel<OP_T, OP_T><<<grid, GT_THREADS, 0, s>>>(d_desc); break;
    case 5: zgemm_desc_kernel<OP_T, OP_C><<<grid, GT_THREADS, 0, s>>>(d_desc); break;
    case 6: zgemm_desc_kernel<OP_C, OP_N><<<grid, GT_THREADS, 0, s>>>(d_desc); break;
    case 7: zgemm_desc_kernel<OP_C, OP_T><<<grid, GT_THREADS, 0, s>>>(d_desc); break;
    default: zgemm_desc_kernel<OP_C, OP_C><<<grid, GT_THREADS, 0, s>>>(d_desc); break;
  }
  launch_check("zgemm_desc");
}

// ============================================================ elementwise kernels
// All take a strided batch of m x n column-major matrices.

__global__ void geam_kernel(int m, int n, cd a, const cd* A, long lda, long sA, cd b, const cd* B, long ldb,
                            long sB, cd* C, long ldc, long sC) {
  const long bz = blockIdx.z;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= m) return;
  cd v = cmul(a, A[bz * sA + r + (long)c * lda]);
  if (B) v = cadd(v, cmul(b, B[bz * sB + r + (long)c * ldb]));
  C[bz * sC + r + (long)c * ldc] = v;
}

void zgeam(hipStream_t s, int m, int n, cd a, const cd* A, long lda, long sA, cd b, const cd* B, long ldb, long sB,
           cd* C, long ldc, long sC, int batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid(cdiv_up(m, 256), n, batch);
  geam_kernel<<<grid, 256, 0, s>>>(m, n, a, A, lda, sA, b, B, ldb, sB, C, ldc, sC);
  launch_check("zgeam");
}

__global__ void set_ident_kernel(int m, int n, cd diag, cd off, cd* C, long ldc, long sC) {
  const long bz = blockIdx.z;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= m) return;
  C[bz * sC + r + (long)c * ldc] = (r == c) ? diag : off;
}

void zset(hipStream_t s, int m, int n, cd diag, cd off, cd* C, long ldc, long sC, int batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid(cdiv_up(m, 256), n, batch);
  set_ident_kernel<<<grid, 256, 0, s>>>(m, n, diag, off, C, ldc, sC);
  launch_check("zset");
}

__global__ void add_diag_kernel(int n, cd v, const cd* vec, long svec, cd* C, long ldc, long sC) {
  const long bz = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cd add = vec ? cmul(v, vec[bz * svec + i]) : v;
  cd* p = C + bz * sC + i + (long)i * ldc;
  *p = cadd(*p, add);
}

void zadd_diag(hipStream_t s, int n, cd v, const cd* vec, long svec, cd* C, long ldc, long sC, int batch) {
  if (n <= 0 || batch <= 0) return;
  dim3 grid(cdiv_up(n, 256), batch);
  add_diag_kernel<<<grid, 256, 0, s>>>(n, v, vec, svec, C, ldc, sC);
  launch_check("zadd_diag");
}

// C = diag(left) * A * diag(right) (either vector may be null); in place allowed.
__global__ void dscale_kernel(int m, int n, const cd* left, long sl, const cd* right, long sr, const cd* A, long lda,
                              long sA, cd* C, long ldc, long sC, int inv_right) {
  const long bz = blockIdx.z;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= m) return;
  cd v = A[bz * sA + r + (long)c * lda];
  if (left) v = cmul(left[bz * sl + r], v);
  if (right) {
    cd f = right[bz * sr + c];
    v = inv_right ? cdiv(v, f) : cmul(v, f);
  }
  C[bz * sC + r + (long)c * ldc] = v;
}

void zdscale(hipStream_t s, int m, int n, const cd* left, long sl, const cd* right, long sr, int inv_right,
             const cd* A, long lda, long sA, cd* C, long ldc, long sC, int batch) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  dim3 grid(cdiv_up(m, 256), n, batch);
  dscale_kernel<<<grid, 256, 0, s>>>(m, n, left, sl, right, sr, A, lda, sA, C, ldc, sC, inv_right);
  launch_check("zdscale");
}

// Max |a_ij| (true modulus) per batch entry -> out[b] (double). One block per entry.
__global__ void maxabs_kernel(int m, int n, const cd* A, long lda, long sA, double* out) {
  const long bz = blockIdx.x;
  __shared__ double red[256];
  double mx = 0.0;
  const long total = (long)m * n;
  for (long e = threadIdx.x; e < total; e += blockDim.x) {
    int r = (int)(e % m), c = (int)(e / m);
    double v = cabsv(A[bz * sA + r + (long)c * lda]);
    mx = (v > mx || v != v) ? v : mx;
  }
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double o = red[threadIdx.x + w];
      double c = red[threadIdx.x];
      red[threadIdx.x] = (o > c || o != o) ? o : c;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[bz] = red[0];
}

void zmaxabs(hipStream_t s, int m, int n, const cd* A, long lda, long sA, double* out, int batch) {
  if (batch <= 0) return;
  maxabs_kernel<<<batch, 256, 0, s>>>(m, n, A, lda, sA, out);
  launch_check("zmaxabs");
}

}  // namespace vrc
