// Batched LU factorization with partial pivoting and triangular solves (complex128).
//
// getrf: right-looking, two levels: LU_NB2-column blocks factored by LU_NB-column
// panels (one workgroup per matrix and panel; pivot = first index of max |re|+|im|, as
// LAPACK izamax), each followed by its row interchanges, a column-parallel unit-lower
// solve and a ZGEMM update inside the block; per block, the interchanges on the other
// columns, the blocked triangular solve for U12 and one trailing ZGEMM with k = LU_NB2.
// getrs: rows gathered through the final permutation, then blocked forward/backward
// substitution (64-row diagonal blocks applied through their inverses, off-diagonal
// updates on ZGEMM).
#include "vrc_common.hpp"
#include "vrc_internal.hpp"

namespace vrc {

constexpr int LU_NB = 32;    // panel width
constexpr int LU_NB2 = 128;  // outer block width
constexpr int TRS_NB = 64;

__global__ void info_zero_kernel(int* info, int batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < batch) info[b] = 0;
}

__global__ __launch_bounds__(256) void getrf_panel_kernel(int n, int k, int kb, cd* A, long lda, long sA, int* ipiv,
                                                          long sP, int* info) {
  const long b = blockIdx.x;
  A += b * sA;
  ipiv += b * sP;
  const int t = threadIdx.x;
  __shared__ double rv[256];
  __shared__ int ri[256];
  __shared__ cd prow[LU_NB];
  __shared__ cd pivval;
  for (int j = 0; j < kb; ++j) {
    const int c = k + j;
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int r = c + t; r < n; r += 256) {
      double v = cabs1(A[r + (long)c * lda]);
      if (v > best) { best = v; bi = r; }
    }
    rv[t] = best;
    ri[t] = bi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) {
        double ov = rv[t + w];
        int oi = ri[t + w];
        if (ov > rv[t] || (ov == rv[t] && oi < ri[t])) { rv[t] = ov; ri[t] = oi; }
      }
      __syncthreads();
    }
    // no comparable candidate (NaN column): keep the diagonal as pivot row
    const int p = (ri[0] >= c && ri[0] < n) ? ri[0] : c;
    const double pv = rv[0];
    __syncthreads();
    if (t == 0) ipiv[c] = p;
    if (!(pv > 0.0)) {
      // exactly singular column (or NaN): record, skip elimination (LAPACK getf2 semantics)
      if (t == 0 && info && info[b] == 0) info[b] = c + 1;
      __syncthreads();
      continue;
    }
    if (p != c && t < kb) {
      cd* a1 = A + c + (long)(k + t) * lda;
      cd* a2 = A + p + (long)(k + t) * lda;
      cd tmp = *a1;
      *a1 = *a2;
      *a2 = tmp;
    }
    __syncthreads();
    if (t < kb) prow[t] = A[c + (long)(k + t) * lda];
    if (t == 0) pivval = A[c + (long)c * lda];
    __syncthreads();
    const cd rinv = crecip(pivval);
    for (int r = c + 1 + t; r < n; r += 256) {
      cd* ar = A + r;
      cd l = cmul(ar[(long)c * lda], rinv);
      ar[(long)c * lda] = l;
      for (int jj = j + 1; jj < kb; ++jj) {
        cd* e = ar + (long)(k + jj) * lda;
        *e = csub(*e, cmul(l, prow[jj]));
      }
    }
    __syncthreads();
  }
}

// Register-resident panel for n - k <= LUR_ROWS: thread t holds row k + t of the
// kb <= LUR_NB panel columns. Per column (one barrier): wave arg-max (first index of the
// largest |re|+|im|) posted with the candidate row, row exchange through LDS, elimination
// in registers. Afterwards the panel's interchanges are applied to the other columns
// [K, K+KB) of the block and, right of the panel, U12 = L11^{-1} A12 is solved (one
// thread per column).
constexpr int LUR_NB = 16;
constexpr int LUR_ROWS = 1024;

// ROWS threads (one row each; ROWS >= n - k): 256 or 512 for small n, so that four or two
// matrices share a CU (at most 128 VGPRs).
template <int ROWS>
__global__ __launch_bounds__(ROWS) __attribute__((amdgpu_waves_per_eu(4))) void getrf_panel_reg_kernel(
    int n, int K, int KB, int k, int kb, cd* A, long lda, long sA, int* ipiv, long sP, int* info) {
  const long b = blockIdx.x;
  A += b * sA;
  ipiv += b * sP;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int r = k + t;
  const bool own = r < n;
  // One barrier per column: every wave posts its best candidate and that candidate's
  // row, the owner of row c posts row c; buffers alternate between columns.
  constexpr int NW = ROWS / 64;
  __shared__ double wv[2][NW];
  __shared__ int wi[2][NW];
  __shared__ cd wrow[2][NW][LUR_NB];
  __shared__ cd crow[2][LUR_NB];
  __shared__ cd L11[LUR_NB][LUR_NB + 1];
  __shared__ int pvs[LUR_NB];
  cd a[LUR_NB];
#pragma unroll
  for (int j = 0; j < LUR_NB; ++j) a[j] = (own && j < kb) ? A[r + (long)(k + j) * lda] : mk(0.0, 0.0);
#pragma unroll
  for (int j = 0; j < LUR_NB; ++j) {
    if (j < kb) {
      const int par = j & 1;
      const int c = k + j;
      double v = -1.0;
      int vi = 0x7fffffff;
      if (own && r >= c) {
        const double m = cabs1(a[j]);
        if (m > v) { v = m; vi = r; }  // NaN never wins
      }
      for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o, 64);
        const int oi = __shfl_xor(vi, o, 64);
        if (ov > v || (ov == v && oi < vi)) { v = ov; vi = oi; }
      }
      if (lane == 0) { wv[par][wave] = v; wi[par][wave] = vi; }
      if (own && r == vi) {
#pragma unroll
        for (int jj = 0; jj < LUR_NB; ++jj) wrow[par][wave][jj] = a[jj];
      }
      if (own && r == c) {
#pragma unroll
        for (int jj = 0; jj < LUR_NB; ++jj) crow[par][jj] = a[jj];
      }
      __syncthreads();
      double best = wv[par][0];
      int bi = wi[par][0], bw = 0;
      for (int w = 1; w < NW; ++w) {
        const double ov = wv[par][w];
        const int oi = wi[par][w];
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; bw = w; }
      }
      const int p = (bi >= c && bi < n) ? bi : c;
      if (t == 0) { ipiv[c] = p; pvs[j] = p; }
      const bool nonsing = best > 0.0;
      if (!nonsing && t == 0 && info && info[b] == 0) info[b] = c + 1;  // getf2 semantics: no elimination
      if (nonsing) {
        const cd* prow = wrow[par][bw];  // row p
        if (p != c && r == c) {
#pragma unroll
          for (int jj = 0; jj < LUR_NB; ++jj) a[jj] = prow[jj];
        } else if (p != c && r == p) {
#pragma unroll
          for (int jj = 0; jj < LUR_NB; ++jj) a[jj] = crow[par][jj];
        }
        if (own && r > c) {
          const cd l = cmul(a[j], crecip(prow[j]));
          a[j] = l;
#pragma unroll
          for (int jj = 0; jj < LUR_NB; ++jj)
            if (jj > j && jj < kb) a[jj] = csub(a[jj], cmul(l, prow[jj]));
        }
      }
    }
  }
  if (own) {
#pragma unroll
    for (int j = 0; j < LUR_NB; ++j)
      if (j < kb) A[r + (long)(k + j) * lda] = a[j];
  }
  if (t < kb) {
#pragma unroll
    for (int j = 0; j < LUR_NB; ++j) L11[t][j] = a[j];
  }
  __syncthreads();
  // the block's other columns: interchanges, and U12 right of the panel
  const int nleft = k - K, nother = KB - kb;
  if (t < nother) {
    const int col = t < nleft ? K + t : k + kb + (t - nleft);
    cd* ac = A + (long)col * lda;
    for (int j = 0; j < kb; ++j) {
      const int p = pvs[j];
      if (p != k + j) {
        const cd tmp = ac[k + j];
        ac[k + j] = ac[p];
        ac[p] = tmp;
      }
    }
    if (col >= k + kb) {
      cd x[LUR_NB];
#pragma unroll
      for (int i = 0; i < LUR_NB; ++i) {
        if (i < kb) {
          cd xi = ac[k + i];
#pragma unroll
          for (int j = 0; j < LUR_NB; ++j)
            if (j < i) xi = csub(xi, cmul(L11[i][j], x[j]));
          x[i] = xi;
          ac[k + i] = xi;
        }
      }
    }
  }
}

// Row interchanges ipiv[k0 .. k0+nswap) (in order) applied to columns [c0, c1); one
// thread per column.
__global__ void laswp_kernel(int k0, int nswap, cd* A, long lda, long sA, const int* ipiv, long sP, int c0, int c1) {
  const long b = blockIdx.y;
  const int col = c0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= c1) return;
  cd* ac = A + b * sA + (long)col * lda;
  const int* pv = ipiv + b * sP;
  for (int j = 0; j < nswap; ++j) {
    const int p = pv[k0 + j];
    if (p != k0 + j) {
      const cd tmp = ac[k0 + j];
      ac[k0 + j] = ac[p];
      ac[p] = tmp;
    }
  }
}

static void laswp(hipStream_t s, int k0, int nswap, cd* A, long lda, long sA, const int* ipiv, long sP, int c0,
                  int c1, int batch) {
  if (c1 <= c0 || nswap <= 0) return;
  dim3 g(cdiv_up(c1 - c0, 256), batch);
  laswp_kernel<<<g, 256, 0, s>>>(k0, nswap, A, lda, sA, ipiv, sP, c0, c1);
  launch_check("laswp");
}

// Block interchanges as one row permutation. Plan (one workgroup per matrix): the rows
// touched by ipiv[k0 .. k0+nswap) in order, with the source row of each (rows below
// k0+nswap that end where they started are left out). Layout per matrix: count, then
// LASWP_MAXROW destination rows, then their source rows.
constexpr int LASWP_MAXROW = 256;  // 2 * the largest block

__global__ __launch_bounds__(256) void laswp_plan_kernel(int n, int k0, int nswap, const int* ipiv, long sP,
                                                         int* plan) {
  const long b = blockIdx.x;
  extern __shared__ int pm[];  // pm[r - k0]: original row now at row r
  __shared__ int cnt;
  const int t = threadIdx.x;
  const int m = n - k0;
  for (int i = t; i < m; i += 256) pm[i] = k0 + i;
  if (t == 0) cnt = nswap;
  __syncthreads();
  if (t == 0) {
    const int* pv = ipiv + b * sP;
    for (int j = 0; j < nswap; ++j) {
      const int p = pv[k0 + j] - k0;
      const int tmp = pm[j];
      pm[j] = pm[p];
      pm[p] = tmp;
    }
  }
  __syncthreads();
  int* pl = plan + b * (1 + 2 * LASWP_MAXROW);
  for (int i = t; i < nswap; i += 256) {
    pl[1 + i] = k0 + i;
    pl[1 + LASWP_MAXROW + i] = pm[i];
  }
  for (int i = nswap + t; i < m; i += 256) {
    if (pm[i] != k0 + i) {
      const int slot = atomicAdd(&cnt, 1);  // LDS atomic; at most nswap rows below
      pl[1 + slot] = k0 + i;
      pl[1 + LASWP_MAXROW + slot] = pm[i];
    }
  }
  __syncthreads();
  if (t == 0) pl[0] = cnt;
}

// Apply a plan to the columns [0, c_skip0) and [c_skip1, n): each workgroup takes
// LASWP_CT columns, every thread one touched row (all reads of a column before its writes).
constexpr int LASWP_CT = 16;

__global__ __launch_bounds__(256) void laswp_apply_kernel(int n, int cs0, int cs1, cd* A, long lda, long sA,
                                                          const int* plan) {
  const long b = blockIdx.y;
  const int* pl = plan + b * (1 + 2 * LASWP_MAXROW);
  const int cnt = pl[0];
  const int t = threadIdx.x;
  const int ncols = n - (cs1 - cs0);
  const int cb = blockIdx.x * LASWP_CT;
  A += b * sA;
  cd v[LASWP_CT];
  int dst = 0;
  if (t < cnt) {
    dst = pl[1 + t];
    const int src = pl[1 + LASWP_MAXROW + t];
#pragma unroll
    for (int q = 0; q < LASWP_CT; ++q) {
      const int ci = cb + q;
      if (ci < ncols) {
        const int col = ci < cs0 ? ci : ci + (cs1 - cs0);
        v[q] = A[src + (long)col * lda];
      }
    }
  }
  __syncthreads();
  if (t < cnt) {
#pragma unroll
    for (int q = 0; q < LASWP_CT; ++q) {
      const int ci = cb + q;
      if (ci < ncols) {
        const int col = ci < cs0 ? ci : ci + (cs1 - cs0);
        A[dst + (long)col * lda] = v[q];
      }
    }
  }
}

// U12 = L11^{-1} A12 for the kb x kb unit-lower panel block at (k, k), columns [c0, c1).
__global__ __launch_bounds__(256) void trsm_panel_kernel(int k, int kb, cd* A, long lda, long sA, int c0, int c1) {
  const long b = blockIdx.y;
  A += b * sA;
  __shared__ cd L[LU_NB][LU_NB];
  const int t = threadIdx.x;
  for (int e = t; e < kb * kb; e += 256) {
    int i = e % kb, j = e / kb;
    L[i][j] = A[(k + i) + (long)(k + j) * lda];
  }
  __syncthreads();
  const int col = c0 + blockIdx.x * 256 + t;
  if (col >= c1) return;
  cd* ac = A + (long)col * lda;
  cd x[LU_NB];
#pragma unroll
  for (int i = 0; i < LU_NB; ++i)
    if (i < kb) x[i] = ac[k + i];
#pragma unroll
  for (int i = 0; i < LU_NB; ++i) {
    if (i < kb) {
#pragma unroll
      for (int j = 0; j < LU_NB; ++j)
        if (j < i) x[i] = csub(x[i], cmul(L[i][j], x[j]));
    }
  }
#pragma unroll
  for (int i = 0; i < LU_NB; ++i)
    if (i < kb) ac[k + i] = x[i];
}

static void trsm_lu_factor(hipStream_t s, int upper, int n, int nrhs, const cd* LU, long lda, long sA, cd* Y,
                           long ldy, long sY, int batch, Workspace& ws);

// Two-level right-looking LU: 128-column blocks, each factored by 32-column panels with
// updates inside the block; the block's interchanges and U12 = L11^{-1} A12 (blocked
// triangular solve) are then applied to the rest, and the trailing matrix is updated
// by one ZGEMM with k = 128.
void zgetrf(hipStream_t s, int n, cd* A, long lda, long sA, int* ipiv, long sP, int* info, int batch, Workspace& ws) {
  if (n <= 0 || batch <= 0) return;
  if (info) {  // LAPACK semantics: info is an output (0, or the first zero pivot's column + 1)
    info_zero_kernel<<<cdiv_up(batch, 256), 256, 0, s>>>(info, batch);
    launch_check("info_zero");
  }
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0);
  const bool reg = n <= LUR_ROWS;
  DevBuf plan = ws.get(sizeof(int) * (size_t)(1 + 2 * LASWP_MAXROW) * batch);
  for (int K = 0; K < n; K += LU_NB2) {
    const int KB = n - K < LU_NB2 ? n - K : LU_NB2;
    if (reg) {
      for (int k = K; k < K + KB; k += LUR_NB) {
        const int kb = K + KB - k < LUR_NB ? K + KB - k : LUR_NB;
        if (n <= 256)
          getrf_panel_reg_kernel<256><<<batch, 256, 0, s>>>(n, K, KB, k, kb, A, lda, sA, ipiv, sP, info);
        else if (n <= 512)
          getrf_panel_reg_kernel<512><<<batch, 512, 0, s>>>(n, K, KB, k, kb, A, lda, sA, ipiv, sP, info);
        else
          getrf_panel_reg_kernel<LUR_ROWS><<<batch, LUR_ROWS, 0, s>>>(n, K, KB, k, kb, A, lda, sA, ipiv, sP, info);
        launch_check("getrf_panel_reg");
        const int c0 = k + kb, c1 = K + KB;
        if (c1 > c0)
          zgemm_strided(s, OP_N, OP_N, n - c0, c1 - c0, kb, mone, A + c0 + (long)k * lda, lda, sA,
                        A + k + (long)c0 * lda, lda, sA, one, A + c0 + (long)c0 * lda, lda, sA, batch);
      }
    }
    for (int k = K; k < K + KB && !reg; k += LU_NB) {
      const int kb = K + KB - k < LU_NB ? K + KB - k : LU_NB;
      getrf_panel_kernel<<<batch, 256, 0, s>>>(n, k, kb, A, lda, sA, ipiv, sP, info);
      launch_check("getrf_panel");
      // the panel's interchanges on the block's other columns
      laswp(s, k, kb, A, lda, sA, ipiv, sP, K, k, batch);
      laswp(s, k, kb, A, lda, sA, ipiv, sP, k + kb, K + KB, batch);
      const int c0 = k + kb, c1 = K + KB;
      if (c1 > c0) {
        dim3 g(cdiv_up(c1 - c0, 256), batch);
        trsm_panel_kernel<<<g, 256, 0, s>>>(k, kb, A, lda, sA, c0, c1);
        launch_check("trsm_panel");
        if (n - c0 > 0)
          zgemm_strided(s, OP_N, OP_N, n - c0, c1 - c0, kb, mone, A + c0 + (long)k * lda, lda, sA,
                        A + k + (long)c0 * lda, lda, sA, one, A + c0 + (long)c0 * lda, lda, sA, batch);
      }
    }
    // the block's interchanges on the columns left and right of the block
    if (n - K > 16000) {  // the plan's row map would not fit in LDS
      laswp(s, K, KB, A, lda, sA, ipiv, sP, 0, K, batch);
      laswp(s, K, KB, A, lda, sA, ipiv, sP, K + KB, n, batch);
    } else if (n - KB > 0) {
      laswp_plan_kernel<<<batch, 256, sizeof(int) * (size_t)(n - K), s>>>(n, K, KB, ipiv, sP, plan.as<int>());
      launch_check("laswp_plan");
      dim3 g(cdiv_up(n - KB, LASWP_CT), batch);
      laswp_apply_kernel<<<g, 256, 0, s>>>(n, K, K + KB, A, lda, sA, plan.as<int>());
      launch_check("laswp_apply");
    }
    const int rest = n - K - KB;
    if (rest > 0) {
      trsm_lu_factor(s, 0, KB, rest, A + K + (long)K * lda, lda, sA, A + K + (long)(K + KB) * lda, lda, sA, batch,
                     ws);
      zgemm_strided(s, OP_N, OP_N, rest, rest, KB, mone, A + (K + KB) + (long)K * lda, lda, sA,
                    A + K + (long)(K + KB) * lda, lda, sA, one, A + (K + KB) + (long)(K + KB) * lda, lda, sA, batch);
    }
  }
  health_note_info(s, info, batch, H_LU);
}

// perm[i] = original row that ends up in row i after applying ipiv[0..n-1] in order.
__global__ void ipiv_to_perm_kernel(int n, const int* ipiv, long sP, int* perm, long sPerm) {
  const long b = blockIdx.x;
  extern __shared__ int pm[];
  for (int i = threadIdx.x; i < n; i += blockDim.x) pm[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int* ip = ipiv + b * sP;
    for (int i = 0; i < n; ++i) {
      int p = ip[i];
      if (p != i) {
        int tmp = pm[i];
        pm[i] = pm[p];
        pm[p] = tmp;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) perm[b * sPerm + i] = pm[i];
}

__global__ void gather_rows_kernel(int n, int nrhs, const int* perm, long sPerm, const cd* B, long ldb, long sB,
                                   cd* Y, long ldy, long sY) {
  const long b = blockIdx.z;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= n) return;
  Y[b * sY + r + (long)c * ldy] = B[b * sB + perm[b * sPerm + r] + (long)c * ldb];
}

// Inverse of every TRS_NB x TRS_NB diagonal block of a triangular factor (UPPER=0: unit
// lower, UPPER=1: non-unit upper) into Dinv[b][blk] (TRS_NB x TRS_NB, zero padded past
// the block size). One workgroup per block: thread (column j, quarter q) substitutes
// the identity column j, rows split over the four quarters.
template <int UPPER>
__global__ __launch_bounds__(256) void trtri_blocks_kernel(int n, const cd* T, long ldt, long sT, cd* Dinv,
                                                           long sD) {
  const long b = blockIdx.y;
  const int blk = blockIdx.x;
  const int k0 = blk * TRS_NB;
  const int bs = n - k0 < TRS_NB ? n - k0 : TRS_NB;
  T += b * sT;
  cd* D = Dinv + b * sD + (long)blk * TRS_NB * TRS_NB;
  __shared__ cd Ls[TRS_NB][TRS_NB + 1];
  __shared__ cd Xs[TRS_NB][TRS_NB + 1];
  const int t = threadIdx.x;
  for (int e = t; e < TRS_NB * TRS_NB; e += 256) {
    const int i = e % TRS_NB, j = e / TRS_NB;
    Ls[i][j] = (i < bs && j < bs) ? T[(k0 + i) + (long)(k0 + j) * ldt] : mk(0.0, 0.0);
    Xs[i][j] = (i == j && i < bs) ? mk(1.0, 0.0) : mk(0.0, 0.0);
  }
  __syncthreads();
  const int col = t & 63, grp = t >> 6;
  if (!UPPER) {
    for (int i = 0; i < bs; ++i) {
      const cd xi = Xs[i][col];
      for (int j = i + 1 + grp; j < bs; j += 4) Xs[j][col] = csub(Xs[j][col], cmul(Ls[j][i], xi));
      __syncthreads();
    }
  } else {
    for (int i = bs - 1; i >= 0; --i) {
      if (grp == 0 && col < bs) Xs[i][col] = cdiv(Xs[i][col], Ls[i][i]);
      __syncthreads();
      const cd xi = Xs[i][col];
      for (int j = grp; j < i; j += 4) Xs[j][col] = csub(Xs[j][col], cmul(Ls[j][i], xi));
      __syncthreads();
    }
  }
  for (int e = t; e < TRS_NB * TRS_NB; e += 256) {
    const int i = e % TRS_NB, j = e / TRS_NB;
    D[i + (long)j * TRS_NB] = Xs[i][j];
  }
}

// In-place triangular solve on Y (n x nrhs): UPPER=0 -> L (unit) ; UPPER=1 -> U (non-unit).
// Blocked by TRS_NB rows; each diagonal block is applied as a product with its inverse
// (computed once per call), the off-diagonal blocks as ZGEMM updates.
// Xo (optional): the solution goes to Xo block by block (each diagonal product written
// there directly) and Y is only the work space of the updates; otherwise in place in Y.
static void trsm_lu_factor(hipStream_t s, int upper, int n, int nrhs, const cd* LU, long lda, long sA, cd* Y,
                           long ldy, long sY, int batch, Workspace& ws, cd* Xo, long ldxo, long sXo) {
  const int nblk = cdiv_up(n, TRS_NB);
  const long sD = (long)nblk * TRS_NB * TRS_NB;
  DevBuf Db = ws.get(sizeof(cd) * (size_t)sD * batch);
  DevBuf Tb;
  cd* Dinv = Db.as<cd>();
  cd* Tmp = nullptr;
  const long sTmp = (long)TRS_NB * nrhs;
  if (!Xo) {
    Tb = ws.get(sizeof(cd) * (size_t)TRS_NB * nrhs * batch);
    Tmp = Tb.as<cd>();
  }
  dim3 g(nblk, batch);
  if (!upper)
    trtri_blocks_kernel<0><<<g, 256, 0, s>>>(n, LU, lda, sA, Dinv, sD);
  else
    trtri_blocks_kernel<1><<<g, 256, 0, s>>>(n, LU, lda, sA, Dinv, sD);
  launch_check("trtri_blocks");
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  // solved rows k0.. live in X (= Xo or Y)
  cd* X = Xo ? Xo : Y;
  const long ldx = Xo ? ldxo : ldy, sX = Xo ? sXo : sY;
  auto diag = [&](int kb, int k0, int bs) {
    const cd* Dk = Dinv + (long)kb * TRS_NB * TRS_NB;
    if (Xo) {
      zgemm_strided(s, OP_N, OP_N, bs, nrhs, bs, one, Dk, TRS_NB, sD, Y + k0, ldy, sY, zero, Xo + k0, ldxo, sXo,
                    batch);
    } else {
      // Y[k0:k0+bs] <- inv(T_kk) Y[k0:k0+bs]  (through Tmp)
      zgemm_strided(s, OP_N, OP_N, bs, nrhs, bs, one, Dk, TRS_NB, sD, Y + k0, ldy, sY, zero, Tmp, TRS_NB, sTmp,
                    batch);
      zgeam(s, bs, nrhs, one, Tmp, TRS_NB, sTmp, zero, nullptr, 0, 0, Y + k0, ldy, sY, batch);
    }
  };
  // Diagonal blocks are taken in pairs: the coupling between the two is one 64-deep
  // update, the rest of Y is updated once per pair (k up to 2*TRS_NB).
  if (!upper) {
    for (int kb = 0; kb < nblk; kb += 2) {
      const int k0 = kb * TRS_NB;
      const int bs0 = n - k0 < TRS_NB ? n - k0 : TRS_NB;
      diag(kb, k0, bs0);
      int w = bs0;
      if (kb + 1 < nblk) {
        const int k1 = k0 + TRS_NB;
        const int bs1 = n - k1 < TRS_NB ? n - k1 : TRS_NB;
        zgemm_strided(s, OP_N, OP_N, bs1, nrhs, bs0, mone, LU + k1 + (long)k0 * lda, lda, sA, X + k0, ldx, sX, one,
                      Y + k1, ldy, sY, batch);
        diag(kb + 1, k1, bs1);
        w += bs1;
      }
      const int rest = n - k0 - w;
      if (rest > 0)
        zgemm_strided(s, OP_N, OP_N, rest, nrhs, w, mone, LU + (k0 + w) + (long)k0 * lda, lda, sA, X + k0, ldx,
                      sX, one, Y + (k0 + w), ldy, sY, batch);
    }
  } else {
    for (int kb = nblk - 1; kb >= 0; kb -= 2) {
      const int k1 = kb * TRS_NB;
      const int bs1 = n - k1 < TRS_NB ? n - k1 : TRS_NB;
      diag(kb, k1, bs1);
      int k0 = k1, w = bs1;
      if (kb >= 1) {
        k0 = k1 - TRS_NB;
        zgemm_strided(s, OP_N, OP_N, TRS_NB, nrhs, bs1, mone, LU + k0 + (long)k1 * lda, lda, sA, X + k1, ldx, sX,
                      one, Y + k0, ldy, sY, batch);
        diag(kb - 1, k0, TRS_NB);
        w += TRS_NB;
      }
      if (k0 > 0)
        zgemm_strided(s, OP_N, OP_N, k0, nrhs, w, mone, LU + (long)k0 * lda, lda, sA, X + k0, ldx, sX, one, Y, ldy,
                      sY, batch);
    }
  }
}

static void trsm_lu_factor(hipStream_t s, int upper, int n, int nrhs, const cd* LU, long lda, long sA, cd* Y,
                           long ldy, long sY, int batch, Workspace& ws) {
  trsm_lu_factor(s, upper, n, nrhs, LU, lda, sA, Y, ldy, sY, batch, ws, nullptr, 0, 0);
}

void ztrsm_upper(hipStream_t s, int n, int nrhs, const cd* U, long ldu, long sU, cd* Y, long ldy, long sY,
                 int batch, Workspace& ws) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return;
  trsm_lu_factor(s, 1, n, nrhs, U, ldu, sU, Y, ldy, sY, batch, ws);
}

void zgetrs(hipStream_t s, int n, int nrhs, const cd* LU, long lda, long sA, const int* ipiv, long sP, const cd* B,
            long ldb, long sB, cd* X, long ldx, long sX, int batch, Workspace& ws) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return;
  DevBuf permb = ws.get(sizeof(int) * (size_t)n * batch);
  int* perm = (int*)permb.ptr;
  ipiv_to_perm_kernel<<<batch, 256, sizeof(int) * n, s>>>(n, ipiv, sP, perm, n);
  launch_check("ipiv_to_perm");
  // X not aliasing B: gather into X, forward substitution X -> T, back substitution
  // T -> X. Aliased: gather into T, forward substitution in place, back substitution
  // T -> X.
  DevBuf Tb = ws.get(sizeof(cd) * (size_t)n * nrhs * batch);
  cd* T = Tb.as<cd>();
  const long sT = (long)n * nrhs;
  dim3 g(cdiv_up(n, 256), nrhs, batch);
  if (X != B) {
    gather_rows_kernel<<<g, 256, 0, s>>>(n, nrhs, perm, n, B, ldb, sB, X, ldx, sX);
    launch_check("gather_rows");
    trsm_lu_factor(s, 0, n, nrhs, LU, lda, sA, X, ldx, sX, batch, ws, T, n, sT);
  } else {
    gather_rows_kernel<<<g, 256, 0, s>>>(n, nrhs, perm, n, B, ldb, sB, T, n, sT);
    launch_check("gather_rows");
    trsm_lu_factor(s, 0, n, nrhs, LU, lda, sA, T, n, sT, batch, ws);
  }
  trsm_lu_factor(s, 1, n, nrhs, LU, lda, sA, T, n, sT, batch, ws, X, ldx, sX);
}

void zgetrs_pair(hipStream_t s, int n, int nrhs, const cd* LU, long lda, long sA, const int* ipiv, long sP,
                 const cd* B1, long ldb1, long sB1, const cd* B2, long ldb2, long sB2, cd* X, long ldx, long sX,
                 int batch, Workspace& ws) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return;
  DevBuf permb = ws.get(sizeof(int) * (size_t)n * batch);
  int* perm = (int*)permb.ptr;
  ipiv_to_perm_kernel<<<batch, 256, sizeof(int) * n, s>>>(n, ipiv, sP, perm, n);
  launch_check("ipiv_to_perm");
  dim3 g(cdiv_up(n, 256), nrhs, batch);
  gather_rows_kernel<<<g, 256, 0, s>>>(n, nrhs, perm, n, B1, ldb1, sB1, X, ldx, sX);
  launch_check("gather_rows");
  gather_rows_kernel<<<g, 256, 0, s>>>(n, nrhs, perm, n, B2, ldb2, sB2, X + (long)nrhs * ldx, ldx, sX);
  launch_check("gather_rows");
  // forward substitution X -> T, back substitution T -> X (no block copies)
  DevBuf Tb = ws.get(sizeof(cd) * (size_t)n * 2 * nrhs * batch);
  const long sT = (long)n * 2 * nrhs;
  trsm_lu_factor(s, 0, n, 2 * nrhs, LU, lda, sA, X, ldx, sX, batch, ws, Tb.as<cd>(), n, sT);
  trsm_lu_factor(s, 1, n, 2 * nrhs, LU, lda, sA, Tb.as<cd>(), n, sT, batch, ws, X, ldx, sX);
}

static size_t ws_round(size_t b) { return b == 0 ? 256 : (b + 255) & ~(size_t)255; }

// Largest workspace zgetrf holds at once: the interchange plan, plus the diagonal-block
// inverses and staging of the first block's U12 solve (the widest).
size_t zgetrf_scratch_bytes(int n, int batch) {
  size_t b = ws_round(sizeof(int) * (size_t)(1 + 2 * LASWP_MAXROW) * batch);
  if (n > LU_NB2)
    b += ws_round(sizeof(cd) * (size_t)2 * TRS_NB * TRS_NB * batch) +
         ws_round(sizeof(cd) * (size_t)TRS_NB * (n - LU_NB2) * batch);
  return b;
}

// Largest workspace zgetrs_pair holds at once: the permutation, the forward-substitution
// result and the diagonal-block inverses.
size_t zgetrs_pair_scratch_bytes(int n, int nrhs, int batch) {
  const size_t nblk = (size_t)cdiv_up(n, TRS_NB);
  return ws_round(sizeof(int) * (size_t)n * batch) + ws_round(sizeof(cd) * (size_t)n * 2 * nrhs * batch) +
         ws_round(sizeof(cd) * nblk * TRS_NB * TRS_NB * batch);
}

void zgetri(hipStream_t s, int n, const cd* LU, long lda, long sA, const int* ipiv, long sP, cd* Ainv, long ldi,
            long sI, int batch, Workspace& ws) {
  DevBuf idb = ws.get(sizeof(cd) * (size_t)n * n * batch);
  cd* I = (cd*)idb.ptr;
  zset(s, n, n, mk(1.0, 0.0), mk(0.0, 0.0), I, n, (long)n * n, batch);
  zgetrs(s, n, n, LU, lda, sA, ipiv, sP, I, n, (long)n * n, Ainv, ldi, sI, batch, ws);
}

}  // namespace vrc
