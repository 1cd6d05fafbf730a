// Reverse-mode (adjoint) derivatives of the S-matrix chain: Fourier assembly, P/Q,
// eigen-decomposition, slab S-matrix, star product and end projection, each as the
// vector-Jacobian product of the forward procedure with the same name. Composed by
// the batched stack tape (stack_tape_vjp) into gradients of any real objective of
// the stack S-matrices with respect to the per-cell permittivity of every slab.
//
// Convention: for a real objective L and a complex array Z the cotangent Zb is the
// array with dL = Re sum(conj(Zb) . dZ) (PyTorch's convention for complex tensors;
// for L = |z|^2, zb = 2z). For products Y = A B: Ab = Yb B^H, Bb = A^H Yb; for
// X = A^{-1} B: Bb = A^{-H} Xb, Ab = -Bb X^H.
//
// There is no reference counterpart: the reference computes no derivatives of the
// S-matrix path (its Parameterization / dSolve* routines, src/rcwa/Parameterization.h,
// src/rcwa/FourierSolver1D.cpp:93-143, serve the legacy solver only).
#include <algorithm>
#include <cstring>

#include "vrc_adjoint.hpp"

namespace vrc {

static const cd ONE = {1.0, 0.0}, MONE = {-1.0, 0.0}, ZERO = {0.0, 0.0};

// ===================================================================== kernels
// Cotangents of the Fourier contraction tables from the cotangents of fe11/fe22/fe33
// (n x n column-major, row index a*ny+b): sums along the (block-)diagonals that the
// forward expansion copied each table entry to.
//   Gb33[p + q M1]           = sum_{a-c = p-hx, b-d = q-hy} fe33b[(a,b),(c,d)]
//   Gb11[(a + c nx) + q nx^2] = sum_{b-d = q-hy} fe11b[(a,b),(c,d)]
//   Gb22[(b + d ny) + m ny^2] = sum_{a-c = m-hx} fe22b[(a,b),(c,d)]
__global__ void fe_bwd_diag_kernel(int nx, int ny, const cd* fe11b, const cd* fe22b, const cd* fe33b, long sfe,
                                   cd* Gb33, cd* Gb11, cd* Gb22) {
  const int l = blockIdx.y;
  const int n = nx * ny, M1 = 2 * nx - 1, M2 = 2 * ny - 1, hx = nx - 1, hy = ny - 1;
  const long nF2 = (long)M1 * M2, nG11 = (long)nx * nx * M2, nG22 = (long)ny * ny * M1;
  const cd* f33 = fe33b + l * sfe;
  const cd* f11 = fe11b + l * sfe;
  const cd* f22 = fe22b + l * sfe;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nF2 + nG11 + nG22;
       e += (long)gridDim.x * blockDim.x) {
    cd acc = mk(0.0, 0.0);
    if (e < nF2) {
      const int dp = (int)(e % M1) - hx, dq = (int)(e / M1) - hy;
      const int a0 = max(0, dp), a1 = min(nx, nx + dp), b0 = max(0, dq), b1 = min(ny, ny + dq);
      for (int a = a0; a < a1; ++a)
        for (int b = b0; b < b1; ++b) {
          const int r = a * ny + b, c = (a - dp) * ny + (b - dq);
          acc = cadd(acc, f33[r + (long)c * n]);
        }
      Gb33[l * nF2 + e] = acc;
    } else if (e < nF2 + nG11) {
      const long ee = e - nF2;
      const int ac = (int)(ee % (nx * nx)), dq = (int)(ee / (nx * nx)) - hy;
      const int a = ac % nx, cc = ac / nx;
      const int b0 = max(0, dq), b1 = min(ny, ny + dq);
      for (int b = b0; b < b1; ++b) {
        const int r = a * ny + b, c = cc * ny + (b - dq);
        acc = cadd(acc, f11[r + (long)c * n]);
      }
      Gb11[l * nG11 + ee] = acc;
    } else {
      const long ee = e - nF2 - nG11;
      const int bd = (int)(ee % (ny * ny)), dm = (int)(ee / (ny * ny)) - hx;
      const int b = bd % ny, d = bd / ny;
      const int a0 = max(0, dm), a1 = min(nx, nx + dm);
      for (int a = a0; a < a1; ++a) {
        const int r = a * ny + b, c = (a - dm) * ny + d;
        acc = cadd(acc, f22[r + (long)c * n]);
      }
      Gb22[l * nG22 + ee] = acc;
    }
  }
}

// Per strip i (row of cells) and per column j of layer l:
//   tbX[i][(a,c)] = sum_q conj(FY1_i[q]) Gb11[(a,c)][q]      (cotangent of inv(Toep(FXi_i)))
//   U[i][p]       = sum_q conj(FY1_i[q]) Gb33[p][q]          (Laurent part per strip)
//   tbY[j][(b,d)] = sum_m conj(FX1_j[m]) Gb22[(b,d)][m]      (cotangent of inv(Toep(FYi_j)))
__global__ void fe_bwd_strip_kernel(FourierDev g, const cd* Gb33, const cd* Gb11, const cd* Gb22, cd* tbX, cd* U,
                                    cd* tbY) {
  const int l = blockIdx.y;
  const int nx = g.nx, ny = g.ny, M1 = 2 * nx - 1, M2 = 2 * ny - 1;
  const int ncx = g.ncx[l], ncy = g.ncy[l];
  const long MX = g.MX, MY = g.MY;
  const long nA = MY * nx * nx, nU = MY * M1, nB = MX * ny * ny;
  const cd* FY1 = g.FY1 + (long)l * MY * M2;
  const cd* FX1 = g.FX1 + (long)l * MX * M1;
  const cd* G33 = Gb33 + (long)l * M1 * M2;
  const cd* G11 = Gb11 + (long)l * nx * nx * M2;
  const cd* G22 = Gb22 + (long)l * ny * ny * M1;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nA + nU + nB; e += (long)gridDim.x * blockDim.x) {
    cd acc = mk(0.0, 0.0);
    if (e < nA) {
      const int i = (int)(e / (nx * nx)), ac = (int)(e % (nx * nx));
      if (i < ncy)
        for (int q = 0; q < M2; ++q) acc = cadd(acc, cmul(cconj(FY1[(long)i * M2 + q]), G11[ac + (long)q * nx * nx]));
      tbX[(long)l * nA + e] = acc;
    } else if (e < nA + nU) {
      const long ee = e - nA;
      const int i = (int)(ee / M1), p = (int)(ee % M1);
      if (i < ncy)
        for (int q = 0; q < M2; ++q) acc = cadd(acc, cmul(cconj(FY1[(long)i * M2 + q]), G33[p + (long)q * M1]));
      U[(long)l * nU + ee] = acc;
    } else {
      const long ee = e - nA - nU;
      const int j = (int)(ee / (ny * ny)), bd = (int)(ee % (ny * ny));
      if (j < ncx)
        for (int m = 0; m < M1; ++m) acc = cadd(acc, cmul(cconj(FX1[(long)j * M1 + m]), G22[bd + (long)m * ny * ny]));
      tbY[(long)l * nB + ee] = acc;
    }
  }
}

// Cotangent of the generating vector v of T(r,c) = v[k-1+r-c] from the cotangent
// Tb of T: vb[p] = sum_{r-c = p-(k-1)} Tb[r + c k]; slots beyond count[layer] -> 0.
__global__ void toep_diag_sum_kernel(int k, const cd* Tb, long sT, cd* vb, long sv, const int* count, int per_layer,
                                     long total) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= total) return;
  const long slot = e / (2 * k - 1);
  const int p = (int)(e % (2 * k - 1));
  const int layer = (int)(slot / per_layer), s = (int)(slot % per_layer);
  cd acc = mk(0.0, 0.0);
  if (s < count[layer]) {
    const int d = p - (k - 1);
    const cd* T = Tb + slot * sT;
    for (int r = max(0, d); r < min(k, k + d); ++r) acc = cadd(acc, T[r + (long)(r - d) * k]);
  }
  vb[slot * sv + p] = acc;
}

// Per cell (strip i, column j) of layer l:
//   epsb = sum_p conj(FX1_j[p]) (U_i[p] + w vbx_i[p]) + w sum_q conj(FY1_i[q]) vby_j[q]
// with w = conj(-1/eps^2) (the inverse-rule tables depend on 1/eps). Written in the
// input layout: offs[l] + i + j*ncy.
__global__ void fe_bwd_cell_kernel(FourierDev g, const cd* U, const cd* vbx, const cd* vby, cd* epsb,
                                   const long* offs) {
  const int l = blockIdx.y;
  const int nx = g.nx, ny = g.ny, M1 = 2 * nx - 1, M2 = 2 * ny - 1;
  const int ncx = g.ncx[l], ncy = g.ncy[l];
  const long MX = g.MX, MY = g.MY;
  const int cells = ncx * ncy;
  const cd* FY1 = g.FY1 + (long)l * MY * M2;
  const cd* FX1 = g.FX1 + (long)l * MX * M1;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < cells; e += gridDim.x * blockDim.x) {
    const int i = e % ncy, j = e / ncy;
    const cd v = g.val[((long)l * MY + i) * MX + j];
    const cd iv = crecip(v);
    const cd w = cconj(cneg(cmul(iv, iv)));
    const cd* Ui = U + ((long)l * MY + i) * M1;
    const cd* vx = vbx + ((long)l * MY + i) * M1;
    const cd* vy = vby + ((long)l * MX + j) * M2;
    const cd* fx = FX1 + (long)j * M1;
    const cd* fy = FY1 + (long)i * M2;
    cd s1 = mk(0.0, 0.0), s2 = mk(0.0, 0.0);
    for (int p = 0; p < M1; ++p) s1 = cadd(s1, cmul(cconj(fx[p]), cadd(Ui[p], cmul(w, vx[p]))));
    for (int q = 0; q < M2; ++q) s2 = cadd(s2, cmul(cconj(fy[q]), vy[q]));
    epsb[offs[l] + e] = cadd(s1, cmul(w, s2));
  }
}

// Cotangents of inv33 = fe33^{-1}, fe11 and fe22 from those of the periodic P and Q
// (transposes the pq_periodic formulas):
//   Eb   = conj(a_r b_c) Pb11 - conj(a_r a_c) Pb12 + conj(b_r b_c) Pb21 - conj(b_r a_c) Pb22
//   fe22b = -k0^2 Qb12 ; fe11b = k0^2 Qb21
__global__ void pq_bwd_kernel(int n, const cd* Pb, const cd* Qb, long sPQ, const cd* alpha, const cd* beta, long sab,
                              const double* k0sq, cd* Eb, cd* fe11b, cd* fe22b, long sfe) {
  const long l = blockIdx.z;
  const int N = 2 * n;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= n) return;
  const cd* al = alpha + l * sab;
  const cd* be = beta + l * sab;
  const double k2 = k0sq[l];
  const cd* P = Pb + l * sPQ;
  const cd* Q = Qb + l * sPQ;
  const long o11 = r + (long)c * N, o12 = r + (long)(c + n) * N, o21 = (r + n) + (long)c * N,
             o22 = (r + n) + (long)(c + n) * N;
  const cd ar = al[r], br = be[r], ac = al[c], bc = be[c];
  cd e = cmul(cconj(cmul(ar, bc)), P[o11]);
  e = csub(e, cmul(cconj(cmul(ar, ac)), P[o12]));
  e = cadd(e, cmul(cconj(cmul(br, bc)), P[o21]));
  e = csub(e, cmul(cconj(cmul(br, ac)), P[o22]));
  const long o = l * sfe + r + (long)c * n;
  Eb[o] = e;
  fe22b[o] = cscale(Q[o12], -k2);
  fe11b[o] = cscale(Q[o21], k2);
}

// out[b][i] = f(x[b][i]) for the small per-mode vectors of the eigen and slab adjoints.
enum VecOp : int {
  VEC_CONJ = 0,        // conj(x)
  VEC_CONJ_INV = 1,    // conj(1/x)
};
__global__ void vec_op_kernel(int n, const cd* x, long sx, cd* out, long so, int op) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cd v = x[b * sx + i];
  out[b * so + i] = (op == VEC_CONJ) ? cconj(v) : cconj(crecip(v));
}

// out_i = sum_k conj(A_ik) B_ik (+ sum_k conj(C_ik) D_ik when C is given), per batch entry.
__global__ void rowdot_kernel(int N, const cd* A, const cd* B, const cd* C, const cd* D, long sM, cd* out, long so) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const long base = b * sM + i;
  cd acc = mk(0.0, 0.0);
  for (int k = 0; k < N; ++k) acc = cadd(acc, cmul(cconj(A[base + (long)k * N]), B[base + (long)k * N]));
  if (C)
    for (int k = 0; k < N; ++k) acc = cadd(acc, cmul(cconj(C[base + (long)k * N]), D[base + (long)k * N]));
  out[b * so + i] = acc;
}

// out_j = sum_k conj(A_kj) B_kj per batch entry: one 256-thread block per column.
__global__ __launch_bounds__(256) void coldot_kernel(int N, const cd* A, const cd* B, long sM, cd* out, long so) {
  const long b = blockIdx.y;
  const int j = blockIdx.x;
  __shared__ cd red[256];
  const cd* a = A + b * sM + (long)j * N;
  const cd* bb = B + b * sM + (long)j * N;
  cd acc = mk(0.0, 0.0);
  for (int k = threadIdx.x; k < N; k += 256) acc = cadd(acc, cmul(cconj(a[k]), bb[k]));
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[b * so + j] = red[0];
}

// Slab phase adjoint: gb_i = conj(i s x_i) xb_i, x_i = exp(i gamma_i s).
__global__ void phase_bwd_kernel(int n, const cd* x, const cd* xb, const double* s, cd* gb, long sv) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ss = s[b];
  const cd f = cconj(cmul(mk(0.0, ss), x[b * sv + i]));
  gb[b * sv + i] = cmul(f, xb[b * sv + i]);
}

// X_i = exp(i gamma_i s) (same arithmetic as the forward slab kernel).
__global__ void phase_fwd_kernel(int n, const cd* gam, const double* s, cd* X, long sv) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cd g = gam[b * sv + i];
  const double ss = s[b];
  X[b * sv + i] = cexp(mk(-g.y * ss, g.x * ss));
}

// Eigen adjoint pieces. gb_i -= conj(1/gamma_i) vd_i (V = Q W / gamma path), then
// lb_i = conj(1/(2 gamma_i)) gb_i (gamma = sqrt(lambda)), lam_i = gamma_i^2.
__global__ void gamma_bwd_kernel(int n, const cd* gam, const cd* vd, cd* gb, cd* lb, cd* lam, long sv) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long o = b * sv + i;
  const cd g = gam[o];
  const cd ig = crecip(g);
  const cd t = csub(gb[o], cmul(cconj(ig), vd[o]));
  gb[o] = t;
  lb[o] = cmul(cconj(cscale(ig, 0.5)), t);
  lam[o] = cmul(g, g);
}

// K_ij = G_ij / conj(lam_j - lam_i) (i != j), K_ii = lb_i; pairs with
// |lam_j - lam_i| <= 1e-13 max(|lam_i|, |lam_j|) (numerically degenerate) get 0: the
// objective is invariant under rotations inside a degenerate eigenspace, so the exact
// contribution of such a pair vanishes.
__global__ void eig_k_kernel(int N, const cd* G, const cd* lam, const cd* lb, long sv, cd* K, long sM) {
  const long b = blockIdx.z;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= N) return;
  const long o = b * sM + i + (long)j * N;
  if (i == j) {
    K[o] = lb[b * sv + i];
    return;
  }
  const cd li = lam[b * sv + i], lj = lam[b * sv + j];
  const cd d = csub(lj, li);
  const double scale = fmax(cabsv(li), cabsv(lj));
  K[o] = (cabsv(d) <= 1e-13 * scale) ? mk(0.0, 0.0) : cdiv(G[o], cconj(d));
}

// ===================================================================== helpers
namespace {
struct Ops {
  hipStream_t s;
  int N;
  int batch;
  long NN;
  Ops(hipStream_t s_, int N_, int b_) : s(s_), N(N_), batch(b_), NN((long)N_ * N_) {}
  // C = alpha op(A) op(B) + beta C
  void mm(int oa, int ob, cd al, const cd* A, long sA, const cd* B, long sB, cd be, cd* C, long sC) const {
    zgemm_strided(s, oa, ob, N, N, N, al, A, N, sA, B, N, sB, be, C, N, sC, batch);
  }
  void add(cd a, const cd* A, long sA, cd b, const cd* B, long sB, cd* C, long sC) const {
    zgeam(s, N, N, a, A, N, sA, b, B, N, sB, C, N, sC, batch);
  }
  void copy(const cd* A, long sA, cd* C, long sC, cd a = ONE) const {
    zgeam(s, N, N, a, A, N, sA, ZERO, nullptr, 0, 0, C, N, sC, batch);
  }
  void eye(cd* C, long sC) const { zset(s, N, N, ONE, ZERO, C, N, sC, batch); }
  void vec(const cd* x, long sx, cd* out, long so, int op) const {
    dim3 g(cdiv_up(N, 256), batch);
    vec_op_kernel<<<g, 256, 0, s>>>(N, x, sx, out, so, op);
    launch_check("vec_op");
  }
};

// Ainv = A^{-1} for a strided batch (A untouched).
void invert(vrc_ctx& c, const Ops& o, const cd* A, long sA, cd* Ainv) {
  DevBuf LU = c.ws.get(sizeof(cd) * (size_t)o.NN * o.batch);
  DevBuf piv = c.ws.get(sizeof(int) * (size_t)o.N * o.batch), inf = c.ws.get(sizeof(int) * (size_t)o.batch);
  fill_int(o.s, inf.as<int>(), 0, o.batch);
  o.copy(A, sA, LU.as<cd>(), o.NN);
  zgetrf(o.s, o.N, LU.as<cd>(), o.N, o.NN, piv.as<int>(), o.N, inf.as<int>(), o.batch, c.ws);
  zgetri(o.s, o.N, LU.as<cd>(), o.N, o.NN, piv.as<int>(), o.N, Ainv, o.N, o.NN, o.batch, c.ws);
}

struct Buf {
  DevBuf b;
  cd* p = nullptr;
  Buf(vrc_ctx& c, size_t count) : b(c.ws.get(sizeof(cd) * count)), p(b.as<cd>()) {}
};
}  // namespace

// ===================================================================== (a)+(b) P/Q -> eps
void assemble_fe_vjp(vrc_ctx& c, int nx, int ny, const std::vector<const HostLayer*>& layers, const cd* fe11b,
                     const cd* fe22b, const cd* fe33b, long sfe, cd* epsb, const std::vector<long>& offs) {
  hipStream_t s = c.stream;
  const int L = (int)layers.size();
  if (L == 0) return;
  for (const HostLayer* l : layers) {
    for (size_t i = 0; i + 1 < l->x.size(); ++i)
      if (l->x[i + 1] < l->x[i]) throw ArgFailure("permittivity gradients need non-decreasing x coordinates");
    for (size_t i = 0; i + 1 < l->y.size(); ++i)
      if (l->y[i + 1] < l->y[i]) throw ArgFailure("permittivity gradients need non-decreasing y coordinates");
  }
  FourierStage fs = fourier_prepare(c, nx, ny, layers);
  const FourierDev& g = fs.g;
  fourier_tables(s, g, L);
  const int M1 = 2 * nx - 1, M2 = 2 * ny - 1;
  const long MX = g.MX, MY = g.MY;
  const long nF2 = (long)M1 * M2, nG11 = (long)nx * nx * M2, nG22 = (long)ny * ny * M1;
  const long nA = MY * nx * nx, nU = MY * M1, nB = MX * ny * ny;
  Buf G(c, (size_t)(nF2 + nG11 + nG22) * L);
  cd* Gb33 = G.p;
  cd* Gb11 = Gb33 + nF2 * L;
  cd* Gb22 = Gb11 + nG11 * L;
  {
    dim3 grid(64, L);
    fe_bwd_diag_kernel<<<grid, 256, 0, s>>>(nx, ny, fe11b, fe22b, fe33b, sfe, Gb33, Gb11, Gb22);
    launch_check("fe_bwd_diag");
  }
  Buf T(c, (size_t)(nA + nU + nB) * L * 2);
  cd* tbX = T.p;
  cd* U = tbX + nA * L;
  cd* tbY = U + nU * L;
  cd* wk = tbY + nB * L;  // GEMM scratch, nA*L + nB*L >= max(nA, nB) * L
  {
    dim3 grid(64, L);
    fe_bwd_strip_kernel<<<grid, 256, 0, s>>>(g, Gb33, Gb11, Gb22, tbX, U, tbY);
    launch_check("fe_bwd_strip");
  }
  // Tb = -t^H tb t^H for every strip (x) and column (y) inverse-Toeplitz block
  auto toep_bwd = [&](int k, const cd* tinv, cd* tb, long slots) {
    const long kk = (long)k * k;
    zgemm_strided(s, OP_N, OP_C, k, k, k, ONE, tb, k, kk, tinv, k, kk, ZERO, wk, k, kk, (int)slots);
    zgemm_strided(s, OP_C, OP_N, k, k, k, MONE, tinv, k, kk, wk, k, kk, ZERO, tb, k, kk, (int)slots);
  };
  toep_bwd(nx, g.tFFx, tbX, MY * L);
  toep_bwd(ny, g.tFFy, tbY, MX * L);
  Buf V(c, (size_t)(MY * M1 + MX * M2) * L);
  cd* vbx = V.p;
  cd* vby = vbx + MY * M1 * L;
  {
    const long tot = MY * L * M1;
    toep_diag_sum_kernel<<<cdiv_up(tot, 256), 256, 0, s>>>(nx, tbX, (long)nx * nx, vbx, M1, g.ncy, (int)MY, tot);
    launch_check("toep_diag_sum_x");
    const long toty = MX * L * M2;
    toep_diag_sum_kernel<<<cdiv_up(toty, 256), 256, 0, s>>>(ny, tbY, (long)ny * ny, vby, M2, g.ncx, (int)MX, toty);
    launch_check("toep_diag_sum_y");
  }
  DevBuf doffs = c.ws.get(sizeof(long) * L);
  VRC_HIP(hipMemcpyAsync(doffs.ptr, offs.data(), sizeof(long) * L, hipMemcpyHostToDevice, s));
  {
    dim3 grid(64, L);
    fe_bwd_cell_kernel<<<grid, 256, 0, s>>>(g, U, vbx, vby, epsb, doffs.as<long>());
    launch_check("fe_bwd_cell");
  }
  VRC_HIP(hipStreamSynchronize(s));  // the offsets vector is the caller's
}

void assemble_pq_vjp(vrc_ctx& c, const AssemblyBatch& a, const cd* Pb, const cd* Qb, long sPQ, cd* epsb,
                     const std::vector<long>& offs) {
  if (!a.periodic) throw ArgFailure("permittivity gradients are implemented for periodic boundaries only");
  hipStream_t s = c.stream;
  const int n = a.nx * a.ny, L = a.nlayers;
  const long nn = (long)n * n;
  // E = fe33^{-1} (recomputed)
  Buf fe(c, (size_t)nn * 3 * L);
  cd* fe11 = fe.p;
  cd* fe22 = fe11 + nn * L;
  cd* fe33 = fe22 + nn * L;
  assemble_fe(c, a, fe11, fe22, fe33, nn);
  Ops o(s, n, L);
  Buf E(c, (size_t)nn * L);
  invert(c, o, fe33, nn, E.p);
  // fe11b, fe22b, Eb into the (now free) fe11, fe22 and fe33 buffers
  cd* Eb = fe33;
  {
    dim3 g(cdiv_up(n, 256), n, L);
    pq_bwd_kernel<<<g, 256, 0, s>>>(n, Pb, Qb, sPQ, a.d_alpha, a.d_beta, a.sab, a.d_k0sq, Eb, fe11, fe22, nn);
    launch_check("pq_bwd");
  }
  // fe33b = -E^H Eb E^H
  Buf t(c, (size_t)nn * L);
  o.mm(OP_N, OP_C, ONE, Eb, nn, E.p, nn, ZERO, t.p, nn);
  o.mm(OP_C, OP_N, MONE, E.p, nn, t.p, nn, ZERO, fe33, nn);
  assemble_fe_vjp(c, a.nx, a.ny, a.layers, fe11, fe22, fe33, nn, epsb, offs);
}

// ===================================================================== (c) eig
void wave_eqn_vjp(vrc_ctx& c, int N, int batch, const cd* P, const cd* Q, long sPQ, const cd* W, const cd* V,
                  const cd* gam, const cd* Wi, cd* Wb, const cd* Vb, cd* gb, cd* Pb, cd* Qb) {
  hipStream_t s = c.stream;
  Ops o(s, N, batch);
  const long NN = o.NN;
  Buf cig(c, (size_t)N * batch), vd(c, (size_t)N * batch), lb(c, (size_t)N * batch), lam(c, (size_t)N * batch);
  Buf T(c, (size_t)NN * batch), K(c, (size_t)NN * batch);
  Buf WiOwn(c, Wi ? 1 : (size_t)NN * batch);
  if (!Wi) {
    invert(c, o, W, NN, WiOwn.p);
    Wi = WiOwn.p;
  }
  // V = Q W diag(1/gamma):  T = Vb diag(conj(1/gamma)); Qb = T W^H; Wb += Q^H T
  o.vec(gam, N, cig.p, N, VEC_CONJ_INV);
  zdscale(s, N, N, nullptr, 0, cig.p, N, 0, Vb, N, NN, T.p, N, NN, batch);
  o.mm(OP_N, OP_C, ONE, T.p, NN, W, NN, ZERO, Qb, NN);
  o.mm(OP_C, OP_N, ONE, Q, sPQ, T.p, NN, ONE, Wb, NN);
  // gb -= conj(1/gamma) diag(V^H Vb); lb = conj(1/(2 gamma)) gb; lam = gamma^2
  {
    dim3 g(N, batch);
    coldot_kernel<<<g, 256, 0, s>>>(N, V, Vb, NN, vd.p, N);
    launch_check("coldot");
    dim3 g2(cdiv_up(N, 256), batch);
    gamma_bwd_kernel<<<g2, 256, 0, s>>>(N, gam, vd.p, gb, lb.p, lam.p, N);
    launch_check("gamma_bwd");
  }
  // PQ adjoint: Ab = W^{-H} K W^H with G = W^H Wb
  o.mm(OP_C, OP_N, ONE, W, NN, Wb, NN, ZERO, T.p, NN);
  {
    dim3 g(cdiv_up(N, 256), N, batch);
    eig_k_kernel<<<g, 256, 0, s>>>(N, T.p, lam.p, lb.p, N, K.p, NN);
    launch_check("eig_k");
  }
  o.mm(OP_N, OP_C, ONE, K.p, NN, W, NN, ZERO, T.p, NN);
  o.mm(OP_C, OP_N, ONE, Wi, NN, T.p, NN, ZERO, K.p, NN);  // K <- Ab
  // PQ = P Q: Pb = Ab Q^H ; Qb += P^H Ab
  o.mm(OP_N, OP_C, ONE, K.p, NN, Q, sPQ, ZERO, Pb, NN);
  o.mm(OP_C, OP_N, ONE, P, sPQ, K.p, NN, ONE, Qb, NN);
}

// ===================================================================== (d) slab S
// Forward of RCWAScatterMatrix::compute with explicit inverses, keeping what the
// reverse pass needs.
void layer_smatrix_fwd_saved(vrc_ctx& c, int N, int batch, const double* d_s, const cd* W, const cd* V, const cd* gam,
                             const cd* W0, const cd* V0, long sRef, SlabSaved& sv) {
  StageScope st(ST_ADJOINT);
  hipStream_t s = c.stream;
  Ops o(s, N, batch);
  const long NN = o.NN;
  const size_t M = sizeof(cd) * (size_t)NN * batch;
  for (DevBuf* b : {&sv.Wi, &sv.Vi, &sv.F0, &sv.G0, &sv.A, &sv.B, &sv.Ai, &sv.H, &sv.Gm, &sv.F, &sv.Mi, &sv.Rud,
                    &sv.Tuu})
    *b = c.ws.get(M);
  sv.x = c.ws.get(sizeof(cd) * (size_t)N * batch);
  Buf XA(c, (size_t)NN * batch), XB(c, (size_t)NN * batch), t1(c, (size_t)NN * batch), t2(c, (size_t)NN * batch);
  cd *Wi = sv.Wi.as<cd>(), *Vi = sv.Vi.as<cd>(), *F0 = sv.F0.as<cd>(), *G0 = sv.G0.as<cd>(), *A = sv.A.as<cd>(),
     *B = sv.B.as<cd>(), *Ai = sv.Ai.as<cd>(), *H = sv.H.as<cd>(), *Gm = sv.Gm.as<cd>(), *F = sv.F.as<cd>(),
     *Mi = sv.Mi.as<cd>(), *x = sv.x.as<cd>();
  invert(c, o, W, NN, Wi);
  invert(c, o, V, NN, Vi);
  o.mm(OP_N, OP_N, ONE, Wi, NN, W0, sRef, ZERO, F0, NN);
  o.mm(OP_N, OP_N, ONE, Vi, NN, V0, sRef, ZERO, G0, NN);
  o.add(ONE, F0, NN, ONE, G0, NN, A, NN);
  o.add(ONE, F0, NN, MONE, G0, NN, B, NN);
  invert(c, o, A, NN, Ai);
  {
    dim3 g(cdiv_up(N, 256), batch);
    phase_fwd_kernel<<<g, 256, 0, s>>>(N, gam, d_s, x, N);
    launch_check("phase_fwd");
  }
  zdscale(s, N, N, x, N, nullptr, 0, 0, A, N, NN, XA.p, N, NN, batch);
  zdscale(s, N, N, x, N, nullptr, 0, 0, B, N, NN, XB.p, N, NN, batch);
  o.mm(OP_N, OP_N, ONE, Ai, NN, B, NN, ZERO, H, NN);
  o.mm(OP_N, OP_N, ONE, Ai, NN, XA.p, NN, ZERO, Gm, NN);
  o.mm(OP_N, OP_N, ONE, Ai, NN, XB.p, NN, ZERO, F, NN);
  o.eye(t1.p, NN);
  o.mm(OP_N, OP_N, MONE, F, NN, F, NN, ONE, t1.p, NN);  // M = I - F F
  invert(c, o, t1.p, NN, Mi);
  o.copy(H, NN, t1.p, NN, MONE);
  o.mm(OP_N, OP_N, ONE, F, NN, Gm, NN, ONE, t1.p, NN);  // R1 = F Gm - H
  o.copy(Gm, NN, t2.p, NN);
  o.mm(OP_N, OP_N, MONE, F, NN, H, NN, ONE, t2.p, NN);  // T1 = Gm - F H
  o.mm(OP_N, OP_N, ONE, Mi, NN, t1.p, NN, ZERO, sv.Rud.as<cd>(), NN);
  o.mm(OP_N, OP_N, ONE, Mi, NN, t2.p, NN, ZERO, sv.Tuu.as<cd>(), NN);
}

// Reverse of layer_smatrix_fwd_saved (consumes sv: several buffers are reused).
void layer_smatrix_bwd_saved(vrc_ctx& c, int N, int batch, const double* d_s, SlabSaved& sv, const cd* tb,
                             const cd* rb, cd* Wb, cd* Vb, cd* gb) {
  StageScope st(ST_ADJOINT);
  hipStream_t s = c.stream;
  Ops o(s, N, batch);
  const long NN = o.NN;
  const size_t M = (size_t)NN * batch;
  Buf t1(c, M), t2(c, M), t3(c, M), XA(c, M), XB(c, M);
  Buf xb(c, (size_t)N * batch), cx(c, (size_t)N * batch);
  cd *Wi = sv.Wi.as<cd>(), *Vi = sv.Vi.as<cd>(), *F0 = sv.F0.as<cd>(), *G0 = sv.G0.as<cd>(), *A = sv.A.as<cd>(),
     *B = sv.B.as<cd>(), *Ai = sv.Ai.as<cd>(), *H = sv.H.as<cd>(), *Gm = sv.Gm.as<cd>(), *F = sv.F.as<cd>(),
     *Mi = sv.Mi.as<cd>(), *x = sv.x.as<cd>(), *R = sv.Rud.as<cd>(), *Tt = sv.Tuu.as<cd>();
  cd* R1b = t1.p;
  cd* T1b = t2.p;
  o.mm(OP_C, OP_N, ONE, Mi, NN, rb, NN, ZERO, R1b, NN);
  o.mm(OP_C, OP_N, ONE, Mi, NN, tb, NN, ZERO, T1b, NN);
  cd* Mb = Mi;  // M^{-1} no longer needed
  o.mm(OP_N, OP_C, MONE, R1b, NN, R, NN, ZERO, Mb, NN);
  o.mm(OP_N, OP_C, MONE, T1b, NN, Tt, NN, ONE, Mb, NN);
  cd* Fb = R;    // Rud no longer needed
  cd* Gmb = Tt;  // Tuu no longer needed
  cd* Hb = t3.p;
  // Fb = -Mb F^H - F^H Mb + R1b Gm^H - T1b H^H
  o.mm(OP_N, OP_C, MONE, Mb, NN, F, NN, ZERO, Fb, NN);
  o.mm(OP_C, OP_N, MONE, F, NN, Mb, NN, ONE, Fb, NN);
  o.mm(OP_N, OP_C, ONE, R1b, NN, Gm, NN, ONE, Fb, NN);
  o.mm(OP_N, OP_C, MONE, T1b, NN, H, NN, ONE, Fb, NN);
  // Gmb = F^H R1b + T1b ; Hb = -R1b - F^H T1b
  o.copy(T1b, NN, Gmb, NN);
  o.mm(OP_C, OP_N, ONE, F, NN, R1b, NN, ONE, Gmb, NN);
  o.copy(R1b, NN, Hb, NN, MONE);
  o.mm(OP_C, OP_N, MONE, F, NN, T1b, NN, ONE, Hb, NN);
  // [B | XA | XB]b = A^{-H} [Hb | Gmb | Fb]; Ab = -(Bb H^H + XAb Gm^H + XBb F^H)
  cd* Bb = t1.p;   // R1b consumed
  cd* XAb = t2.p;  // T1b consumed
  cd* XBb = Mb;    // Mb consumed
  o.mm(OP_C, OP_N, ONE, Ai, NN, Hb, NN, ZERO, Bb, NN);
  o.mm(OP_C, OP_N, ONE, Ai, NN, Gmb, NN, ZERO, XAb, NN);
  o.mm(OP_C, OP_N, ONE, Ai, NN, Fb, NN, ZERO, XBb, NN);
  cd* Ab = Hb;
  o.mm(OP_N, OP_C, MONE, Bb, NN, H, NN, ZERO, Ab, NN);
  o.mm(OP_N, OP_C, MONE, XAb, NN, Gm, NN, ONE, Ab, NN);
  o.mm(OP_N, OP_C, MONE, XBb, NN, F, NN, ONE, Ab, NN);
  // phases: xb_i = sum_k conj(A_ik) XAb_ik + conj(B_ik) XBb_ik ; Ab += conj(X) XAb ; Bb += conj(X) XBb
  {
    dim3 g(cdiv_up(N, 256), batch);
    rowdot_kernel<<<g, 256, 0, s>>>(N, A, XAb, B, XBb, NN, xb.p, N);
    launch_check("rowdot");
    phase_bwd_kernel<<<g, 256, 0, s>>>(N, x, xb.p, d_s, gb, N);
    launch_check("phase_bwd");
  }
  o.vec(x, N, cx.p, N, VEC_CONJ);
  zdscale(s, N, N, cx.p, N, nullptr, 0, 0, XAb, N, NN, XAb, N, NN, batch);
  zdscale(s, N, N, cx.p, N, nullptr, 0, 0, XBb, N, NN, XBb, N, NN, batch);
  o.add(ONE, Ab, NN, ONE, XAb, NN, Ab, NN);
  o.add(ONE, Bb, NN, ONE, XBb, NN, Bb, NN);
  // F0b = Ab + Bb ; G0b = Ab - Bb ; Wb = -W^{-H} F0b F0^H ; Vb = -V^{-H} G0b G0^H
  cd* F0b = Gm;
  cd* G0b = H;
  o.add(ONE, Ab, NN, ONE, Bb, NN, F0b, NN);
  o.add(ONE, Ab, NN, MONE, Bb, NN, G0b, NN);
  o.mm(OP_N, OP_C, ONE, F0b, NN, F0, NN, ZERO, XA.p, NN);
  o.mm(OP_C, OP_N, MONE, Wi, NN, XA.p, NN, ZERO, Wb, NN);
  o.mm(OP_N, OP_C, ONE, G0b, NN, G0, NN, ZERO, XB.p, NN);
  o.mm(OP_C, OP_N, MONE, Vi, NN, XB.p, NN, ZERO, Vb, NN);
}

void layer_smatrix_vjp(vrc_ctx& c, int N, int batch, const double* d_s, const cd* W, const cd* V, const cd* gam,
                       const cd* W0, const cd* V0, long sRef, const cd* tb, const cd* rb, cd* Wb, cd* Vb, cd* gb,
                       cd* Wi_out) {
  SlabSaved sv;
  layer_smatrix_fwd_saved(c, N, batch, d_s, W, V, gam, W0, V0, sRef, sv);
  if (Wi_out)
    zgeam(c.stream, N, N, ONE, sv.Wi.as<cd>(), N, (long)N * N, ZERO, nullptr, 0, 0, Wi_out, N, (long)N * N, batch);
  layer_smatrix_bwd_saved(c, N, batch, d_s, sv, tb, rb, Wb, Vb, gb);
}

// ===================================================================== (d) star product
void star_product_vjp(vrc_ctx& c, int N, int batch, const SMat& S0, const cd* tdd, const cd* tuu, const cd* rud,
                      const cd* rdu, long ss, SMat& Sb, cd* tddb, cd* tuub, cd* rudb, cd* rdub, long sb,
                      const cd* saved) {
  StageScope st(ST_ADJOINT);
  hipStream_t s = c.stream;
  Ops o(s, N, batch);
  const long NN = o.NN;
  const size_t M = (size_t)NN * batch;
  Buf C1o(c, saved ? 1 : M), RC1o(c, saved ? 1 : M), C2o(c, saved ? 1 : M);
  Buf U1(c, M), U2(c, M), Wk(c, M), C1b(c, M), C2b(c, M), RC1b(c, M);
  Buf Tddn(c, M), Tuun(c, M), Rudn(c, M);
  const cd* C1 = saved ? saved : C1o.p;
  const cd* RC1 = saved ? saved + M : RC1o.p;
  const cd* C2 = saved ? saved + 2 * M : C2o.p;
  if (!saved) {
    // recompute C1 = (I - rdu Rud)^{-1}, RC1 = Rud C1, C2 = I + RC1 rdu
    o.eye(Wk.p, NN);
    o.mm(OP_N, OP_N, MONE, rdu, ss, S0.Rud, NN, ONE, Wk.p, NN);
    invert(c, o, Wk.p, NN, C1o.p);
    o.mm(OP_N, OP_N, ONE, S0.Rud, NN, C1o.p, NN, ZERO, RC1o.p, NN);
    o.eye(C2o.p, NN);
    o.mm(OP_N, OP_N, ONE, RC1o.p, NN, rdu, ss, ONE, C2o.p, NN);
  }
  // Rdu' = Rdu + (Tdd rdu) C2 Tuu
  o.mm(OP_N, OP_N, ONE, S0.Tdd, NN, rdu, ss, ZERO, U1.p, NN);  // U1 = Tdd rdu
  o.mm(OP_N, OP_N, ONE, U1.p, NN, C2, NN, ZERO, U2.p, NN);    // U2 = U1 C2
  o.mm(OP_C, OP_N, ONE, U2.p, NN, Sb.Rdu, NN, ZERO, Tuun.p, NN);      // Tuub = U2^H Rdu'b
  o.mm(OP_N, OP_C, ONE, Sb.Rdu, NN, S0.Tuu, NN, ZERO, Wk.p, NN);      // U2b = Rdu'b Tuu^H
  o.mm(OP_C, OP_N, ONE, U1.p, NN, Wk.p, NN, ZERO, C2b.p, NN);         // C2b = U1^H U2b
  o.mm(OP_N, OP_C, ONE, Wk.p, NN, C2, NN, ZERO, U2.p, NN);          // U1b = U2b C2^H (into U2)
  o.mm(OP_N, OP_C, ONE, U2.p, NN, rdu, ss, ZERO, Tddn.p, NN);         // Tddb = U1b rdu^H
  o.mm(OP_C, OP_N, ONE, S0.Tdd, NN, U2.p, NN, ZERO, rdub, sb);        // rdub = Tdd^H U1b
  // Tdd' = (Tdd C1) tdd
  o.mm(OP_N, OP_N, ONE, S0.Tdd, NN, C1, NN, ZERO, U1.p, NN);        // U3 = Tdd C1
  o.mm(OP_C, OP_N, ONE, U1.p, NN, Sb.Tdd, NN, ZERO, tddb, sb);        // tddb = U3^H Tdd'b
  o.mm(OP_N, OP_C, ONE, Sb.Tdd, NN, tdd, ss, ZERO, Wk.p, NN);         // U3b = Tdd'b tdd^H
  o.mm(OP_N, OP_C, ONE, Wk.p, NN, C1, NN, ONE, Tddn.p, NN);         // Tddb += U3b C1^H
  o.mm(OP_C, OP_N, ONE, S0.Tdd, NN, Wk.p, NN, ZERO, C1b.p, NN);       // C1b = Tdd^H U3b
  // Rud' = rud + (tuu RC1) tdd
  o.copy(Sb.Rud, NN, rudb, sb);                                        // rudb = Rud'b
  o.mm(OP_N, OP_N, ONE, tuu, ss, RC1, NN, ZERO, U1.p, NN);          // U5 = tuu RC1
  o.mm(OP_C, OP_N, ONE, U1.p, NN, Sb.Rud, NN, ONE, tddb, sb);         // tddb += U5^H Rud'b
  o.mm(OP_N, OP_C, ONE, Sb.Rud, NN, tdd, ss, ZERO, Wk.p, NN);         // U5b = Rud'b tdd^H
  o.mm(OP_C, OP_N, ONE, tuu, ss, Wk.p, NN, ZERO, RC1b.p, NN);         // RC1b = tuu^H U5b
  o.mm(OP_N, OP_C, ONE, Wk.p, NN, RC1, NN, ZERO, tuub, sb);         // tuub = U5b RC1^H
  // Tuu' = (tuu C2) Tuu
  o.mm(OP_N, OP_N, ONE, tuu, ss, C2, NN, ZERO, U1.p, NN);           // U6 = tuu C2
  o.mm(OP_C, OP_N, ONE, U1.p, NN, Sb.Tuu, NN, ONE, Tuun.p, NN);       // Tuub += U6^H Tuu'b
  o.mm(OP_N, OP_C, ONE, Sb.Tuu, NN, S0.Tuu, NN, ZERO, Wk.p, NN);      // U6b = Tuu'b Tuu^H
  o.mm(OP_N, OP_C, ONE, Wk.p, NN, C2, NN, ONE, tuub, sb);           // tuub += U6b C2^H
  o.mm(OP_C, OP_N, ONE, tuu, ss, Wk.p, NN, ONE, C2b.p, NN);           // C2b += tuu^H U6b
  // C2 = I + RC1 rdu
  o.mm(OP_N, OP_C, ONE, C2b.p, NN, rdu, ss, ONE, RC1b.p, NN);         // RC1b += C2b rdu^H
  o.mm(OP_C, OP_N, ONE, RC1, NN, C2b.p, NN, ONE, rdub, sb);         // rdub += RC1^H C2b
  // RC1 = Rud C1
  o.mm(OP_N, OP_C, ONE, RC1b.p, NN, C1, NN, ZERO, Rudn.p, NN);      // Rudb = RC1b C1^H
  o.mm(OP_C, OP_N, ONE, S0.Rud, NN, RC1b.p, NN, ONE, C1b.p, NN);      // C1b += Rud^H RC1b
  // C1 = D^{-1}, D = I - rdu Rud: Db = -C1^H C1b C1^H ; rdub -= Db Rud^H ; Rudb -= rdu^H Db
  o.mm(OP_N, OP_C, ONE, C1b.p, NN, C1, NN, ZERO, Wk.p, NN);
  o.mm(OP_C, OP_N, MONE, C1, NN, Wk.p, NN, ZERO, U1.p, NN);         // Db
  o.mm(OP_N, OP_C, MONE, U1.p, NN, S0.Rud, NN, ONE, rdub, sb);
  o.mm(OP_C, OP_N, MONE, rdu, ss, U1.p, NN, ONE, Rudn.p, NN);
  // Rdub of the prefix is Rdu'b itself (unchanged)
  o.copy(Tddn.p, NN, Sb.Tdd, NN);
  o.copy(Tuun.p, NN, Sb.Tuu, NN);
  o.copy(Rudn.p, NN, Sb.Rud, NN);
}

// ===================================================================== (d) left projection
void project_left_vjp(vrc_ctx& c, int N, int batch, const SMat& S, const LUPair& LU1, const cd* W0, const cd* V0,
                      long sRef, SMat& Sb) {
  StageScope st(ST_ADJOINT);
  hipStream_t s = c.stream;
  Ops o(s, N, batch);
  const long NN = o.NN;
  const size_t M = (size_t)NN * batch;
  Buf F(c, M), G(c, M), A(c, M), B(c, M), Ki(c, M), Rn(c, M), Tn(c, M), Y(c, M), t1(c, M), t2(c, M);
  zgetrs(s, N, N, LU1.W, N, LU1.sM, LU1.pW, N, W0, N, sRef, F.p, N, NN, batch, c.ws);
  zgetrs(s, N, N, LU1.V, N, LU1.sM, LU1.pV, N, V0, N, sRef, G.p, N, NN, batch, c.ws);
  const cd h = mk(0.5, 0.0), mh = mk(-0.5, 0.0);
  o.add(h, F.p, NN, h, G.p, NN, A.p, NN);
  o.add(h, F.p, NN, mh, G.p, NN, B.p, NN);
  // K = A - Rdu B ; Ki = K^{-1} ; Rdu_n = Ki (Rdu A - B) ; Tdd_n = Ki Tdd ; Y = B Rdu_n + A
  o.copy(A.p, NN, t1.p, NN);
  o.mm(OP_N, OP_N, MONE, S.Rdu, NN, B.p, NN, ONE, t1.p, NN);
  invert(c, o, t1.p, NN, Ki.p);
  o.copy(B.p, NN, t1.p, NN, MONE);
  o.mm(OP_N, OP_N, ONE, S.Rdu, NN, A.p, NN, ONE, t1.p, NN);
  o.mm(OP_N, OP_N, ONE, Ki.p, NN, t1.p, NN, ZERO, Rn.p, NN);
  o.mm(OP_N, OP_N, ONE, Ki.p, NN, S.Tdd, NN, ZERO, Tn.p, NN);
  o.copy(A.p, NN, Y.p, NN);
  o.mm(OP_N, OP_N, ONE, B.p, NN, Rn.p, NN, ONE, Y.p, NN);
  // backward
  cd* Tuub = F.p;
  cd* Rdunb = G.p;
  o.mm(OP_N, OP_C, ONE, Sb.Tuu, NN, Y.p, NN, ZERO, Tuub, NN);         // Tuub = Tuu_nb Y^H
  o.mm(OP_C, OP_N, ONE, S.Tuu, NN, Sb.Tuu, NN, ZERO, t1.p, NN);       // Yb = Tuu^H Tuu_nb
  o.copy(Sb.Rdu, NN, Rdunb, NN);
  o.mm(OP_C, OP_N, ONE, B.p, NN, t1.p, NN, ONE, Rdunb, NN);           // Rdu_nb += B^H Yb
  o.mm(OP_N, OP_N, ONE, S.Tuu, NN, B.p, NN, ZERO, t1.p, NN);          // U = Tuu B
  o.copy(Sb.Tdd, NN, t2.p, NN);
  o.mm(OP_C, OP_N, ONE, t1.p, NN, Sb.Rud, NN, ONE, t2.p, NN);         // Tdd_nb += U^H Rud_nb
  o.mm(OP_N, OP_C, ONE, Sb.Rud, NN, Tn.p, NN, ZERO, t1.p, NN);
  o.mm(OP_N, OP_C, ONE, t1.p, NN, B.p, NN, ONE, Tuub, NN);            // Tuub += Rud_nb Tdd_n^H B^H
  // Zb = Ki^H Rdu_nb ; Tddb = Ki^H Tdd_nb ; Kb = -Zb Rdu_n^H - Tddb Tdd_n^H ; Rdub = Zb A^H - Kb B^H
  o.mm(OP_C, OP_N, ONE, Ki.p, NN, Rdunb, NN, ZERO, t1.p, NN);         // Zb
  o.mm(OP_C, OP_N, ONE, Ki.p, NN, t2.p, NN, ZERO, Sb.Tdd, NN);        // Tddb (prefix)
  o.mm(OP_N, OP_C, MONE, t1.p, NN, Rn.p, NN, ZERO, Y.p, NN);
  o.mm(OP_N, OP_C, MONE, Sb.Tdd, NN, Tn.p, NN, ONE, Y.p, NN);         // Kb
  o.mm(OP_N, OP_C, ONE, t1.p, NN, A.p, NN, ZERO, Sb.Rdu, NN);
  o.mm(OP_N, OP_C, MONE, Y.p, NN, B.p, NN, ONE, Sb.Rdu, NN);          // Rdub
  o.copy(Tuub, NN, Sb.Tuu, NN);
  // Rudb = Rud_nb (unchanged)
}

}  // namespace vrc
