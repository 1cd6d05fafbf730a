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
// Ebar[2b] = (rb + tb) / 2, Ebar[2b+1] = (rb - tb) / 2 (reverse of evenodd_combine).
__global__ void evenodd_combine_bwd_kernel(long NN, const cd* tb, const cd* rb, cd* Eb) {
  const long b = blockIdx.y;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < NN; i += (long)gridDim.x * blockDim.x) {
    const cd t = tb[b * NN + i], r = rb[b * NN + i];
    Eb[(2 * b) * NN + i] = mk(0.5 * (r.x + t.x), 0.5 * (r.y + t.y));
    Eb[(2 * b + 1) * NN + i] = mk(0.5 * (r.x - t.x), 0.5 * (r.y - t.y));
  }
}

// Reverse of evenodd_system: from Lbar, Rbar (interleaved as L, R) and x, with
// a = Lb+ - Rb+, p = Lb+ + Rb+, q = Lb- - Rb-, d = Lb- + Rb-:
//   F0b = conj(1-x) a + conj(1+x) q,  G0b = conj(1+x) p + conj(1-x) d,
//   D1 = q - a, D2 = p - d  (xbar_i = sum_k conj(F0_ik) D1_ik + conj(G0_ik) D2_ik).
__global__ void evenodd_system_bwd_kernel(int N, const cd* Lb, const cd* Rb, const cd* x, cd* F0b, cd* G0b, cd* D1,
                                          cd* D2) {
  const long b = blockIdx.z;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= N) return;
  const long NN = (long)N * N, o = r + (long)c * N;
  const cd lp = Lb[(2 * b) * NN + o], lm = Lb[(2 * b + 1) * NN + o];
  const cd rp = Rb[(2 * b) * NN + o], rm = Rb[(2 * b + 1) * NN + o];
  const cd av = csub(lp, rp), pv = cadd(lp, rp), qv = csub(lm, rm), dv = cadd(lm, rm);
  const cd xr = x[b * N + r];
  const cd omc = mk(1.0 - xr.x, xr.y), opc = mk(1.0 + xr.x, -xr.y);  // conj(1-x), conj(1+x)
  F0b[b * NN + o] = cadd(cmul(omc, av), cmul(opc, qv));
  G0b[b * NN + o] = cadd(cmul(opc, pv), cmul(omc, dv));
  D1[b * NN + o] = csub(qv, av);
  D2[b * NN + o] = csub(pv, dv);
}

// Forward of the slab S-matrix in the even/odd form of slab_evenodd (ops.hip), which
// computes the quantities of RCWAScatterMatrix::compute (src/rcwa/RCWAScatterMatrix.cpp:
// 26-85), with the inverses W^{-1}, V^{-1}, L^{-1} formed explicitly for the reverse pass.
void layer_smatrix_fwd_saved(vrc_ctx& c, int N, int batch, const double* d_s, const cd* W, const cd* V, const cd* gam,
                             const cd* W0, const cd* V0, long sRef, SlabSaved& sv) {
  StageScope st(ST_ADJOINT);
  hipStream_t s = c.stream;
  Ops o(s, N, batch), o2(s, N, 2 * batch);
  const long NN = o.NN;
  const size_t M = sizeof(cd) * (size_t)NN * batch;
  for (DevBuf* b : {&sv.Wi, &sv.Vi, &sv.F0, &sv.G0, &sv.Rud, &sv.Tuu}) *b = c.ws.get(M);
  sv.Li = c.ws.get(2 * M);
  sv.E = c.ws.get(2 * M);
  sv.x = c.ws.get(sizeof(cd) * (size_t)N * batch);
  cd *Wi = sv.Wi.as<cd>(), *Vi = sv.Vi.as<cd>(), *F0 = sv.F0.as<cd>(), *G0 = sv.G0.as<cd>(), *Li = sv.Li.as<cd>(),
     *E = sv.E.as<cd>(), *x = sv.x.as<cd>();
  invert(c, o, W, NN, Wi);
  invert(c, o, V, NN, Vi);
  o.mm(OP_N, OP_N, ONE, Wi, NN, W0, sRef, ZERO, F0, NN);
  o.mm(OP_N, OP_N, ONE, Vi, NN, V0, sRef, ZERO, G0, NN);
  {
    dim3 g(cdiv_up(N, 256), batch);
    phase_fwd_kernel<<<g, 256, 0, s>>>(N, gam, d_s, x, N);
    launch_check("phase_fwd");
  }
  Buf Lm(c, 2 * (size_t)NN * batch), Rm(c, 2 * (size_t)NN * batch);
  evenodd_system(s, N, batch, F0, NN, G0, NN, nullptr, x, Lm.p, Rm.p);
  invert(c, o2, Lm.p, NN, Li);
  o2.mm(OP_N, OP_N, ONE, Li, NN, Rm.p, NN, ZERO, E, NN);
  evenodd_combine(s, N, batch, E, sv.Tuu.as<cd>(), sv.Rud.as<cd>());
}

// Reverse of layer_smatrix_fwd_saved (consumes sv: several buffers are reused; Wi kept).
void layer_smatrix_bwd_saved(vrc_ctx& c, int N, int batch, const double* d_s, SlabSaved& sv, const cd* tb,
                             const cd* rb, cd* Wb, cd* Vb, cd* gb) {
  StageScope st(ST_ADJOINT);
  hipStream_t s = c.stream;
  Ops o(s, N, batch), o2(s, N, 2 * batch);
  const long NN = o.NN;
  const size_t M = (size_t)NN * batch;
  cd *Wi = sv.Wi.as<cd>(), *Vi = sv.Vi.as<cd>(), *F0 = sv.F0.as<cd>(), *G0 = sv.G0.as<cd>(), *Li = sv.Li.as<cd>(),
     *E = sv.E.as<cd>(), *x = sv.x.as<cd>();
  Buf Eb(c, 2 * M), Rbar(c, 2 * M), xb(c, (size_t)N * batch);
  {
    dim3 g(cdiv_up(NN, 256) < 1024 ? cdiv_up(NN, 256) : 1024, batch);
    evenodd_combine_bwd_kernel<<<g, 256, 0, s>>>(NN, tb, rb, Eb.p);
    launch_check("evenodd_combine_bwd");
  }
  // E = L^{-1} R:  Rbar = L^{-H} Ebar ; Lbar = -Rbar E^H (into Eb)
  o2.mm(OP_C, OP_N, ONE, Li, NN, Eb.p, NN, ZERO, Rbar.p, NN);
  o2.mm(OP_N, OP_C, MONE, Rbar.p, NN, E, NN, ZERO, Eb.p, NN);
  // F0b, G0b into E (consumed); D1, D2 into Li (consumed)
  cd *F0b = E, *G0b = E + M, *D1 = Li, *D2 = Li + M;
  {
    dim3 g(cdiv_up(N, 256), N, batch);
    evenodd_system_bwd_kernel<<<g, 256, 0, s>>>(N, Eb.p, Rbar.p, x, F0b, G0b, D1, D2);
    launch_check("evenodd_system_bwd");
  }
  {
    dim3 g(cdiv_up(N, 256), batch);
    rowdot_kernel<<<g, 256, 0, s>>>(N, F0, D1, G0, D2, NN, xb.p, N);
    launch_check("rowdot");
    phase_bwd_kernel<<<g, 256, 0, s>>>(N, x, xb.p, d_s, gb, N);
    launch_check("phase_bwd");
  }
  // F0 = W^{-1} W0: Wb = -W^{-H} F0b F0^H ; likewise Vb
  cd* T = Rbar.p;
  o.mm(OP_N, OP_C, ONE, F0b, NN, F0, NN, ZERO, T, NN);
  o.mm(OP_C, OP_N, MONE, Wi, NN, T, NN, ZERO, Wb, NN);
  o.mm(OP_N, OP_C, ONE, G0b, NN, G0, NN, ZERO, T, NN);
  o.mm(OP_C, OP_N, MONE, Vi, NN, T, NN, ZERO, Vb, NN);
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
// Reverse of the factored star product (ops.hip, star_product):
//   M = I - Rud rdu, [Yt | Yr] = M^{-1} [Tuu | Rud], K = Yr tdd, Rud' = rud + tuu K,
//   T1 = tdd + rdu K, Z2 = rdu Yt, Rdu' = Rdu + Tdd Z2, Tdd' = Tdd T1, Tuu' = tuu Yt.
void star_product_vjp(vrc_ctx& c, int N, int batch, const SMat& S0, const cd* tdd, const cd* tuu, const cd* rud,
                      const cd* rdu, long ss, SMat& Sb, cd* tddb, cd* tuub, cd* rudb, cd* rdub, long sb,
                      const cd* saved) {
  StageScope st(ST_ADJOINT);
  hipStream_t s = c.stream;
  Ops o(s, N, batch);
  const long NN = o.NN;
  const size_t M = (size_t)NN * batch;
  Buf own(c, saved ? 1 : STAR_SAVE * M);
  if (!saved) {
    // recompute the forward intermediates on a copy of the prefix
    Buf S4(c, 4 * M);
    SMat T;
    T.Tdd = S4.p;
    T.Tuu = S4.p + M;
    T.Rud = S4.p + 2 * M;
    T.Rdu = S4.p + 3 * M;
    o.copy(S0.Tdd, NN, T.Tdd, NN);
    o.copy(S0.Tuu, NN, T.Tuu, NN);
    o.copy(S0.Rud, NN, T.Rud, NN);
    o.copy(S0.Rdu, NN, T.Rdu, NN);
    star_product(c, N, batch, T, tdd, tuu, rud, rdu, ss, own.p);
    saved = own.p;
  }
  const cd *Mi = saved, *Yt = saved + M, *Yr = saved + 2 * M, *K = saved + 3 * M, *T1 = saved + 4 * M,
           *Z2 = saved + 5 * M;
  Buf Ytb(c, M), T1b(c, M), Z2b(c, M), Tn(c, M), Kb(c, M);
  // Tuu' = tuu Yt
  o.mm(OP_N, OP_C, ONE, Sb.Tuu, NN, Yt, NN, ZERO, tuub, sb);
  o.mm(OP_C, OP_N, ONE, tuu, ss, Sb.Tuu, NN, ZERO, Ytb.p, NN);
  // Tdd' = Tdd T1 ; Rdu' = Rdu + Tdd Z2 (Rdu's cotangent passes through unchanged)
  o.mm(OP_C, OP_N, ONE, S0.Tdd, NN, Sb.Tdd, NN, ZERO, T1b.p, NN);
  o.mm(OP_C, OP_N, ONE, S0.Tdd, NN, Sb.Rdu, NN, ZERO, Z2b.p, NN);
  o.mm(OP_N, OP_C, ONE, Sb.Tdd, NN, T1, NN, ZERO, Tn.p, NN);
  o.mm(OP_N, OP_C, ONE, Sb.Rdu, NN, Z2, NN, ONE, Tn.p, NN);
  // Z2 = rdu Yt ; T1 = tdd + rdu K
  o.mm(OP_C, OP_N, ONE, rdu, ss, Z2b.p, NN, ONE, Ytb.p, NN);
  o.mm(OP_C, OP_N, ONE, rdu, ss, T1b.p, NN, ZERO, Kb.p, NN);
  o.mm(OP_N, OP_C, ONE, T1b.p, NN, K, NN, ZERO, rdub, sb);
  o.mm(OP_N, OP_C, ONE, Z2b.p, NN, Yt, NN, ONE, rdub, sb);
  o.copy(T1b.p, NN, tddb, sb);
  // Rud' = rud + tuu K
  o.copy(Sb.Rud, NN, rudb, sb);
  o.mm(OP_N, OP_C, ONE, Sb.Rud, NN, K, NN, ONE, tuub, sb);
  o.mm(OP_C, OP_N, ONE, tuu, ss, Sb.Rud, NN, ONE, Kb.p, NN);
  // K = Yr tdd
  cd* Yrb = Z2b.p;  // consumed
  o.mm(OP_N, OP_C, ONE, Kb.p, NN, tdd, ss, ZERO, Yrb, NN);
  o.mm(OP_C, OP_N, ONE, Yr, NN, Kb.p, NN, ONE, tddb, sb);
  // [Yt | Yr] = M^{-1} [Tuu | Rud]: [Tuub | Rudb] = M^{-H} [Ytb | Yrb], Mb = -[Tuub | Rudb] [Yt | Yr]^H
  o.mm(OP_C, OP_N, ONE, Mi, NN, Ytb.p, NN, ZERO, Sb.Tuu, NN);
  o.mm(OP_C, OP_N, ONE, Mi, NN, Yrb, NN, ZERO, Sb.Rud, NN);
  cd* Mb = T1b.p;  // consumed
  o.mm(OP_N, OP_C, MONE, Sb.Tuu, NN, Yt, NN, ZERO, Mb, NN);
  o.mm(OP_N, OP_C, MONE, Sb.Rud, NN, Yr, NN, ONE, Mb, NN);
  // M = I - Rud rdu
  o.mm(OP_N, OP_C, MONE, Mb, NN, rdu, ss, ONE, Sb.Rud, NN);
  o.mm(OP_C, OP_N, MONE, S0.Rud, NN, Mb, NN, ONE, rdub, sb);
  o.copy(Tn.p, NN, Sb.Tdd, NN);
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

// ===================================================================== shape-parameter contraction
// part[p][blk] = Re sum over this block's share of conj(A[i]) B[p][i] (i < count).
__global__ void dot_re_partial_kernel(long count, const cd* A, const cd* B, long sB, double* part) {
  const int p = blockIdx.y;
  const cd* b = B + (long)p * sB;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x) {
    const cd x = A[i], y = b[i];
    acc += x.x * y.x + x.y * y.y;
  }
  __shared__ double red[256];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(long)p * gridDim.x + blockIdx.x] = red[0];
}

// out[p] += sum of the nblk partials of parameter p (in order: deterministic).
__global__ void dot_re_finish_kernel(int npara, int nblk, const double* part, double* out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npara) return;
  double acc = 0.0;
  for (int k = 0; k < nblk; ++k) acc += part[(long)p * nblk + k];
  out[p] += acc;
}

void dot_re_accumulate(vrc_ctx& c, long count, const cd* A, const cd* B, long sB, int npara, double* out) {
  if (count <= 0 || npara <= 0) return;
  const int nblk = (int)std::min<long>(256, (count + 255) / 256);
  DevBuf part = c.ws.get(sizeof(double) * (size_t)nblk * npara);
  dot_re_partial_kernel<<<dim3(nblk, npara), 256, 0, c.stream>>>(count, A, B, sB, part.as<double>());
  launch_check("dot_re_partial");
  dot_re_finish_kernel<<<cdiv_up(npara, 64), 64, 0, c.stream>>>(npara, nblk, part.as<double>(), out);
  launch_check("dot_re_finish");
}

}  // namespace vrc
