// Parts (b) and (d): operator assembly, eigen-system post-processing, layer S-matrix,
// Redheffer star product, end projections and the varRCWA perturbation, as batched
// device procedures built on the dense kernels (ZGEMM, LU) plus fused elementwise
// kernels defined here.
#include "vrc_ops.hpp"

namespace vrc {

// ===================================================================== small kernels
// Periodic-boundary P and Q (Kx = diag(alpha), Ky = diag(beta)), src/rcwa/Layer.cpp:95-100,134-138:
//  P = [[a inv33 b, k0^2 I - a inv33 a], [b inv33 b - k0^2 I, -b inv33 a]]
//  Q = [[-a b, a^2 - k0^2 fe22], [-b^2 + k0^2 fe11, b a]]
// with (a X b)[r][c] = alpha_r X[r][c] beta_c; per-layer alpha/beta/k0^2.
__global__ void pq_periodic_kernel(int n, const cd* inv33, const cd* fe11, const cd* fe22, long sfe,
                                   const cd* alpha, const cd* beta, long sab, const double* k0sq, cd* P, cd* Q,
                                   long sPQ) {
  const long l = blockIdx.z;
  const int N = 2 * n;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= N) return;
  const int rb = r / n, ri = r % n, cb = c / n, ci = c % n;
  const cd* al = alpha + l * sab;
  const cd* be = beta + l * sab;
  const double k2 = k0sq[l];
  const long e = l * sfe + ri + (long)ci * n;
  const cd inv = inv33[e];
  const cd ar = al[ri], br = be[ri], ac = al[ci], bc = be[ci];
  const bool dg = (ri == ci);
  cd p, q;
  if (rb == 0 && cb == 0) {
    p = cmul(cmul(ar, inv), bc);
    q = dg ? cneg(cmul(ar, br)) : mk(0.0, 0.0);
  } else if (rb == 0) {
    p = cneg(cmul(cmul(ar, inv), ac));
    if (dg) p.x += k2;
    q = cscale(fe22[e], -k2);
    if (dg) q = cadd(q, cmul(ar, ar));
  } else if (cb == 0) {
    p = cmul(cmul(br, inv), bc);
    if (dg) p.x -= k2;
    q = cscale(fe11[e], k2);
    if (dg) q = csub(q, cmul(br, br));
  } else {
    p = cneg(cmul(cmul(br, inv), ac));
    q = dg ? cmul(br, ar) : mk(0.0, 0.0);
  }
  const long o = l * sPQ + r + (long)c * N;
  P[o] = p;
  Q[o] = q;
}

// Periodic P Q without the N x N product. With a = diag(alpha), b = diag(beta), E = fe33^{-1}
// and the blocks of pq_periodic_kernel (P11 = a E b, P12 = -a E a + k^2 I, P21 = b E b - k^2 I,
// P22 = -b E a; Q11 = -a b, Q12 = a^2 - k^2 fe22, Q21 = k^2 fe11 - b^2, Q22 = b a):
//   PQ11 = a M1 + k^2 Q21, PQ21 = b M1 - k^2 Q11, PQ12 = a M2 + k^2 Q22, PQ22 = b M2 - k^2 Q12
// with [M1 | M2] = E [b Q11 - a Q21 | b Q12 - a Q22] (one n x 2n product per layer).
// First kernel: R = [b Q11 - a Q21 | b Q12 - a Q22] (n x 2n per layer).
__global__ void pq_rhs_kernel(int n, const cd* fe11, const cd* fe22, long sfe, const cd* alpha, const cd* beta,
                              long sab, const double* k0sq, cd* R) {
  const long l = blockIdx.z;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;  // 0 .. 2n-1
  if (r >= n) return;
  const int cb = c / n, ci = c % n;
  const cd* al = alpha + l * sab;
  const cd* be = beta + l * sab;
  const double k2 = k0sq[l];
  const cd ar = al[r], br = be[r];
  const bool dg = (r == ci);
  const long e = l * sfe + r + (long)ci * n;
  cd v;
  if (cb == 0) {
    // b Q11 - a Q21 = -b a b (diagonal) - a (k^2 fe11 - b^2)
    cd q21 = cscale(fe11[e], k2);
    if (dg) q21 = csub(q21, cmul(br, br));
    v = cneg(cmul(ar, q21));
    if (dg) v = csub(v, cmul(br, cmul(ar, br)));
  } else {
    // b Q12 - a Q22 = b (a^2 - k^2 fe22) - a b a (diagonal)
    cd q12 = cscale(fe22[e], -k2);
    if (dg) q12 = cadd(q12, cmul(ar, ar));
    v = cmul(br, q12);
    if (dg) v = csub(v, cmul(ar, cmul(br, ar)));
  }
  R[l * 2 * (long)n * n + r + (long)c * n] = v;
}

// Second kernel: PQ from M = [M1 | M2] and Q (N x N per layer, stride sPQ).
__global__ void pq_combine_kernel(int n, const cd* M, const cd* alpha, const cd* beta, long sab, const double* k0sq,
                                  const cd* Q, long sQ, cd* PQ, long sPQ) {
  const long l = blockIdx.z;
  const int N = 2 * n;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= N) return;
  const int rb = r / n, ri = r % n, cb = c / n, ci = c % n;
  const double k2 = k0sq[l];
  const cd m = M[l * 2 * (long)n * n + ri + (long)c * n];  // M1 for cb = 0, M2 for cb = 1
  const cd d = rb == 0 ? alpha[l * sab + ri] : beta[l * sab + ri];
  // k^2 Q21 / k^2 Q22 (top) and -k^2 Q11 / -k^2 Q12 (bottom): the Q block one row block
  // below (top) or above (bottom) in the same column
  const int qr = rb == 0 ? n + ri : ri;
  const cd q = Q[l * sQ + qr + (long)c * N];
  const cd add = rb == 0 ? cscale(q, k2) : cscale(q, -k2);
  PQ[l * sPQ + r + (long)c * N] = cadd(cmul(d, m), add);
}

// X_i = exp(i * gamma_i * s) with s = dz/k0 per batch entry (src/rcwa/RCWAScatterMatrix.cpp:36-40).
__global__ void phase_kernel(int n, const cd* gam, long sg, const double* s, cd* X, long sx) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const cd g = gam[b * sg + i];
  const double ss = s[b];
  // exp(i*g*ss) = exp(-Im(g)*ss) * (cos(Re(g)*ss) + i sin(Re(g)*ss))
  X[b * sx + i] = cexp(mk(-g.y * ss, g.x * ss));
}

void phases(hipStream_t s, int n, const cd* gam, long sg, const double* d_s, cd* X, long sx, int batch) {
  if (n <= 0 || batch <= 0) return;
  dim3 g(cdiv_up(n, 256), batch);
  phase_kernel<<<g, 256, 0, s>>>(n, gam, sg, d_s, X, sx);
  launch_check("phase");
}

// dst[b] = src[idx[b]] (n x n, ld n), used to give every slice its instance's reference basis.
__global__ void gather_mats_kernel(long nn, const cd* src, long ss, const int* idx, cd* dst, long sd) {
  const long b = blockIdx.y;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= nn) return;
  dst[b * sd + e] = src[(long)idx[b] * ss + e];
}

// Simpson combination of the varRCWA segment (src/core/RedhefferIntegrator.cpp:336-397):
// per entry (r,c): with ph, pf the half/full phases and coefficient c0 = i*h/(12 k0)
//   tuu = c0 (pf_r dA0 + 4 ph_r dA1 ph_c + dA2 pf_c) (+ pf on the diagonal)
//   rud = -c0 (pf_r dB0 pf_c + 4 ph_r dB1 ph_c + dB2)
//   rdu = -c0 (dB0 + 4 ph_r dB1 ph_c + pf_r dB2 pf_c)
//   tdd = c0 (dA0 pf_c + 4 ph_r dA1 ph_c + pf_r dA2) (+ pf on the diagonal)
// and the per-block max |perturbation| (before the diagonal is added) into err[4*blocks].
__global__ __launch_bounds__(256) void simpson_kernel(int n, const cd* dA0, const cd* dA1, const cd* dA2,
                                                      const cd* dB0, const cd* dB1, const cd* dB2, const cd* ph,
                                                      const cd* pf, cd c0, cd* tdd, cd* tuu, cd* rud, cd* rdu,
                                                      double* err) {
  __shared__ double red[4][256];
  const long total = (long)n * n;
  double m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256L) {
    const int r = (int)(e % n), c = (int)(e / n);
    const cd phr = ph[r], phc = ph[c], pfr = pf[r], pfc = pf[c];
    const cd h4 = cscale(cmul(phr, phc), 4.0);
    cd a = cadd(cadd(cmul(pfr, dA0[e]), cmul(h4, dA1[e])), cmul(dA2[e], pfc));
    a = cmul(c0, a);
    m0 = fmax(m0, cabsv(a));
    if (r == c) a = cadd(a, pfr);
    tuu[e] = a;
    cd b = cadd(cadd(cmul(cmul(pfr, dB0[e]), pfc), cmul(h4, dB1[e])), dB2[e]);
    b = cneg(cmul(c0, b));
    m1 = fmax(m1, cabsv(b));
    rud[e] = b;
    cd d = cadd(cadd(dB0[e], cmul(h4, dB1[e])), cmul(cmul(pfr, dB2[e]), pfc));
    d = cneg(cmul(c0, d));
    m2 = fmax(m2, cabsv(d));
    rdu[e] = d;
    cd f = cadd(cadd(cmul(dA0[e], pfc), cmul(h4, dA1[e])), cmul(pfr, dA2[e]));
    f = cmul(c0, f);
    m3 = fmax(m3, cabsv(f));
    if (r == c) f = cadd(f, pfr);
    tdd[e] = f;
  }
  const int t = threadIdx.x;
  red[0][t] = fmax(fmax(m0, m1), fmax(m2, m3));
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[0][t] = fmax(red[0][t], red[0][t + w]);
    __syncthreads();
  }
  if (t == 0) err[blockIdx.x] = red[0][0];
}

// ===================================================================== helpers
static const cd ONE = {1.0, 0.0}, MONE = {-1.0, 0.0}, ZERO = {0.0, 0.0};

void gather_mats(hipStream_t s, long nn, const cd* src, long ss, const int* d_idx, cd* dst, long sd, int batch) {
  if (batch <= 0) return;
  dim3 g(cdiv_up(nn, 256), batch);
  gather_mats_kernel<<<g, 256, 0, s>>>(nn, src, ss, d_idx, dst, sd);
  launch_check("gather_mats");
}

static void mm(hipStream_t s, int N, cd alpha, const cd* A, const cd* B, cd beta, cd* C, long stride, int batch) {
  zgemm_strided(s, OP_N, OP_N, N, N, N, alpha, A, N, stride, B, N, stride, beta, C, N, stride, batch);
}

static void copy(hipStream_t s, int m, int n, const cd* A, cd* C, long stride, int batch, cd a = ONE) {
  zgeam(s, m, n, a, A, m, stride, ZERO, nullptr, 0, 0, C, m, stride, batch);
}

// ===================================================================== (b) assembly
void assemble_pq(vrc_ctx& c, const AssemblyBatch& a, cd* P, cd* Q, cd* PQ) {
  StageScope st(ST_PQ);
  hipStream_t s = c.stream;
  const int n = a.nx * a.ny, N = 2 * n, L = a.nlayers;
  const long nn = (long)n * n, NN = (long)N * N;
  DevBuf fe = c.ws.get(sizeof(cd) * 3 * (size_t)nn * L);
  cd* fe11 = fe.as<cd>();
  cd* fe22 = fe11 + nn * L;
  cd* fe33 = fe22 + nn * L;
  assemble_fe(c, a, fe11, fe22, fe33, nn);
  DevBuf piv = c.ws.get(sizeof(int) * (size_t)n * L);
  DevBuf inf = c.ws.get(sizeof(int) * (size_t)L);
  fill_int(s, inf.as<int>(), 0, L);
  zgetrf(s, n, fe33, n, nn, piv.as<int>(), n, inf.as<int>(), L, c.ws);
  if (a.periodic) {
    DevBuf inv = c.ws.get(sizeof(cd) * (size_t)nn * L);
    zgetri(s, n, fe33, n, nn, piv.as<int>(), n, inv.as<cd>(), n, nn, L, c.ws);
    dim3 g(cdiv_up(N, 256), N, L);
    pq_periodic_kernel<<<g, 256, 0, s>>>(n, inv.as<cd>(), fe11, fe22, nn, a.d_alpha, a.d_beta, a.sab, a.d_k0sq, P,
                                         Q, NN);
    launch_check("pq_periodic");
    if (PQ) {
      DevBuf Rb = c.ws.get(sizeof(cd) * (size_t)nn * 2 * L), Mb = c.ws.get(sizeof(cd) * (size_t)nn * 2 * L);
      dim3 g1(cdiv_up(n, 256), 2 * n, L);
      pq_rhs_kernel<<<g1, 256, 0, s>>>(n, fe11, fe22, nn, a.d_alpha, a.d_beta, a.sab, a.d_k0sq, Rb.as<cd>());
      launch_check("pq_rhs");
      zgemm_strided(s, OP_N, OP_N, n, 2 * n, n, ONE, inv.as<cd>(), n, nn, Rb.as<cd>(), n, 2 * nn, ZERO, Mb.as<cd>(), n,
                    2 * nn, L);
      pq_combine_kernel<<<g, 256, 0, s>>>(n, Mb.as<cd>(), a.d_alpha, a.d_beta, a.sab, a.d_k0sq, Q, NN, PQ, NN);
      launch_check("pq_combine");
    }
  } else {
    // F = fe33^{-1} [Kx | Ky]  (n x 2n), then P blocks by ZGEMM against the dense K
    DevBuf Fb = c.ws.get(sizeof(cd) * (size_t)nn * 2 * L);
    cd* F = Fb.as<cd>();
    DevBuf Kb = c.ws.get(sizeof(cd) * (size_t)nn * 2);
    VRC_HIP(hipMemcpyAsync(Kb.as<cd>(), a.d_Kx, sizeof(cd) * nn, hipMemcpyDeviceToDevice, s));
    VRC_HIP(hipMemcpyAsync(Kb.as<cd>() + nn, a.d_Ky, sizeof(cd) * nn, hipMemcpyDeviceToDevice, s));
    zgetrs(s, n, 2 * n, fe33, n, nn, piv.as<int>(), n, Kb.as<cd>(), n, 0, F, n, 2 * nn, L, c.ws);
    const cd* F1h = F;
    const cd* F2h = F + nn;
    const long sF = 2 * nn;
    // P11 = Kx F2h ; P12 = -Kx F1h + k0^2 I ; P21 = Ky F2h - k0^2 I ; P22 = -Ky F1h
    zgemm_strided(s, OP_N, OP_N, n, n, n, ONE, a.d_Kx, n, 0, F2h, n, sF, ZERO, P, N, NN, L);
    zgemm_strided(s, OP_N, OP_N, n, n, n, MONE, a.d_Kx, n, 0, F1h, n, sF, ZERO, P + (long)n * N, N, NN, L);
    zgemm_strided(s, OP_N, OP_N, n, n, n, ONE, a.d_Ky, n, 0, F2h, n, sF, ZERO, P + n, N, NN, L);
    zgemm_strided(s, OP_N, OP_N, n, n, n, MONE, a.d_Ky, n, 0, F1h, n, sF, ZERO, P + n + (long)n * N, N, NN, L);
    const double k2 = a.k0sq_uniform;
    zadd_diag(s, n, mk(k2, 0.0), nullptr, 0, P + (long)n * N, N, NN, L);
    zadd_diag(s, n, mk(-k2, 0.0), nullptr, 0, P + n, N, NN, L);
    // Q11 = -KxKy ; Q12 = KxKx - k0^2 fe22 ; Q21 = -KyKy + k0^2 fe11 ; Q22 = KyKx
    zgeam(s, n, n, MONE, a.d_KxKy, n, 0, ZERO, nullptr, 0, 0, Q, N, NN, L);
    zgeam(s, n, n, ONE, a.d_KxKx, n, 0, mk(-k2, 0.0), fe22, n, nn, Q + (long)n * N, N, NN, L);
    zgeam(s, n, n, MONE, a.d_KyKy, n, 0, mk(k2, 0.0), fe11, n, nn, Q + n, N, NN, L);
    zgeam(s, n, n, ONE, a.d_KyKx, n, 0, ZERO, nullptr, 0, 0, Q + n + (long)n * N, N, NN, L);
    if (PQ) zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, P, N, NN, Q, N, NN, ZERO, PQ, N, NN, L);
  }
}

// ===================================================================== (c) wave equation
void wave_eqn(vrc_ctx& c, int N, const cd* P, const cd* Q, long sPQ, cd* W, cd* V, cd* gam, int* info, int batch) {
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  DevBuf PQb = c.ws.get(sizeof(cd) * (size_t)NN * batch);
  DevBuf lamb = c.ws.get(sizeof(cd) * (size_t)N * batch);
  {
    StageScope st(ST_EIGPOST);
    zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, P, N, sPQ, Q, N, sPQ, ZERO, PQb.as<cd>(), N, NN, batch);
  }
  zgeev_batched(s, N, PQb.as<cd>(), N, NN, lamb.as<cd>(), N, W, N, NN, info, batch, c.ws);
  StageScope st(ST_EIGPOST);
  gamma_sqrt(s, N, lamb.as<cd>(), N, gam, N, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, Q, N, sPQ, W, N, NN, ZERO, V, N, NN, batch);
  zdscale(s, N, N, nullptr, 0, gam, N, 1, V, N, NN, V, N, NN, batch);
}

// ===================================================================== (d) layer S-matrix
// Systems of the mirror-symmetric slab: per instance b and thread (row r, column c),
// with f = F0(r,c), g = G0(r,c) (G given as Gamma G0 when gam is set), x = X(r):
//   L[2b] = (1-x) f + (1+x) g,  R[2b] = (1+x) g - (1-x) f,
//   L[2b+1] = (1+x) f + (1-x) g,  R[2b+1] = (1-x) g - (1+x) f.
// R may alias F0/G when R[2b] is F0[b] and R[2b+1] is G[b] (each thread reads before it writes).
__global__ void evenodd_system_kernel(int N, const cd* F0, long sF, const cd* G, long sG, const cd* gam,
                                      const cd* x, cd* L, cd* R) {
  const long b = blockIdx.z;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= N) return;
  const long o = r + (long)c * N, NN = (long)N * N;
  const cd f = F0[b * sF + o];
  cd g = G[b * sG + o];
  if (gam) g = cdiv(g, gam[b * N + r]);
  const cd xr = x[b * N + r];
  const cd om = mk(1.0 - xr.x, -xr.y), op = mk(1.0 + xr.x, xr.y);
  const cd P = cmul(om, f), Q = cmul(op, g), Pp = cmul(op, f), Qp = cmul(om, g);
  L[(2 * b) * NN + o] = cadd(P, Q);
  L[(2 * b + 1) * NN + o] = cadd(Pp, Qp);
  R[(2 * b) * NN + o] = csub(Q, P);
  R[(2 * b + 1) * NN + o] = csub(Qp, Pp);
}

// Rud = (E[2b] + E[2b+1]) / 2, Tuu = (E[2b] - E[2b+1]) / 2.
__global__ void evenodd_combine_kernel(long NN, const cd* E, cd* Tuu, cd* Rud) {
  const long b = blockIdx.y;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < NN; i += (long)gridDim.x * blockDim.x) {
    const cd ep = E[(2 * b) * NN + i], em = E[(2 * b + 1) * NN + i];
    Rud[b * NN + i] = mk(0.5 * (ep.x + em.x), 0.5 * (ep.y + em.y));
    Tuu[b * NN + i] = mk(0.5 * (ep.x - em.x), 0.5 * (ep.y - em.y));
  }
}

void evenodd_system(hipStream_t s, int N, int batch, const cd* F0, long sF, const cd* G, long sG, const cd* gam,
                    const cd* x, cd* L, cd* R) {
  dim3 g(cdiv_up(N, 256), N, batch);
  evenodd_system_kernel<<<g, 256, 0, s>>>(N, F0, sF, G, sG, gam, x, L, R);
  launch_check("evenodd_system");
}

void evenodd_combine(hipStream_t s, int N, int batch, const cd* E, cd* Tuu, cd* Rud) {
  const long NN = (long)N * N;
  dim3 g(cdiv_up(NN, 256) < 1024 ? cdiv_up(NN, 256) : 1024, batch);
  evenodd_combine_kernel<<<g, 256, 0, s>>>(NN, E, Tuu, Rud);
  launch_check("evenodd_combine");
}

// Slab S-matrix from F0 = W^{-1} W0 and G0 = V^{-1} V0 (G = Gamma G0 when gam_scale is
// set). The reference (src/rcwa/RCWAScatterMatrix.cpp:26-85) computes, with A = F0 + G0,
// B = F0 - G0, X = diag(exp(i gamma s)) and D = A - X B A^{-1} X B,
//   Rud = D^{-1} (X B A^{-1} X A - B),  Tuu = D^{-1} X (A - B A^{-1} B).
// D factors both as (I + X B A^{-1})(A - X B) and as (I - X B A^{-1})(A + X B), which gives
//   Rud + Tuu = (A - X B)^{-1} (X A - B),  Rud - Tuu = -(A + X B)^{-1} (X A + B),
// and A -+ X B, X A -+ B are the row-scaled sums in evenodd_system_kernel. Two LU
// factorizations with N right-hand sides each (batched together) replace the
// reference's two factorizations with 2N right-hand sides and four products.
// R (2 N x N per instance) may be the storage of [F0 | G] (stride sF = 2 N^2, G = F0 + N^2).
static void slab_evenodd(vrc_ctx& c, int N, int batch, const double* d_s, const cd* gam, const cd* F0, long sF,
                         const cd* G, long sG, bool gam_scale, cd* R, cd* Tuu, cd* Rud) {
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  const size_t B2 = sizeof(cd) * (size_t)NN * 2 * batch;
  DevBuf xb = c.ws.get(sizeof(cd) * (size_t)N * batch);
  {
    dim3 g(cdiv_up(N, 256), batch);
    phase_kernel<<<g, 256, 0, s>>>(N, gam, N, d_s, xb.as<cd>(), N);
    launch_check("phase");
  }
  DevBuf Lb = c.ws.get(B2), Eb = c.ws.get(B2);
  DevBuf pivb = c.ws.get(sizeof(int) * (size_t)N * 2 * batch), infb = c.ws.get(sizeof(int) * (size_t)2 * batch);
  evenodd_system(s, N, batch, F0, sF, G, sG, gam_scale ? gam : nullptr, xb.as<cd>(), Lb.as<cd>(), R);
  fill_int(s, infb.as<int>(), 0, 2 * batch);
  zgetrf(s, N, Lb.as<cd>(), N, NN, pivb.as<int>(), N, infb.as<int>(), 2 * batch, c.ws);
  zgetrs(s, N, N, Lb.as<cd>(), N, NN, pivb.as<int>(), N, R, N, NN, Eb.as<cd>(), N, NN, 2 * batch, c.ws);
  evenodd_combine(s, N, batch, Eb.as<cd>(), Tuu, Rud);
}

void layer_smatrix(vrc_ctx& c, int N, int batch, const double* d_s, const cd* W, const cd* V, const cd* gam,
                   const cd* W0, const cd* V0, long sRef, cd* Tuu, cd* Rud) {
  StageScope st(ST_LAYER_S);
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  DevBuf LUb = c.ws.get(sizeof(cd) * (size_t)NN * batch);
  DevBuf pivb = c.ws.get(sizeof(int) * (size_t)N * batch);
  DevBuf infb = c.ws.get(sizeof(int) * (size_t)batch);
  DevBuf F0b = c.ws.get(sizeof(cd) * (size_t)NN * batch);
  DevBuf G0b = c.ws.get(sizeof(cd) * (size_t)NN * batch);
  cd* LU = LUb.as<cd>();
  int* piv = pivb.as<int>();
  cd* F0 = F0b.as<cd>();
  cd* G0 = G0b.as<cd>();
  fill_int(s, infb.as<int>(), 0, batch);
  // F0 = W^{-1} W0 ; G0 = V^{-1} V0
  copy(s, N, N, W, LU, NN, batch);
  zgetrf(s, N, LU, N, NN, piv, N, infb.as<int>(), batch, c.ws);
  zgetrs(s, N, N, LU, N, NN, piv, N, W0, N, sRef, F0, N, NN, batch, c.ws);
  copy(s, N, N, V, LU, NN, batch);
  zgetrf(s, N, LU, N, NN, piv, N, infb.as<int>(), batch, c.ws);
  zgetrs(s, N, N, LU, N, NN, piv, N, V0, N, sRef, G0, N, NN, batch, c.ws);
  DevBuf Rb = c.ws.get(sizeof(cd) * (size_t)NN * 2 * batch);
  slab_evenodd(c, N, batch, d_s, gam, F0, NN, G0, NN, false, Rb.as<cd>(), Tuu, Rud);
}

// C = diag(1/d) A (row scaling by the reciprocal of d), batched.
__global__ void rowscale_inv_kernel(int m, int n, const cd* d, long sd, const cd* A, long lda, long sA, cd* C,
                                    long ldc, long sC) {
  const long b = blockIdx.z;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int col = blockIdx.y;
  if (r >= m) return;
  C[b * sC + r + (long)col * ldc] = cdiv(A[b * sA + r + (long)col * lda], d[b * sd + r]);
}

static void rowscale_inv(hipStream_t s, int m, int n, const cd* d, long sd, const cd* A, long lda, long sA, cd* C,
                         long ldc, long sC, int batch) {
  dim3 g(cdiv_up(m, 256), n, batch);
  rowscale_inv_kernel<<<g, 256, 0, s>>>(m, n, d, sd, A, lda, sA, C, ldc, sC);
  launch_check("rowscale_inv");
}

// (c) as wave_eqn, but the eigenvectors stay factored: PQ = D Z T Z^H D^{-1},
// W = D Z X (unnormalized; every S-matrix is invariant under the column scaling).
void wave_eqn_factors(vrc_ctx& c, int N, const cd* P, const cd* Q, long sPQ, cd* Z, cd* X, cd* dbal, cd* gam,
                      int* info, int batch, cd* PQ) {
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  DevBuf PQb;
  DevBuf lamb = c.ws.get(sizeof(cd) * (size_t)N * batch);
  if (!PQ) {
    PQb = c.ws.get(sizeof(cd) * (size_t)NN * batch);
    PQ = PQb.as<cd>();
    StageScope st(ST_EIGPOST);
    zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, P, N, sPQ, Q, N, sPQ, ZERO, PQ, N, NN, batch);
  }
  zgeev_factors(s, N, PQ, N, NN, lamb.as<cd>(), N, Z, X, dbal, info, batch, c.ws);
  StageScope st(ST_EIGPOST);
  gamma_sqrt(s, N, lamb.as<cd>(), N, gam, N, batch);
}

// Normalized W (zgeev convention) and V = Q W diag(1/gamma) from the factors.
void wave_eqn_vectors(vrc_ctx& c, int N, const cd* Q, long sQ, const cd* Z, const cd* X, const cd* dbal,
                      const cd* gam, cd* W, cd* V, int batch) {
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  zgeev_vectors(s, N, Z, X, dbal, W, N, NN, batch);
  StageScope st(ST_EIGPOST);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, Q, N, sQ, W, N, NN, ZERO, V, N, NN, batch);
  zdscale(s, N, N, nullptr, 0, gam, N, 1, V, N, NN, V, N, NN, batch);
}

// Reference basis of a homogeneous medium (W0 = I, V0 coupling only harmonic pairs
// (i, n+i): V0[k, c] != 0 only for k == c mod n). out = D^{-1} P V0 with two terms per
// entry (d: the balancing diagonal of the slab).
__global__ void pv0_pair_kernel(int n, const cd* P, long sP, const cd* V0, long sV, const cd* d, cd* out, long ldo,
                                long so) {
  const long b = blockIdx.z;
  const int N = 2 * n;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= N) return;
  const int k1 = c % n, k2 = k1 + n;
  const cd* Pb = P + b * sP;
  const cd* Vb = V0 + b * sV;
  const cd v = cadd(cmul(Pb[r + (long)k1 * N], Vb[k1 + (long)c * N]), cmul(Pb[r + (long)k2 * N], Vb[k2 + (long)c * N]));
  out[b * so + r + (long)c * ldo] = cdiv(v, d[b * N + r]);
}

// Y[i, j] = conj(Z[j, i]) / d[j] (Z^H D^{-1}), 32 x 32 tiles through LDS.
__global__ __launch_bounds__(256) void ctrans_scale_kernel(int N, const cd* Z, long sZ, const cd* d, cd* Y, long ldy,
                                                           long sY) {
  const long b = blockIdx.z;
  __shared__ cd tile[32][33];
  const int i0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int row = j0 + tx, col = i0 + k;  // Z[row = j, col = i]
    if (row < N && col < N) tile[k][tx] = cdiv(cconj(Z[b * sZ + row + (long)col * N]), d[b * N + row]);
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int i = i0 + tx, j = j0 + k;
    if (i < N && j < N) Y[b * sY + i + (long)j * ldy] = tile[tx][k];
  }
}

// (d) Slab S-matrix from the eigen factors, without forming W or V:
//   W^{-1} = X^{-1} Z^H D^{-1}, and since PQ W = W Gamma^2 gives V = Q W / Gamma =
//   P^{-1} W Gamma, V^{-1} V0 = Gamma^{-1} W^{-1} (P V0). So
//   [F0 | Gamma G0] = X^{-1} Z^H D^{-1} [W0 | P V0]  (one triangular solve, no LU),
// then the even/odd split of slab_evenodd.
void layer_smatrix_factors(vrc_ctx& c, int N, int batch, const double* d_s, const cd* P, long sP, const cd* Z,
                           const cd* X, const cd* dbal, const cd* gam, const cd* W0, const cd* V0, long sRef,
                           cd* Tuu, cd* Rud, bool ref_homog) {
  StageScope st(ST_LAYER_S);
  hipStream_t s = c.stream;
  const long NN = (long)N * N, s2 = 2 * NN;
  DevBuf Yb = c.ws.get(sizeof(cd) * (size_t)NN * 2 * batch);
  cd* Y = Yb.as<cd>();
  if (ref_homog) {
    // W0 = I: Y[:, :N] = Z^H D^{-1}; V0 pair-sparse: Y[:, N:] = Z^H (D^{-1} P V0)
    DevBuf R2b = c.ws.get(sizeof(cd) * (size_t)NN * batch);
    {
      dim3 g(cdiv_up(N, 256), N, batch);
      pv0_pair_kernel<<<g, 256, 0, s>>>(N / 2, P, sP, V0, sRef, dbal, R2b.as<cd>(), N, NN);
      launch_check("pv0_pair");
    }
    {
      dim3 g(cdiv_up(N, 32), cdiv_up(N, 32), batch);
      ctrans_scale_kernel<<<g, 256, 0, s>>>(N, Z, NN, dbal, Y, N, s2);
      launch_check("ctrans_scale");
    }
    zgemm_strided(s, OP_C, OP_N, N, N, N, ONE, Z, N, NN, R2b.as<cd>(), N, NN, ZERO, Y + NN, N, s2, batch);
    ztrsm_upper(s, N, 2 * N, X, N, NN, Y, N, s2, batch, c.ws);
  } else {
    DevBuf R2b = c.ws.get(sizeof(cd) * (size_t)NN * 2 * batch);
    cd* R2 = R2b.as<cd>();
    // R2 = D^{-1} [W0 | P V0]
    zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, P, N, sP, V0, N, sRef, ZERO, Y + NN, N, s2, batch);
    zgeam(s, N, N, ONE, W0, N, sRef, ZERO, nullptr, 0, 0, Y, N, s2, batch);
    rowscale_inv(s, N, 2 * N, dbal, N, Y, N, s2, R2, N, s2, batch);
    // Y = Z^H R2 ; Y <- X^{-1} Y
    zgemm_strided(s, OP_C, OP_N, N, 2 * N, N, ONE, Z, N, NN, R2, N, s2, ZERO, Y, N, s2, batch);
    ztrsm_upper(s, N, 2 * N, X, N, NN, Y, N, s2, batch, c.ws);
  }
  // F0 = Y[:, :N], Gamma G0 = Y[:, N:]; the systems' right-hand sides overwrite Y
  slab_evenodd(c, N, batch, d_s, gam, Y, s2, Y + NN, s2, true, Y, Tuu, Rud);
}

// Closed-form eigen-system of a homogeneous periodic medium (one permittivity value):
// P and Q only couple the harmonic pair (i, n+i), and on it PQ = lambda_i I_2 with
// lambda_i = k0^2 (k0^2 eps - alpha_i^2 - beta_i^2) (exactly, from the block formulas of
// src/rcwa/Layer.cpp:95-100,134-138 with fe = eps I). So W = I, gamma_i from lambda_i
// and V = Q diag(1/gamma). Per batch entry b: alpha/beta at stride sab, k0sq[b], eps[b].
__global__ void homogeneous_eig_kernel(int n, const cd* alpha, const cd* beta, long sab, const double* k0sq,
                                       const cd* eps, cd* W, cd* V, cd* gam, long sM) {
  const long b = blockIdx.z;
  const int N = 2 * n;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= N) return;
  const int ic = c % n;
  const cd a = alpha[b * sab + ic], be = beta[b * sab + ic];
  const double k2 = k0sq[b];
  const cd e = eps[b];
  const cd lam = cscale(csub(cscale(e, k2), cadd(cmul(a, a), cmul(be, be))), k2);
  const cd g = gamma_branch(lam);
  cd q = mk(0.0, 0.0);
  if (r % n == ic) {
    const int rb = r / n, cb = c / n;
    if (rb == 0 && cb == 0) q = cneg(cmul(a, be));
    else if (rb == 0) q = csub(cmul(a, a), cscale(e, k2));
    else if (cb == 0) q = csub(cscale(e, k2), cmul(be, be));
    else q = cmul(be, a);
  }
  const long o = b * sM + r + (long)c * N;
  W[o] = (r == c) ? mk(1.0, 0.0) : mk(0.0, 0.0);
  V[o] = cdiv(q, g);
  if (r == 0) gam[b * N + c] = g;
}

void homogeneous_eig(hipStream_t s, int n, int batch, const cd* alpha, const cd* beta, long sab, const double* k0sq,
                     const cd* eps, cd* W, cd* V, cd* gam) {
  if (batch <= 0) return;
  const int N = 2 * n;
  dim3 g(cdiv_up(N, 256), N, batch);
  homogeneous_eig_kernel<<<g, 256, 0, s>>>(n, alpha, beta, sab, k0sq, eps, W, V, gam, (long)N * N);
  launch_check("homogeneous_eig");
  health_note_gamma(s, gam, (long)N * batch);
}

// ===================================================================== (d) star product
// The factored star product with M^{-1} formed explicitly, keeping the intermediates the
// reverse pass needs (save: STAR_SAVE N x N per instance, in the order of vrc_ops.hpp).
static void star_product_saved(vrc_ctx& c, int N, int batch, SMat& S, const cd* tdd, const cd* tuu,
                               const cd* rud, const cd* rdu, long ss, cd* save) {
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  const long sv = (long)batch * NN;
  cd *Mi = save, *Yt = save + sv, *Yr = save + 2 * sv, *K = save + 3 * sv, *T1 = save + 4 * sv,
     *Z2 = save + 5 * sv;
  DevBuf Mb = c.ws.get(sizeof(cd) * (size_t)sv);
  DevBuf pivb = c.ws.get(sizeof(int) * (size_t)N * batch), infb = c.ws.get(sizeof(int) * (size_t)batch);
  cd* M = Mb.as<cd>();
  fill_int(s, infb.as<int>(), 0, batch);
  // M = I - Rud rdu ; Mi = M^{-1}
  zset(s, N, N, ONE, ZERO, M, N, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, MONE, S.Rud, N, NN, rdu, N, ss, ONE, M, N, NN, batch);
  zgetrf(s, N, M, N, NN, pivb.as<int>(), N, infb.as<int>(), batch, c.ws);
  zgetri(s, N, M, N, NN, pivb.as<int>(), N, Mi, N, NN, batch, c.ws);
  // Yt = Mi Tuu ; Yr = Mi Rud ; K = Yr tdd
  mm(s, N, ONE, Mi, S.Tuu, ZERO, Yt, NN, batch);
  mm(s, N, ONE, Mi, S.Rud, ZERO, Yr, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, Yr, N, NN, tdd, N, ss, ZERO, K, N, NN, batch);
  // Rud' = rud + tuu K ; T1 = tdd + rdu K ; Z2 = rdu Yt
  zgeam(s, N, N, ONE, rud, N, ss, ZERO, nullptr, 0, 0, S.Rud, N, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, tuu, N, ss, K, N, NN, ONE, S.Rud, N, NN, batch);
  zgeam(s, N, N, ONE, tdd, N, ss, ZERO, nullptr, 0, 0, T1, N, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, rdu, N, ss, K, N, NN, ONE, T1, N, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, rdu, N, ss, Yt, N, NN, ZERO, Z2, N, NN, batch);
  // Rdu' = Rdu + Tdd Z2 ; Tdd' = Tdd T1 ; Tuu' = tuu Yt
  mm(s, N, ONE, S.Tdd, Z2, ONE, S.Rdu, NN, batch);
  mm(s, N, ONE, S.Tdd, T1, ZERO, M, NN, batch);
  copy(s, N, N, M, S.Tdd, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, tuu, N, ss, Yt, N, NN, ZERO, S.Tuu, N, NN, batch);
}

// Redheffer star product S <- S * s (src/core/RCWAIntegrator.cpp:57-72):
//   C1 = (I - rdu Rud)^{-1}, C2 = (I - Rud rdu)^{-1},
//   Rdu' = Rdu + Tdd rdu C2 Tuu, Tdd' = Tdd C1 tdd, Rud' = rud + tuu Rud C1 tdd, Tuu' = tuu C2 Tuu.
// With Rud C1 = C2 Rud and C1 = I + rdu C2 Rud, everything follows from one factorization
// of M = I - Rud rdu and the left solves [Yt | Yr] = M^{-1} [Tuu | Rud]:
//   K = Yr tdd, Rud' = rud + tuu K, Tdd' = Tdd (tdd + rdu K), Rdu' = Rdu + Tdd (rdu Yt),
//   Tuu' = tuu Yt
// (8 products, one LU, 2N right-hand sides; the reference form needs 12 products, one LU
// and an inverse).
void star_product(vrc_ctx& c, int N, int batch, SMat& S, const cd* tdd, const cd* tuu, const cd* rud,
                  const cd* rdu, long ss, cd* save) {
  StageScope st(ST_STAR);
  if (save) {
    star_product_saved(c, N, batch, S, tdd, tuu, rud, rdu, ss, save);
    return;
  }
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  const size_t B = sizeof(cd) * (size_t)NN * batch;
  DevBuf Mb = c.ws.get(B), Yb = c.ws.get(2 * B), Kb = c.ws.get(B), T1b = c.ws.get(B);
  DevBuf pivb = c.ws.get(sizeof(int) * (size_t)N * batch), infb = c.ws.get(sizeof(int) * (size_t)batch);
  cd *M = Mb.as<cd>(), *Y = Yb.as<cd>(), *K = Kb.as<cd>(), *T1 = T1b.as<cd>();
  cd* Yt = Y;
  cd* Yr = Y + NN;
  const long s2 = 2 * NN;
  int* piv = pivb.as<int>();
  fill_int(s, infb.as<int>(), 0, batch);
  // M = I - Rud rdu
  zset(s, N, N, ONE, ZERO, M, N, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, MONE, S.Rud, N, NN, rdu, N, ss, ONE, M, N, NN, batch);
  zgetrf(s, N, M, N, NN, piv, N, infb.as<int>(), batch, c.ws);
  zgetrs_pair(s, N, N, M, N, NN, piv, N, S.Tuu, N, NN, S.Rud, N, NN, Y, N, s2, batch, c.ws);
  // K = Yr tdd ; Rud' = rud + tuu K
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, Yr, N, s2, tdd, N, ss, ZERO, K, N, NN, batch);
  zgeam(s, N, N, ONE, rud, N, ss, ZERO, nullptr, 0, 0, S.Rud, N, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, tuu, N, ss, K, N, NN, ONE, S.Rud, N, NN, batch);
  // T1 = tdd + rdu K ; Yr <- rdu Yt (Yr consumed)
  zgeam(s, N, N, ONE, tdd, N, ss, ZERO, nullptr, 0, 0, T1, N, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, rdu, N, ss, K, N, NN, ONE, T1, N, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, rdu, N, ss, Yt, N, s2, ZERO, Yr, N, s2, batch);
  // Rdu' = Rdu + Tdd (rdu Yt) ; Tdd' = Tdd T1 ; Tuu' = tuu Yt
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, S.Tdd, N, NN, Yr, N, s2, ONE, S.Rdu, N, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, S.Tdd, N, NN, T1, N, NN, ZERO, K, N, NN, batch);
  copy(s, N, N, K, S.Tdd, NN, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, tuu, N, ss, Yt, N, s2, ZERO, S.Tuu, N, NN, batch);
}

// Peak workspace bytes of star_product (without the saved intermediates): its own five
// N x N blocks, pivots and info, plus the larger of the LU's and the solve's peaks.
size_t star_product_scratch_bytes(int N, int batch) {
  auto r = [](size_t b) { return b == 0 ? (size_t)256 : (b + 255) & ~(size_t)255; };
  const size_t B = sizeof(cd) * (size_t)N * N * batch;
  const size_t own = r(B) + r(2 * B) + r(B) + r(B) + r(sizeof(int) * (size_t)N * batch) + r(sizeof(int) * (size_t)batch);
  const size_t f = zgetrf_scratch_bytes(N, batch), g = zgetrs_pair_scratch_bytes(N, N, batch);
  return own + (f > g ? f : g);
}

// ===================================================================== plane-wave efficiencies
// Per instance b and harmonic i (RCWASolver::scatterPlaneWave, src/rcwa/RCWASolver.cpp:1671-1732):
// Kz = sqrt(k0^2 eps - kx^2 - ky^2), Ez = (kx Ex + ky Ey) / (-/+ Kz) for the reflected /
// transmitted field, efficiency = |Re Kz| / (sqrt(Re eps) k0) (|Ex|^2 + |Ey|^2 + |Ez|^2).
// A component with Kz == 0 (a plane wave moving transversely, which the reference rejects)
// makes the instance's efficiencies NaN.
__global__ void efficiency_kernel(int n, const cd* rxy, const cd* txy, const cd* ab, const double* k0v,
                                  const cd* eps2, double* REF, double* TRN) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double k0 = k0v[b];
  const cd kx = ab[b * 2 * n + i], ky = ab[b * 2 * n + n + i];
  const cd t2 = cadd(cmul(kx, kx), cmul(ky, ky));
  for (int side = 0; side < 2; ++side) {
    const cd eps = eps2[2 * b + side];
    const cd* f = (side == 0 ? rxy : txy) + b * 2 * n;
    const cd kz = csqrt_principal(csub(cscale(eps, k0 * k0), t2));
    double out;
    if (kz.x == 0.0 && kz.y == 0.0) {
      out = nan("");
    } else {
      const cd ex = f[i], ey = f[n + i];
      const cd ez = cdiv(cadd(cmul(kx, ex), cmul(ky, ey)), kz);  // the sign of Kz drops out of |Ez|
      out = fabs(kz.x) / (sqrt(eps.x) * k0) * (cnorm2(ex) + cnorm2(ey) + cnorm2(ez));
    }
    (side == 0 ? REF : TRN)[b * n + i] = out;
  }
}

void efficiencies(hipStream_t s, int n, int batch, const cd* rxy, const cd* txy, const cd* ab, const double* k0v,
                  const cd* eps2, double* REF, double* TRN) {
  dim3 g(cdiv_up(n, 256), batch);
  efficiency_kernel<<<g, 256, 0, s>>>(n, rxy, txy, ab, k0v, eps2, REF, TRN);
  launch_check("efficiency");
}

// ===================================================================== modal power through a window
// The sampled flux sum of evaluateEnergy (src/tests/test_cpu.cpp:335-412),
//   sum_{p,q} E(x_p, y_q) conj(H(x_p, y_q)),  E = sum_{i,j} K(i,j) a_i(x) b_j(y),
// with a_i(x) = exp(-i 2 pi m_i x / Lx) at midpoints x_p, factors exactly through the window
// Gram matrices Gx(i,i') = sum_p a_i(x_p) conj(a_i'(x_p)) and Gy likewise:
//   sum = sum_{i,j} K(i,j) (Gx conj(L) Gy^T)(i,j).
// G[i + i' nh] = sum_p exp(sgn i 2 pi (m_i - m_i') x_p / L), x_p = x0 + step (p + 1/2);
// sgn = -1 gives the Gram matrix, +1 its transpose.
constexpr double PI_D = 3.1415926535897932;  // src/rcwa/defns.h:29
__global__ void window_gram_kernel(int nh, int nres, double x0, double step, double L, double sgn, cd* G) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nh * nh) return;
  const int i = t % nh, ip = t / nh;
  const double w = sgn * 2.0 * PI_D * (double)(i - ip) / L;
  double re = 0.0, im = 0.0;
  for (int p = 0; p < nres; ++p) {
    double sn, cs;
    sincos(w * (x0 + step * (p + 0.5)), &sn, &cs);
    re += cs;
    im += sn;
  }
  G[t] = mk(re, im);
}

// energy[b] = 0.5 dx dy Re sum_r (Ex_b[r] conj(Z_b[n + r]) - Ey_b[r] conj(Z_b[r])), where
// Z_b = [Gy^T Hx_b Gx; Gy^T Hy_b Gx] in the (ny x nx) layout of a mode column.
__global__ void __launch_bounds__(256) modal_power_kernel(int n, const cd* W, long ldw, const cd* Z, double dxdy,
                                                          double* energy) {
  __shared__ double red[256];
  const long b = blockIdx.x;
  const cd* e = W + b * ldw;
  const cd* z = Z + b * 2 * n;
  double acc = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const cd ex = e[r], ey = e[n + r], zx = z[r], zy = z[n + r];
    acc += (ex.x * zy.x + ex.y * zy.y) - (ey.x * zx.x + ey.y * zx.y);  // Re(ex conj(zy)) - Re(ey conj(zx))
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) energy[b] = 0.5 * dxdy * red[0];
}

void window_gram(hipStream_t s, int nh, int nres, double x0, double x1, double L, double sgn, cd* G) {
  window_gram_kernel<<<cdiv_up((long)nh * nh, 256), 256, 0, s>>>(nh, nres, x0, (x1 - x0) / nres, L, sgn, G);
  launch_check("window_gram");
}

void modal_power_reduce(hipStream_t s, int n, int nmodes, const cd* W, long ldw, const cd* Z, double dxdy,
                        double* energy) {
  modal_power_kernel<<<nmodes, 256, 0, s>>>(n, W, ldw, Z, dxdy, energy);
  launch_check("modal_power");
}

// ===================================================================== (d) projections
// Homogeneous end media (W0 = W1 = I, V0 and V1 coupling only the harmonic pair (i, i+n)):
// every matrix of the projection built from the end bases is "pair-sparse", a 2x2 block
// per harmonic i on the indices (i, i+n), stored compactly as 4 values per harmonic
// (s00, s01, s10, s11). ends_pair_kernel: G_i = V1_i^{-1} V0_i, A_i = (I + G_i)/2,
// B_i = (I - G_i)/2 (the A, B of project_left with F = W1^{-1} W0 = I).
__global__ void ends_pair_kernel(int n, const cd* V0, const cd* V1, long sV, cd* Ap, cd* Bp) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int N = 2 * n;
  const cd* v0 = V0 + b * sV;
  const cd* v1 = V1 + b * sV;
  auto at = [&](const cd* M, int r, int c) { return M[(r ? i + n : i) + (long)(c ? i + n : i) * N]; };
  const cd a = at(v1, 0, 0), bb = at(v1, 0, 1), cc = at(v1, 1, 0), d = at(v1, 1, 1);
  const cd det = csub(cmul(a, d), cmul(bb, cc));
  const cd idet = crecip(det);
  // V1_i^{-1} = [[d, -b], [-c, a]] / det
  const cd i00 = cmul(d, idet), i01 = cneg(cmul(bb, idet)), i10 = cneg(cmul(cc, idet)), i11 = cmul(a, idet);
  const cd p00 = at(v0, 0, 0), p01 = at(v0, 0, 1), p10 = at(v0, 1, 0), p11 = at(v0, 1, 1);
  const cd g00 = cadd(cmul(i00, p00), cmul(i01, p10)), g01 = cadd(cmul(i00, p01), cmul(i01, p11));
  const cd g10 = cadd(cmul(i10, p00), cmul(i11, p10)), g11 = cadd(cmul(i10, p01), cmul(i11, p11));
  cd* A = Ap + (b * n + i) * 4;
  cd* B = Bp + (b * n + i) * 4;
  A[0] = mk(0.5 * (1.0 + g00.x), 0.5 * g00.y);
  A[1] = cscale(g01, 0.5);
  A[2] = cscale(g10, 0.5);
  A[3] = mk(0.5 * (1.0 + g11.x), 0.5 * g11.y);
  B[0] = mk(0.5 * (1.0 - g00.x), -0.5 * g00.y);
  B[1] = cscale(g01, -0.5);
  B[2] = cscale(g10, -0.5);
  B[3] = mk(0.5 * (1.0 - g11.x), -0.5 * g11.y);
}

// out = s1 * (X P) + s2 * Y (X dense N x N, P, Y pair-sparse), or with left = true
// out = s1 * (P X) + s2 * Y.
__global__ void pair_mul_kernel(int n, bool left, const cd* X, long sX, const cd* P, cd s1, const cd* Y, cd s2,
                                cd* out, long so) {
  const long b = blockIdx.z;
  const int N = 2 * n;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= N) return;
  const cd* x = X + b * sX;
  cd v;
  if (!left) {
    const int i = c % n, cb = c / n;
    const cd* p = P + (b * n + i) * 4;
    v = cadd(cmul(x[r + (long)i * N], p[cb]), cmul(x[r + (long)(i + n) * N], p[2 + cb]));
  } else {
    const int i = r % n, rb = r / n;
    const cd* p = P + (b * n + i) * 4;
    v = cadd(cmul(p[2 * rb], x[i + (long)c * N]), cmul(p[2 * rb + 1], x[(i + n) + (long)c * N]));
  }
  v = cmul(s1, v);
  if (Y && (r % n) == (c % n)) v = cadd(v, cmul(s2, Y[(b * n + r % n) * 4 + 2 * (r / n) + c / n]));
  out[b * so + r + (long)c * N] = v;
}

static void pair_mul(hipStream_t s, int n, int batch, bool left, const cd* X, long sX, const cd* P, cd s1,
                     const cd* Y, cd s2, cd* out, long so) {
  const int N = 2 * n;
  dim3 g(cdiv_up(N, 256), N, batch);
  pair_mul_kernel<<<g, 256, 0, s>>>(n, left, X, sX, P, s1, Y, s2, out, so);
  launch_check("pair_mul");
}

// project_left for homogeneous end media (same formulas, pair-sparse A and B).
void project_left_homog(vrc_ctx& c, int N, int batch, SMat& S, const cd* V0, const cd* V1, long sRef) {
  StageScope st(ST_PROJECT);
  hipStream_t s = c.stream;
  const int n = N / 2;
  const long NN = (long)N * N, s2 = 2 * NN;
  const size_t B = sizeof(cd) * (size_t)NN * batch;
  DevBuf Apb = c.ws.get(sizeof(cd) * (size_t)4 * n * batch), Bpb = c.ws.get(sizeof(cd) * (size_t)4 * n * batch);
  DevBuf Kb = c.ws.get(B), R2b = c.ws.get(2 * B), Tb = c.ws.get(B);
  DevBuf pivb = c.ws.get(sizeof(int) * (size_t)N * batch), infb = c.ws.get(sizeof(int) * (size_t)batch);
  cd *Ap = Apb.as<cd>(), *Bp = Bpb.as<cd>(), *K = Kb.as<cd>(), *R2 = R2b.as<cd>(), *T = Tb.as<cd>();
  {
    dim3 g(cdiv_up(n, 256), batch);
    ends_pair_kernel<<<g, 256, 0, s>>>(n, V0, V1, sRef, Ap, Bp);
    launch_check("ends_pair");
  }
  fill_int(s, infb.as<int>(), 0, batch);
  // K = A - Rdu B ; RHS = [Rdu A - B | Tdd]
  pair_mul(s, n, batch, false, S.Rdu, NN, Bp, MONE, Ap, ONE, K, NN);
  pair_mul(s, n, batch, false, S.Rdu, NN, Ap, ONE, Bp, MONE, R2, s2);
  zgeam(s, N, N, ONE, S.Tdd, N, NN, ZERO, nullptr, 0, 0, R2 + NN, N, s2, batch);
  zgetrf(s, N, K, N, NN, pivb.as<int>(), N, infb.as<int>(), batch, c.ws);
  zgetrs(s, N, 2 * N, K, N, NN, pivb.as<int>(), N, R2, N, s2, R2, N, s2, batch, c.ws);
  zgeam(s, N, N, ONE, R2, N, s2, ZERO, nullptr, 0, 0, S.Rdu, N, NN, batch);
  zgeam(s, N, N, ONE, R2 + NN, N, s2, ZERO, nullptr, 0, 0, S.Tdd, N, NN, batch);
  // Rud = (Tuu B) Tdd + Rud
  pair_mul(s, n, batch, false, S.Tuu, NN, Bp, ONE, nullptr, ZERO, T, NN);
  mm(s, N, ONE, T, S.Tdd, ONE, S.Rud, NN, batch);
  // Tuu = Tuu (B Rdu + A)
  pair_mul(s, n, batch, true, S.Rdu, NN, Bp, ONE, Ap, ONE, K, NN);
  mm(s, N, ONE, S.Tuu, K, ZERO, T, NN, batch);
  copy(s, N, N, T, S.Tuu, NN, batch);
}


// Left end (src/core/RCWAIntegrator.cpp:73-86): S expressed with basis (W1,V1) on the left
// is re-expressed with (W0,V0). LU1 holds the factorizations of W1 and V1.
void project_left(vrc_ctx& c, int N, int batch, SMat& S, const LUPair& LU1, const cd* W0, const cd* V0,
                  long sRef) {
  StageScope st(ST_PROJECT);
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  const size_t B = sizeof(cd) * (size_t)NN * batch;
  DevBuf Fb = c.ws.get(B), Gb = c.ws.get(B), Kb = c.ws.get(B), R2b = c.ws.get(2 * B), Tb = c.ws.get(B);
  DevBuf pivb = c.ws.get(sizeof(int) * (size_t)N * batch), infb = c.ws.get(sizeof(int) * (size_t)batch);
  cd *F = Fb.as<cd>(), *G = Gb.as<cd>(), *K = Kb.as<cd>(), *R2 = R2b.as<cd>(), *T = Tb.as<cd>();
  const long s2 = 2 * NN;
  fill_int(s, infb.as<int>(), 0, batch);
  zgetrs(s, N, N, LU1.W, N, LU1.sM, LU1.pW, N, W0, N, sRef, F, N, NN, batch, c.ws);
  zgetrs(s, N, N, LU1.V, N, LU1.sM, LU1.pV, N, V0, N, sRef, G, N, NN, batch, c.ws);
  // A = (F + G)/2 -> F ; B = (F - G)/2 -> G   (in place, element-wise)
  zgeam(s, N, N, mk(0.5, 0), F, N, NN, mk(0.5, 0), G, N, NN, K, N, NN, batch);   // K = A
  zgeam(s, N, N, mk(0.5, 0), F, N, NN, mk(-0.5, 0), G, N, NN, G, N, NN, batch);  // G = B
  copy(s, N, N, K, F, NN, batch);                                                // F = A
  // K = A - Rdu B
  mm(s, N, MONE, S.Rdu, G, ONE, K, NN, batch);
  // RHS = [Rdu A - B | Tdd]
  zgeam(s, N, N, MONE, G, N, NN, ZERO, nullptr, 0, 0, R2, N, s2, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, S.Rdu, N, NN, F, N, NN, ONE, R2, N, s2, batch);
  zgeam(s, N, N, ONE, S.Tdd, N, NN, ZERO, nullptr, 0, 0, R2 + NN, N, s2, batch);
  zgetrf(s, N, K, N, NN, pivb.as<int>(), N, infb.as<int>(), batch, c.ws);
  zgetrs(s, N, 2 * N, K, N, NN, pivb.as<int>(), N, R2, N, s2, R2, N, s2, batch, c.ws);
  zgeam(s, N, N, ONE, R2, N, s2, ZERO, nullptr, 0, 0, S.Rdu, N, NN, batch);
  zgeam(s, N, N, ONE, R2 + NN, N, s2, ZERO, nullptr, 0, 0, S.Tdd, N, NN, batch);
  // Rud = Tuu B Tdd + Rud
  mm(s, N, ONE, S.Tuu, G, ZERO, T, NN, batch);
  mm(s, N, ONE, T, S.Tdd, ONE, S.Rud, NN, batch);
  // Tuu = Tuu (B Rdu + A)
  mm(s, N, ONE, G, S.Rdu, ONE, F, NN, batch);
  mm(s, N, ONE, S.Tuu, F, ZERO, T, NN, batch);
  copy(s, N, N, T, S.Tuu, NN, batch);
}

// Right end (src/core/RedhefferIntegrator.cpp:211-223): S whose right basis is
// (Wl, Vl) (factorized in LUl) re-expressed with (W1, V1).
void project_right(vrc_ctx& c, int N, int batch, SMat& S, const LUPair& LUl, const cd* W1, const cd* V1,
                   long sRef) {
  StageScope st(ST_PROJECT);
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  const size_t B = sizeof(cd) * (size_t)NN * batch;
  DevBuf Fb = c.ws.get(B), Gb = c.ws.get(B), Kb = c.ws.get(B), R2b = c.ws.get(2 * B), Tb = c.ws.get(B);
  DevBuf pivb = c.ws.get(sizeof(int) * (size_t)N * batch), infb = c.ws.get(sizeof(int) * (size_t)batch);
  cd *F = Fb.as<cd>(), *G = Gb.as<cd>(), *K = Kb.as<cd>(), *R2 = R2b.as<cd>(), *T = Tb.as<cd>();
  const long s2 = 2 * NN;
  fill_int(s, infb.as<int>(), 0, batch);
  zgetrs(s, N, N, LUl.W, N, LUl.sM, LUl.pW, N, W1, N, sRef, F, N, NN, batch, c.ws);
  zgetrs(s, N, N, LUl.V, N, LUl.sM, LUl.pV, N, V1, N, sRef, G, N, NN, batch, c.ws);
  zgeam(s, N, N, mk(0.5, 0), F, N, NN, mk(0.5, 0), G, N, NN, K, N, NN, batch);   // A
  zgeam(s, N, N, mk(0.5, 0), F, N, NN, mk(-0.5, 0), G, N, NN, G, N, NN, batch);  // B
  copy(s, N, N, K, F, NN, batch);                                                // F = A
  // K = A - Rud B
  mm(s, N, MONE, S.Rud, G, ONE, K, NN, batch);
  // [Tuu | Rud A - B]
  zgeam(s, N, N, ONE, S.Tuu, N, NN, ZERO, nullptr, 0, 0, R2, N, s2, batch);
  zgeam(s, N, N, MONE, G, N, NN, ZERO, nullptr, 0, 0, R2 + NN, N, s2, batch);
  zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, S.Rud, N, NN, F, N, NN, ONE, R2 + NN, N, s2, batch);
  zgetrf(s, N, K, N, NN, pivb.as<int>(), N, infb.as<int>(), batch, c.ws);
  zgetrs(s, N, 2 * N, K, N, NN, pivb.as<int>(), N, R2, N, s2, R2, N, s2, batch, c.ws);
  zgeam(s, N, N, ONE, R2, N, s2, ZERO, nullptr, 0, 0, S.Tuu, N, NN, batch);
  zgeam(s, N, N, ONE, R2 + NN, N, s2, ZERO, nullptr, 0, 0, S.Rud, N, NN, batch);
  // Rdu = Rdu + Tdd B Tuu
  mm(s, N, ONE, S.Tdd, G, ZERO, T, NN, batch);
  mm(s, N, ONE, T, S.Tuu, ONE, S.Rdu, NN, batch);
  // Tdd = Tdd (B Rud + A)
  mm(s, N, ONE, G, S.Rud, ONE, F, NN, batch);
  mm(s, N, ONE, S.Tdd, F, ZERO, T, NN, batch);
  copy(s, N, N, T, S.Tdd, NN, batch);
}

// ===================================================================== (d) varRCWA segment
// Perturbation of one segment [z0, z1] around the reference (Pr, Qr, W, V, gamma):
// produces (tdd, tuu, rud, rdu) and the max-modulus error estimate (returned on host).
// dst[b] = the nbytes at src[b] (one pointer per batch entry; nbytes a multiple of 8).
__global__ void gather_ptr_kernel(long nwords, const double* const* src, double* dst) {
  const long b = blockIdx.y;
  const double* a = src[b];
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nwords; e += (long)gridDim.x * blockDim.x)
    dst[b * nwords + e] = a[e];
}

static void gather_ptrs(vrc_ctx& c, const std::vector<const void*>& src, size_t nbytes, void* dst) {
  const int B = (int)src.size();
  if (B == 0) return;
  DevBuf p = c.ws.get(sizeof(void*) * B);
  c.pinned.upload(c.stream, p.ptr, src.data(), sizeof(void*) * B);
  const long nw = (long)(nbytes / 8);
  const long blocks = std::min<long>((nw + 255) / 256, 1024);
  gather_ptr_kernel<<<dim3((unsigned)blocks, B), 256, 0, c.stream>>>(nw, (const double* const*)p.ptr, (double*)dst);
  launch_check("gather_ptr");
}

// simpson_kernel over a batch of segments (blockIdx.y): inputs at stride NN (phases at
// stride n), outputs T[b] = [tdd | tuu | rud | rdu] at stride 4 NN, c0[b] per segment,
// RH_ERR_BLOCKS maxima per segment.
__global__ __launch_bounds__(256) void simpson_batch_kernel(int n, const cd* dA, const cd* dB, const cd* ph,
                                                            const cd* pf, const cd* c0v, cd* T, double* err) {
  const long b = blockIdx.y;
  const long NN = (long)n * n;
  const cd* dA0 = dA + b * 3 * NN;
  const cd* dA1 = dA0 + NN;
  const cd* dA2 = dA1 + NN;
  const cd* dB0 = dB + b * 3 * NN;
  const cd* dB1 = dB0 + NN;
  const cd* dB2 = dB1 + NN;
  const cd* phb = ph + b * n;
  const cd* pfb = pf + b * n;
  const cd c0 = c0v[b];
  cd* tdd = T + b * 4 * NN;
  cd* tuu = tdd + NN;
  cd* rud = tuu + NN;
  cd* rdu = rud + NN;
  __shared__ double red[256];
  double m = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < NN; e += (long)gridDim.x * 256L) {
    const int r = (int)(e % n), cc = (int)(e / n);
    const cd phr = phb[r], phc = phb[cc], pfr = pfb[r], pfc = pfb[cc];
    const cd h4 = cscale(cmul(phr, phc), 4.0);
    cd a = cmul(c0, cadd(cadd(cmul(pfr, dA0[e]), cmul(h4, dA1[e])), cmul(dA2[e], pfc)));
    m = fmax(m, cabsv(a));
    if (r == cc) a = cadd(a, pfr);
    tuu[e] = a;
    cd q = cneg(cmul(c0, cadd(cadd(cmul(cmul(pfr, dB0[e]), pfc), cmul(h4, dB1[e])), dB2[e])));
    m = fmax(m, cabsv(q));
    rud[e] = q;
    cd d = cneg(cmul(c0, cadd(cadd(dB0[e], cmul(h4, dB1[e])), cmul(cmul(pfr, dB2[e]), pfc))));
    m = fmax(m, cabsv(d));
    rdu[e] = d;
    cd f = cmul(c0, cadd(cadd(cmul(dA0[e], pfc), cmul(h4, dA1[e])), cmul(pfr, dA2[e])));
    m = fmax(m, cabsv(f));
    if (r == cc) f = cadd(f, pfr);
    tdd[e] = f;
  }
  const int t = threadIdx.x;
  red[t] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] = fmax(red[t], red[t + w]);
    __syncthreads();
  }
  if (t == 0) err[b * gridDim.x + blockIdx.x] = red[0];
}

// redheffer_perturbation_enqueue of a batch of segments with their own references and
// samples: the operands are gathered into strided batches, then every step is one
// batched launch. T: [tdd | tuu | rud | rdu] per segment (stride 4 N^2); err:
// RH_ERR_BLOCKS maxima per segment.
void redheffer_perturbation_batch(vrc_ctx& c, int N, const std::vector<RhJob>& jobs, cd* T, double* err) {
  hipStream_t s = c.stream;
  const int B = (int)jobs.size();
  if (B == 0) return;
  const long NN = (long)N * N;
  const size_t M = sizeof(cd) * NN, Vn = sizeof(cd) * N, Pn = sizeof(int) * N;
  auto col = [&](auto f) {
    std::vector<const void*> v(B);
    for (int b = 0; b < B; ++b) v[b] = f(jobs[b]);
    return v;
  };
  DevBuf Wb = c.ws.get(M * B), Vb = c.ws.get(M * B), LWb = c.ws.get(M * B), LVb = c.ws.get(M * B);
  DevBuf pWb = c.ws.get(std::max<size_t>(Pn, 8) * B + 8), pVb = c.ws.get(std::max<size_t>(Pn, 8) * B + 8);
  DevBuf gb = c.ws.get(Vn * B);
  gather_ptrs(c, col([](const RhJob& j) { return (const void*)j.R.W; }), M, Wb.ptr);
  gather_ptrs(c, col([](const RhJob& j) { return (const void*)j.R.V; }), M, Vb.ptr);
  gather_ptrs(c, col([](const RhJob& j) { return (const void*)j.R.LU.W; }), M, LWb.ptr);
  gather_ptrs(c, col([](const RhJob& j) { return (const void*)j.R.LU.V; }), M, LVb.ptr);
  gather_ptrs(c, col([](const RhJob& j) { return (const void*)j.R.gam; }), Vn, gb.ptr);
  // pivots: N ints (gathered as whole 8-byte words; N even as N = 2 nx ny)
  gather_ptrs(c, col([](const RhJob& j) { return (const void*)j.R.LU.pW; }), Pn, pWb.ptr);
  gather_ptrs(c, col([](const RhJob& j) { return (const void*)j.R.LU.pV; }), Pn, pVb.ptr);
  DevBuf Prb = c.ws.get(M * B), Qrb = c.ws.get(M * B);
  gather_ptrs(c, col([](const RhJob& j) { return (const void*)j.Pr; }), M, Prb.ptr);
  gather_ptrs(c, col([](const RhJob& j) { return (const void*)j.Qr; }), M, Qrb.ptr);
  DevBuf dAb = c.ws.get(3 * M * B), dBb = c.ws.get(3 * M * B), Xb = c.ws.get(M * B), Yb = c.ws.get(M * B);
  cd *dA = dAb.as<cd>(), *dB = dBb.as<cd>(), *X = Xb.as<cd>(), *Y = Yb.as<cd>();
  for (int k = 0; k < 3; ++k) {
    // a = W^{-1} (P_k - Pr) V ; b = V^{-1} (Q_k - Qr) W   (per segment, stride 3 NN in dA, dB)
    gather_ptrs(c, col([k](const RhJob& j) { return (const void*)j.P3[k]; }), M, X);
    zgeam(s, N, N, ONE, X, N, NN, MONE, Prb.as<cd>(), N, NN, X, N, NN, B);
    zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, X, N, NN, Vb.as<cd>(), N, NN, ZERO, Y, N, NN, B);
    zgetrs(s, N, N, LWb.as<cd>(), N, NN, pWb.as<int>(), N, Y, N, NN, dA + k * NN, N, 3 * NN, B, c.ws);
    gather_ptrs(c, col([k](const RhJob& j) { return (const void*)j.Q3[k]; }), M, X);
    zgeam(s, N, N, ONE, X, N, NN, MONE, Qrb.as<cd>(), N, NN, X, N, NN, B);
    zgemm_strided(s, OP_N, OP_N, N, N, N, ONE, X, N, NN, Wb.as<cd>(), N, NN, ZERO, Y, N, NN, B);
    zgetrs(s, N, N, LVb.as<cd>(), N, NN, pVb.as<int>(), N, Y, N, NN, dB + k * NN, N, 3 * NN, B, c.ws);
    // dA = a + b ; dB = a - b
    zgeam(s, N, N, ONE, dA + k * NN, N, 3 * NN, ONE, dB + k * NN, N, 3 * NN, X, N, NN, B);
    zgeam(s, N, N, ONE, dA + k * NN, N, 3 * NN, MONE, dB + k * NN, N, 3 * NN, dB + k * NN, N, 3 * NN, B);
    zgeam(s, N, N, ONE, X, N, NN, ZERO, nullptr, 0, 0, dA + k * NN, N, 3 * NN, B);
  }
  std::vector<double> hh(B), hf(B);
  std::vector<cd> c0(B);
  for (int b = 0; b < B; ++b) {
    const RhJob& j = jobs[b];
    hh[b] = 0.5 * (j.z1 - j.z0) / j.k0;
    hf[b] = (j.z1 - j.z0) / j.k0;
    c0[b] = mk(0.0, (j.z1 - j.z0) / (6.0 * 2.0 * j.k0));
  }
  DevBuf dh = c.ws.get(sizeof(double) * 2 * B), dc0 = c.ws.get(sizeof(cd) * B);
  c.pinned.upload(s, dh.ptr, hh.data(), sizeof(double) * B);
  c.pinned.upload(s, dh.as<double>() + B, hf.data(), sizeof(double) * B);
  c.pinned.upload(s, dc0.ptr, c0.data(), sizeof(cd) * B);
  DevBuf phb = c.ws.get(2 * Vn * B);
  {
    dim3 g(cdiv_up(N, 256), B);
    phase_kernel<<<g, 256, 0, s>>>(N, gb.as<cd>(), N, dh.as<double>(), phb.as<cd>(), N);
    phase_kernel<<<g, 256, 0, s>>>(N, gb.as<cd>(), N, dh.as<double>() + B, phb.as<cd>() + (long)N * B, N);
    launch_check("phase");
  }
  simpson_batch_kernel<<<dim3(RH_ERR_BLOCKS, B), 256, 0, s>>>(N, dA, dB, phb.as<cd>(), phb.as<cd>() + (long)N * B,
                                                            dc0.as<cd>(), T, err);
  launch_check("simpson_batch");
}

void redheffer_perturbation_enqueue(vrc_ctx& c, int N, double k0, double z0, double z1, const cd* Pr,
                                    const cd* Qr, const RefSys& R, const cd* P3[3], const cd* Q3[3], cd* tdd,
                                    cd* tuu, cd* rud, cd* rdu, const double* d_hs, double* err) {
  hipStream_t s = c.stream;
  const long NN = (long)N * N;
  DevBuf dAb = c.ws.get(sizeof(cd) * 6 * (size_t)NN);
  DevBuf tmpb = c.ws.get(sizeof(cd) * 2 * (size_t)NN);
  cd* dA = dAb.as<cd>();       // dA0 dA1 dA2
  cd* dB = dA + 3 * NN;        // dB0 dB1 dB2
  cd* tmp = tmpb.as<cd>();
  for (int k = 0; k < 3; ++k) {
    // a = W^{-1} (P_k - Pr) V ; b = V^{-1} (Q_k - Qr) W
    zgeam(s, N, N, ONE, P3[k], N, 0, MONE, Pr, N, 0, tmp, N, 0, 1);
    mm(s, N, ONE, tmp, R.V, ZERO, tmp + NN, 0, 1);
    zgetrs(s, N, N, R.LU.W, N, 0, R.LU.pW, N, tmp + NN, N, 0, dA + k * NN, N, 0, 1, c.ws);
    zgeam(s, N, N, ONE, Q3[k], N, 0, MONE, Qr, N, 0, tmp, N, 0, 1);
    mm(s, N, ONE, tmp, R.W, ZERO, tmp + NN, 0, 1);
    zgetrs(s, N, N, R.LU.V, N, 0, R.LU.pV, N, tmp + NN, N, 0, dB + k * NN, N, 0, 1, c.ws);
    // dA = a + b ; dB = a - b
    zgeam(s, N, N, ONE, dA + k * NN, N, 0, ONE, dB + k * NN, N, 0, tmp, N, 0, 1);
    zgeam(s, N, N, ONE, dA + k * NN, N, 0, MONE, dB + k * NN, N, 0, dB + k * NN, N, 0, 1);
    copy(s, N, N, tmp, dA + k * NN, 0, 1);
  }
  DevBuf phb = c.ws.get(sizeof(cd) * 2 * (size_t)N);
  {
    dim3 g(cdiv_up(N, 256), 1);
    phase_kernel<<<g, 256, 0, s>>>(N, R.gam, 0, d_hs, phb.as<cd>(), 0);
    phase_kernel<<<g, 256, 0, s>>>(N, R.gam, 0, d_hs + 1, phb.as<cd>() + N, 0);
    launch_check("phase");
  }
  const cd c0 = mk(0.0, (z1 - z0) / (6.0 * 2.0 * k0));
  simpson_kernel<<<RH_ERR_BLOCKS, 256, 0, s>>>(N, dA, dA + NN, dA + 2 * NN, dB, dB + NN, dB + 2 * NN, phb.as<cd>(),
                                              phb.as<cd>() + N, c0, tdd, tuu, rud, rdu, err);
  launch_check("simpson");
}

double redheffer_perturbation(vrc_ctx& c, int N, double k0, double z0, double z1, const cd* Pr, const cd* Qr,
                              const RefSys& R, const cd* P3[3], const cd* Q3[3], cd* tdd, cd* tuu, cd* rud,
                              cd* rdu) {
  hipStream_t s = c.stream;
  DevBuf sb = c.ws.get(sizeof(double) * 2);
  const double hs[2] = {0.5 * (z1 - z0) / k0, (z1 - z0) / k0};
  VRC_HIP(hipMemcpyAsync(sb.as<double>(), hs, sizeof(hs), hipMemcpyHostToDevice, s));
  DevBuf errb = c.ws.get(sizeof(double) * RH_ERR_BLOCKS);
  redheffer_perturbation_enqueue(c, N, k0, z0, z1, Pr, Qr, R, P3, Q3, tdd, tuu, rud, rdu, sb.as<double>(),
                                 errb.as<double>());
  std::vector<double> h(RH_ERR_BLOCKS);
  VRC_HIP(hipMemcpyAsync(h.data(), errb.as<double>(), sizeof(double) * RH_ERR_BLOCKS, hipMemcpyDeviceToHost, s));
  VRC_HIP(hipStreamSynchronize(s));
  double e = 0.0;
  for (double v : h) e = v > e ? v : e;
  return e;
}

}  // namespace vrc
