// Host-side geometry preparation: painter-order refill, K matrices (periodic and
// PML-stretched), and upload of a layer batch for the Fourier assembly kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "vrc_ops.hpp"

namespace vrc {

static constexpr double PI_REF = 3.1415926535897932;  // src/rcwa/defns.h:29

static double sinc_ref(double x) { return x == 0.0 ? 1.0 : std::sin(PI_REF * x) / (PI_REF * x); }
static cd cexp_i(double a) { return mk(std::cos(a), std::sin(a)); }
static cd hmul(cd a, cd b) { return cmul(a, b); }

PreparedLayer prepare_layer(const HostLayer& L) {
  const int nCx = (int)L.x.size(), nCy = (int)L.y.size();
  if (nCx < 2 || nCy < 2) throw ArgFailure("layer needs at least 2 coordinates per axis");
  if ((int)L.eps.size() != (nCx - 1) * (nCy - 1)) throw ArgFailure("eps size does not match coordinates");
  for (const cd& e : L.eps) {
    if (!std::isfinite(e.x) || !std::isfinite(e.y)) throw ArgFailure("non-finite permittivity (NaN or infinity)");
    // the inverse Fourier rules (fe11, fe22) are built from 1/eps of every cell
    if (e.x == 0.0 && e.y == 0.0) throw ArgFailure("zero permittivity in a cell: the inverse Fourier rules need 1/eps");
  }
  PreparedLayer p;
  p.ncx = nCx - 1;
  p.ncy = nCy - 1;
  p.Lx = L.x[nCx - 1] - L.x[0];
  p.Ly = L.y[nCy - 1] - L.y[0];
  p.xu = L.x;
  p.ys = L.y;
  // FourierSolver1D::reorder: indices sorted by coordinate (stable for ties)
  std::vector<int> index(nCx), rindex(nCx);
  std::iota(index.begin(), index.end(), 0);
  std::stable_sort(index.begin(), index.end(), [&](int a, int b) { return L.x[a] < L.x[b]; });
  for (int i = 0; i < nCx; ++i) rindex[index[i]] = i;
  p.xs.resize(nCx);
  for (int i = 0; i < nCx; ++i) p.xs[i] = L.x[index[i]];
  bool need = false;
  for (int i = 0; i + 1 < nCx; ++i)
    if (rindex[i] != i) { need = true; break; }
  p.val.resize((size_t)p.ncy * p.ncx);
  for (int s = 0; s < p.ncy; ++s) {
    for (int j = 0; j < p.ncx; ++j) p.val[(size_t)s * p.ncx + j] = L.at(s, j);
    if (!need) continue;
    // painter's order: cell i covers sorted intervals rindex[i] .. rindex[i+1]-1
    for (int i = 0; i < p.ncx; ++i) {
      if (L.x[i] >= L.x[i + 1]) continue;
      for (int j = 0; j < p.ncx; ++j)
        if (rindex[i] <= j && rindex[i + 1] >= j + 1) p.val[(size_t)s * p.ncx + j] = L.at(s, i);
    }
  }
  return p;
}

void alpha_beta(int nx, int ny, double Lx, double Ly, std::vector<cd>& alpha, std::vector<cd>& beta) {
  const int Nx = (nx - 1) / 2, Ny = (ny - 1) / 2;
  const double Kx = 2 * PI_REF / Lx, Ky = 2 * PI_REF / Ly;
  alpha.assign((size_t)nx * ny, mk(0, 0));
  beta.assign((size_t)nx * ny, mk(0, 0));
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      alpha[(size_t)ny * i + j] = mk(Kx * (double)(i - Nx), 0.0);
      beta[(size_t)ny * i + j] = mk(Ky * (double)(j - Ny), 0.0);
    }
}

// PMLSolver1D::solve (src/rcwa/PMLSolver1D.cpp:15-30)
static cd pml1d(double b0, double b1, double L, cd gamma, int n, double coeff) {
  const double bm = (b1 - b0) / L, bp = (b1 + b0) / L;
  auto ex = [](double x) { return cexp_i(PI_REF * x); };
  cd r = cscale(ex(-(n * bp + coeff * 0.5)), 4.0 * sinc_ref(n * bm - 0.5));
  r = cadd(r, cscale(ex(-(n * bp)), 8.0 * sinc_ref(n * bm)));
  r = cadd(r, cscale(hmul(gamma, ex(-(n * bp))), -2.0 * sinc_ref(n * bm)));
  r = cadd(r, cscale(ex(-(n * bp - coeff * 0.5)), 4.0 * sinc_ref(n * bm + 0.5)));
  r = cadd(r, cscale(hmul(gamma, ex(-(n * bp + coeff * 1.0))), sinc_ref(n * bm - 1.0)));
  r = cadd(r, cscale(hmul(gamma, ex(-(n * bp - coeff * 1.0))), sinc_ref(n * bm + 1.0)));
  return cscale(r, bm / 16.0);
}

// VarLayerSampler::assignSimulationRegion (src/core/VarLayerSampler.cpp:23-39): the cell
// and the absorbing regions from the first/last two coordinates of each axis.
Region region_from_coords(const std::vector<double>& x, const std::vector<double>& y) {
  const int nCx = (int)x.size(), nCy = (int)y.size();
  Region r;
  r.Lx = x[nCx - 1] - x[0];
  r.Ly = y[nCy - 1] - y[0];
  r.b0 = y[0];
  r.b1 = nCy > 1 ? y[1] : y[0];
  r.u0 = nCy > 1 ? y[nCy - 2] : y[0];
  r.u1 = y[nCy - 1];
  r.l0 = x[0];
  r.l1 = nCx > 1 ? x[1] : x[0];
  r.r0 = nCx > 1 ? x[nCx - 2] : x[0];
  r.r1 = x[nCx - 1];
  return r;
}

void pml_k_matrices(int nx, int ny, const std::vector<double>& x, const std::vector<double>& y,
                    std::vector<cd>& Kx, std::vector<cd>& Ky) {
  pml_k_matrices(nx, ny, region_from_coords(x, y), Kx, Ky);
}

// LayerSampler::evaluateKMatrices + evaluatePML (src/core/LayerSampler.cpp:34-132) for a
// region given by its scalars.
void pml_k_matrices(int nx, int ny, const Region& R, std::vector<cd>& Kx, std::vector<cd>& Ky) {
  const double Lx = R.Lx, Ly = R.Ly;
  const double b0 = R.b0, b1 = R.b1, u0 = R.u0, u1 = R.u1;
  const double l0 = R.l0, l1 = R.l1, r0 = R.r0, r1 = R.r1;
  std::vector<cd> alpha, beta;
  alpha_beta(nx, ny, Lx, Ly, alpha, beta);
  const cd gamma = crecip(mk(1.0, -1.0));
  const int nPx = 2 * nx - 1, nPy = 2 * ny - 1, maxPx = (nPx - 1) / 2, maxPy = (nPy - 1) / 2;
  std::vector<cd> fyy(nPy), fxx(nPx), fyx(nPx), fxy(nPy);
  for (int i = 0; i < nPy; ++i) {
    const int n = i - maxPy;
    cd v = cdiv(mk((u0 - b1) / Ly * sinc_ref(n * (u0 - b1) / Ly), 0.0), cexp_i(PI_REF * (n * (b1 + u0) / Ly)));
    v = cadd(v, pml1d(b0, b1, Ly, gamma, n, 1.0));
    fyy[i] = cadd(v, pml1d(u0, u1, Ly, gamma, n, -1.0));
  }
  for (int i = 0; i < nPx; ++i) {
    const int m = i - maxPx;
    cd v = cdiv(mk((r0 - l1) * sinc_ref(m * (r0 - l1) / Lx) / Lx, 0.0), cexp_i(PI_REF * m * (l1 + r0) / Lx));
    v = cadd(v, pml1d(l0, l1, Lx, gamma, m, 1.0));
    fxx[i] = cadd(v, pml1d(r0, r1, Lx, gamma, m, -1.0));
  }
  for (int i = 0; i < nPx; ++i) {
    const int m = i - maxPx;
    fyx[i] = (m == 0) ? mk(r1 - l0, 0.0)
                      : cmul(csub(cexp_i(-2 * PI_REF / Lx * m * r1), cexp_i(-2 * PI_REF / Lx * m * l0)),
                             mk(0.0, Lx * Lx / (m * 2 * PI_REF)));
  }
  for (int i = 0; i < nPy; ++i) {
    const int n = i - maxPy;
    fxy[i] = (n == 0) ? mk(u1 - b0, 0.0)
                      : cmul(csub(cexp_i(-2 * PI_REF / Ly * n * u1), cexp_i(-2 * PI_REF / Ly * n * b0)),
                             mk(0.0, Ly * Ly / (2 * n * PI_REF)));
  }
  // fy = fyx fyy^T / Lx ; fx = fxx fxy^T / Ly ; K = BlockToeplitz(f) * diag(alpha/beta)
  const int n = nx * ny;
  Kx.assign((size_t)n * n, mk(0, 0));
  Ky.assign((size_t)n * n, mk(0, 0));
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      const int m = r / ny, nn = r % ny, j = c / ny, l = c % ny;
      const int p = nx - 1 + (m - j), q = ny - 1 + (nn - l);
      const cd fx = cscale(cmul(fxx[p], fxy[q]), 1.0 / Ly);
      const cd fy = cscale(cmul(fyx[p], fyy[q]), 1.0 / Lx);
      Kx[(size_t)r + (size_t)c * n] = cmul(fx, alpha[c]);
      Ky[(size_t)r + (size_t)c * n] = cmul(fy, beta[c]);
    }
}

__global__ void fill_int_kernel(int* p, int v, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void fill_int(hipStream_t s, int* p, int v, long n) {
  if (n <= 0) return;
  fill_int_kernel<<<cdiv_up(n, 256), 256, 0, s>>>(p, v, n);
  launch_check("fill_int");
}

FourierStage fourier_prepare(vrc_ctx& c, int nx, int ny, const std::vector<const HostLayer*>& layers) {
  hipStream_t s = c.stream;
  const int L = (int)layers.size();
  FourierStage fs;
  fs.L = L;
  std::vector<PreparedLayer> pl;
  pl.reserve(L);
  int MX = 1, MY = 1;
  for (int l = 0; l < L; ++l) {
    pl.push_back(prepare_layer(*layers[l]));
    MX = std::max(MX, pl.back().ncx);
    MY = std::max(MY, pl.back().ncy);
  }
  const int M1 = 2 * nx - 1, M2 = 2 * ny - 1;
  // uniform power-of-two grids along x (sorted, so no refill) and y take the FFT path
  auto fft_log2 = [](const std::vector<double>& c, const std::vector<double>* sorted) -> int {
    const int n = (int)c.size() - 1;
    if (n < 2 || n > FOURIER_FFT_MAX || (n & (n - 1)) != 0) return 0;
    if (sorted && *sorted != c) return 0;
    const double d = c[n] - c[0];
    if (!(d > 0)) return 0;
    for (int j = 0; j <= n; ++j)
      if (std::fabs(c[j] - (c[0] + d * j / n)) > 1e-12 * d) return 0;
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    return lg;
  };
  std::vector<int> fx(L), fy(L);
  int fft_max = 0;
  for (int l = 0; l < L; ++l) {
    fx[l] = fft_log2(pl[l].xu, &pl[l].xs);
    fy[l] = fft_log2(pl[l].ys, nullptr);
    fft_max = std::max(fft_max, std::max(fx[l] ? 1 << fx[l] : 0, fy[l] ? 1 << fy[l] : 0));
  }
  // host staging (one upload)
  std::vector<int> ncx(L), ncy(L);
  std::vector<double> Lx(L), Ly(L), xu((size_t)L * (MX + 1), 0.0), xs((size_t)L * (MX + 1), 0.0),
      ys((size_t)L * (MY + 1), 0.0);
  std::vector<cd> val((size_t)L * MY * MX, mk(1.0, 0.0));
  for (int l = 0; l < L; ++l) {
    const PreparedLayer& p = pl[l];
    ncx[l] = p.ncx;
    ncy[l] = p.ncy;
    Lx[l] = p.Lx;
    Ly[l] = p.Ly;
    std::copy(p.xu.begin(), p.xu.end(), xu.begin() + (size_t)l * (MX + 1));
    std::copy(p.xs.begin(), p.xs.end(), xs.begin() + (size_t)l * (MX + 1));
    std::copy(p.ys.begin(), p.ys.end(), ys.begin() + (size_t)l * (MY + 1));
    for (int i = 0; i < p.ncy; ++i)
      for (int j = 0; j < p.ncx; ++j) val[((size_t)l * MY + i) * MX + j] = p.val[(size_t)i * p.ncx + j];
  }
  const size_t bytes_int = sizeof(int) * 4 * L, bytes_d = sizeof(double) * (2 * L + 2 * L * (MX + 1) + L * (MY + 1)),
               bytes_c = sizeof(cd) * val.size();
  std::vector<char> host(bytes_int + bytes_d + bytes_c);
  char* h = host.data();
  size_t o = 0;
  auto put = [&](const void* src, size_t b) {
    std::memcpy(h + o, src, b);
    o += b;
  };
  put(ncx.data(), sizeof(int) * L);
  put(ncy.data(), sizeof(int) * L);
  put(fx.data(), sizeof(int) * L);
  put(fy.data(), sizeof(int) * L);
  const size_t od = o;
  put(Lx.data(), sizeof(double) * L);
  put(Ly.data(), sizeof(double) * L);
  put(xu.data(), sizeof(double) * xu.size());
  put(xs.data(), sizeof(double) * xs.size());
  put(ys.data(), sizeof(double) * ys.size());
  const size_t oc = (o + 15) & ~(size_t)15;
  host.resize(oc + bytes_c);
  h = host.data();
  std::memcpy(h + oc, val.data(), bytes_c);
  fs.geo = c.ws.get(host.size());
  c.pinned.upload(s, fs.geo.ptr, host.data(), host.size());
  char* d = (char*)fs.geo.ptr;
  FourierDev& g = fs.g;
  g.nx = nx;
  g.ny = ny;
  g.MX = MX;
  g.MY = MY;
  g.ncx = (const int*)d;
  g.ncy = (const int*)d + L;
  g.fftx = (const int*)d + 2 * L;
  g.ffty = (const int*)d + 3 * L;
  g.fft_max = fft_max;
  const double* dd = (const double*)(d + od);
  g.Lx = dd;
  g.Ly = dd + L;
  g.xu = dd + 2 * L;
  g.xs = g.xu + (size_t)L * (MX + 1);
  g.ys = g.xs + (size_t)L * (MX + 1);
  g.val = (const cd*)(d + oc);
  const size_t nFX = (size_t)MY * M1, nFY1 = (size_t)MY * M2, nFYi = (size_t)MX * M2, nFX1 = (size_t)MX * M1,
               ntx = (size_t)MY * nx * nx, nty = (size_t)MX * ny * ny, nF2 = (size_t)M1 * M2,
               nG11 = (size_t)nx * nx * M2, nG22 = (size_t)ny * ny * M1;
  const size_t per = 2 * nFX + nFY1 + nFYi + nFX1 + ntx + nty + nF2 + nG11 + nG22;
  fs.wk = c.ws.get(sizeof(cd) * per * L);
  cd* w = fs.wk.as<cd>();
  g.FX = w;
  g.FXi = g.FX + nFX * L;
  g.FY1 = g.FXi + nFX * L;
  g.FYi = g.FY1 + nFY1 * L;
  g.FX1 = g.FYi + nFYi * L;
  g.tFFx = g.FX1 + nFX1 * L;
  g.tFFy = g.tFFx + ntx * L;
  g.F2 = g.tFFy + nty * L;
  g.G11 = g.F2 + nF2 * L;
  g.G22 = g.G11 + nG11 * L;
  return fs;
}

void assemble_fe(vrc_ctx& c, const AssemblyBatch& a, cd* fe11, cd* fe22, cd* fe33, long sfe) {
  FourierStage fs = fourier_prepare(c, a.nx, a.ny, a.layers);
  StageScope st(ST_FOURIER);
  fourier_assemble(c.stream, fs.g, fs.L, fe11, fe22, fe33, sfe);
}

}  // namespace vrc
