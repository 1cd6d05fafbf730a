// Host-side samplers and integrators with the reference's compute-S-matrix API
// (src/core/LayerSampler.*, VarLayerSampler.*, RCWAIntegrator.*, RedhefferIntegrator.*,
// DifferentialIntegrator.*), driving the batched device procedures.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "vrc_adjoint.hpp"

namespace vrc {

static constexpr double PI_REF = 3.1415926535897932;
static const cd ONE = {1.0, 0.0}, MONE = {-1.0, 0.0}, ZERO = {0.0, 0.0};

// Host -> device copy of a small array on the context stream (kept alive until done).
// Instances whose star products and projections run in one batch (batch_stack_rcwa).
static int g_cascade_group = 512;
void stack_set_cascade_group(int n) { g_cascade_group = n < 1 ? 1 : n; }
// Instances per cascade group: g_cascade_group, fewer when their slab S-matrices and end
// systems would take more than 96 GiB (deep stacks at large N; 128 instances of 10 slabs at
// N = 1250 take 77 GB, so their 1280 slab eigensolves still fill five 256-matrix launches).
static int cascade_group_cap(int N, int L) {
  const size_t per_inst = sizeof(cd) * (size_t)N * N * (2 * (size_t)L + 4);
  const size_t cap = std::max<size_t>(1, ((size_t)96 << 30) / per_inst);
  return (int)std::max<size_t>(1, std::min<size_t>((size_t)g_cascade_group, cap));
}

static DevBuf upload(vrc_ctx& c, const void* host, size_t bytes) {
  DevBuf b = c.ws.get(bytes);
  c.pinned.upload(c.stream, b.ptr, host, bytes);
  return b;
}

// ===================================================================== samplers
Sampler::Sampler(vrc_ctx* c, double lambda, int nx, int ny, bool pml)
    : c_(c), lambda_(lambda), nx_(nx), ny_(ny), pml_(pml) {
  if (nx < 1 || ny < 1) throw ArgFailure("harmonic counts must be positive");
  if (nx % 2 == 0 || ny % 2 == 0) throw ArgFailure("harmonic counts per axis must be odd (2*order+1)");
  if (!(lambda > 0)) throw ArgFailure("wavelength must be positive");
}

double Sampler::k0() const { return 2 * PI_REF / lambda_; }

void Sampler::post_init(const std::vector<double>& x, const std::vector<double>& y) {
  if (pml_ && (x.size() < 4 || y.size() < 4)) throw ArgFailure("PML needs at least 4 coordinates per axis");
  post_init(region_from_coords(x, y));
}

// LayerSampler::postInitialization (src/core/LayerSampler.cpp:145-153): K matrices of the
// region, periodic (diagonal) or PML-stretched (dense, with the products P/Q assembly needs).
void Sampler::post_init(const Region& r) {
  const int n = nx_ * ny_;
  const long nn = (long)n * n;
  const double Lx = r.Lx, Ly = r.Ly;
  if (!(Lx > 0) || !(Ly > 0)) throw ArgFailure("cell size must be positive");
  region_ = r;
  if (pml_) {
    pml_k_matrices(nx_, ny_, r, hKx_, hKy_);
    std::vector<cd> both(2 * nn);
    std::copy(hKx_.begin(), hKx_.end(), both.begin());
    std::copy(hKy_.begin(), hKy_.end(), both.begin() + nn);
    d_K_ = c_->ws.get(sizeof(cd) * 6 * nn);
    cd* K = d_K_.as<cd>();
    VRC_HIP(hipMemcpyAsync(K, both.data(), sizeof(cd) * 2 * nn, hipMemcpyHostToDevice, c_->stream));
    VRC_HIP(hipStreamSynchronize(c_->stream));
    const cd* Kx = K;
    const cd* Ky = K + nn;
    zgemm_strided(c_->stream, OP_N, OP_N, n, n, n, ONE, Kx, n, 0, Ky, n, 0, ZERO, K + 2 * nn, n, 0, 1);
    zgemm_strided(c_->stream, OP_N, OP_N, n, n, n, ONE, Kx, n, 0, Kx, n, 0, ZERO, K + 3 * nn, n, 0, 1);
    zgemm_strided(c_->stream, OP_N, OP_N, n, n, n, ONE, Ky, n, 0, Ky, n, 0, ZERO, K + 4 * nn, n, 0, 1);
    zgemm_strided(c_->stream, OP_N, OP_N, n, n, n, ONE, Ky, n, 0, Kx, n, 0, ZERO, K + 5 * nn, n, 0, 1);
  } else {
    alpha_beta(nx_, ny_, Lx, Ly, alpha_, beta_);
    hKx_.assign(nn, mk(0, 0));
    hKy_.assign(nn, mk(0, 0));
    for (int i = 0; i < n; ++i) {
      hKx_[(size_t)i * (n + 1)] = alpha_[i];
      hKy_[(size_t)i * (n + 1)] = beta_[i];
    }
    std::vector<cd> ab(2 * n);
    std::copy(alpha_.begin(), alpha_.end(), ab.begin());
    std::copy(beta_.begin(), beta_.end(), ab.begin() + n);
    d_alpha_ = upload(*c_, ab.data(), sizeof(cd) * 2 * n);
  }
}

void Sampler::sample_gpu(const std::vector<double>& zs, cd* P, cd* Q) {
  std::vector<HostLayer> ls;
  ls.reserve(zs.size());
  for (double z : zs) ls.push_back(layer_at(z));
  sample_layers(ls, P, Q);
}

void Sampler::sample_host(double z, cd* Ph, cd* Qh) {
  const long NN = (long)nDim() * nDim();
  DevBuf P = c_->ws.get(sizeof(cd) * NN), Q = c_->ws.get(sizeof(cd) * NN);
  sample_gpu({z}, P.as<cd>(), Q.as<cd>());
  VRC_HIP(hipMemcpyAsync(Ph, P.ptr, sizeof(cd) * NN, hipMemcpyDeviceToHost, c_->stream));
  VRC_HIP(hipMemcpyAsync(Qh, Q.ptr, sizeof(cd) * NN, hipMemcpyDeviceToHost, c_->stream));
  VRC_HIP(hipStreamSynchronize(c_->stream));
}

CallbackSampler::CallbackSampler(vrc_ctx* c, double lambda, int nx, int ny, const Region& r, bool pml,
                                 SampleHostFn hf, SampleDeviceFn df, void* user)
    : Sampler(c, lambda, nx, ny, pml), hf_(hf), df_(df), user_(user) {
  if (!hf && !df) throw ArgFailure("a callback sampler needs a host or a device sampling function");
  if (pml && !(r.l0 <= r.l1 && r.l1 <= r.r0 && r.r0 <= r.r1 && r.b0 <= r.b1 && r.b1 <= r.u0 && r.u0 <= r.u1))
    throw ArgFailure("PML region boundaries must be ordered l0 <= l1 <= r0 <= r1 and b0 <= b1 <= u0 <= u1");
  post_init(r);
}

HostLayer CallbackSampler::layer_at(double) const {
  throw ArgFailure("a callback sampler has no layer description (its P and Q come from the caller)");
}

static void callback_failed(int rc, const char* which) {
  throw std::runtime_error(std::string("sampler ") + which + " callback returned " + std::to_string(rc));
}

void CallbackSampler::sample_gpu(const std::vector<double>& zs, cd* P, cd* Q) {
  const long NN = (long)nDim() * nDim();
  std::vector<cd> hP, hQ;
  for (size_t k = 0; k < zs.size(); ++k) {
    cd* Pk = P + (long)k * NN;
    cd* Qk = Q + (long)k * NN;
    if (df_) {
      const int rc = df_(zs[k], Pk, Qk, (void*)c_->stream, user_);
      if (rc == 0) continue;
      if (rc != 1 || !hf_) callback_failed(rc, "device");
    }
    hP.resize(NN);
    hQ.resize(NN);
    const int rc = hf_(zs[k], (double*)hP.data(), (double*)hQ.data(), user_);
    if (rc != 0) callback_failed(rc, "host");
    VRC_HIP(hipMemcpyAsync(Pk, hP.data(), sizeof(cd) * NN, hipMemcpyHostToDevice, c_->stream));
    VRC_HIP(hipMemcpyAsync(Qk, hQ.data(), sizeof(cd) * NN, hipMemcpyHostToDevice, c_->stream));
    VRC_HIP(hipStreamSynchronize(c_->stream));  // the host buffers are reused
  }
}

void CallbackSampler::sample_host(double z, cd* Ph, cd* Qh) {
  if (hf_) {
    const int rc = hf_(z, (double*)Ph, (double*)Qh, user_);
    if (rc != 0) callback_failed(rc, "host");
    return;
  }
  Sampler::sample_host(z, Ph, Qh);
}

void Sampler::sample_layers(const std::vector<HostLayer>& ls, cd* P, cd* Q) {
  const int L = (int)ls.size();
  if (L == 0) return;
  AssemblyBatch a;
  a.nx = nx_;
  a.ny = ny_;
  a.nlayers = L;
  for (auto& l : ls) a.layers.push_back(&l);
  a.periodic = !pml_;
  const double k2 = k0() * k0();
  a.k0sq_uniform = k2;
  std::vector<double> k2v(L, k2);
  DevBuf dk = upload(*c_, k2v.data(), sizeof(double) * L);
  a.d_k0sq = dk.as<double>();
  if (pml_) {
    const long nn = (long)n() * n();
    const cd* K = d_K_.as<cd>();
    a.d_Kx = K;
    a.d_Ky = K + nn;
    a.d_KxKy = K + 2 * nn;
    a.d_KxKx = K + 3 * nn;
    a.d_KyKy = K + 4 * nn;
    a.d_KyKx = K + 5 * nn;
  } else {
    a.d_alpha = d_alpha_.as<cd>();
    a.d_beta = d_alpha_.as<cd>() + n();
    a.sab = 0;
  }
  assemble_pq(*c_, a, P, Q);
}

// Derivative outputs of Layer::waveEqnCoeff (src/rcwa/Layer.cpp:36-162) with the
// tangents of FourierSolver2D::solve (src/rcwa/FourierSolver2D.cpp:12-351):
//   dP = [[Kx Y1, Kx Y2], [Ky Y1, Ky Y2]],  [Y1 | Y2] = fe33^-1 dfe33 fe33^-1 [-Ky | Kx]
//   dQ = [[0, -k0^2 dfe22], [k0^2 dfe11, 0]],  dPl, dQl with d(k0^2)/d lambda = -8 pi^2 / lambda^3.
void shape_tangents(vrc_ctx& c, int nx, int ny, double lambda, bool pml, const cd* d_alpha, const cd* d_K,
                    const HostLayer& L, int nvar, const int* var_index, const double* var_coeff, int npara, cd* dP,
                    cd* dQ, cd* dfe11, cd* dfe22, cd* dfe33, cd* dPl, cd* dQl) {
  hipStream_t st = c.stream;
  vrc_ctx* c_ = &c;
  const int nx_ = nx, ny_ = ny;
  const double lambda_ = lambda;
  const bool pml_ = pml;
  auto k0 = [&] { return 2 * PI_REF / lambda; };
  const int nCx = (int)L.x.size();
  if (npara < 0 || nvar < 0 || (nvar > 0 && (!var_index || !var_coeff))) throw ArgFailure("bad parameterization");
  for (int i = 0; i + 1 < nCx; ++i)
    if (L.x[i] > L.x[i + 1])
      throw ArgFailure("shape derivatives need non-decreasing x coordinates (crossing coordinates have none)");
  std::vector<char> seen(nCx, 0);
  for (int k = 0; k < nvar; ++k) {
    const int v = var_index[k];
    if (v <= 0 || v >= nCx - 1) throw ArgFailure("only interior x coordinates can be varied");
    if (seen[v]) throw ArgFailure("a coordinate appears twice in the parameterization");
    seen[v] = 1;
  }
  const int n = nx_ * ny_, N = 2 * n;
  const long nn = (long)n * n, NN = (long)N * N;
  const HostLayer* lp = &L;
  FourierStage fs = fourier_prepare(*c_, nx_, ny_, {lp});
  DevBuf fe = c_->ws.get(sizeof(cd) * 3 * nn);
  cd* fe11 = fe.as<cd>();
  cd* fe22 = fe11 + nn;
  cd* fe33 = fe22 + nn;
  fourier_assemble(st, fs.g, 1, fe11, fe22, fe33, nn);
  const double k2 = k0() * k0();
  const double dk02 = -8.0 * PI_REF * PI_REF / (lambda_ * lambda_ * lambda_);
  if (dPl) {
    zset(st, N, N, ZERO, ZERO, dPl, N, 0, 1);
    zset(st, n, n, mk(dk02, 0.0), ZERO, dPl + (long)n * N, N, 0, 1);
    zset(st, n, n, mk(-dk02, 0.0), ZERO, dPl + n, N, 0, 1);
  }
  if (dQl) {
    zset(st, N, N, ZERO, ZERO, dQl, N, 0, 1);
    zgeam(st, n, n, mk(-dk02, 0.0), fe22, n, 0, ZERO, nullptr, 0, 0, dQl + (long)n * N, N, 0, 1);
    zgeam(st, n, n, mk(dk02, 0.0), fe11, n, 0, ZERO, nullptr, 0, 0, dQl + n, N, 0, 1);
  }
  if (npara == 0) return;
  // coefficient of every coordinate (sorted = given order for non-decreasing coordinates)
  const long MX = fs.g.MX, MY = fs.g.MY;
  std::vector<double> cfs((size_t)(MX + 1) * npara, 0.0);
  for (int k = 0; k < nvar; ++k)
    for (int p = 0; p < npara; ++p) cfs[(size_t)var_index[k] * npara + p] = var_coeff[(size_t)k * npara + p];
  DevBuf dcfs = upload(*c_, cfs.data(), sizeof(double) * cfs.size());
  const int M1 = 2 * nx_ - 1, M2 = 2 * ny_ - 1;
  const long nxx = (long)nx_ * nx_;
  DevBuf wk = c_->ws.get(sizeof(cd) * ((size_t)npara * (2 * MY * M1 + MX * M1 + 3 * MY * nxx) + M1 * M2 +
                                       nxx * M2 + (long)ny_ * ny_ * M1));
  ShapeWork w;
  w.dFX = wk.as<cd>();
  w.dFXi = w.dFX + npara * MY * M1;
  w.dFX1 = w.dFXi + npara * MY * M1;
  w.D = w.dFX1 + npara * MX * M1;
  w.Y = w.D + npara * MY * nxx;
  w.dtFF = w.Y + npara * MY * nxx;
  w.F2 = w.dtFF + npara * MY * nxx;
  w.G11 = w.F2 + (long)M1 * M2;
  w.G22 = w.G11 + nxx * M2;
  DevBuf dfe;
  if (!dfe11 || !dfe22 || !dfe33) {
    dfe = c_->ws.get(sizeof(cd) * 3 * nn * npara);
    if (!dfe11) dfe11 = dfe.as<cd>();
    if (!dfe22) dfe22 = dfe.as<cd>() + nn * npara;
    if (!dfe33) dfe33 = dfe.as<cd>() + 2 * nn * npara;
  }
  fourier_tangents(st, fs.g, npara, dcfs.as<double>(), w, dfe11, dfe22, dfe33, nn);
  // dQ
  zset(st, N, N, ZERO, ZERO, dQ, N, NN, npara);
  zgeam(st, n, n, mk(-k2, 0.0), dfe22, n, nn, ZERO, nullptr, 0, 0, dQ + (long)n * N, N, NN, npara);
  zgeam(st, n, n, mk(k2, 0.0), dfe11, n, nn, ZERO, nullptr, 0, 0, dQ + n, N, NN, npara);
  // F = fe33^-1 [-Ky | Kx]; X = fe33^-1 dfe33 (all parameters as one right-hand side block)
  DevBuf lu = c_->ws.get(sizeof(cd) * nn), piv = c_->ws.get(sizeof(int) * n), inf = c_->ws.get(sizeof(int));
  DevBuf F = c_->ws.get(sizeof(cd) * 2 * nn), X = c_->ws.get(sizeof(cd) * nn * npara),
         Y = c_->ws.get(sizeof(cd) * 2 * nn * npara);
  fill_int(st, inf.as<int>(), 0, 1);
  zgeam(st, n, n, ONE, fe33, n, 0, ZERO, nullptr, 0, 0, lu.as<cd>(), n, 0, 1);
  zgetrf(st, n, lu.as<cd>(), n, 0, piv.as<int>(), 0, inf.as<int>(), 1, c_->ws);
  if (pml_) {
    const cd* K = d_K;
    zgeam(st, n, n, mk(-1.0, 0.0), K + nn, n, 0, ZERO, nullptr, 0, 0, F.as<cd>(), n, 0, 1);
    zgeam(st, n, n, ONE, K, n, 0, ZERO, nullptr, 0, 0, F.as<cd>() + nn, n, 0, 1);
  } else {
    zset(st, n, 2 * n, ZERO, ZERO, F.as<cd>(), n, 0, 1);
    zadd_diag(st, n, mk(-1.0, 0.0), d_alpha + n, 0, F.as<cd>(), n, 0, 1);
    zadd_diag(st, n, ONE, d_alpha, 0, F.as<cd>() + nn, n, 0, 1);
  }
  zgetrs(st, n, 2 * n, lu.as<cd>(), n, 0, piv.as<int>(), 0, F.as<cd>(), n, 0, F.as<cd>(), n, 0, 1, c_->ws);
  zgetrs(st, n, n * npara, lu.as<cd>(), n, 0, piv.as<int>(), 0, dfe33, n, 0, X.as<cd>(), n, 0, 1, c_->ws);
  zgemm_strided(st, OP_N, OP_N, n, 2 * n, n, ONE, X.as<cd>(), n, nn, F.as<cd>(), n, 0, ZERO, Y.as<cd>(), n, 2 * nn,
                npara);
  if (pml_) {
    const cd* K = d_K;
    zgemm_strided(st, OP_N, OP_N, n, 2 * n, n, ONE, K, n, 0, Y.as<cd>(), n, 2 * nn, ZERO, dP, N, NN, npara);
    zgemm_strided(st, OP_N, OP_N, n, 2 * n, n, ONE, K + nn, n, 0, Y.as<cd>(), n, 2 * nn, ZERO, dP + n, N, NN,
                  npara);
  } else {
    zdscale(st, n, 2 * n, d_alpha, 0, nullptr, 0, 0, Y.as<cd>(), n, 2 * nn, dP, N, NN, npara);
    zdscale(st, n, 2 * n, d_alpha + n, 0, nullptr, 0, 0, Y.as<cd>(), n, 2 * nn, dP + n, N, NN, npara);
  }
}

void Sampler::shape_derivatives(const HostLayer& L, int nvar, const int* var_index, const double* var_coeff,
                                int npara, cd* dP, cd* dQ, cd* dfe11, cd* dfe22, cd* dfe33, cd* dPl, cd* dQl) {
  shape_tangents(*c_, nx_, ny_, lambda_, pml_, d_alpha_.as<cd>(), pml_ ? d_K_.as<cd>() : nullptr, L, nvar, var_index,
                 var_coeff, npara, dP, dQ, dfe11, dfe22, dfe33, dPl, dQl);
}

VarSampler::VarSampler(vrc_ctx* c, double lambda, int nx, int ny, HostLayer base, bool pml)
    : Sampler(c, lambda, nx, ny, pml), base_(std::move(base)) {
  prepare_layer(base_);  // validates shapes
  post_init(base_.x, base_.y);
}

void VarSampler::add_rule(int axis, int layout, RuleFn fn, void* user) {
  const int lim = axis == 0 ? (int)base_.x.size() : (int)base_.y.size();
  if (layout < 0 || layout >= lim) throw ArgFailure("rule layout index out of range");
  if (!fn) throw ArgFailure("null rule");
  (axis == 0 ? xr_ : yr_)[layout] = {fn, user};
}

HostLayer VarSampler::layer_at(double z) const {
  HostLayer l = base_;
  for (auto& kv : xr_) l.x[kv.first] = kv.second.first(z, kv.second.second);
  for (auto& kv : yr_) l.y[kv.first] = kv.second.first(z, kv.second.second);
  return l;
}

StackSampler::StackSampler(vrc_ctx* c, double lambda, int nx, int ny, std::vector<HostLayer> layers,
                           std::vector<double> z_edges, HostLayer bottom, HostLayer top, bool pml)
    : Sampler(c, lambda, nx, ny, pml), layers_(std::move(layers)), ze_(std::move(z_edges)),
      bottom_(std::move(bottom)), top_(std::move(top)) {
  if (ze_.size() != layers_.size() + 1) throw ArgFailure("need nlayers+1 z edges");
  for (size_t i = 0; i + 1 < ze_.size(); ++i)
    if (!(ze_[i + 1] > ze_[i])) throw ArgFailure("z edges must increase");
  for (auto& l : layers_) prepare_layer(l);
  prepare_layer(bottom_);
  prepare_layer(top_);
  post_init(bottom_.x, bottom_.y);
}

HostLayer StackSampler::layer_at(double z) const {
  if (z <= ze_.front()) return bottom_;
  if (z >= ze_.back()) return top_;
  size_t k = std::upper_bound(ze_.begin(), ze_.end(), z) - ze_.begin() - 1;
  return layers_[k];
}

// ===================================================================== integrators
Integrator::Integrator(vrc_ctx* c, double z0, double z1, Sampler* s) : c_(c), z0_(z0), z1_(z1), s_(s) {
  if (!s) throw ArgFailure("null sampler");
  Nd_ = s->nDim();
  alloc();
}

void Integrator::alloc() {
  const long NN = (long)Nd_ * Nd_;
  S_ = c_->ws.get(sizeof(cd) * 4 * NN);
  W0_ = c_->ws.get(sizeof(cd) * NN);
  V0_ = c_->ws.get(sizeof(cd) * NN);
  W1_ = c_->ws.get(sizeof(cd) * NN);
  V1_ = c_->ws.get(sizeof(cd) * NN);
  g0_ = c_->ws.get(sizeof(cd) * Nd_);
  g1_ = c_->ws.get(sizeof(cd) * Nd_);
  info_ = c_->ws.get(sizeof(int) * 64);
}

SMat Integrator::S() const {
  const long NN = (long)Nd_ * Nd_;
  cd* p = S_.as<cd>();
  SMat s;
  s.Tdd = p;
  s.Tuu = p + NN;
  s.Rud = p + 2 * NN;
  s.Rdu = p + 3 * NN;
  return s;
}

void Integrator::get(int which, cd* host) const {
  const long NN = (long)Nd_ * Nd_;
  const cd* src = nullptr;
  size_t cnt = NN;
  switch (which) {
    case 0: src = S().Tdd; break;
    case 1: src = S().Tuu; break;
    case 2: src = S().Rud; break;
    case 3: src = S().Rdu; break;
    case 4: src = W0_.as<cd>(); break;
    case 5: src = V0_.as<cd>(); break;
    case 6: src = g0_.as<cd>(); cnt = Nd_; break;
    case 7: src = W1_.as<cd>(); break;
    case 8: src = V1_.as<cd>(); break;
    case 9: src = g1_.as<cd>(); cnt = Nd_; break;
    default: throw ArgFailure("unknown output selector");
  }
  VRC_HIP(hipMemcpyAsync(host, src, sizeof(cd) * cnt, hipMemcpyDeviceToHost, c_->stream));
  VRC_HIP(hipStreamSynchronize(c_->stream));
}

void Integrator::eig_at(double z, cd* P, cd* Q, cd* W, cd* V, cd* gam) {
  s_->sample_gpu({z}, P, Q);
  wave_eqn(*c_, Nd_, P, Q, (long)Nd_ * Nd_, W, V, gam, info_.as<int>(), 1);
}

namespace {
struct LUStore {
  DevBuf W, V, pW, pV, inf;
  LUPair pair;
  void make(vrc_ctx& c, int N, const cd* Wsrc, const cd* Vsrc, int batch, long sSrc) {
    const long NN = (long)N * N;
    W = c.ws.get(sizeof(cd) * NN * batch);
    V = c.ws.get(sizeof(cd) * NN * batch);
    pW = c.ws.get(sizeof(int) * (size_t)N * batch);
    pV = c.ws.get(sizeof(int) * (size_t)N * batch);
    inf = c.ws.get(sizeof(int) * (size_t)batch);
    fill_int(c.stream, inf.as<int>(), 0, batch);
    zgeam(c.stream, N, N, ONE, Wsrc, N, sSrc, ZERO, nullptr, 0, 0, W.as<cd>(), N, NN, batch);
    zgeam(c.stream, N, N, ONE, Vsrc, N, sSrc, ZERO, nullptr, 0, 0, V.as<cd>(), N, NN, batch);
    zgetrf(c.stream, N, W.as<cd>(), N, NN, pW.as<int>(), N, inf.as<int>(), batch, c.ws);
    zgetrf(c.stream, N, V.as<cd>(), N, NN, pV.as<int>(), N, inf.as<int>(), batch, c.ws);
    pair.W = W.as<cd>();
    pair.V = V.as<cd>();
    pair.pW = pW.as<int>();
    pair.pV = pV.as<int>();
    pair.sM = NN;
  }
};

void init_identity_S(vrc_ctx& c, int N, SMat& S, int batch) {
  const long NN = (long)N * N;
  zset(c.stream, N, N, ONE, ZERO, S.Tdd, N, NN, batch);
  zset(c.stream, N, N, ONE, ZERO, S.Tuu, N, NN, batch);
  zset(c.stream, N, N, ZERO, ZERO, S.Rud, N, NN, batch);
  zset(c.stream, N, N, ZERO, ZERO, S.Rdu, N, NN, batch);
}
}  // namespace

// RCWAIntegrator::compute (src/core/RCWAIntegrator.cpp:20-87) over explicit slices
// (midpoint z, width). The slices are independent until the cascade, so their samples,
// eigensolves and slab S-matrices run in groups of up to g_rcwa_group per launch (one
// eigensolve per workgroup), the two end media riding with the first group; the star
// products then follow in slice order, as the reference cascades them. Each matrix's
// result does not depend on the group it is solved in.
static int g_rcwa_group = 64;
void rcwa_set_group(int g) { g_rcwa_group = g < 1 ? 1 : g; }
void Integrator::rcwa(const std::vector<std::pair<double, double>>& slices, bool) {
  vrc_ctx& c = *c_;
  const int N = Nd_;
  tune_apply(N);
  const long NN = (long)N * N;
  tree_.clear();
  const int ns = (int)slices.size();
  const int G = std::max(1, std::min(ns, g_rcwa_group));
  const int GE = G + 2;  // + the two end media in the first launch
  DevBuf Pb = c.ws.get(sizeof(cd) * NN * GE), Qb = c.ws.get(sizeof(cd) * NN * GE);
  DevBuf Wb = c.ws.get(sizeof(cd) * NN * GE), Vb = c.ws.get(sizeof(cd) * NN * GE), gb = c.ws.get(sizeof(cd) * N * GE);
  DevBuf tu = c.ws.get(sizeof(cd) * NN * G), ru = c.ws.get(sizeof(cd) * NN * G);
  DevBuf infb = c.ws.get(sizeof(int) * GE);
  SMat S = this->S();
  init_identity_S(c, N, S, 1);
  const double k0 = s_->k0();
  LUStore lu1;
  for (int i0 = 0; i0 < ns; i0 += G) {
    const int g = std::min(G, ns - i0);
    const int e = i0 == 0 ? 2 : 0;
    std::vector<double> zs, sv(g);
    if (e) zs = {z1_, z0_};
    for (int j = 0; j < g; ++j) {
      zs.push_back(slices[i0 + j].first);
      sv[j] = slices[i0 + j].second / k0;
    }
    s_->sample_gpu(zs, Pb.as<cd>(), Qb.as<cd>());
    wave_eqn(c, N, Pb.as<cd>(), Qb.as<cd>(), NN, Wb.as<cd>(), Vb.as<cd>(), gb.as<cd>(), infb.as<int>(), g + e);
    if (e) {
      auto cp = [&](int rows, int cols, const cd* src, DevBuf& dst) {
        zgeam(c.stream, rows, cols, ONE, src, rows, 0, ZERO, nullptr, 0, 0, dst.as<cd>(), rows, 0, 1);
      };
      cp(N, N, Wb.as<cd>(), W1_);
      cp(N, N, Vb.as<cd>(), V1_);
      cp(N, N, Wb.as<cd>() + NN, W0_);
      cp(N, N, Vb.as<cd>() + NN, V0_);
      cp(N, 1, gb.as<cd>(), g1_);
      cp(N, 1, gb.as<cd>() + N, g0_);
      lu1.make(c, N, W1_.as<cd>(), V1_.as<cd>(), 1, 0);
    }
    DevBuf ds = upload(c, sv.data(), sizeof(double) * g);
    layer_smatrix(c, N, g, ds.as<double>(), Wb.as<cd>() + e * NN, Vb.as<cd>() + e * NN, gb.as<cd>() + e * N,
                  W1_.as<cd>(), V1_.as<cd>(), 0, tu.as<cd>(), ru.as<cd>());
    for (int j = 0; j < g; ++j) {
      const cd* t = tu.as<cd>() + j * NN;
      const cd* r = ru.as<cd>() + j * NN;
      star_product(c, N, 1, S, t, t, r, r, 0);
    }
  }
  project_left(c, N, 1, S, lu1.pair, W0_.as<cd>(), V0_.as<cd>(), 0);
  N_ = ns;
  nEig_ = ns + 2;
  VRC_HIP(hipStreamSynchronize(c.stream));
}

// ===================================================================== varRCWA
namespace {
// One batched group of samples P, Q (count matrices); the first neig of them also hold
// their eigen systems W, V, gamma and the LU factors of W and V (reference systems).
struct Level {
  int count = 0, neig = 0;
  DevBuf P, Q, W, V, g, inf;
  LUStore lu;
};
using LevelP = std::shared_ptr<Level>;
struct Smp {
  LevelP lv;
  int j = 0;
  const cd* P(long NN) const { return lv->P.as<cd>() + j * NN; }
  const cd* Q(long NN) const { return lv->Q.as<cd>() + j * NN; }
};
struct Ref {
  LevelP lv;
  int j = 0;
  RefSys sys(int N) const {
    const long NN = (long)N * N;
    RefSys r;
    r.W = lv->W.as<cd>() + j * NN;
    r.V = lv->V.as<cd>() + j * NN;
    r.gam = lv->g.as<cd>() + (long)j * N;
    r.LU = lv->lu.pair;
    r.LU.W += j * NN;
    r.LU.V += j * NN;
    r.LU.pW += (long)j * N;
    r.LU.pV += (long)j * N;
    return r;
  }
};
struct Seg {
  int inst, depth;
  double za, zb;
  Ref R;
  Smp a, m, b;
};
struct Done {
  double za, zb;
  Ref R;
  std::shared_ptr<DevBuf> T;  // the level's segment S-matrices
  long off;                   // this segment's Tdd, Tuu, Rud, Rdu (basis of R) at T + off
};

LevelP new_level(vrc_ctx& c, int N, int count) {
  const long NN = (long)N * N;
  auto L = std::make_shared<Level>();
  L->count = count;
  L->P = c.ws.get(sizeof(cd) * NN * count);
  L->Q = c.ws.get(sizeof(cd) * NN * count);
  return L;
}

// Eigen systems of the first neig samples of L (one batched launch) and the LU factors
// of W and V for the first nlu of them.
void solve_level(vrc_ctx& c, int N, Level& L, int neig, int nlu) {
  const long NN = (long)N * N;
  L.neig = neig;
  if (neig <= 0) return;
  L.W = c.ws.get(sizeof(cd) * NN * neig);
  L.V = c.ws.get(sizeof(cd) * NN * neig);
  L.g = c.ws.get(sizeof(cd) * (size_t)N * neig);
  L.inf = c.ws.get(sizeof(int) * (size_t)neig);
  wave_eqn(c, N, L.P.as<cd>(), L.Q.as<cd>(), NN, L.W.as<cd>(), L.V.as<cd>(), L.g.as<cd>(), L.inf.as<int>(), neig);
  if (nlu > 0) L.lu.make(c, N, L.W.as<cd>(), L.V.as<cd>(), nlu, NN);
}
}  // namespace

// RedhefferIntegrator::compute, trisection branch (src/core/RedhefferIntegrator.cpp:174-314),
// for one integrator.
void Integrator::redheffer(double pe, int max_depth) { redheffer_many({this}, pe, max_depth); }

// The reference recurses depth first and solves each new reference eigensystem on its
// own. Whether a segment is accepted depends only on the segment (its reference system
// and its three samples), so the same tree grows level by level: every pending segment
// of a level is evaluated (perturbation and error, one readback for all), the accepted
// ones keep their S-matrices, and the rejected ones are trisected with their new
// samples taken together and their new reference eigensystems solved in one batched
// launch. The accepted segments are then projected and star-multiplied in z order,
// exactly as the recursion does (src/core/RedhefferIntegrator.cpp:248-276), and the
// right projection follows (:211-223).
void Integrator::redheffer_many(const std::vector<Integrator*>& its, double pe, int max_depth) {
  if (its.empty()) return;
  vrc_ctx& c = *its[0]->c_;
  const int N = its[0]->Nd_;
  tune_apply(N);
  for (auto* it : its)
    if (it->c_ != &c || it->Nd_ != N) throw ArgFailure("integrators of one batch need one context and equal N");
  const long NN = (long)N * N;
  const int I = (int)its.size();
  // ---- level 0: the midpoint references (with LU), then both ends, of every integrator
  LevelP L0 = new_level(c, N, 3 * I);
  for (int i = 0; i < I; ++i) {
    Integrator& it = *its[i];
    it.tree_.clear();
    it.N_ = 0;
    it.nEig_ = 2;
    it.s_->sample_gpu({0.5 * (it.z0_ + it.z1_)}, L0->P.as<cd>() + i * NN, L0->Q.as<cd>() + i * NN);
    it.s_->sample_gpu({it.z0_}, L0->P.as<cd>() + (I + i) * NN, L0->Q.as<cd>() + (I + i) * NN);
    it.s_->sample_gpu({it.z1_}, L0->P.as<cd>() + (2 * I + i) * NN, L0->Q.as<cd>() + (2 * I + i) * NN);
  }
  solve_level(c, N, *L0, 3 * I, I);
  std::vector<Seg> pending;
  for (int i = 0; i < I; ++i) {
    Integrator& it = *its[i];
    auto cp = [&](int rows, int cols, const cd* src, DevBuf& dst) {
      zgeam(c.stream, rows, cols, ONE, src, rows, 0, ZERO, nullptr, 0, 0, dst.as<cd>(), rows, 0, 1);
    };
    cp(N, N, L0->W.as<cd>() + (I + i) * NN, it.W0_);
    cp(N, N, L0->V.as<cd>() + (I + i) * NN, it.V0_);
    cp(N, 1, L0->g.as<cd>() + (long)(I + i) * N, it.g0_);
    cp(N, N, L0->W.as<cd>() + (2 * I + i) * NN, it.W1_);
    cp(N, N, L0->V.as<cd>() + (2 * I + i) * NN, it.V1_);
    cp(N, 1, L0->g.as<cd>() + (long)(2 * I + i) * N, it.g1_);
    pending.push_back(Seg{i, 0, it.z0_, it.z1_, Ref{L0, i}, Smp{L0, I + i}, Smp{L0, i}, Smp{L0, 2 * I + i}});
  }
  std::vector<std::vector<Done>> done(I);
  while (!pending.empty()) {
    // ---- evaluate every pending segment in one batch, one readback
    const int P = (int)pending.size();
    std::vector<RhJob> jobs(P);
    for (int q = 0; q < P; ++q) {
      const Seg& sg = pending[q];
      RhJob& j = jobs[q];
      j.k0 = its[sg.inst]->s_->k0();
      j.z0 = sg.za;
      j.z1 = sg.zb;
      const Smp pr{sg.R.lv, sg.R.j};
      j.Pr = pr.P(NN);
      j.Qr = pr.Q(NN);
      j.R = sg.R.sys(N);
      j.P3[0] = sg.a.P(NN);
      j.P3[1] = sg.m.P(NN);
      j.P3[2] = sg.b.P(NN);
      j.Q3[0] = sg.a.Q(NN);
      j.Q3[1] = sg.m.Q(NN);
      j.Q3[2] = sg.b.Q(NN);
    }
    auto Tl = std::make_shared<DevBuf>(c.ws.get(sizeof(cd) * 4 * NN * P));
    DevBuf errb = c.ws.get(sizeof(double) * RH_ERR_BLOCKS * (size_t)P);
    for (int q0 = 0; q0 < P; q0 += 128) {  // bounded work space per batch
      const int q1 = std::min(P, q0 + 128);
      redheffer_perturbation_batch(c, N, std::vector<RhJob>(jobs.begin() + q0, jobs.begin() + q1),
                                   Tl->as<cd>() + 4 * NN * q0, errb.as<double>() + (size_t)RH_ERR_BLOCKS * q0);
    }
    std::vector<double> eh(RH_ERR_BLOCKS * (size_t)P);
    VRC_HIP(hipMemcpyAsync(eh.data(), errb.as<double>(), sizeof(double) * eh.size(), hipMemcpyDeviceToHost,
                           c.stream));
    VRC_HIP(hipStreamSynchronize(c.stream));
    // ---- accept, or trisect
    std::vector<int> split;
    for (int q = 0; q < P; ++q) {
      const Seg& sg = pending[q];
      double err = 0.0;
      for (int k = 0; k < RH_ERR_BLOCKS; ++k) err = std::max(err, eh[(size_t)RH_ERR_BLOCKS * q + k]);
      if (err <= pe || sg.depth >= max_depth) {
        if (err > pe && !c.host_code) {
          c.host_code = H_DEPTH;
          c.host_msg = std::string(health_text(H_DEPTH)) + " (segment [" + std::to_string(sg.za) + ", " +
                       std::to_string(sg.zb) + "], error " + std::to_string(err) + ", depth " +
                       std::to_string(max_depth) + ")";
        }
        done[sg.inst].push_back(Done{sg.za, sg.zb, sg.R, Tl, 4 * NN * q});
      } else {
        split.push_back(q);
      }
    }
    Tl.reset();
    const int S = (int)split.size();
    std::vector<Seg> next;
    if (S > 0) {
      // slots: [0, S) new left references (midpoint of the first third), [S, 2S) new right
      // references (midpoint of the last third), [2S, 3S) first trisection point,
      // [3S, 4S) second trisection point
      LevelP L = new_level(c, N, 4 * S);
      // one sampler call per slot group and run of segments of one integrator (the split
      // segments keep the integrators' order)
      for (int k = 0; k < 4; ++k) {
        for (int u0 = 0; u0 < S;) {
          const int inst = pending[split[u0]].inst;
          int u1 = u0;
          std::vector<double> zs;
          while (u1 < S && pending[split[u1]].inst == inst) {
            const Seg& sg = pending[split[u1]];
            const double zm0 = 2. / 3. * sg.za + 1. / 3. * sg.zb;
            const double zm1 = 1. / 3. * sg.za + 2. / 3. * sg.zb;
            const double zz[4] = {0.5 * (sg.za + zm0), 0.5 * (zm1 + sg.zb), zm0, zm1};
            zs.push_back(zz[k]);
            ++u1;
          }
          its[inst]->s_->sample_gpu(zs, L->P.as<cd>() + (k * S + u0) * NN, L->Q.as<cd>() + (k * S + u0) * NN);
          u0 = u1;
        }
      }
      solve_level(c, N, *L, 2 * S, 2 * S);
      for (int u = 0; u < S; ++u) {
        const Seg& sg = pending[split[u]];
        const double zm0 = 2. / 3. * sg.za + 1. / 3. * sg.zb;
        const double zm1 = 1. / 3. * sg.za + 2. / 3. * sg.zb;
        const Smp st{L, u}, sk2{L, S + u}, sk{L, 2 * S + u}, st2{L, 3 * S + u};
        next.push_back(Seg{sg.inst, sg.depth + 1, sg.za, zm0, Ref{L, u}, sg.a, st, sk});
        next.push_back(Seg{sg.inst, sg.depth + 1, zm0, zm1, sg.R, sk, sg.m, st2});
        next.push_back(Seg{sg.inst, sg.depth + 1, zm1, sg.zb, Ref{L, S + u}, st2, sk2, sg.b});
      }
    }
    pending.swap(next);
  }
  // ---- per integrator: project each accepted segment to the left basis and cascade, in z order
  for (int i = 0; i < I; ++i) {
    Integrator& it = *its[i];
    auto& D = done[i];
    std::sort(D.begin(), D.end(), [](const Done& x, const Done& y) { return x.za < y.za; });
    SMat S = it.S();
    init_identity_S(c, N, S, 1);
    const cd* leftW = it.W0_.as<cd>();
    const cd* leftV = it.V0_.as<cd>();
    for (auto& d : D) {
      cd* T0 = d.T->as<cd>() + d.off;
      SMat T;
      T.Tdd = T0;
      T.Tuu = T0 + NN;
      T.Rud = T0 + 2 * NN;
      T.Rdu = T0 + 3 * NN;
      const RefSys R = d.R.sys(N);
      project_left(c, N, 1, T, R.LU, leftW, leftV, 0);
      leftW = R.W;
      leftV = R.V;
      star_product(c, N, 1, S, T.Tdd, T.Tuu, T.Rud, T.Rdu, 0);
      it.tree_.push_back({d.za, d.zb});
      ++it.N_;
      d.T.reset();
    }
    LUStore lul;
    lul.make(c, N, leftW, leftV, 1, 0);
    project_right(c, N, 1, S, lul.pair, it.W1_.as<cd>(), it.V1_.as<cd>(), 0);
    D.clear();  // releases the segments' references once their work is queued
  }
  VRC_HIP(hipStreamSynchronize(c.stream));
}

// ===================================================================== differential method
// DifferentialIntegrator::compute (src/core/DifferentialIntegrator.cpp:21-134).
void Integrator::differential(int Nsteps) {
  vrc_ctx& c = *c_;
  hipStream_t st = c.stream;
  const int N = Nd_;
  tune_apply(N);
  const long NN = (long)N * N;
  const double k0 = s_->k0();
  tree_.clear();
  DevBuf P = c.ws.get(sizeof(cd) * NN), Q = c.ws.get(sizeof(cd) * NN);
  eig_at(z1_, P.as<cd>(), Q.as<cd>(), W1_.as<cd>(), V1_.as<cd>(), g1_.as<cd>());
  LUStore lu1;
  lu1.make(c, N, W1_.as<cd>(), V1_.as<cd>(), 1, 0);
  SMat S = this->S();
  init_identity_S(c, N, S, 1);
  const double dz = (z1_ - z0_) / (double)Nsteps;
  DevBuf buf = c.ws.get(sizeof(cd) * NN * 30);
  cd* b = buf.as<cd>();
  cd *P0 = b, *Q0 = b + NN, *P1 = b + 2 * NN, *Q1 = b + 3 * NN, *P2 = b + 4 * NN, *Q2 = b + 5 * NN;
  cd *X1 = b + 6 * NN, *X2 = b + 7 * NN, *X3 = b + 8 * NN, *Y1 = b + 9 * NN, *Y2 = b + 10 * NN, *Y3 = b + 11 * NN;
  cd *L11 = b + 12 * NN, *L12 = b + 13 * NN, *L21 = b + 14 * NN, *L22 = b + 15 * NN;
  cd *iW11 = b + 16 * NN, *iW12 = b + 17 * NN, *iV21 = b + 18 * NN, *iV22 = b + 19 * NN;
  cd *A11 = b + 20 * NN, *A12 = b + 21 * NN, *A21 = b + 22 * NN, *A22 = b + 23 * NN;
  cd *T11 = b + 24 * NN, *T12 = b + 25 * NN, *T21 = b + 26 * NN, *T22 = b + 27 * NN, *tdd = b + 28 * NN,
     *rdu = b + 29 * NN;
  DevBuf extra = c.ws.get(sizeof(cd) * NN * 2);
  cd *tuu = extra.as<cd>(), *rud = extra.as<cd>() + NN;
  DevBuf piv = c.ws.get(sizeof(int) * N), inf = c.ws.get(sizeof(int));
  auto mm = [&](cd al, const cd* A, const cd* B, cd be, cd* C) {
    zgemm_strided(st, OP_N, OP_N, N, N, N, al, A, N, 0, B, N, 0, be, C, N, 0, 1);
  };
  auto axpy = [&](cd al, const cd* A, cd* C) { zgeam(st, N, N, al, A, N, 0, ONE, C, N, 0, C, N, 0, 1); };
  s_->sample_gpu({z0_}, P0, Q0);
  const double a = dz * dz / (6. * k0 * k0), bb = dz * dz / (4. * k0 * k0), cc = dz * dz / (2. * k0 * k0);
  const cd ci = mk(0.0, dz / (6. * k0));
  for (int i = 0; i < Nsteps; ++i) {
    const double zl = z0_ + i * dz, zr = z0_ + (i + 1) * dz, zm = 0.5 * (zl + zr);
    s_->sample_gpu({zm}, P1, Q1);
    s_->sample_gpu({zr}, P2, Q2);
    mm(ONE, P1, Q0, ZERO, X1);
    mm(ONE, P1, Q1, ZERO, X2);
    mm(ONE, P2, Q1, ZERO, X3);
    mm(ONE, Q1, P0, ZERO, Y1);
    mm(ONE, Q1, P1, ZERO, Y2);
    mm(ONE, Q2, P1, ZERO, Y3);
    // L11 = I - a (X1 + X2 + X3) + a b X3 X1
    zset(st, N, N, ONE, ZERO, L11, N, 0, 1);
    axpy(mk(-a, 0), X1, L11);
    axpy(mk(-a, 0), X2, L11);
    axpy(mk(-a, 0), X3, L11);
    mm(mk(a * bb, 0), X3, X1, ONE, L11);
    // L22 = I - a (Y1 + Y2 + Y3) + a b Y3 Y1
    zset(st, N, N, ONE, ZERO, L22, N, 0, 1);
    axpy(mk(-a, 0), Y1, L22);
    axpy(mk(-a, 0), Y2, L22);
    axpy(mk(-a, 0), Y3, L22);
    mm(mk(a * bb, 0), Y3, Y1, ONE, L22);
    // L12 = ci (P0 + 4 P1 + P2) - ci cc (X2 P0 + X3 P1)
    zgeam(st, N, N, ci, P0, N, 0, cscale(ci, 4.0), P1, N, 0, L12, N, 0, 1);
    axpy(ci, P2, L12);
    mm(cscale(ci, -cc), X2, P0, ONE, L12);
    mm(cscale(ci, -cc), X3, P1, ONE, L12);
    // L21 = ci (Q0 + 4 Q1 + Q2) - ci cc (Y2 Q0 + Y3 Q1)
    zgeam(st, N, N, ci, Q0, N, 0, cscale(ci, 4.0), Q1, N, 0, L21, N, 0, 1);
    axpy(ci, Q2, L21);
    mm(cscale(ci, -cc), Y2, Q0, ONE, L21);
    mm(cscale(ci, -cc), Y3, Q1, ONE, L21);
    zgetrs(st, N, N, lu1.pair.W, N, 0, lu1.pair.pW, N, L11, N, 0, iW11, N, 0, 1, c.ws);
    zgetrs(st, N, N, lu1.pair.W, N, 0, lu1.pair.pW, N, L12, N, 0, iW12, N, 0, 1, c.ws);
    zgetrs(st, N, N, lu1.pair.V, N, 0, lu1.pair.pV, N, L21, N, 0, iV21, N, 0, 1, c.ws);
    zgetrs(st, N, N, lu1.pair.V, N, 0, lu1.pair.pV, N, L22, N, 0, iV22, N, 0, 1, c.ws);
    const cd h = mk(0.5, 0), mh = mk(-0.5, 0);
    zgeam(st, N, N, h, iW11, N, 0, h, iV21, N, 0, A11, N, 0, 1);
    zgeam(st, N, N, h, iW12, N, 0, h, iV22, N, 0, A12, N, 0, 1);
    zgeam(st, N, N, h, iW11, N, 0, mh, iV21, N, 0, A21, N, 0, 1);
    zgeam(st, N, N, h, iW12, N, 0, mh, iV22, N, 0, A22, N, 0, 1);
    const cd* W1 = W1_.as<cd>();
    const cd* V1 = V1_.as<cd>();
    mm(ONE, A11, W1, ZERO, T11);
    mm(ONE, A12, V1, ONE, T11);
    mm(ONE, A11, W1, ZERO, T12);
    mm(MONE, A12, V1, ONE, T12);
    mm(ONE, A21, W1, ZERO, T21);
    mm(ONE, A22, V1, ONE, T21);
    mm(ONE, A21, W1, ZERO, T22);
    mm(MONE, A22, V1, ONE, T22);
    // tdd = T22^{-1}; rud = T12 tdd; rdu = -T22^{-1} T21; tuu = T11 + T12 rdu
    fill_int(st, inf.as<int>(), 0, 1);
    zgetrf(st, N, T22, N, 0, piv.as<int>(), N, inf.as<int>(), 1, c.ws);
    zgetri(st, N, T22, N, 0, piv.as<int>(), N, tdd, N, 0, 1, c.ws);
    mm(ONE, T12, tdd, ZERO, rud);
    zgetrs(st, N, N, T22, N, 0, piv.as<int>(), N, T21, N, 0, rdu, N, 0, 1, c.ws);
    zgeam(st, N, N, MONE, rdu, N, 0, ZERO, nullptr, 0, 0, rdu, N, 0, 1);
    zgeam(st, N, N, ONE, T11, N, 0, ZERO, nullptr, 0, 0, tuu, N, 0, 1);
    mm(ONE, T12, rdu, ONE, tuu);
    star_product(c, N, 1, S, tdd, tuu, rud, rdu, 0);
    std::swap(P0, P2);
    std::swap(Q0, Q2);
  }
  eig_at(z0_, P.as<cd>(), Q.as<cd>(), W0_.as<cd>(), V0_.as<cd>(), g0_.as<cd>());
  project_left(c, N, 1, S, lu1.pair, W0_.as<cd>(), V0_.as<cd>(), 0);
  N_ = Nsteps;
  nEig_ = 2;
  VRC_HIP(hipStreamSynchronize(st));
}

// ===================================================================== batched stacks
// Cascade of one group of GB instances (slab S-matrices tu/ru [instance][slab], end systems
// We/Ve/ge [bottoms | tops]) and the end projection, batched over the group; results into
// the outputs at instance offset I0.
// tape (optional): receives the prefix before every K-th layer and the cascade before the
// projection of the group's instances.
static void cascade_group(vrc_ctx& c, int N, int L, int GB, int I0, bool homog, const cd* tu, const cd* ru,
                          const cd* We, const cd* Ve, const cd* ge, cd* S_out, cd* W0o, cd* V0o, cd* g0o, cd* W1o,
                          cd* V1o, cd* g1o, StackTape* tape = nullptr) {
  hipStream_t st = c.stream;
  const long NN = (long)N * N;
  const int B = GB, i0 = I0;
  const cd* W0 = We;
  const cd* V0 = Ve;
  const cd* W1 = We + (long)GB * NN;
  const cd* V1 = Ve + (long)GB * NN;
  const cd* g0 = ge;
  const cd* g1 = ge + (long)GB * N;
  LUStore lu1;
  if (!homog) lu1.make(c, N, W1, V1, B, NN);
  DevBuf Sb = c.ws.get(sizeof(cd) * 4 * NN * B);
  SMat S;
  S.Tdd = Sb.as<cd>();
  S.Tuu = S.Tdd + B * NN;
  S.Rud = S.Tdd + 2 * B * NN;
  S.Rdu = S.Tdd + 3 * B * NN;
  // identity * slab 0 = slab 0: the cascade starts from the first slab's S-matrix
  zgeam(st, N, N, ONE, tu, N, (long)L * NN, ZERO, nullptr, 0, 0, S.Tdd, N, NN, B);
  zgeam(st, N, N, ONE, tu, N, (long)L * NN, ZERO, nullptr, 0, 0, S.Tuu, N, NN, B);
  zgeam(st, N, N, ONE, ru, N, (long)L * NN, ZERO, nullptr, 0, 0, S.Rud, N, NN, B);
  zgeam(st, N, N, ONE, ru, N, (long)L * NN, ZERO, nullptr, 0, 0, S.Rdu, N, NN, B);
  // tape layout: 4 blocks [Tdd|Tuu|Rud|Rdu] per instance
  auto keep_S = [&](cd* dst, long per_inst) {
    const SMat* src = &S;
    const cd* blk[4] = {src->Tdd, src->Tuu, src->Rud, src->Rdu};
    for (int q = 0; q < 4; ++q)
      zgeam(st, N, N, ONE, blk[q], N, NN, ZERO, nullptr, 0, 0, dst + (long)i0 * per_inst + q * NN, N, per_inst, B);
  };
  const int nb = tape ? tape->nblocks() : 0;
  for (int l = 1; l < L; ++l) {
    if (tape && l % tape->K == 0)
      keep_S(tape->ckpt.as<cd>() + (long)(l / tape->K - 1) * 4 * NN, (long)(nb - 1) * 4 * NN);
    const cd* t = tu + l * NN;
    const cd* r = ru + l * NN;
    star_product(c, N, B, S, t, t, r, r, (long)L * NN);
  }
  if (tape) keep_S(tape->Sfin.as<cd>(), 4 * NN);
  if (homog)
    project_left_homog(c, N, B, S, V0, V1, NN);  // W0 = W1 = I
  else
    project_left(c, N, B, S, lu1.pair, W0, V0, NN);
  // ---- outputs
  const long so = 4 * NN;
  cd* So = S_out + (long)i0 * so;
  zgeam(st, N, N, ONE, S.Tdd, N, NN, ZERO, nullptr, 0, 0, So, N, so, B);
  zgeam(st, N, N, ONE, S.Tuu, N, NN, ZERO, nullptr, 0, 0, So + NN, N, so, B);
  zgeam(st, N, N, ONE, S.Rud, N, NN, ZERO, nullptr, 0, 0, So + 2 * NN, N, so, B);
  zgeam(st, N, N, ONE, S.Rdu, N, NN, ZERO, nullptr, 0, 0, So + 3 * NN, N, so, B);
  auto out = [&](cd* dst, const cd* src, long sz) {
    if (dst) zgeam(st, (int)sz, 1, ONE, src, sz, sz, ZERO, nullptr, 0, 0, dst + (long)i0 * sz, sz, sz, B);
  };
  out(W0o, W0, NN);
  out(V0o, V0, NN);
  out(g0o, g0, N);
  out(W1o, W1, NN);
  out(V1o, V1, NN);
  out(g1o, g1, N);
}

// Forward path for instances whose end media are single-valued (closed-form eigen-systems):
// the slab eigensolves of a cascade group are split evenly into launches of at most
// `per_launch` slabs (0: zgeev_launch_plan, full launches of the QR kernels and one launch of
// the rest), regardless of instance boundaries (the QR kernels run one to four matrices per
// CU, so launches that fill the kernels' slots matter more than whole instances).
static void batch_stack_homog(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst, int per_launch,
                              cd* S_out, cd* W0o, cd* V0o, cd* g0o, cd* W1o, cd* V1o, cd* g1o, int* flags) {
  hipStream_t st = c.stream;
  const int n = nx * ny, N = 2 * n;
  const long NN = (long)N * N;
  const int ninst = (int)inst.size();
  const int L = (int)inst[0].layers.size();
  const int gsz = cascade_group_cap(N, L);
  const int ngroups = (ninst + gsz - 1) / gsz;
  for (int g = 0; g < ngroups; ++g) {
    const int I0 = (int)((long)ninst * g / ngroups), GB = (int)((long)ninst * (g + 1) / ngroups) - I0;
    DevBuf tuG = c.ws.get(sizeof(cd) * NN * GB * L), ruG = c.ws.get(sizeof(cd) * NN * GB * L);
    DevBuf WeG = c.ws.get(sizeof(cd) * NN * 2 * GB), VeG = c.ws.get(sizeof(cd) * NN * 2 * GB),
           geG = c.ws.get(sizeof(cd) * (size_t)N * 2 * GB);
    // per instance: alpha/beta, k0^2; end media [bottoms | tops] in closed form
    std::vector<cd> abI((size_t)GB * 2 * n), ev(2 * GB);
    std::vector<double> k0I(GB), k2I(GB), k2E(2 * GB);
    for (int b = 0; b < GB; ++b) {
      const StackInstance& I = inst[I0 + b];
      const double k0 = 2 * PI_REF / I.lambda;
      std::vector<cd> al, be;
      alpha_beta(nx, ny, I.bottom->x.back() - I.bottom->x.front(), I.bottom->y.back() - I.bottom->y.front(), al, be);
      std::copy(al.begin(), al.end(), abI.begin() + (size_t)b * 2 * n);
      std::copy(be.begin(), be.end(), abI.begin() + (size_t)b * 2 * n + n);
      k0I[b] = k0;
      k2I[b] = k0 * k0;
      k2E[b] = k2E[GB + b] = k0 * k0;
      ev[b] = I.bottom->eps[0];
      ev[GB + b] = I.top->eps[0];
    }
    {
      std::vector<cd> abE((size_t)2 * GB * 2 * n);
      std::copy(abI.begin(), abI.end(), abE.begin());
      std::copy(abI.begin(), abI.end(), abE.begin() + (size_t)GB * 2 * n);
      DevBuf dabE = upload(c, abE.data(), sizeof(cd) * abE.size());
      DevBuf dk2E = upload(c, k2E.data(), sizeof(double) * k2E.size());
      DevBuf dev = upload(c, ev.data(), sizeof(cd) * ev.size());
      homogeneous_eig(st, n, 2 * GB, dabE.as<cd>(), dabE.as<cd>() + n, 2 * n, dk2E.as<double>(), dev.as<cd>(),
                      WeG.as<cd>(), VeG.as<cd>(), geG.as<cd>());
    }
    const cd* V1 = VeG.as<cd>() + (long)GB * NN;
    // slab launches over the group's GB * L slabs (instance-major)
    const int SLall = GB * L;
    std::vector<int> lstart{0};
    if (per_launch > 0) {
      const int nch = (SLall + per_launch - 1) / per_launch;
      for (int ch = 1; ch <= nch; ++ch) lstart.push_back((int)((long)SLall * ch / nch));
    } else {
      for (int sz : zgeev_launch_plan(N, SLall, 1)) lstart.push_back(lstart.back() + sz);
    }
    for (size_t ch = 0; ch + 1 < lstart.size(); ++ch) {
      const int s0 = lstart[ch], SL = lstart[ch + 1] - s0;
      std::vector<const HostLayer*> jl(SL);
      std::vector<double> k2(SL), sv(SL);
      std::vector<cd> ab((size_t)SL * 2 * n);
      std::vector<int> owner(SL);
      for (int q = 0; q < SL; ++q) {
        const int b = (s0 + q) / L, l = (s0 + q) % L;
        const StackInstance& I = inst[I0 + b];
        jl[q] = I.layers[l];
        k2[q] = k2I[b];
        sv[q] = I.thickness[l] / k0I[b];
        owner[q] = b;
        std::copy(abI.begin() + (size_t)b * 2 * n, abI.begin() + (size_t)(b + 1) * 2 * n, ab.begin() + (size_t)q * 2 * n);
      }
      DevBuf dk2 = upload(c, k2.data(), sizeof(double) * SL);
      DevBuf dab = upload(c, ab.data(), sizeof(cd) * ab.size());
      DevBuf dsv = upload(c, sv.data(), sizeof(double) * SL);
      DevBuf down = upload(c, owner.data(), sizeof(int) * SL);
      AssemblyBatch a;
      a.nx = nx;
      a.ny = ny;
      a.nlayers = SL;
      a.layers = jl;
      a.periodic = true;
      a.d_alpha = dab.as<cd>();
      a.d_beta = a.d_alpha + n;
      a.sab = 2 * n;
      a.d_k0sq = dk2.as<double>();
      DevBuf P = c.ws.get(sizeof(cd) * NN * SL), Q = c.ws.get(sizeof(cd) * NN * SL),
             PQ = c.ws.get(sizeof(cd) * NN * SL);
      assemble_pq(c, a, P.as<cd>(), Q.as<cd>(), PQ.as<cd>());
      DevBuf gam = c.ws.get(sizeof(cd) * (size_t)N * SL), inf = c.ws.get(sizeof(int) * SL);
      fill_int(st, inf.as<int>(), 0, SL);
      DevBuf Zb = c.ws.get(sizeof(cd) * NN * SL), Xb = c.ws.get(sizeof(cd) * NN * SL),
             Db = c.ws.get(sizeof(cd) * (size_t)N * SL);
      wave_eqn_factors(c, N, P.as<cd>(), Q.as<cd>(), NN, Zb.as<cd>(), Xb.as<cd>(), Db.as<cd>(), gam.as<cd>(),
                       inf.as<int>(), SL, PQ.as<cd>());
      if (flags) info_to_owner(st, inf.as<int>(), down.as<int>(), SL, H_EIG, flags + I0);
      PQ = DevBuf();
      DevBuf Vr = c.ws.get(sizeof(cd) * NN * SL);
      gather_mats(st, NN, V1, NN, down.as<int>(), Vr.as<cd>(), NN, SL);  // W1 = I
      layer_smatrix_factors(c, N, SL, dsv.as<double>(), P.as<cd>(), NN, Zb.as<cd>(), Xb.as<cd>(), Db.as<cd>(),
                            gam.as<cd>(), nullptr, Vr.as<cd>(), NN, tuG.as<cd>() + (long)s0 * NN,
                            ruG.as<cd>() + (long)s0 * NN, true);
    }
    cascade_group(c, N, L, GB, I0, true, tuG.as<cd>(), ruG.as<cd>(), WeG.as<cd>(), VeG.as<cd>(), geG.as<cd>(), S_out,
                  W0o, V0o, g0o, W1o, V1o, g1o);
  }
}

static void batch_stack_solve(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst, int chunk,
                              cd* S_out, cd* W0o, cd* V0o, cd* g0o, cd* W1o, cd* V1o, cd* g1o, StackTape* tape,
                              int* flags);

void batch_stack_rcwa(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst, int chunk, cd* S_out,
                      cd* W0o, cd* V0o, cd* g0o, cd* W1o, cd* V1o, cd* g1o, StackTape* tape, int* flags) {
  tune_apply(2 * nx * ny);
  batch_stack_solve(c, nx, ny, inst, chunk, S_out, W0o, V0o, g0o, W1o, V1o, g1o, tape, flags);
  const long NN = 4L * nx * ny * nx * ny;
  if (flags) nonfinite_flags(c.stream, S_out, 4 * NN, 4 * NN, (int)inst.size(), flags, H_NONFINITE);
}

static void batch_stack_solve(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst, int chunk,
                              cd* S_out, cd* W0o, cd* V0o, cd* g0o, cd* W1o, cd* V1o, cd* g1o, StackTape* tape,
                              int* flags) {
  hipStream_t st = c.stream;
  const int n = nx * ny, N = 2 * n;
  const long NN = (long)N * N;
  const int ninst = (int)inst.size();
  if (ninst == 0) return;
  const int L = (int)inst[0].layers.size();
  for (auto& I : inst) {
    if ((int)I.layers.size() != L || (int)I.thickness.size() != L)
      throw ArgFailure("all instances of one batch must have the same layer count");
    if (!I.bottom || !I.top) throw ArgFailure("instance without bottom/top medium");
    if (!(I.lambda > 0)) throw ArgFailure("wavelength must be positive");
    prepare_layer(*I.bottom);  // validates the half-space media (their eps may take the closed form)
    prepare_layer(*I.top);
  }
  // End media given by one permittivity value get their eigen-system in closed form.
  bool homog = true;
  for (auto& I : inst)
    homog = homog && I.bottom->x.size() == 2 && I.bottom->y.size() == 2 && I.top->x.size() == 2 &&
            I.top->y.size() == 2;
  if (homog && !tape) {
    batch_stack_homog(c, nx, ny, inst, chunk > 0 ? chunk * L : 0, S_out, W0o, V0o, g0o, W1o, V1o, g1o, flags);
    return;
  }
  // Chunks of instances, one eigensolver launch each (L slices, + 2 end media per instance
  // unless closed form): by default full launches (one matrix per QR workgroup slot of the
  // GPU) and one launch of the rest; a given chunk size splits the instances evenly.
  std::vector<int> starts{0};
  if (chunk <= 0) {
    for (int sz : zgeev_launch_plan(N, ninst, L + (homog ? 0 : 2))) starts.push_back(starts.back() + sz);
    chunk = starts[1];
  } else {
    const int nch = (ninst + chunk - 1) / chunk;
    for (int ci = 0; ci < nch; ++ci) starts.push_back(starts.back() + ninst / nch + (ci < ninst % nch ? 1 : 0));
  }
  const int nchunks = (int)starts.size() - 1;
  // Cascade groups: the slab S-matrices and end systems of consecutive chunks are kept
  // and cascaded together, so that the star products and projections run batched over
  // up to g_cascade_group instances (chunks split evenly over the groups).
  const int gmax = std::max(1, cascade_group_cap(N, L) / std::max(1, chunk));
  const int ngroups = (nchunks + gmax - 1) / gmax;
  std::vector<int> gstart(ngroups + 1, 0);
  for (int g = 0; g < ngroups; ++g) gstart[g + 1] = gstart[g] + nchunks / ngroups + (g < nchunks % ngroups ? 1 : 0);
  for (int g = 0; g < ngroups; ++g) {
  const int I0 = starts[gstart[g]];
  const int GB = starts[gstart[g + 1]] - I0;
  // group buffers: slab S [instance][slab], end systems [bottoms GB | tops GB]
  DevBuf tuG = c.ws.get(sizeof(cd) * NN * GB * L), ruG = c.ws.get(sizeof(cd) * NN * GB * L);
  DevBuf WeG = c.ws.get(sizeof(cd) * NN * 2 * GB), VeG = c.ws.get(sizeof(cd) * NN * 2 * GB),
         geG = c.ws.get(sizeof(cd) * (size_t)N * 2 * GB);
  for (int ci = gstart[g]; ci < gstart[g + 1]; ++ci) {
    const int i0 = starts[ci];
    const int B = starts[ci + 1] - i0;
    // ---- jobs: [bottoms(B) | tops(B) | slices(B*L, instance-major)]
    const int J = 2 * B + B * L;
    std::vector<const HostLayer*> jl(J);
    std::vector<double> k2(J), sv(B * L);
    std::vector<cd> ab((size_t)J * 2 * n);
    std::vector<int> owner(B * L);
    for (int b = 0; b < B; ++b) {
      const StackInstance& I = inst[i0 + b];
      const double k0 = 2 * PI_REF / I.lambda;
      std::vector<cd> al, be;
      const double Lx = I.bottom->x.back() - I.bottom->x.front(), Ly = I.bottom->y.back() - I.bottom->y.front();
      alpha_beta(nx, ny, Lx, Ly, al, be);
      auto put_job = [&](int j, const HostLayer* l) {
        jl[j] = l;
        k2[j] = k0 * k0;
        std::copy(al.begin(), al.end(), ab.begin() + (size_t)j * 2 * n);
        std::copy(be.begin(), be.end(), ab.begin() + (size_t)j * 2 * n + n);
      };
      put_job(b, I.bottom);
      put_job(B + b, I.top);
      for (int l = 0; l < L; ++l) {
        put_job(2 * B + b * L + l, I.layers[l]);
        sv[b * L + l] = I.thickness[l] / k0;
        owner[b * L + l] = b;
      }
    }
    std::vector<int> job_owner(J);  // instance (within the chunk) of every job
    for (int j = 0; j < J; ++j) job_owner[j] = j < B ? j : (j < 2 * B ? j - B : owner[j - 2 * B]);
    DevBuf dk2 = upload(c, k2.data(), sizeof(double) * J);
    DevBuf dab = upload(c, ab.data(), sizeof(cd) * ab.size());
    DevBuf dsv = upload(c, sv.data(), sizeof(double) * sv.size());
    DevBuf down = upload(c, owner.data(), sizeof(int) * owner.size());
    DevBuf djo = upload(c, job_owner.data(), sizeof(int) * J);
    // ---- one eigen-solve launch for end media (unless closed form) and slices
    const int SL = B * L;
    const int je = homog ? 2 * B : 0;  // first job solved by the dense eigensolver
    const int JD = J - je;
    AssemblyBatch a;
    a.nx = nx;
    a.ny = ny;
    a.nlayers = JD;
    a.layers.assign(jl.begin() + je, jl.end());
    a.periodic = true;
    a.d_alpha = dab.as<cd>() + (size_t)je * 2 * n;
    a.d_beta = a.d_alpha + n;
    a.sab = 2 * n;
    a.d_k0sq = dk2.as<double>() + je;
    DevBuf P = c.ws.get(sizeof(cd) * NN * JD), Q = c.ws.get(sizeof(cd) * NN * JD);
    DevBuf PQ;  // P Q formed with P and Q (periodic cells: without the N x N product)
    if (!tape) PQ = c.ws.get(sizeof(cd) * NN * JD);
    assemble_pq(c, a, P.as<cd>(), Q.as<cd>(), PQ.as<cd>());
    const int JW = tape ? J : 2 * B;  // jobs whose W, V are formed explicitly
    DevBuf Wall = c.ws.get(sizeof(cd) * NN * JW), Vall = c.ws.get(sizeof(cd) * NN * JW),
           gall = c.ws.get(sizeof(cd) * (size_t)N * J), infall = c.ws.get(sizeof(int) * J);
    fill_int(st, infall.as<int>(), 0, J);
    if (homog) {
      std::vector<cd> ev(2 * B);
      for (int j = 0; j < 2 * B; ++j) ev[j] = jl[j]->eps[0];
      DevBuf dev = upload(c, ev.data(), sizeof(cd) * ev.size());
      homogeneous_eig(st, n, 2 * B, dab.as<cd>(), dab.as<cd>() + n, 2 * n, dk2.as<double>(), dev.as<cd>(),
                      Wall.as<cd>(), Vall.as<cd>(), gall.as<cd>());
    }
    cd* tu = tuG.as<cd>() + (long)(i0 - I0) * L * NN;  // this chunk's slab S-matrices
    cd* ru = ruG.as<cd>() + (long)(i0 - I0) * L * NN;
    DevBuf Wr = c.ws.get(sizeof(cd) * NN * SL), Vr = c.ws.get(sizeof(cd) * NN * SL);
    const cd* W0 = Wall.as<cd>();
    const cd* V0 = Vall.as<cd>();
    const cd* W1 = Wall.as<cd>() + B * NN;
    const cd* V1 = Vall.as<cd>() + B * NN;
    const long s0 = (long)(2 * B - je);  // first slab within the dense jobs
    if (tape) {
      wave_eqn(c, N, P.as<cd>(), Q.as<cd>(), NN, Wall.as<cd>() + (long)je * NN, Vall.as<cd>() + (long)je * NN,
               gall.as<cd>() + (long)je * N, infall.as<int>() + je, JD);
      // eigen-systems in the tape's order: bottoms [inst], tops [inst], slabs [inst][L]
      auto keep = [&](int rows, const cd* src, cd* dst, int cnt) {
        zgeam(st, rows, 1, ONE, src, rows, rows, ZERO, nullptr, 0, 0, dst, rows, rows, cnt);
      };
      const int nn2 = (int)NN;
      keep(nn2, Wall.as<cd>(), tape->Wb.as<cd>() + (long)i0 * NN, B);
      keep(nn2, Vall.as<cd>(), tape->Vb.as<cd>() + (long)i0 * NN, B);
      keep(N, gall.as<cd>(), tape->gb.as<cd>() + (long)i0 * N, B);
      keep(nn2, Wall.as<cd>() + B * NN, tape->Wt.as<cd>() + (long)i0 * NN, B);
      keep(nn2, Vall.as<cd>() + B * NN, tape->Vt.as<cd>() + (long)i0 * NN, B);
      keep(N, gall.as<cd>() + (long)B * N, tape->gt.as<cd>() + (long)i0 * N, B);
      if (!tape->recompute) {
        keep(nn2, Wall.as<cd>() + 2 * B * NN, tape->Ws.as<cd>() + (long)i0 * L * NN, B * L);
        keep(nn2, Vall.as<cd>() + 2 * B * NN, tape->Vs.as<cd>() + (long)i0 * L * NN, B * L);
        keep(N, gall.as<cd>() + (long)2 * B * N, tape->gs.as<cd>() + (long)i0 * L * N, B * L);
      }
      gather_mats(st, NN, W1, NN, down.as<int>(), Wr.as<cd>(), NN, SL);
      gather_mats(st, NN, V1, NN, down.as<int>(), Vr.as<cd>(), NN, SL);
      layer_smatrix(c, N, SL, dsv.as<double>(), Wall.as<cd>() + 2 * B * NN, Vall.as<cd>() + 2 * B * NN,
                    gall.as<cd>() + (long)2 * B * N, Wr.as<cd>(), Vr.as<cd>(), NN, tu, ru);
    } else {
      // slabs keep their eigenvectors factored (W = D Z X); only the end media form W, V
      DevBuf Zb = c.ws.get(sizeof(cd) * NN * JD), Xb = c.ws.get(sizeof(cd) * NN * JD),
             Db = c.ws.get(sizeof(cd) * (size_t)N * JD);
      wave_eqn_factors(c, N, P.as<cd>(), Q.as<cd>(), NN, Zb.as<cd>(), Xb.as<cd>(), Db.as<cd>(),
                       gall.as<cd>() + (long)je * N, infall.as<int>() + je, JD, PQ.as<cd>());
      PQ = DevBuf();
      if (!homog)
        wave_eqn_vectors(c, N, Q.as<cd>(), NN, Zb.as<cd>(), Xb.as<cd>(), Db.as<cd>(), gall.as<cd>(), Wall.as<cd>(),
                         Vall.as<cd>(), 2 * B);
      if (!homog) gather_mats(st, NN, W1, NN, down.as<int>(), Wr.as<cd>(), NN, SL);  // homogeneous: W1 = I
      gather_mats(st, NN, V1, NN, down.as<int>(), Vr.as<cd>(), NN, SL);
      layer_smatrix_factors(c, N, SL, dsv.as<double>(), P.as<cd>() + s0 * NN, NN, Zb.as<cd>() + s0 * NN,
                            Xb.as<cd>() + s0 * NN, Db.as<cd>() + s0 * N, gall.as<cd>() + (long)2 * B * N,
                            Wr.as<cd>(), Vr.as<cd>(), NN, tu, ru, homog);
    }
    if (flags) info_to_owner(st, infall.as<int>(), djo.as<int>(), J, H_EIG, flags + i0);
    Wr = DevBuf();
    Vr = DevBuf();
    P = DevBuf();
    Q = DevBuf();
    // end systems into the group buffers
    const long og = i0 - I0;
    zgeam(st, N, N, ONE, W0, N, NN, ZERO, nullptr, 0, 0, WeG.as<cd>() + og * NN, N, NN, B);
    zgeam(st, N, N, ONE, V0, N, NN, ZERO, nullptr, 0, 0, VeG.as<cd>() + og * NN, N, NN, B);
    zgeam(st, N, N, ONE, W1, N, NN, ZERO, nullptr, 0, 0, WeG.as<cd>() + (GB + og) * NN, N, NN, B);
    zgeam(st, N, N, ONE, V1, N, NN, ZERO, nullptr, 0, 0, VeG.as<cd>() + (GB + og) * NN, N, NN, B);
    zgeam(st, N, 1, ONE, gall.as<cd>(), N, N, ZERO, nullptr, 0, 0, geG.as<cd>() + og * N, N, N, B);
    zgeam(st, N, 1, ONE, gall.as<cd>() + (long)B * N, N, N, ZERO, nullptr, 0, 0, geG.as<cd>() + (GB + og) * N, N, N,
          B);
  }
  cascade_group(c, N, L, GB, I0, homog, tuG.as<cd>(), ruG.as<cd>(), WeG.as<cd>(), VeG.as<cd>(), geG.as<cd>(), S_out,
                W0o, V0o, g0o, W1o, V1o, g1o, tape);
  }
}

// The power bookkeeping of RCWASolver::scatterPlaneWave (src/rcwa/RCWASolver.cpp:1636-1733)
// on the batched stack S-matrices: c = W0^{-1} e_inc, reflected W0 Rdu c, transmitted W1 Tuu c.
void stack_efficiencies(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst, const cd* S,
                        const cd* W0, const cd* W1, const cd* e_inc, double* REF, double* TRN) {
  hipStream_t st = c.stream;
  const int n = nx * ny, N = 2 * n, B = (int)inst.size();
  if (B == 0) return;
  const long NN = (long)N * N;
  std::vector<cd> ab((size_t)B * 2 * n), eps2((size_t)2 * B);
  std::vector<double> k0v(B);
  for (int b = 0; b < B; ++b) {
    const StackInstance& I = inst[b];
    if (!(I.lambda > 0)) throw ArgFailure("wavelength must be positive");
    std::vector<cd> al, be;
    alpha_beta(nx, ny, I.bottom->x.back() - I.bottom->x.front(), I.bottom->y.back() - I.bottom->y.front(), al, be);
    std::copy(al.begin(), al.end(), ab.begin() + (size_t)b * 2 * n);
    std::copy(be.begin(), be.end(), ab.begin() + (size_t)b * 2 * n + n);
    k0v[b] = 2 * PI_REF / I.lambda;
    eps2[2 * b] = I.bottom->eps[0];  // eps(0, 0) of the medium, as the reference reads it
    eps2[2 * b + 1] = I.top->eps[0];
  }
  DevBuf dab = upload(c, ab.data(), sizeof(cd) * ab.size());
  DevBuf deps = upload(c, eps2.data(), sizeof(cd) * eps2.size());
  DevBuf dk0 = upload(c, k0v.data(), sizeof(double) * k0v.size());
  // c = W0^{-1} e_inc
  DevBuf LU = c.ws.get(sizeof(cd) * NN * B), piv = c.ws.get(sizeof(int) * (size_t)N * B),
         inf = c.ws.get(sizeof(int) * (size_t)B);
  DevBuf cb = c.ws.get(sizeof(cd) * (size_t)N * B), cr = c.ws.get(sizeof(cd) * (size_t)N * B),
         ctr = c.ws.get(sizeof(cd) * (size_t)N * B), fr = c.ws.get(sizeof(cd) * (size_t)N * B),
         ft = c.ws.get(sizeof(cd) * (size_t)N * B);
  fill_int(st, inf.as<int>(), 0, B);
  zgeam(st, N, N, ONE, W0, N, NN, ZERO, nullptr, 0, 0, LU.as<cd>(), N, NN, B);
  zgetrf(st, N, LU.as<cd>(), N, NN, piv.as<int>(), N, inf.as<int>(), B, c.ws);
  zgetrs(st, N, 1, LU.as<cd>(), N, NN, piv.as<int>(), N, e_inc, N, N, cb.as<cd>(), N, N, B, c.ws);
  // reflected Rdu c, transmitted Tuu c; fields W0 (Rdu c), W1 (Tuu c)
  const long sS = 4 * NN;
  zgemm_strided(st, OP_N, OP_N, N, 1, N, ONE, S + 3 * NN, N, sS, cb.as<cd>(), N, N, ZERO, cr.as<cd>(), N, N, B);
  zgemm_strided(st, OP_N, OP_N, N, 1, N, ONE, S + NN, N, sS, cb.as<cd>(), N, N, ZERO, ctr.as<cd>(), N, N, B);
  zgemm_strided(st, OP_N, OP_N, N, 1, N, ONE, W0, N, NN, cr.as<cd>(), N, N, ZERO, fr.as<cd>(), N, N, B);
  zgemm_strided(st, OP_N, OP_N, N, 1, N, ONE, W1, N, NN, ctr.as<cd>(), N, N, ZERO, ft.as<cd>(), N, N, B);
  efficiencies(st, n, B, fr.as<cd>(), ft.as<cd>(), dab.as<cd>(), dk0.as<double>(), deps.as<cd>(), REF, TRN);
}

// evaluateEnergy (src/tests/test_cpu.cpp:335-412) for many modes: with the window Gram
// matrices Gx (nx x nx) and Gy (ny x ny) of the sampled wavelets, the grid sum of
// E conj(H) of one mode is sum(K o conj(Gy^T L Gx)) in the (ny x nx) layout of a mode
// column, so each mode costs two small products and a dot product instead of a field
// reconstruction on the whole grid.
void modal_power(vrc_ctx& c, int nx, int ny, int nmodes, const cd* W, long ldw, const cd* V, long ldv, double Lx,
                 double Ly, int nres_x, int nres_y, double x0, double x1, double y0, double y1, double* energy) {
  hipStream_t st = c.stream;
  const int n = nx * ny;
  if (nmodes == 0) return;
  DevBuf gx = c.ws.get(sizeof(cd) * (size_t)nx * nx), gyt = c.ws.get(sizeof(cd) * (size_t)ny * ny);
  DevBuf T = c.ws.get(sizeof(cd) * (size_t)2 * n * nmodes), Z = c.ws.get(sizeof(cd) * (size_t)2 * n * nmodes);
  window_gram(st, nx, nres_x, x0, x1, Lx, -1.0, gx.as<cd>());
  window_gram(st, ny, nres_y, y0, y1, Ly, +1.0, gyt.as<cd>());
  // T_b = Gy^T [Hx_b | Hy_b] (ny x 2nx per mode), Z = T Gx per (ny x nx) block
  zgemm_strided(st, OP_N, OP_N, ny, 2 * nx, ny, ONE, gyt.as<cd>(), ny, 0, V, ny, ldv, ZERO, T.as<cd>(), ny, 2 * n,
                nmodes);
  zgemm_strided(st, OP_N, OP_N, ny, nx, nx, ONE, T.as<cd>(), ny, n, gx.as<cd>(), nx, 0, ZERO, Z.as<cd>(), ny, n,
                2 * nmodes);
  modal_power_reduce(st, n, nmodes, W, ldw, Z.as<cd>(), (x1 - x0) / nres_x * ((y1 - y0) / nres_y), energy);
}

}  // namespace vrc

// ===================================================================== adjoint of batched stacks
namespace vrc {

long StackTape::cells() const {
  long n = 0;
  for (int i = 0; i < ninst; ++i)
    for (int l = 0; l < L; ++l) n += (long)slab(i, l)->eps.size();
  return n;
}

size_t StackTape::device_bytes() const {
  return Wb.bytes + Vb.bytes + gb.bytes + Wt.bytes + Vt.bytes + gt.bytes + Ws.bytes + Vs.bytes + gs.bytes +
         ckpt.bytes + Sfin.bytes;
}

std::unique_ptr<StackTape> make_stack_tape(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst,
                                           bool recompute, int K) {
  auto t = std::make_unique<StackTape>();
  t->c = &c;
  t->nx = nx;
  t->ny = ny;
  t->ninst = (int)inst.size();
  t->L = inst.empty() ? 0 : (int)inst[0].layers.size();
  const int n = nx * ny, N = 2 * n, L = t->L, I = t->ninst;
  const size_t NN = (size_t)N * N;
  t->recompute = recompute;
  t->K = std::max(1, std::min(L, K > 0 ? K : 16));
  t->layers.reserve((size_t)I * (L + 2));
  for (auto& s : inst) {
    if ((int)s.layers.size() != L) throw ArgFailure("all instances of one batch must have the same layer count");
    for (auto* l : s.layers) t->layers.push_back(*l);
    t->lambda.push_back(s.lambda);
    t->thickness.insert(t->thickness.end(), s.thickness.begin(), s.thickness.end());
  }
  for (auto& s : inst) t->layers.push_back(*s.bottom);
  for (auto& s : inst) t->layers.push_back(*s.top);
  t->Wb = c.ws.get(sizeof(cd) * NN * I);
  t->Vb = c.ws.get(sizeof(cd) * NN * I);
  t->gb = c.ws.get(sizeof(cd) * (size_t)N * I);
  t->Wt = c.ws.get(sizeof(cd) * NN * I);
  t->Vt = c.ws.get(sizeof(cd) * NN * I);
  t->gt = c.ws.get(sizeof(cd) * (size_t)N * I);
  if (!recompute) {
    t->Ws = c.ws.get(sizeof(cd) * NN * I * L);
    t->Vs = c.ws.get(sizeof(cd) * NN * I * L);
    t->gs = c.ws.get(sizeof(cd) * (size_t)N * I * L);
  }
  if (t->nblocks() > 1) t->ckpt = c.ws.get(sizeof(cd) * 4 * NN * I * (size_t)(t->nblocks() - 1));
  t->Sfin = c.ws.get(sizeof(cd) * 4 * NN * I);
  return t;
}

// Shape-parameter gradients of a block's slabs: dJ/dtheta = Re sum conj(Pbar) dP/dtheta +
// conj(Qbar) dQ/dtheta, the tangents dP, dQ from the closed-form edge terms of the Fourier
// tables (shape_tangents) of each parameterized slab.
static void shape_vjp(vrc_ctx& c, const StackTape& t, const AssemblyBatch& a, const std::vector<int>& sidx,
                      const cd* Pbar, const cd* Qbar, long sPQ, const ShapeParams& sp, double* shape_bar) {
  const int nx = t.nx, ny = t.ny, n = nx * ny, N = 2 * n;
  const long NN = (long)N * N;
  for (size_t j = 0; j < sidx.size(); ++j) {
    const int s = sidx[j];
    const int np = sp.npara[s];
    if (np <= 0) continue;
    const int i = s / t.L;
    std::vector<cd> al, be;
    const HostLayer* bot = t.bottom(i);
    alpha_beta(nx, ny, bot->x.back() - bot->x.front(), bot->y.back() - bot->y.front(), al, be);
    al.insert(al.end(), be.begin(), be.end());
    DevBuf dab = upload(c, al.data(), sizeof(cd) * al.size());
    DevBuf dP = c.ws.get(sizeof(cd) * NN * np), dQ = c.ws.get(sizeof(cd) * NN * np);
    shape_tangents(c, nx, ny, t.lambda[i], false, dab.as<cd>(), nullptr, *a.layers[j], sp.nvar[s],
                   sp.var_index.data() + sp.vofs[s], sp.coeff.data() + sp.cofs[s], np, dP.as<cd>(),
                   dQ.as<cd>(), nullptr, nullptr, nullptr, nullptr, nullptr);
    dot_re_accumulate(c, NN, Pbar + (long)j * sPQ, dP.as<cd>(), NN, np, shape_bar + sp.pofs[s]);
    dot_re_accumulate(c, NN, Qbar + (long)j * sPQ, dQ.as<cd>(), NN, np, shape_bar + sp.pofs[s]);
  }
}

// Reverse pass of batch_stack_rcwa, per chunk of instances: the projection's adjoint from
// the taped cascade, then the layers block by block from the top. Per block of K layers:
// P/Q, the eigen-systems (taped, or solved again in one batched launch), the slab
// S-matrices with their intermediates, the cascade prefixes from the block's checkpoint,
// then the cotangents back through the block's star products (last layer first), the
// slab S-matrices, the eigen-decompositions and the Fourier assembly. Device memory
// grows with K, not with the layer count.
void stack_tape_vjp(StackTape& t, const cd* Sb_in, cd* epsb, int chunk, const ShapeParams* shape,
                    double* shape_bar) {
  vrc_ctx& c = *t.c;
  hipStream_t st = c.stream;
  const int nx = t.nx, ny = t.ny, n = nx * ny, N = 2 * n, L = t.L, I = t.ninst, K = t.K, NB = t.nblocks();
  const long NN = (long)N * N;
  if (I == 0 || L == 0) return;
  tune_apply(N);
  std::vector<long> off((size_t)I * L + 1, 0);
  for (int i = 0; i < I; ++i)
    for (int l = 0; l < L; ++l) off[(size_t)i * L + l + 1] = off[(size_t)i * L + l] + (long)t.slab(i, l)->eps.size();
  if (chunk <= 0) chunk = std::max(1, std::min(I, 128 / K));
  for (int i0 = 0; i0 < I; i0 += chunk) {
    const int B = std::min(chunk, I - i0);
    const cd* W0 = t.Wb.as<cd>() + (long)i0 * NN;
    const cd* V0 = t.Vb.as<cd>() + (long)i0 * NN;
    const cd* W1 = t.Wt.as<cd>() + (long)i0 * NN;
    const cd* V1 = t.Vt.as<cd>() + (long)i0 * NN;
    LUStore lu1;
    lu1.make(c, N, W1, V1, B, NN);
    auto blocks = [&](cd* base) {
      SMat m;
      m.Tdd = base;
      m.Tuu = base + B * NN;
      m.Rud = base + 2 * B * NN;
      m.Rdu = base + 3 * B * NN;
      return m;
    };
    // tape cascades ([inst][4 N x N]) -> block-major SMat buffers ([4][B] N x N)
    auto load_S = [&](cd* dst, const cd* src, long per_inst) {
      for (int q = 0; q < 4; ++q)
        zgeam(st, N, N, ONE, src + (long)i0 * per_inst + q * NN, N, per_inst, ZERO, nullptr, 0, 0,
              dst + (long)q * B * NN, N, NN, B);
    };
    // ---- cotangent of the cascade before the projection
    DevBuf Sbb = c.ws.get(sizeof(cd) * 4 * NN * B);
    SMat Sb = blocks(Sbb.as<cd>());
    load_S(Sbb.as<cd>(), Sb_in, 4 * NN);
    {
      DevBuf Sf = c.ws.get(sizeof(cd) * 4 * NN * B);
      load_S(Sf.as<cd>(), t.Sfin.as<cd>(), 4 * NN);
      SMat S = blocks(Sf.as<cd>());
      project_left_vjp(c, N, B, S, lu1.pair, W0, V0, NN, Sb);
    }
    for (int bk = NB - 1; bk >= 0; --bk) {
      const int l0 = bk * K, l1 = std::min(L, l0 + K), KB = l1 - l0, SL = B * KB;
      // ---- the block's slabs, instance-major: P, Q (per-instance alpha, beta, k0^2)
      std::vector<double> k2(SL), sv(SL);
      std::vector<cd> ab((size_t)SL * 2 * n);
      std::vector<int> owner(SL), sidx(SL);
      AssemblyBatch a;
      a.nx = nx;
      a.ny = ny;
      a.nlayers = SL;
      a.periodic = true;
      for (int b = 0; b < B; ++b) {
        const int i = i0 + b;
        const double k0 = 2 * PI_REF / t.lambda[i];
        const HostLayer* bot = t.bottom(i);
        std::vector<cd> al, be;
        alpha_beta(nx, ny, bot->x.back() - bot->x.front(), bot->y.back() - bot->y.front(), al, be);
        for (int l = l0; l < l1; ++l) {
          const int j = b * KB + (l - l0);
          a.layers.push_back(t.slab(i, l));
          k2[j] = k0 * k0;
          sv[j] = t.thickness[(size_t)i * L + l] / k0;
          owner[j] = b;
          sidx[j] = i * L + l;
          std::copy(al.begin(), al.end(), ab.begin() + (size_t)j * 2 * n);
          std::copy(be.begin(), be.end(), ab.begin() + (size_t)j * 2 * n + n);
        }
      }
      DevBuf dk2 = upload(c, k2.data(), sizeof(double) * SL);
      DevBuf dab = upload(c, ab.data(), sizeof(cd) * ab.size());
      DevBuf dsv = upload(c, sv.data(), sizeof(double) * SL);
      DevBuf down = upload(c, owner.data(), sizeof(int) * SL);
      a.d_alpha = dab.as<cd>();
      a.d_beta = a.d_alpha + n;
      a.sab = 2 * n;
      a.d_k0sq = dk2.as<double>();
      DevBuf P = c.ws.get(sizeof(cd) * NN * SL), Q = c.ws.get(sizeof(cd) * NN * SL);
      assemble_pq(c, a, P.as<cd>(), Q.as<cd>());
      // ---- eigen-systems of the block: taped, or solved again
      DevBuf Ws = c.ws.get(sizeof(cd) * NN * SL), Vs = c.ws.get(sizeof(cd) * NN * SL),
             gs = c.ws.get(sizeof(cd) * (size_t)N * SL);
      if (t.recompute) {
        DevBuf inf = c.ws.get(sizeof(int) * SL);
        wave_eqn(c, N, P.as<cd>(), Q.as<cd>(), NN, Ws.as<cd>(), Vs.as<cd>(), gs.as<cd>(), inf.as<int>(), SL);
      } else {
        DevBuf dsi = upload(c, sidx.data(), sizeof(int) * SL);
        gather_mats(st, NN, t.Ws.as<cd>(), NN, dsi.as<int>(), Ws.as<cd>(), NN, SL);
        gather_mats(st, NN, t.Vs.as<cd>(), NN, dsi.as<int>(), Vs.as<cd>(), NN, SL);
        gather_mats(st, N, t.gs.as<cd>(), N, dsi.as<int>(), gs.as<cd>(), N, SL);
      }
      DevBuf Wr = c.ws.get(sizeof(cd) * NN * SL), Vr = c.ws.get(sizeof(cd) * NN * SL);
      gather_mats(st, NN, W1, NN, down.as<int>(), Wr.as<cd>(), NN, SL);
      gather_mats(st, NN, V1, NN, down.as<int>(), Vr.as<cd>(), NN, SL);
      SlabSaved slab;
      layer_smatrix_fwd_saved(c, N, SL, dsv.as<double>(), Ws.as<cd>(), Vs.as<cd>(), gs.as<cd>(), Wr.as<cd>(),
                              Vr.as<cd>(), NN, slab);
      Wr = DevBuf();
      Vr = DevBuf();
      const cd* tuv = slab.Tuu.as<cd>();
      const cd* ruv = slab.Rud.as<cd>();
      // ---- prefixes of the block from its checkpoint (identity before the first layer)
      DevBuf pre = c.ws.get(sizeof(cd) * 4 * NN * B * KB);
      DevBuf sstar = c.ws.get(sizeof(cd) * STAR_SAVE * NN * B * KB);
      DevBuf Sbuf = c.ws.get(sizeof(cd) * 4 * NN * B);
      SMat S = blocks(Sbuf.as<cd>());
      if (bk == 0)
        init_identity_S(c, N, S, B);
      else
        load_S(Sbuf.as<cd>(), t.ckpt.as<cd>() + (long)(bk - 1) * 4 * NN, (long)(NB - 1) * 4 * NN);
      const long ss = (long)KB * NN;
      for (int l = 0; l < KB; ++l) {
        VRC_HIP(hipMemcpyAsync(pre.as<cd>() + (long)l * 4 * B * NN, Sbuf.ptr, sizeof(cd) * 4 * NN * B,
                               hipMemcpyDeviceToDevice, st));
        const cd* tl = tuv + l * NN;
        const cd* rl = ruv + l * NN;
        star_product(c, N, B, S, tl, tl, rl, rl, ss, sstar.as<cd>() + (long)l * STAR_SAVE * B * NN);
      }
      Sbuf = DevBuf();
      // ---- cotangents back through the block's star products
      DevBuf tb = c.ws.get(sizeof(cd) * NN * SL), rb = c.ws.get(sizeof(cd) * NN * SL);
      {
        DevBuf q = c.ws.get(sizeof(cd) * 4 * NN * B);
        cd* q0 = q.as<cd>();
        for (int l = KB - 1; l >= 0; --l) {
          SMat S0 = blocks(pre.as<cd>() + (long)l * 4 * B * NN);
          const cd* tl = tuv + l * NN;
          const cd* rl = ruv + l * NN;
          star_product_vjp(c, N, B, S0, tl, tl, rl, rl, ss, Sb, q0, q0 + B * NN, q0 + 2 * B * NN, q0 + 3 * B * NN,
                           NN, sstar.as<cd>() + (long)l * STAR_SAVE * B * NN);
          // the slab's Tdd and Tuu (Rud and Rdu) are one matrix: their cotangents add
          zgeam(st, N, N, ONE, q0, N, NN, ONE, q0 + B * NN, N, NN, tb.as<cd>() + l * NN, N, ss, B);
          zgeam(st, N, N, ONE, q0 + 2 * B * NN, N, NN, ONE, q0 + 3 * B * NN, N, NN, rb.as<cd>() + l * NN, N, ss,
                B);
        }
      }
      pre = DevBuf();
      sstar = DevBuf();
      // ---- slab S -> (W, V, gamma) -> (P, Q) -> eps (and shape parameters)
      DevBuf Wbar = c.ws.get(sizeof(cd) * NN * SL), Vbar = c.ws.get(sizeof(cd) * NN * SL),
             gbar = c.ws.get(sizeof(cd) * (size_t)N * SL);
      layer_smatrix_bwd_saved(c, N, SL, dsv.as<double>(), slab, tb.as<cd>(), rb.as<cd>(), Wbar.as<cd>(),
                              Vbar.as<cd>(), gbar.as<cd>());
      const cd* Wi = slab.Wi.as<cd>();  // W^{-1}, unchanged by the slab reverse pass
      DevBuf& Pbar = tb;  // slab cotangents consumed
      DevBuf& Qbar = rb;
      {
        StageScope sc(ST_ADJOINT);
        wave_eqn_vjp(c, N, SL, P.as<cd>(), Q.as<cd>(), NN, Ws.as<cd>(), Vs.as<cd>(), gs.as<cd>(), Wi, Wbar.as<cd>(),
                     Vbar.as<cd>(), gbar.as<cd>(), Pbar.as<cd>(), Qbar.as<cd>());
        if (epsb) {
          std::vector<long> o(SL);
          for (int j = 0; j < SL; ++j) o[j] = off[sidx[j]];
          assemble_pq_vjp(c, a, Pbar.as<cd>(), Qbar.as<cd>(), NN, epsb, o);
        }
        if (shape && shape_bar) shape_vjp(c, t, a, sidx, Pbar.as<cd>(), Qbar.as<cd>(), NN, *shape, shape_bar);
      }
    }
  }
  VRC_HIP(hipStreamSynchronize(st));
}

}  // namespace vrc
