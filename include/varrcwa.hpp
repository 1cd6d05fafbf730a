// varrcwa.hpp -- header-only C++ API over the C ABI of libvarrcwa.so, with the
// class names, constructors, compute signatures and getters of the reference's
// CPU path, so that a caller of the reference can switch by changing includes:
//
//   reference                                  here
//   VarLayerSampler (src/core/VarLayerSampler.h:8-39)      varrcwa::VarLayerSampler
//   RCWAIntegrator (src/core/RCWAIntegrator.h:8-40)        varrcwa::RCWAIntegrator
//   RedhefferIntegrator (src/core/RedhefferIntegrator.h:10-90) varrcwa::RedhefferIntegrator
//   DifferentialIntegrator (src/core/DifferentialIntegrator.h:8-40) varrcwa::DifferentialIntegrator
//
// Matrices are column-major complex128 (`MatrixXcs`, `VectorXcs` below; the memory
// layout of Eigen::MatrixXcd, so `Eigen::Map<const Eigen::MatrixXcd>(m.data(),
// m.rows(), m.cols())` views one without a copy). Errors throw varrcwa::Error with
// the library's message. All objects of one thread share a per-device context on the
// device's default stream unless a Context is passed explicitly.
#pragma once

#include <algorithm>
#include <complex>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "varrcwa.h"

namespace varrcwa {

using scalar = double;
using scalex = std::complex<double>;

class Error : public std::runtime_error {
 public:
  Error(const std::string& what, int code) : std::runtime_error(what), code(code) {}
  int code;
};

// Column-major dense complex matrix (vector = one column).
class MatrixXcs {
 public:
  MatrixXcs() = default;
  MatrixXcs(long rows, long cols) : r_(rows), c_(cols), v_((size_t)(rows * cols)) {}
  long rows() const { return r_; }
  long cols() const { return c_; }
  long size() const { return r_ * c_; }
  void resize(long rows, long cols) {
    r_ = rows;
    c_ = cols;
    v_.assign((size_t)(rows * cols), scalex(0.0, 0.0));
  }
  scalex& operator()(long i, long j) { return v_[(size_t)(i + j * r_)]; }
  const scalex& operator()(long i, long j) const { return v_[(size_t)(i + j * r_)]; }
  scalex& operator()(long i) { return v_[(size_t)i]; }
  const scalex& operator()(long i) const { return v_[(size_t)i]; }
  scalex* data() { return v_.data(); }
  const scalex* data() const { return v_.data(); }

 private:
  long r_ = 0, c_ = 0;
  std::vector<scalex> v_;
};
using VectorXcs = MatrixXcs;
using VectorXs = std::vector<double>;

namespace detail {
inline void check(int rc, const vrc_ctx* c, const char* what) {
  if (rc != 0) {
    std::string msg = std::string(what) + " failed";
    if (c) msg += std::string(": ") + vrc_last_error(c);
    throw Error(msg, rc);
  }
}
}  // namespace detail

// Owns a vrc_ctx (device + stream + device memory pool).
class Context {
 public:
  explicit Context(int device = 0, void* stream = nullptr) {
    vrc_ctx* h = nullptr;
    detail::check(vrc_ctx_create(device, stream, &h), nullptr, "vrc_ctx_create");
    h_ = h;
  }
  ~Context() {
    if (h_) vrc_ctx_destroy(h_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  vrc_ctx* get() const { return h_; }
  // Per-thread default context of device 0 on the default stream.
  static std::shared_ptr<Context> thread_default() {
    thread_local std::shared_ptr<Context> c = std::make_shared<Context>(0, nullptr);
    return c;
  }

 private:
  vrc_ctx* h_ = nullptr;
};

// One Layer(x, y, eps) slab (src/rcwa/Layer.cpp:18-34): eps is (ny-1) x (nx-1),
// row i = y strip, column j = x cell.
struct LayerDesc {
  VectorXs x, y;
  MatrixXcs eps;
  vrc_layer view() const {
    if (eps.rows() != (long)y.size() - 1 || eps.cols() != (long)x.size() - 1)
      throw Error("eps must be (y.size()-1) x (x.size()-1)", -1);
    return vrc_layer{(int)x.size(), x.data(), (int)y.size(), y.data(), reinterpret_cast<const double*>(eps.data())};
  }
};

// LayerSampler (src/core/LayerSampler.h:5-77). The library's samplers (VarLayerSampler,
// StackLayerSampler, GDSIISampler below) derive from it; so can a caller's own sampler, as
// with the reference: derive, override sample() (and, optionally, sampleOnGPU()) and
// assignSimulationRegion() -- which sets Lx_, Ly_, b0_ .. r1_ -- and call
// postInitialization() at the end of the constructor. The integrators then call back into
// the subclass for every P, Q they need (sampleOnGPU when overridden, else sample() and a
// host-to-device copy).
class LayerSampler {
 public:
  virtual ~LayerSampler() {
    if (s_) vrc_sampler_destroy(s_);
  }
  LayerSampler(const LayerSampler&) = delete;
  LayerSampler& operator=(const LayerSampler&) = delete;

  scalar lambda() const { return lambda_; }
  scalar k0() const { return 2 * 3.1415926535897932 / lambda_; }
  scalar Lx() const { return Lx_; }
  scalar Ly() const { return Ly_; }
  int nDim() const { return 2 * nx_ * ny_; }
  const MatrixXcs& Kx() const { return Kx_; }
  const MatrixXcs& Ky() const { return Ky_; }

  // LayerSampler::sample (src/core/LayerSampler.h:49-51): host P, Q (resized).
  virtual void sample(scalar z, MatrixXcs& P, MatrixXcs& Q) {
    if (user_) throw Error("LayerSampler subclass must override sample()", -1);
    const int N = nDim();
    P.resize(N, N);
    Q.resize(N, N);
    detail::check(vrc_sampler_sample(s_, z, reinterpret_cast<double*>(P.data()), reinterpret_cast<double*>(Q.data())),
                  ctx_->get(), "sample");
  }
  // LayerSampler::sampleOnGPU (src/core/LayerSampler.h:53-55): device N x N buffers, work on
  // the context's stream. A subclass that does not override it is sampled on the host.
  virtual void sampleOnGPU(scalar z, void* P, void* Q) {
    if (user_) {
      gpu_default_ = true;
      return;
    }
    detail::check(vrc_sampler_sample_gpu(s_, z, P, Q), ctx_->get(), "sampleOnGPU");
  }

  vrc_sampler* handle() const { return s_; }
  const std::shared_ptr<Context>& context() const { return ctx_; }

 protected:
  // A caller's sampler (src/core/LayerSampler.h:31-37).
  LayerSampler(scalar lambda, int nx, int ny, bool enablePML = true, std::shared_ptr<Context> ctx = nullptr)
      : lambda_(lambda), ctx_(ctx ? std::move(ctx) : Context::thread_default()), nx_(nx), ny_(ny),
        enablePML_(enablePML), user_(true) {}
  // Put this at the end of the subclass constructor (src/core/LayerSampler.cpp:145-153).
  void postInitialization() {
    assignSimulationRegion();
    const vrc_region r{Lx_, Ly_, b0_, b1_, u0_, u1_, l0_, l1_, r0_, r1_};
    vrc_sampler* s = nullptr;
    detail::check(vrc_callback_sampler_create(ctx_->get(), lambda_, nx_, ny_, &r, enablePML_ ? 1 : 0, &host_cb,
                                              &device_cb, this, &s),
                  ctx_->get(), "postInitialization");
    finish(s, nx_, ny_);
  }
  // Assign Lx_, Ly_, b0_, b1_, u0_, u1_, l0_, l1_, r0_, r1_ (src/core/LayerSampler.h:57-62).
  virtual void assignSimulationRegion() {}

  // The library's own samplers.
  LayerSampler(scalar lambda, std::shared_ptr<Context> ctx)
      : lambda_(lambda), ctx_(ctx ? std::move(ctx) : Context::thread_default()) {}
  void finish(vrc_sampler* s, int nx, int ny) {
    s_ = s;
    nx_ = nx;
    ny_ = ny;
    const int n = nx * ny;
    Kx_.resize(n, n);
    Ky_.resize(n, n);
    detail::check(vrc_sampler_k(s_, reinterpret_cast<double*>(Kx_.data()), reinterpret_cast<double*>(Ky_.data())),
                  ctx_->get(), "Kx/Ky");
    vrc_region r{};
    detail::check(vrc_sampler_region(s_, &r), ctx_->get(), "region");
    Lx_ = r.Lx, Ly_ = r.Ly, b0_ = r.b0, b1_ = r.b1, u0_ = r.u0, u1_ = r.u1, l0_ = r.l0, l1_ = r.l1, r0_ = r.r0,
    r1_ = r.r1;
  }

  scalar lambda_;
  std::shared_ptr<Context> ctx_;
  vrc_sampler* s_ = nullptr;
  MatrixXcs Kx_, Ky_;
  int nx_ = 0, ny_ = 0;
  bool enablePML_ = true;
  scalar Lx_ = 0, Ly_ = 0, b0_ = 0, b1_ = 0, u0_ = 0, u1_ = 0, l0_ = 0, l1_ = 0, r0_ = 0, r1_ = 0;

 private:
  // C-ABI trampolines into the subclass; exceptions become non-zero returns.
  static int host_cb(double z, double* P, double* Q, void* user) {
    auto* self = static_cast<LayerSampler*>(user);
    try {
      MatrixXcs p, q;
      self->sample(z, p, q);
      const long N = self->nDim();
      if (p.rows() != N || p.cols() != N || q.rows() != N || q.cols() != N) return 3;
      std::copy(p.data(), p.data() + N * N, reinterpret_cast<scalex*>(P));
      std::copy(q.data(), q.data() + N * N, reinterpret_cast<scalex*>(Q));
      return 0;
    } catch (...) {
      return 2;
    }
  }
  static int device_cb(double z, void* P, void* Q, void* stream, void* user) {
    (void)stream;
    auto* self = static_cast<LayerSampler*>(user);
    try {
      self->gpu_default_ = false;
      self->sampleOnGPU(z, P, Q);
      return self->gpu_default_ ? 1 : 0;  // 1: not overridden, sample on the host
    } catch (...) {
      return 2;
    }
  }
  bool user_ = false;
  bool gpu_default_ = false;
};

// VarLayerSampler (src/core/VarLayerSampler.h:8-39).
class VarLayerSampler : public LayerSampler {
 public:
  VarLayerSampler(scalar lambda, int nx, int ny, const VectorXs& x, const VectorXs& y, const MatrixXcs& eps,
                  bool enablePML = true, std::shared_ptr<Context> ctx = nullptr)
      : LayerSampler(lambda, std::move(ctx)) {
    base_.x = x;
    base_.y = y;
    base_.eps = eps;
    vrc_layer l = base_.view();
    vrc_sampler* s = nullptr;
    detail::check(vrc_var_sampler_create(ctx_->get(), lambda, nx, ny, &l, enablePML ? 1 : 0, &s), ctx_->get(),
                  "VarLayerSampler");
    finish(s, nx, ny);
  }
  // x(layout) = rule(z) at every sample (VarLayerSampler::addXRule, src/core/VarLayerSampler.cpp:42-50).
  void addXRule(int layout, std::function<scalar(scalar)> rule) { add(0, layout, std::move(rule)); }
  void addYRule(int layout, std::function<scalar(scalar)> rule) { add(1, layout, std::move(rule)); }

 private:
  static double trampoline(double z, void* user) { return (*static_cast<std::function<scalar(scalar)>*>(user))(z); }
  void add(int axis, int layout, std::function<scalar(scalar)> rule) {
    rules_.push_back(std::make_unique<std::function<scalar(scalar)>>(std::move(rule)));
    detail::check(vrc_sampler_add_rule(s_, axis, layout, &trampoline, rules_.back().get()), ctx_->get(),
                  axis == 0 ? "addXRule" : "addYRule");
  }
  LayerDesc base_;
  std::vector<std::unique_ptr<std::function<scalar(scalar)>>> rules_;
};

// Piecewise-constant stack of layers (no reference counterpart; extension):
// layer k occupies (z_edges[k], z_edges[k+1]); bottom below, top above.
class StackLayerSampler : public LayerSampler {
 public:
  StackLayerSampler(scalar lambda, int nx, int ny, const std::vector<LayerDesc>& layers, const VectorXs& z_edges,
                    const LayerDesc& bottom, const LayerDesc& top, bool enablePML = false,
                    std::shared_ptr<Context> ctx = nullptr)
      : LayerSampler(lambda, std::move(ctx)) {
    std::vector<vrc_layer> ls;
    for (auto& l : layers) ls.push_back(l.view());
    vrc_layer b = bottom.view(), t = top.view();
    vrc_sampler* s = nullptr;
    detail::check(vrc_stack_sampler_create(ctx_->get(), lambda, nx, ny, (int)ls.size(), ls.data(), z_edges.data(), &b,
                                           &t, enablePML ? 1 : 0, &s),
                  ctx_->get(), "StackLayerSampler");
    finish(s, nx, ny);
  }
};

// Material (src/core/Material.h:5-18): a constant permittivity or a function of wavelength.
class Material {
 public:
  Material(scalex permittivity) : f_([permittivity](scalar) { return permittivity; }) {}
  Material(double permittivity) : Material(scalex(permittivity, 0.0)) {}
  Material(std::function<scalex(scalar)> permittivity) : f_(std::move(permittivity)) {}
  scalex permittivity(scalar lambda) const { return f_(lambda); }

 private:
  std::function<scalex(scalar)> f_;
};
inline const Material m_SiO2{1.445 * 1.445};
inline const Material m_Si{3.48 * 3.48};

// GDSIISampler (src/core/GDSIISampler.h:9-57): x-z cross-sections of the first cell of a
// GDSII file (micrometers), PML always on; shift = {x, y, z}.
class GDSIISampler : public LayerSampler {
 public:
  GDSIISampler(scalar lambda, int nx, int ny, const std::string& filename, const std::vector<scalar>& shift = {0, 0, 0},
               scalar lPML = 0.5, scalar rPML = 0.5, scalar lspace = 1.5, scalar rspace = 1.5, scalar bPML = 0.5,
               scalar uPML = 0.5, scalar bspace = 1.08, scalar uspace = 1., const Material& permBackground = m_SiO2,
               const Material& permDevice = m_Si, scalar thickness = 0.22, std::shared_ptr<Context> ctx = nullptr)
      : LayerSampler(lambda, std::move(ctx)) {
    if (shift.size() != 3) throw std::invalid_argument("GDSIISampler: shift needs x, y, z");
    const double ps[8] = {lPML, rPML, lspace, rspace, bPML, uPML, bspace, uspace};
    const scalex eb = permBackground.permittivity(lambda), ed = permDevice.permittivity(lambda);
    const double b[2] = {eb.real(), eb.imag()}, d[2] = {ed.real(), ed.imag()};
    vrc_sampler* s = nullptr;
    detail::check(vrc_gds_sampler_create(ctx_->get(), lambda, nx, ny, filename.c_str(), shift.data(), ps, b, d,
                                         thickness, &s),
                  ctx_->get(), "GDSIISampler");
    finish(s, nx, ny);
    detail::check(vrc_gds_sampler_bounds(s_, bounds_), ctx_->get(), "GDSIISampler bounds");
    VectorXs x;
    MatrixXcs eps;
    sampleLayout(minZ(), x, eps);  // y does not depend on z
  }
  scalar minX() const { return bounds_[0]; }
  scalar maxX() const { return bounds_[1]; }
  scalar minZ() const { return bounds_[2]; }
  scalar maxZ() const { return bounds_[3]; }
  // GDSIISampler::sampleLayout (src/core/GDSIISampler.cpp:94-146); eps is (ny-1) x (nx-1).
  void sampleLayout(scalar z, VectorXs& x, MatrixXcs& eps) {
    int ncx = 0, ncy = 0;
    detail::check(vrc_sampler_layout(s_, z, 0, nullptr, &ncx, 0, nullptr, &ncy, nullptr), ctx_->get(),
                  "sampleLayout");
    x.resize(ncx);
    y_.resize(ncy);
    eps.resize(ncy - 1, ncx - 1);
    detail::check(vrc_sampler_layout(s_, z, ncx, x.data(), &ncx, ncy, y_.data(), &ncy,
                                     reinterpret_cast<double*>(eps.data())),
                  ctx_->get(), "sampleLayout");
  }
  const VectorXs& y() const { return y_; }

 private:
  double bounds_[4] = {0, 0, 0, 0};
  VectorXs y_;
};

// The legacy stacked-layer solver RCWASolver (src/rcwa/RCWASolver.h:60-80): nX, nY are
// harmonic orders; layers are LayerDesc (the reference's Layer(x, y, eps)).
class RCWASolver {
 public:
  RCWASolver(int nX, int nY, std::shared_ptr<Context> ctx = nullptr) : ctx_(ctx ? std::move(ctx) : Context::thread_default()) {
    detail::check(vrc_rcwa_solver_create(ctx_->get(), nX, nY, &h_), ctx_->get(), "RCWASolver");
  }
  ~RCWASolver() {
    if (h_) vrc_rcwa_solver_destroy(h_);
  }
  RCWASolver(const RCWASolver&) = delete;
  RCWASolver& operator=(const RCWASolver&) = delete;

  void addLayer(const LayerDesc& layer) {
    vrc_layer l = layer.view();
    detail::check(vrc_rcwa_solver_add_layer(h_, &l), ctx_->get(), "addLayer");
    ++ntypes_;
  }
  int nLayerTypes() const { return ntypes_; }
  void solve(scalar lambda, const std::vector<int>& layerStack, const std::vector<scalar>& thickness,
             const std::vector<scalar>* translation = nullptr) {
    if (translation && translation->size() != layerStack.size())
      throw Error("translation needs one value per stack entry", -1);
    detail::check(vrc_rcwa_solver_solve(h_, lambda, (int)layerStack.size(), layerStack.data(), thickness.data(),
                                        translation ? translation->data() : nullptr),
                  ctx_->get(), "solve");
    int N = 0, t = 0;
    vrc_rcwa_solver_dims(h_, &N, &t);
    Tdd_.resize(N, N);
    Rud_.resize(N, N);
    detail::check(vrc_rcwa_solver_get(h_, 0, 0, reinterpret_cast<double*>(Tdd_.data())), ctx_->get(), "Tdd");
    detail::check(vrc_rcwa_solver_get(h_, 1, 0, reinterpret_cast<double*>(Rud_.data())), ctx_->get(), "Rud");
  }
  const MatrixXcs& Tdd() const { return Tdd_; }
  const MatrixXcs& Rud() const { return Rud_; }
  scalex tdd(int row, int col) const { return Tdd_(row, col); }
  scalex rud(int row, int col) const { return Rud_(row, col); }
  // src/rcwa/RCWASolver.cpp:582-670
  scalar transmittance(int inputMode, int outputMode) const { return std::norm(Tdd_(outputMode, inputMode)); }
  scalar reflectance(int inputMode, int outputMode) const { return std::norm(Rud_(outputMode, inputMode)); }
  scalar transmittance(int inputMode) const { return colsum(Tdd_, inputMode); }
  scalar reflectance(int inputMode) const { return colsum(Rud_, inputMode); }

  // RCWASolver::evaluateIntermediateField (src/rcwa/RCWASolver.cpp:1373-1621): modal
  // amplitudes of every inner layer of the solved stack (layerStack and thickness must be
  // those of the last solve) for an input mode or a list of (mode, coefficient) pairs.
  void evaluateIntermediateField(std::vector<VectorXcs>& c_m, std::vector<VectorXcs>& c_p, int inputMode,
                                 const std::vector<int>& layerStack, const std::vector<scalar>& thickness) const {
    std::vector<std::pair<int, scalex>> one{{inputMode, scalex(1.0, 0.0)}};
    evaluateIntermediateField(c_m, c_p, one, layerStack, thickness);
  }
  void evaluateIntermediateField(std::vector<VectorXcs>& c_m, std::vector<VectorXcs>& c_p,
                                 const std::vector<std::pair<int, scalex>>& inputCoeffs,
                                 const std::vector<int>& layerStack, const std::vector<scalar>& thickness) const {
    if (layerStack.size() != thickness.size()) throw Error("thickness needs one value per stack entry", -1);
    const long N = Tdd_.rows();
    const int nin = (int)layerStack.size() - 2;
    if (nin < 1) throw Error("the stack has no inner layer", -1);
    VectorXcs in(N, 1);
    for (auto& [m, v] : inputCoeffs) {
      if (m < 0 || m >= N) throw Error("input mode out of range", -1);
      in(m, 0) += v;
    }
    std::vector<scalex> bm((size_t)N * nin), bp((size_t)N * nin);
    detail::check(vrc_rcwa_solver_intermediate_field(h_, 1, reinterpret_cast<const double*>(in.data()),
                                                     reinterpret_cast<double*>(bm.data()),
                                                     reinterpret_cast<double*>(bp.data())),
                  ctx_->get(), "evaluateIntermediateField");
    c_m.assign(nin, VectorXcs(N, 1));
    c_p.assign(nin, VectorXcs(N, 1));
    for (int j = 0; j < nin; ++j)
      for (long r = 0; r < N; ++r) {
        c_m[j](r, 0) = bm[(size_t)j * N + r];
        c_p[j](r, 0) = bp[(size_t)j * N + r];
      }
  }

 private:
  static scalar colsum(const MatrixXcs& M, int c) {
    scalar s = 0;
    for (long r = 0; r < M.rows(); ++r) s += std::norm(M(r, c));
    return s;
  }
  std::shared_ptr<Context> ctx_;
  vrc_rcwa_solver* h_ = nullptr;
  int ntypes_ = 0;
  MatrixXcs Tdd_, Rud_;
};

namespace detail {
// Common state of the three integrators: handle, sampler reference, host copies of
// the results (filled after each compute).
class IntegratorBase {
 public:
  virtual ~IntegratorBase() {
    if (it_) vrc_integrator_destroy(it_);
  }
  IntegratorBase(const IntegratorBase&) = delete;
  IntegratorBase& operator=(const IntegratorBase&) = delete;

  const MatrixXcs& Tdd() const { return m_[0]; }
  const MatrixXcs& Tuu() const { return m_[1]; }
  const MatrixXcs& Rud() const { return m_[2]; }
  const MatrixXcs& Rdu() const { return m_[3]; }
  const MatrixXcs& W0() const { return m_[4]; }
  const MatrixXcs& V0() const { return m_[5]; }
  const VectorXcs& gam0() const { return m_[6]; }
  const MatrixXcs& W1() const { return m_[7]; }
  const MatrixXcs& V1() const { return m_[8]; }
  const VectorXcs& gam1() const { return m_[9]; }
  vrc_integrator* handle() const { return it_; }

 protected:
  IntegratorBase(int kind, scalar z0, scalar z1, std::shared_ptr<LayerSampler> sampler)
      : sampler_(std::move(sampler)) {
    if (!sampler_) throw Error("null sampler", -1);
    vrc_integrator* it = nullptr;
    check(vrc_integrator_create(ctx(), kind, z0, z1, sampler_->handle(), &it), ctx(), "integrator");
    it_ = it;
  }
  vrc_ctx* ctx() const { return sampler_->context()->get(); }
  void fetch() {
    const int N = sampler_->nDim();
    for (int w = 0; w < 10; ++w) {
      const bool vec = (w == 6 || w == 9);
      m_[w].resize(N, vec ? 1 : N);
      check(vrc_integrator_get(it_, w, reinterpret_cast<double*>(m_[w].data())), ctx(), "result read-back");
    }
    check(vrc_integrator_counts(it_, &N_, &nEig_), ctx(), "counts");
  }
  std::shared_ptr<LayerSampler> sampler_;
  vrc_integrator* it_ = nullptr;
  MatrixXcs m_[10];
  int N_ = 0, nEig_ = 0;
};
}  // namespace detail

// RCWAIntegrator (src/core/RCWAIntegrator.h:8-40): N midpoint slices.
class RCWAIntegrator : public detail::IntegratorBase {
 public:
  RCWAIntegrator(scalar z0, scalar z1, std::shared_ptr<LayerSampler> sampler)
      : IntegratorBase(0, z0, z1, std::move(sampler)) {}
  void compute(int N) {
    detail::check(vrc_rcwa_compute(it_, N), ctx(), "RCWAIntegrator::compute");
    fetch();
  }
  // Same result; the slices' eigenproblems are solved in batches.
  void computeBatched(int N) {
    detail::check(vrc_rcwa_compute_batched(it_, N), ctx(), "RCWAIntegrator::computeBatched");
    fetch();
  }
};

// RedhefferIntegrator (src/core/RedhefferIntegrator.h:10-90): adaptive trisection.
class RedhefferIntegrator : public detail::IntegratorBase {
 public:
  RedhefferIntegrator(scalar z0, scalar z1, std::shared_ptr<LayerSampler> sampler)
      : IntegratorBase(1, z0, z1, std::move(sampler)) {}
  void compute(int p, int maxN = 10, scalar pe = 1e-2) {
    detail::check(vrc_redheffer_compute(it_, p, maxN, pe), ctx(), "RedhefferIntegrator::compute");
    fetch();
  }
  // Several integrators (one context, equal N; e.g. one per wavelength) at once: the
  // results of compute() on each, with the new reference eigensystems of each
  // subdivision level solved in one batched launch.
  static void computeMany(const std::vector<RedhefferIntegrator*>& its, int p, int maxN = 10, scalar pe = 1e-2) {
    if (its.empty()) return;
    std::vector<vrc_integrator*> h;
    for (auto* x : its) h.push_back(x->it_);
    detail::check(vrc_redheffer_compute_many((int)h.size(), h.data(), p, maxN, pe), its[0]->ctx(),
                  "RedhefferIntegrator::computeMany");
    for (auto* x : its) x->fetch();
  }
  int N() const { return N_; }
  int nEigVal() const { return nEig_; }
};

// DifferentialIntegrator (src/core/DifferentialIntegrator.h:8-40).
class DifferentialIntegrator : public detail::IntegratorBase {
 public:
  DifferentialIntegrator(scalar z0, scalar z1, std::shared_ptr<LayerSampler> sampler)
      : IntegratorBase(2, z0, z1, std::move(sampler)) {}
  void compute(int N) {
    detail::check(vrc_differential_compute(it_, N), ctx(), "DifferentialIntegrator::compute");
    fetch();
  }
};

}  // namespace varrcwa
