// Batched device procedures of the S-matrix path and the host-side sampler /
// integrator classes that keep the reference's API.
#pragma once

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "vrc_context.hpp"
#include "vrc_geom.hpp"
#include "vrc_internal.hpp"

namespace vrc {

void fill_int(hipStream_t s, int* p, int v, long n);
// Looks up the device address of the allocation-free entry point's scratch array on this
// device (current) once per device.
void noalloc_init_device(int device);

// ---------------------------------------------------------------- numerical failures
// (codes and the per-launch notes in vrc_internal.hpp)
void health_reset(vrc_ctx& c);
// Synchronizes; returns the first recorded code (0 none) and its description.
int health_read(vrc_ctx& c, std::string& msg);
const char* health_text(int code);
// flags[i] = code (if still 0) for each instance i whose `per` values at A + i*stride hold a
// NaN or an infinity; flags[owner[q]] = code for each job q with info[q] != 0.
void nonfinite_flags(hipStream_t s, const cd* A, long per, long stride, int ninst, int* flags, int code);
void info_to_owner(hipStream_t s, const int* info, const int* owner, int njobs, int code, int* flags);

// ---------------------------------------------------------------- geometry (host)
// FourierSolver1D::reorder/refill semantics (src/rcwa/FourierSolver1D.cpp:145-193)
// applied to one layer: painter-order values and the two interval lists.
struct PreparedLayer {
  std::vector<double> xu, xs, ys;
  std::vector<cd> val;  // [strip][cell], ncy x ncx
  double Lx, Ly;
  int ncx, ncy;
};
PreparedLayer prepare_layer(const HostLayer& L);

// alpha/beta of LayerSampler::evaluateAlphaBeta (src/core/LayerSampler.cpp:9-32).
void alpha_beta(int nx, int ny, double Lx, double Ly, std::vector<cd>& alpha, std::vector<cd>& beta);
// Simulation region of a sampler (LayerSampler's Lx_, Ly_, b0_ .. r1_,
// src/core/LayerSampler.h:7-28).
struct Region {
  double Lx = 0, Ly = 0, b0 = 0, b1 = 0, u0 = 0, u1 = 0, l0 = 0, l1 = 0, r0 = 0, r1 = 0;
};
Region region_from_coords(const std::vector<double>& x, const std::vector<double>& y);
// Dense PML-stretched K matrices (src/core/LayerSampler.cpp:34-131), n x n column-major.
void pml_k_matrices(int nx, int ny, const std::vector<double>& x, const std::vector<double>& y,
                    std::vector<cd>& Kx, std::vector<cd>& Ky);
void pml_k_matrices(int nx, int ny, const Region& r, std::vector<cd>& Kx, std::vector<cd>& Ky);

// ---------------------------------------------------------------- (a) + (b)
struct AssemblyBatch {
  int nx = 0, ny = 0, nlayers = 0;
  std::vector<const HostLayer*> layers;
  bool periodic = true;
  const cd* d_alpha = nullptr;  // periodic: [l][n] (stride sab; 0 = shared)
  const cd* d_beta = nullptr;
  long sab = 0;
  const double* d_k0sq = nullptr;  // [l]
  // PML: shared dense K and products
  const cd *d_Kx = nullptr, *d_Ky = nullptr, *d_KxKy = nullptr, *d_KxKx = nullptr, *d_KyKy = nullptr,
           *d_KyKx = nullptr;
  double k0sq_uniform = 0.0;
};
void assemble_fe(vrc_ctx& c, const AssemblyBatch& a, cd* fe11, cd* fe22, cd* fe33, long sfe);

// Uploaded layer batch and workspace of the Fourier kernels (device view in g).
struct FourierStage {
  DevBuf geo, wk;
  FourierDev g{};
  int L = 0;
};
FourierStage fourier_prepare(vrc_ctx& c, int nx, int ny, const std::vector<const HostLayer*>& layers);
// P, Q of every layer (N x N, stride N^2); PQ (optional) also receives P Q, for periodic
// cells without the N x N product.
void assemble_pq(vrc_ctx& c, const AssemblyBatch& a, cd* P, cd* Q, cd* PQ = nullptr);

// ---------------------------------------------------------------- (c)
// PQ = P Q; eig; gamma = sqrt branch; V = Q W diag(1/gamma)  (src/rcwa/WaveEquationSolver.cpp:24-37)
void wave_eqn(vrc_ctx& c, int N, const cd* P, const cd* Q, long sPQ, cd* W, cd* V, cd* gam, int* info, int batch);
// X[b][i] = exp(i gam[b][i] s[b]) (d_s: device, one value per batch entry).
void phases(hipStream_t s, int n, const cd* gam, long sg, const double* d_s, cd* X, long sx, int batch);

// The legacy stacked-layer solver RCWASolver (src/rcwa/RCWASolver.cpp:73-580): layer types,
// interface matrices and the RTCM recursion; Tdd and Rud stay on the device.
class LegacySolver {
 public:
  LegacySolver(vrc_ctx* c, int nX, int nY);
  void add_layer(const HostLayer& L);
  void solve(double lambda, const std::vector<int>& stack, const std::vector<double>& thickness,
             const double* translation);
  int N() const { return 2 * nx_ * ny_; }
  int ntypes() const { return (int)layers_.size(); }
  // which: 0 Tdd, 1 Rud (N x N); layer type k: 2 E, 3 H (N x N), 4 gamma (N) (gamma = sqrt(eig) / k0)
  void get(int which, int k, cd* host) const;
  bool solved() const { return solved_; }
  int nstack() const { return (int)stack_.size(); }
  // RCWASolver::evaluateIntermediateField (src/rcwa/RCWASolver.cpp:1373-1621) for the stack of
  // the last solve: inputs (N x ncols, host, column-major) -> per inner layer j (nstack - 2
  // of them) c_m[j], c_p[j] (N x ncols each, host, layer-major).
  void intermediate_field(int ncols, const cd* input, cd* c_m, cd* c_p);

 private:
  vrc_ctx* c_;
  int nx_, ny_;
  std::vector<HostLayer> layers_;
  DevBuf W_, V_, gam_, Tdd_, Rud_;
  // kept for the intermediate fields: interface matrices (unique pairs), the stack's
  // interface map, the Rud of every RTCM step, the phase scale of every stack entry
  DevBuf IE_, IH_, REF_, ds_;
  std::vector<int> stack_, imap_;
  double k0_ = 0.0;
  bool solved_ = false;
};

// Eigen factors instead of W, V: W = diag(dbal) Z X (unnormalized), and the vectors.
// PQ (optional): P Q already formed (stride N^2); consumed as the eigensolver's work matrix.
void wave_eqn_factors(vrc_ctx& c, int N, const cd* P, const cd* Q, long sPQ, cd* Z, cd* X, cd* dbal, cd* gam,
                      int* info, int batch, cd* PQ = nullptr);
void wave_eqn_vectors(vrc_ctx& c, int N, const cd* Q, long sQ, const cd* Z, const cd* X, const cd* dbal,
                      const cd* gam, cd* W, cd* V, int batch);

// Closed-form W (= I), V, gamma of homogeneous periodic media (eps per batch entry).
void homogeneous_eig(hipStream_t s, int n, int batch, const cd* alpha, const cd* beta, long sab, const double* k0sq,
                     const cd* eps, cd* W, cd* V, cd* gam);

// ---------------------------------------------------------------- (d)
struct SMat {
  cd *Tdd = nullptr, *Tuu = nullptr, *Rud = nullptr, *Rdu = nullptr;
};
struct LUPair {
  const cd* W = nullptr;  // LU factors of W
  const cd* V = nullptr;  // LU factors of V
  const int* pW = nullptr;
  const int* pV = nullptr;
  long sM = 0;  // batch stride of the factors (pivot stride is N)
};
struct RefSys {
  const cd *W, *V, *gam;
  LUPair LU;
};

void gather_mats(hipStream_t s, long nn, const cd* src, long ss, const int* d_idx, cd* dst, long sd, int batch);
void stack_set_cascade_group(int n);  // instances per batched cascade of batch_stack_rcwa
void rcwa_set_group(int g);           // slices per eigensolve launch of RCWAIntegrator (1: one at a time)
// Slab S-matrix (src/rcwa/RCWAScatterMatrix.cpp:26-85); d_s[b] = dz_b / k0_b.
void layer_smatrix(vrc_ctx& c, int N, int batch, const double* d_s, const cd* W, const cd* V, const cd* gam,
                   const cd* W0, const cd* V0, long sRef, cd* Tuu, cd* Rud);
// Slab S-matrix from the eigen factors of the slab (P per slab, stride sP). ref_homog: the
// reference basis is that of a homogeneous medium (W0 = I, V0 nonzero only where the row
// and column harmonics agree), used without products by W0 and V0.
void layer_smatrix_factors(vrc_ctx& c, int N, int batch, const double* d_s, const cd* P, long sP, const cd* Z,
                           const cd* X, const cd* dbal, const cd* gam, const cd* W0, const cd* V0, long sRef,
                           cd* Tuu, cd* Rud, bool ref_homog = false);
// Systems of the mirror-symmetric slab (see slab_evenodd in ops.hip): per instance b,
// L[2b] = (1-x) F0 + (1+x) G0, R[2b] = (1+x) G0 - (1-x) F0, L[2b+1] = (1+x) F0 + (1-x) G0,
// R[2b+1] = (1-x) G0 - (1+x) F0 (x = exp(i gamma s) scaling rows; G = Gamma G0 when gam is set).
void evenodd_system(hipStream_t s, int N, int batch, const cd* F0, long sF, const cd* G, long sG, const cd* gam,
                    const cd* x, cd* L, cd* R);
// Rud = (E[2b] + E[2b+1]) / 2, Tuu = (E[2b] - E[2b+1]) / 2.
void evenodd_combine(hipStream_t s, int N, int batch, const cd* E, cd* Tuu, cd* Rud);
// Star product of src/core/RCWAIntegrator.cpp:57-67 (S updated in place; slice stride ss).
// save (optional, STAR_SAVE * batch * N*N): receives M^{-1}, Yt, Yr, K, T1, Z2 of the
// factored form (M = I - Rud rdu, [Yt | Yr] = M^{-1} [Tuu | Rud], K = Yr tdd,
// T1 = tdd + rdu K, Z2 = rdu Yt) for the reverse pass.
constexpr int STAR_SAVE = 6;
void star_product(vrc_ctx& c, int N, int batch, SMat& S, const cd* tdd, const cd* tuu, const cd* rud,
                  const cd* rdu, long ss, cd* save = nullptr);
// Peak workspace bytes of star_product without save.
size_t star_product_scratch_bytes(int N, int batch);
void project_left(vrc_ctx& c, int N, int batch, SMat& S, const LUPair& LU1, const cd* W0, const cd* V0, long sRef);
// project_left for homogeneous end media (W0 = W1 = I; V0, V1 couple only harmonic pairs).
void project_left_homog(vrc_ctx& c, int N, int batch, SMat& S, const cd* V0, const cd* V1, long sRef);
void project_right(vrc_ctx& c, int N, int batch, SMat& S, const LUPair& LUl, const cd* W1, const cd* V1,
                   long sRef);
double redheffer_perturbation(vrc_ctx& c, int N, double k0, double z0, double z1, const cd* Pr, const cd* Qr,
                              const RefSys& R, const cd* P3[3], const cd* Q3[3], cd* tdd, cd* tuu, cd* rud,
                              cd* rdu);
// The same without the readback: d_hs (device) holds {(z1 - z0) / (2 k0), (z1 - z0) / k0};
// the error estimate is the largest of the RH_ERR_BLOCKS values written to err (device).
constexpr int RH_ERR_BLOCKS = 256;
// One varRCWA segment for redheffer_perturbation_batch.
struct RhJob {
  double k0, z0, z1;
  const cd *Pr, *Qr;
  RefSys R;
  const cd* P3[3];
  const cd* Q3[3];
};
void redheffer_perturbation_batch(vrc_ctx& c, int N, const std::vector<RhJob>& jobs, cd* T, double* err);
void redheffer_perturbation_enqueue(vrc_ctx& c, int N, double k0, double z0, double z1, const cd* Pr,
                                    const cd* Qr, const RefSys& R, const cd* P3[3], const cd* Q3[3], cd* tdd,
                                    cd* tuu, cd* rud, cd* rdu, const double* d_hs, double* err);

// Tangents of P and Q (and fe11, fe22, fe33; dP/dlambda, dQ/dlambda) of one layer with
// respect to npara parameters of its interior x coordinates (Sampler::shape_derivatives
// without a sampler object): d_alpha holds alpha then beta (n each, periodic cells), d_K
// the six dense K products (absorbing regions, else null).
void shape_tangents(vrc_ctx& c, int nx, int ny, double lambda, bool pml, const cd* d_alpha, const cd* d_K,
                    const HostLayer& L, int nvar, const int* var_index, const double* var_coeff, int npara, cd* dP,
                    cd* dQ, cd* dfe11, cd* dfe22, cd* dfe33, cd* dPl, cd* dQl);

// ---------------------------------------------------------------- samplers
typedef double (*RuleFn)(double z, void* user);

class Sampler {
 public:
  Sampler(vrc_ctx* c, double lambda, int nx, int ny, bool pml);
  virtual ~Sampler() = default;
  // The layer at z (geometric samplers; others throw).
  virtual HostLayer layer_at(double z) const = 0;
  int nDim() const { return 2 * nx_ * ny_; }
  int n() const { return nx_ * ny_; }
  double lambda() const { return lambda_; }
  double k0() const;
  vrc_ctx* ctx() const { return c_; }
  // P, Q (N x N, stride N*N) for every z (device).
  virtual void sample_gpu(const std::vector<double>& zs, cd* P, cd* Q);
  // P, Q (N x N) into host buffers.
  virtual void sample_host(double z, cd* P, cd* Q);
  const Region& region() const { return region_; }
  // P, Q of explicit layers with this sampler's wavelength and K matrices.
  void sample_layers(const std::vector<HostLayer>& ls, cd* P, cd* Q);
  const std::vector<cd>& Kx() const { return hKx_; }
  const std::vector<cd>& Ky() const { return hKy_; }
  // Tangents of P and Q of layer L (this sampler's wavelength and K, held fixed) with
  // respect to npara parameters of its x coordinates: x[var_index[k]] moves with
  // var_coeff[k * npara + p]. Device outputs, npara matrices each: dP, dQ (N x N);
  // dfe11/dfe22/dfe33 (n x n, may be null); dPl, dQl: d/d lambda (N x N, may be null).
  void shape_derivatives(const HostLayer& L, int nvar, const int* var_index, const double* var_coeff, int npara,
                         cd* dP, cd* dQ, cd* dfe11, cd* dfe22, cd* dfe33, cd* dPl, cd* dQl);

 protected:
  void post_init(const std::vector<double>& x, const std::vector<double>& y);
  void post_init(const Region& r);
  vrc_ctx* c_;
  Region region_;
  double lambda_;
  int nx_, ny_;
  bool pml_;
  std::vector<cd> hKx_, hKy_, alpha_, beta_;
  DevBuf d_alpha_, d_beta_, d_K_;  // d_K_: Kx, Ky, KxKy, KxKx, KyKy, KyKx (PML)
};

class VarSampler : public Sampler {
 public:
  VarSampler(vrc_ctx* c, double lambda, int nx, int ny, HostLayer base, bool pml);
  void add_rule(int axis, int layout, RuleFn fn, void* user);
  HostLayer layer_at(double z) const override;

 private:
  HostLayer base_;
  std::map<int, std::pair<RuleFn, void*>> xr_, yr_;
};

// A sampler whose P and Q come from the caller (the reference's user subclasses of
// LayerSampler, src/core/LayerSampler.h:5-77): host_fn fills host N x N column-major P, Q;
// device_fn fills device buffers on the given stream (returns 1 to defer to host_fn).
// Either may be null, not both. K matrices from the given region (postInitialization).
typedef int (*SampleHostFn)(double z, double* P, double* Q, void* user);
typedef int (*SampleDeviceFn)(double z, void* P, void* Q, void* stream, void* user);
class CallbackSampler : public Sampler {
 public:
  CallbackSampler(vrc_ctx* c, double lambda, int nx, int ny, const Region& r, bool pml, SampleHostFn hf,
                  SampleDeviceFn df, void* user);
  HostLayer layer_at(double z) const override;
  void sample_gpu(const std::vector<double>& zs, cd* P, cd* Q) override;
  void sample_host(double z, cd* P, cd* Q) override;

 private:
  SampleHostFn hf_;
  SampleDeviceFn df_;
  void* user_;
};

class StackSampler : public Sampler {
 public:
  StackSampler(vrc_ctx* c, double lambda, int nx, int ny, std::vector<HostLayer> layers, std::vector<double> z_edges,
               HostLayer bottom, HostLayer top, bool pml);
  HostLayer layer_at(double z) const override;

 private:
  std::vector<HostLayer> layers_;
  std::vector<double> ze_;
  HostLayer bottom_, top_;
};

// GDSII polygons (gdstk::read_gds subset: BOUNDARY and BOX elements of every cell,
// src/gdstk/library.cpp:160-420); coordinates in units of `unit` meters.
struct GdsPolygon {
  std::vector<double> x, y;
  int layer = 0, datatype = 0;
};
struct GdsCell {
  std::string name;
  std::vector<GdsPolygon> polygons;
};
struct GdsLibrary {
  std::string name;
  double unit = 0, precision = 0;
  std::vector<GdsCell> cells;
};
double gdsii_real(uint64_t r);
GdsLibrary read_gds(const std::string& path, double unit);

// GDSIISampler (src/core/GDSIISampler.cpp:17-146): x-z cross-sections of the first cell's
// polygons, always with PML. pml_space: lPML, rPML, lspace, rspace, bPML, uPML, bspace, uspace.
class GdsSampler : public Sampler {
 public:
  GdsSampler(vrc_ctx* c, double lambda, int nx, int ny, const std::string& file, const double shift[3],
             const double pml_space[8], cd eps_background, cd eps_device, double thickness);
  HostLayer layer_at(double z) const override;
  double minX() const { return minX_; }
  double maxX() const { return maxX_; }
  double minZ() const { return minZ_; }
  double maxZ() const { return maxZ_; }

 private:
  std::vector<GdsPolygon> polys_;
  double shift_[3];
  cd eps_bg_, eps_dev_;
  double thickness_;
  double minX_, maxX_, minZ_, maxZ_;
  std::vector<double> xr_, y_;
};

// ---------------------------------------------------------------- integrators
class Integrator {
 public:
  Integrator(vrc_ctx* c, double z0, double z1, Sampler* s);
  virtual ~Integrator() = default;
  // which: 0 Tdd, 1 Tuu, 2 Rud, 3 Rdu, 4 W0, 5 V0, 6 gam0, 7 W1, 8 V1, 9 gam1
  void get(int which, cd* host) const;
  int N_ = 0, nEig_ = 0;
  std::vector<std::pair<double, double>> tree_;

  void rcwa(const std::vector<std::pair<double, double>>& slices, bool batched);  // (mid z, dz)
  double it_z0() const { return z0_; }
  double it_z1() const { return z1_; }
  void redheffer(double pe, int max_depth);
  // RedhefferIntegrator::compute of several integrators (one context, equal N) at once:
  // the subdivision trees are grown breadth-first, with every new reference eigensystem of
  // a level (across segments and integrators) solved in one batched launch.
  static void redheffer_many(const std::vector<Integrator*>& its, double pe, int max_depth);
  void differential(int N);

 protected:
  void alloc();
  void eig_at(double z, cd* P, cd* Q, cd* W, cd* V, cd* gam);
  vrc_ctx* c_;
  double z0_, z1_;
  Sampler* s_;
  int Nd_ = 0;
  DevBuf S_, W0_, V0_, g0_, W1_, V1_, g1_, info_;
  SMat S() const;
};

// ---------------------------------------------------------------- batched stacks
struct StackInstance {
  double lambda;
  std::vector<const HostLayer*> layers;
  std::vector<double> thickness;
  const HostLayer* bottom;
  const HostLayer* top;
};
struct StackTape;
// Autotuner (autotune.cpp): the best switches for N = 2 nx ny on this sample of a batch
// (returned, not stored), and a one-line report; the table and its cache file.
std::string autotune_stack(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& sample, double budget_s,
                           TuneKnobs* best_out);
void tune_store(int N, const TuneKnobs& k);
void tune_clear();
void tune_save(const std::string& path, int N, const TuneKnobs& k, const std::string& note);
int tune_load(const std::string& path);
// Plane-wave diffraction efficiencies per instance (n per instance) from the reflected and
// transmitted field harmonics rxy, txy ([Ex; Ey], 2n per instance).
void efficiencies(hipStream_t s, int n, int batch, const cd* rxy, const cd* txy, const cd* ab, const double* k0v,
                  const cd* eps2, double* REF, double* TRN);
// Window Gram matrix of the harmonic wavelets over nres midpoints of [x0, x1] (sgn -1), or
// its transpose (sgn +1); nh x nh column-major.
void window_gram(hipStream_t s, int nh, int nres, double x0, double x1, double L, double sgn, cd* G);
void modal_power_reduce(hipStream_t s, int n, int nmodes, const cd* W, long ldw, const cd* Z, double dxdy,
                        double* energy);
// Power of nmodes modes (field harmonics W [Ex; Ey], V [Hx; Hy] per column) through the window
// [x0, x1] x [y0, y1] sampled at nres_x x nres_y midpoints.
void modal_power(vrc_ctx& c, int nx, int ny, int nmodes, const cd* W, long ldw, const cd* V, long ldv, double Lx,
                 double Ly, int nres_x, int nres_y, double x0, double x1, double y0, double y1, double* energy);
// For instances solved by batch_stack_rcwa (S, end bases W0, W1 in its output layout): the
// wave e_inc ([Ex; Ey] harmonics, 2n per instance) incident from the bottom medium;
// REF, TRN: efficiency per harmonic of the reflected / transmitted orders.
void stack_efficiencies(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst, const cd* S,
                        const cd* W0, const cd* W1, const cd* e_inc, double* REF, double* TRN);
// Out: S[inst] = [Tdd|Tuu|Rud|Rdu] (4 N x N); optional W0/V0/gam0/W1/V1/gam1 per instance.
// With a tape, the eigen-systems of every instance are kept for the reverse pass.
// flags (device, one int per instance, zeroed by the caller, or NULL): H_EIG for an instance
// whose eigensolves failed, H_NONFINITE for one whose S-matrix holds NaN or infinity.
void batch_stack_rcwa(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst, int chunk, cd* S_out,
                      cd* W0, cd* V0, cd* g0, cd* W1, cd* V1, cd* g1, StackTape* tape = nullptr,
                      int* flags = nullptr);

}  // namespace vrc
