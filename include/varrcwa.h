/*
 * varrcwa.h -- C ABI of the MI355X (gfx950) scattering-matrix library.
 *
 * Conventions (every entry point):
 *  - complex128 values are two doubles (re, im), interleaved; matrices are
 *    column-major with an explicit leading dimension; strided batches give the
 *    distance in complex elements between consecutive matrices.
 *  - pointers are device pointers unless the parameter name ends in _host.
 *  - all work is enqueued on the context's stream; calls return without waiting
 *    unless documented (results read back to _host buffers synchronize).
 *  - status: 0 ok, -1 bad argument, -2 HIP runtime failure, -3 other failure;
 *    vrc_last_error() gives the text. Per-matrix numerical failures are reported
 *    LAPACK-style through info arrays (0 ok, >0 failing index) by the dense kernels;
 *    the calls that solve whole problems (integrator computes, batched stacks, the
 *    legacy solver) read them back once at the end and return a positive status on a
 *    numerical failure, with the results written as computed:
 *      1 the eigensolver did not converge, 2 an LU met an exactly zero pivot,
 *      3 a zero propagation constant (Rayleigh/Wood anomaly: the reference divides by
 *        gamma in V = Q W / gamma, src/rcwa/WaveEquationSolver.cpp:36),
 *      4 non-finite results, 5 varRCWA reached its subdivision depth limit (40 levels of
 *        trisection) with the segment error still above pe.
 *    Non-finite or zero permittivity values are bad arguments (-1): the inverse Fourier
 *    rules need 1/eps of every cell.
 *  - one context per host thread; calls on one context are not re-entrant.
 *
 * Harmonic ordering follows the reference: index m*ny + n (x order outer), field
 * rows [x harmonics; y harmonics] (src/core/LayerSampler.cpp:9-32).
 */
#ifndef VARRCWA_H
#define VARRCWA_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vrc_ctx vrc_ctx;
typedef struct vrc_sampler vrc_sampler;
typedef struct vrc_integrator vrc_integrator;

/* ------------------------------------------------------------------ context */
/* No reference counterpart: the reference creates cuBLAS/cuSOLVER/MAGMA handles per
 * integrator (src/gpu/RedhefferIntegratorGPU.cpp:99-146). stream is an existing
 * hipStream_t or NULL for the device's default (null) stream. */
int vrc_ctx_create(int device, void* stream, vrc_ctx** out);
int vrc_ctx_destroy(vrc_ctx* ctx);
const char* vrc_last_error(const vrc_ctx* ctx);
int vrc_sync(vrc_ctx* ctx);
/* Bytes currently held by the context's device memory pool. */
long long vrc_pool_bytes(const vrc_ctx* ctx);
int vrc_pool_trim(vrc_ctx* ctx);
const char* vrc_version(void);
/* Stage timers (no reference counterpart; the reference's timers are commented out,
 * src/core/VarLayerSampler.cpp:62-68): when enabled, HIP events bracket each stage on
 * the context stream. vrc_timing_read synchronizes, returns the accumulated
 * milliseconds and launch counts per stage (0 Fourier assembly, 1 P/Q assembly
 * incl. stage 0, 2 Hessenberg + Q, 3 QR iteration, 4 eigenvectors, 5 PQ product and
 * V formation, 6 slab S-matrices, 7 star products, 8 end projections, 9 reverse
 * pass) and resets. */
int vrc_timing_enable(vrc_ctx* ctx, int on);
/* Kernel tuning switches (process-wide; no reference counterpart). Keys:
 * "zgemm.large_tile" (default 1): 128x64-tile ZGEMM for m >= 128, n >= 64, else the
 * 64x64 tile; "zgemm.large_tile_min_k" (default 129): smallest k for the 128x64 tile;
 * "zgemm.big_waves", "zgemm.small_waves" (default 8, or 4): wavefronts per workgroup of the
 * 128x64 and 64x64 tiles; "zgemm.big_n" (default 128, or 64): a 128x128 tile for n >= 128; "zgeev.aed_window" (default 24, 0 = off, at most 64): deflation-window
 * size of the aggressive early deflation in the QR iteration; "zgeev.bulges" (default 12,
 * 1..16): double-shift bulges chased per QR sweep; "rcwa.group" (default 64): slices of
 * RCWAIntegrator::compute solved per eigensolve launch (1: one slice at a time, as the
 * reference's GPU integrator); "stack.cascade_group" (default 512, fewer when their slab
 * S-matrices would exceed 96 GiB): instances of vrc_batch_stack_rcwa whose star products
 * and projections run in one batch;
 * "zgeev.qr_variant" (default -1: automatic, variant 4 for N <= 256 in launches of at least
 * 384 matrices and for N <= 512 in launches of at least 768, variant 2 for N <= 1024 in
 * launches of at least 384, else 0): the QR (Schur) kernel, 0: 64-wide window on 512 threads,
 * window products formed from registers; 1: the same with the window products staged
 * through LDS; 2: 48-wide window on 256 threads, two matrices per CU, register products;
 * 4: 32-wide window on 128 threads, four matrices per CU (small N); batched calls size their
 * eigensolver launches to 256, 512 or 1024 matrices accordingly;
 * "zgeev.bulges_w48" (default 8, 1..12) and "zgeev.bulges_w32" (default 6, 1..7): bulges per
 * sweep of variants 2 and 4;
 * "zgeev.max_sweeps" (default 0 = 30 * max(N, 10)): QR sweeps per matrix before the
 * eigensolver gives up and reports info > 0; "redheffer.max_depth" (default 40, 0..60):
 * trisection depth at which vrc_redheffer_compute accepts a segment regardless of pe
 * (status 5).
 * Returns -1 for an unknown key. */
int vrc_tune_set(const char* key, long long value);
int vrc_timing_read(vrc_ctx* ctx, double* ms, int* counts, int nstages);
/* Autotuner (no reference counterpart). vrc_autotune_stack times vrc_batch_stack_rcwa on
 * the given instances (a sample of the batch to be solved; a few hundred slabs make full
 * launches) under candidate kernel switch settings, one switch at a time (the QR kernel
 * (zgeev.qr_variant 0, 2 or 4), the QR iteration's deflation window, the bulge count of
 * each QR kernel the sample launches, the ZGEMM tile shape, wavefronts and large-tile
 * threshold; settings that cannot change what runs are skipped), keeping a candidate only
 * when it is >= 1 % faster, until
 * budget_s seconds have been spent. The winners are kept for N = 2 nx ny and put in
 * place before every later call that solves N x N systems (other N keep the values set
 * by vrc_tune_set); with cache_path the entry (and the report as a comment) is appended
 * to that text file. report (optional) receives a one-line summary. Synchronizes.
 * vrc_autotune_load applies a cache file's entries (returns how many, -1 on error);
 * vrc_autotune_clear forgets every entry. The switches are process-wide. */
struct vrc_stack_instance;
int vrc_autotune_stack(vrc_ctx* ctx, int nx, int ny, int ninst, const struct vrc_stack_instance* inst,
                       double budget_s, const char* cache_path, char* report, int report_len);
int vrc_autotune_load(const char* path);
int vrc_autotune_clear(void);
/* QR-iteration statistics gathered while timing is enabled, then reset: out[6] =
 * {matrices, sweeps, windows, window-chase ticks, off-window-apply ticks,
 * small-block ticks} (ticks of the device's constant-rate clock, 100 MHz). */
int vrc_qr_stats(vrc_ctx* ctx, long long* out);

/* ------------------------------------------------- dense complex128 kernels */
/* op: 0 = N, 1 = T, 2 = C (conjugate transpose).
 * C[b] = alpha*op(A[b])*op(B[b]) + beta*C[b], b < batch.
 * Replaces cublasZgemm3m (src/gpu/types.cuh:33) at every site listed for P/Q
 * assembly, PQ, QW and the star products (src/gpu/CudaLayerSampler.cu:88-195,
 * src/gpu/RedhefferIntegratorGPU.cpp:195-201,801-912); computed as a 4M product. */
int vrc_zgemm(vrc_ctx* ctx, int opA, int opB, int m, int n, int k, const double* alpha_host, const void* A,
              long lda, long strideA, const void* B, long ldb, long strideB, const double* beta_host, void* C,
              long ldc, long strideC, int batch);

/* LU with partial pivoting (0-based ipiv). Replaces cusolverDnZgetrf
 * (src/gpu/types.cuh:31) and Eigen PartialPivLU::compute
 * (src/core/RCWAIntegrator.cpp:30-34, src/rcwa/RCWAScatterMatrix.cpp:50-78). */
int vrc_zgetrf(vrc_ctx* ctx, int n, void* A, long lda, long strideA, int* ipiv, long strideP, int* info,
               int batch);

/* X = A^{-1} B from a vrc_zgetrf factorization (X may equal B). Replaces
 * cusolverDnZgetrs (src/gpu/types.cuh:32) and PartialPivLU::solve. */
int vrc_zgetrs(vrc_ctx* ctx, int n, int nrhs, const void* LU, long lda, long strideA, const int* ipiv,
               long strideP, const void* B, long ldb, long strideB, void* X, long ldx, long strideX, int batch);

/* Right eigenvectors of general matrices (A destroyed): eigenvalues w (n per
 * matrix), VR unit 2-norm with the largest component real. Replaces
 * MKLEigenSolver::compute -> LAPACKE_zgeev(ROW_MAJOR,'N','V')
 * (src/rcwa/MKLEigenSolver.hpp:24-74) and magma_zgeev
 * (src/gpu/RedhefferIntegratorGPU.cpp:216-229). */
int vrc_zgeev(vrc_ctx* ctx, int n, void* A, long lda, long strideA, void* w, long strideW, void* VR, long ldv,
              long strideV, int* info, int batch);


/* ------------------------------------------------------- layer geometry */
/* One Layer(x, y, eps) slab (src/rcwa/Layer.cpp:18-34): ncoord_x coordinates x,
 * ncoord_y coordinates y (host arrays), eps = (ncoord_y-1) x (ncoord_x-1) complex,
 * column-major (row i = y strip [y_i, y_i+1], column j = x cell [x_j, x_j+1]). */
typedef struct vrc_layer {
  int ncoord_x;
  const double* x;
  int ncoord_y;
  const double* y;
  const double* eps;
} vrc_layer;

/* (a) Fourier convolution matrices fe11 (inverse rule along x), fe22 (inverse
 * rule along y), fe33 (Laurent) of nlayers layers, n x n each (n = nx*ny), stride
 * in complex elements between layers. Replaces FourierSolver2D::solve with
 * DiscontinuousX / DiscontinuousY / ContinuousXY (src/rcwa/Layer.cpp:63-80,
 * src/rcwa/FourierSolver2D.cpp:12-351) and CudaFourierSolver2D
 * (src/gpu/CudaFourierSolver2D.cu:98-394). */
int vrc_assemble_toeplitz(vrc_ctx* ctx, int nx, int ny, int nlayers, const vrc_layer* layers, void* fe11,
                          void* fe22, void* fe33, long stride);

/* (a)+(b) P and Q (N x N, N = 2*nx*ny) of nlayers layers at one wavelength.
 * region: the layer whose coordinates define the cell and, with enable_pml, the
 * PML stretch (LayerSampler::postInitialization, src/core/LayerSampler.cpp:145-153).
 * Replaces Layer::waveEqnCoeff (src/rcwa/Layer.cpp:36-171) and
 * CudaLayerSampler::sample (src/gpu/CudaLayerSampler.cu:10-203). */
int vrc_assemble_pq(vrc_ctx* ctx, int nx, int ny, int nlayers, const vrc_layer* layers, double lambda,
                    int enable_pml, const vrc_layer* region, void* P, void* Q, long stride);

/* Shape and wavelength derivatives of P and Q for one layer: the derivative outputs of
 * Layer::waveEqnCoeff (src/rcwa/Layer.cpp:36-162) with the tangents of FourierSolver2D::solve
 * (src/rcwa/FourierSolver2D.cpp:12-351; FourierSolver1D::dSolve / dSolve_v2,
 * src/rcwa/FourierSolver1D.cpp:93-321). Parameterization (src/rcwa/Parameterization.h:5-25):
 * x[var_index[k]] = sum_p var_coeff[k*npara + p] * param_p + const for nvar interior x
 * coordinates (y coordinates are not varied, as in the reference); x must be non-decreasing.
 * K (from region, with enable_pml as in vrc_assemble_pq) is held fixed. Device outputs:
 * dP, dQ: npara matrices N x N (stride N*N); dfe11, dfe22, dfe33: npara matrices n x n
 * (stride n*n) or NULL; dPl, dQl: d/d lambda, N x N, or NULL. */
int vrc_wave_eqn_derivatives(vrc_ctx* ctx, int nx, int ny, const vrc_layer* layer, double lambda, int enable_pml,
                             const vrc_layer* region, int nvar, const int* var_index, const double* var_coeff,
                             int npara, void* dP, void* dQ, void* dfe11, void* dfe22, void* dfe33, void* dPl,
                             void* dQl);

/* (c) PQ = P Q; eigen-decomposition; gamma = sqrt(lambda) with the reference branch
 * rule; V = Q W diag(1/gamma). Replaces WaveEquationSolver::solve
 * (src/rcwa/WaveEquationSolver.cpp:24-37) and evaluateEigenvalueAndSetReference
 * (src/gpu/RedhefferIntegratorGPU.cpp:188-284). gam: N per matrix. */
int vrc_eig_pq(vrc_ctx* ctx, int N, int batch, const void* P, const void* Q, long stride, void* W, void* V,
               void* gam, int* info);

/* (d) Slab S-matrix of each slice in the reference basis (W0, V0):
 * dtmm::RCWAScatterMatrix::compute (src/rcwa/RCWAScatterMatrix.cpp:26-85) after its
 * eigen step. lambda_host/dz_host: per slice. ref_stride 0 = one shared basis.
 * Outputs Tuu (= Tdd) and Rud (= Rdu) (src/rcwa/RCWAScatterMatrix.h:44-47). */
int vrc_layer_smatrix(vrc_ctx* ctx, int N, int batch, const double* lambda_host, const double* dz_host,
                      const void* W, const void* V, const void* gam, const void* W0, const void* V0,
                      long ref_stride, void* Tuu, void* Rud);

/* (d) Redheffer star product S <- S * s in place, batched; S blocks (Tdd, Tuu, Rud,
 * Rdu) each N x N with batch stride N*N; slice blocks with batch stride
 * slice_stride (src/core/RCWAIntegrator.cpp:57-67; redhefferStarRightProduct,
 * src/gpu/RedhefferIntegratorGPU.cpp:777-932). */
int vrc_star_product(vrc_ctx* ctx, int N, int batch, void* Tdd, void* Tuu, void* Rud, void* Rdu,
                     const void* tdd, const void* tuu, const void* rud, const void* rdu, long slice_stride);

/* Star product without device allocation. No counterpart in the reference: this is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments, kernels and results
 * as vrc_star_product, but every temporary comes from a fixed device array of
 * VRC_NOALLOC_SCRATCH_BYTES that the library holds on each GPU it runs on until the
 * process ends; the call makes no hipMalloc/hipFree. Limit: the temporaries, about
 * 7 N^2 batch complex values, must fit (N = 882 up to batch 1, N = 450 up to batch 5,
 * N = 242 up to batch 19); larger requests return -1 with the size in the error text.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, contexts or threads) race and must be serialized by the caller. */
#define VRC_NOALLOC_SCRATCH_BYTES (128ull << 20)
int vrc_star_product_noalloc(vrc_ctx* ctx, int N, int batch, void* Tdd, void* Tuu, void* Rud, void* Rdu,
                             const void* tdd, const void* tuu, const void* rud, const void* rdu,
                             long slice_stride);

/* (d) End projection of the left side from basis (W1, V1) to (W0, V0)
 * (src/core/RCWAIntegrator.cpp:73-86; reprojectLeftEnd,
 * src/gpu/RedhefferIntegratorGPU.cpp:500-627). */
int vrc_project_left(vrc_ctx* ctx, int N, int batch, void* Tdd, void* Tuu, void* Rud, void* Rdu,
                     const void* W1, const void* V1, const void* W0, const void* V0, long ref_stride);

/* (d) varRCWA segment [z0, z1] around the reference (Pr, Qr, W, V, gam): Simpson
 * perturbation with samples at z0, (z0+z1)/2, z1 and the max-modulus error estimate
 * (RedhefferIntegrator::evaluatePerturbationAndDifference,
 * src/core/RedhefferIntegrator.cpp:317-399). Synchronizes (err_host). */
int vrc_redheffer_segment(vrc_ctx* ctx, int N, double lambda, double z0, double z1, const void* Pr,
                          const void* Qr, const void* W, const void* V, const void* gam, const void* P0,
                          const void* Pm, const void* P1, const void* Q0, const void* Qm, const void* Q1,
                          void* tdd, void* tuu, void* rud, void* rdu, double* err_host);

/* ------------------------------------------------------------- samplers */
/* VarLayerSampler (src/core/VarLayerSampler.cpp:9-69): base layer whose x/y
 * coordinates move with z through rules (addXRule/addYRule). */
int vrc_var_sampler_create(vrc_ctx* ctx, double lambda, int nx, int ny, const vrc_layer* base, int enable_pml,
                           vrc_sampler** out);
/* axis 0 = x (addXRule), 1 = y (addYRule). rule(z, user) returns the coordinate. */
int vrc_sampler_add_rule(vrc_sampler* s, int axis, int layout, double (*rule)(double z, void* user), void* user);
/* The simulation region of a sampler: LayerSampler's Lx_, Ly_ (periods) and the absorbing
 * regions [b0, b1], [u0, u1] along y and [l0, l1], [r0, r1] along x that a subclass's
 * assignSimulationRegion() sets (src/core/LayerSampler.h:7-28, src/core/LayerSampler.h:57-62). */
typedef struct vrc_region {
  double Lx, Ly, b0, b1, u0, u1, l0, l1, r0, r1;
} vrc_region;
/* The sampler plugin boundary: a user subclass of the abstract LayerSampler
 * (src/core/LayerSampler.h:5-77) overriding sample(z, P, Q) and/or sampleOnGPU(z, P, Q).
 * host_fn(z, P_host, Q_host, user) fills host N x N column-major complex P and Q (N = 2 nx ny);
 * device_fn(z, P, Q, stream, user) fills device N x N buffers, enqueuing its work on `stream`
 * (the context's hipStream_t); it may return 1 to defer to host_fn. Any other non-zero return
 * is a failure (status -3). Either function may be NULL, not both; device_fn is preferred when
 * the integrators sample. Kx, Ky come from region and enable_pml as
 * LayerSampler::postInitialization computes them (src/core/LayerSampler.cpp:145-153). The
 * callbacks run on the calling thread, inside the integrator calls. */
typedef int (*vrc_sample_host_fn)(double z, double* P_host, double* Q_host, void* user);
typedef int (*vrc_sample_device_fn)(double z, void* P, void* Q, void* stream, void* user);
int vrc_callback_sampler_create(vrc_ctx* ctx, double lambda, int nx, int ny, const vrc_region* region,
                                int enable_pml, vrc_sample_host_fn host_fn, vrc_sample_device_fn device_fn,
                                void* user, vrc_sampler** out);
/* LayerSampler::Lx(), Ly() and the region scalars of any sampler (src/core/LayerSampler.h:42-43). */
int vrc_sampler_region(const vrc_sampler* s, vrc_region* out);
/* Piecewise-constant stack (no reference counterpart; extension): layer k occupies
 * (z_edges[k], z_edges[k+1]); bottom at z <= z_edges[0], top at z >= z_edges[nlayers]. */
int vrc_stack_sampler_create(vrc_ctx* ctx, double lambda, int nx, int ny, int nlayers, const vrc_layer* layers,
                             const double* z_edges, const vrc_layer* bottom, const vrc_layer* top, int enable_pml,
                             vrc_sampler** out);
/* GDSIISampler (src/core/GDSIISampler.cpp:17-146): x-z cross-sections of the polygons
 * (BOUNDARY/BOX elements) of the first cell of a GDSII file, read as gdstk::read_gds does
 * with coordinates in micrometers (src/gdstk/library.cpp:160-420); PML always on. shift:
 * x, y, z (NULL = 0); pml_space: lPML, rPML, lspace, rspace, bPML, uPML, bspace, uspace
 * (NULL = the reference defaults 0.5, 0.5, 1.5, 1.5, 0.5, 0.5, 1.08, 1.0); eps_background,
 * eps_device: complex permittivities (re, im) at this wavelength (Material::permittivity). */
int vrc_gds_sampler_create(vrc_ctx* ctx, double lambda, int nx, int ny, const char* filename, const double* shift,
                           const double* pml_space, const double* eps_background, const double* eps_device,
                           double thickness, vrc_sampler** out);
/* The polygons (BOUNDARY/BOX elements, closing point dropped) of cell `cell` of a GDSII file
 * as gdstk::read_gds reads them (src/gdstk/library.cpp:160-420), coordinates scaled to `unit`
 * meters. Host only, no context. Writes *npolygons always; counts (points per polygon) when
 * npolygons <= max_polygons; xy (x, y pairs, polygon after polygon) when the total point count
 * is <= max_points. */
int vrc_gds_polygons(const char* filename, double unit, int cell, int max_polygons, int* npolygons, int* counts,
                     int max_points, double* xy);
/* GDSIISampler::minX/maxX/minZ/maxZ (src/core/GDSIISampler.h:41-44) into out4. */
int vrc_gds_sampler_bounds(const vrc_sampler* s, double* out4);
/* The layer a sampler sees at z (GDSIISampler::sampleLayout, src/core/GDSIISampler.cpp:94-146;
 * VarLayerSampler's moved coordinates for the others). Host only. Writes the counts always;
 * x (ncoord_x), y (ncoord_y) and eps ((ncoord_y-1) x (ncoord_x-1) complex, column-major)
 * only when they fit max_coord_x / max_coord_y. */
int vrc_sampler_layout(vrc_sampler* s, double z, int max_coord_x, double* x, int* ncoord_x, int max_coord_y,
                       double* y, int* ncoord_y, double* eps);
int vrc_sampler_destroy(vrc_sampler* s);

/* ----------------------------------------------------------- legacy stacked-layer solver */
/* RCWASolver (src/rcwa/RCWASolver.h:60-80): layer types added once, then a stack of type
 * indices with thicknesses (the first and last entries are semi-infinite; their thickness is
 * not used). nX, nY: harmonic orders (nx = 2 nX + 1). PML always on, region from the first
 * layer. */
typedef struct vrc_rcwa_solver vrc_rcwa_solver;
int vrc_rcwa_solver_create(vrc_ctx* ctx, int nX, int nY, vrc_rcwa_solver** out);
/* RCWASolver::addLayer (src/rcwa/RCWASolver.cpp:73-102): same cell and PML region as the
 * first layer, or an error. */
int vrc_rcwa_solver_add_layer(vrc_rcwa_solver* s, const vrc_layer* layer);
/* RCWASolver::solve (src/rcwa/RCWASolver.cpp:104-113): Layer::solve of every type,
 * interfaces (:460-580; translation: NULL or one x shift per stack entry,
 * Layer::permuteEigVecX) and the RTCM recursion (:343-457) for Tdd and Rud. */
int vrc_rcwa_solver_solve(vrc_rcwa_solver* s, double lambda, int nstack, const int* stack, const double* thickness,
                          const double* translation);
/* Host copies: which 0 Tdd, 1 Rud (N x N, column-major); for layer type layer_type: 2 eigvecE,
 * 3 eigvecH (N x N), 4 gamma (N, Layer::gamma = sqrt(eig) / k0). */
int vrc_rcwa_solver_get(vrc_rcwa_solver* s, int which, int layer_type, double* host);
/* RCWASolver::evaluateIntermediateField (src/rcwa/RCWASolver.cpp:1373-1621) for the stack
 * of the last solve: for ncols incident amplitude vectors of the first stack entry
 * (input_host: N x ncols, column-major; a unit vector e_m is the inputMode variant),
 * c_m_host and c_p_host (host, (nstack - 2) x N x ncols, inner layer by inner layer, each
 * N x ncols column-major) receive the reference's c_m and c_p: c_p[j] the forward
 * amplitudes of inner layer j + 1 at its first interface, c_m[j] its backward amplitudes at
 * its second interface. The stack needs an inner layer (nstack >= 3). */
int vrc_rcwa_solver_intermediate_field(vrc_rcwa_solver* s, int ncols, const double* input_host, double* c_m_host,
                                       double* c_p_host);
int vrc_rcwa_solver_dims(const vrc_rcwa_solver* s, int* N, int* ntypes);
int vrc_rcwa_solver_destroy(vrc_rcwa_solver* s);
/* LayerSampler::nDim (src/core/LayerSampler.h:44) */
int vrc_sampler_ndim(const vrc_sampler* s);
/* LayerSampler::sample (src/core/LayerSampler.h:49-51): P, Q into host N x N buffers. */
int vrc_sampler_sample(vrc_sampler* s, double z, double* P_host, double* Q_host);
/* LayerSampler::sampleOnGPU (src/core/LayerSampler.h:53-55): device N x N buffers. */
int vrc_sampler_sample_gpu(vrc_sampler* s, double z, void* P, void* Q);
/* LayerSampler::Kx()/Ky() (src/core/LayerSampler.h:46-47): host n x n buffers. */
int vrc_sampler_k(vrc_sampler* s, double* Kx_host, double* Ky_host);

/* ----------------------------------------------------------- integrators */
/* kind: 0 RCWAIntegrator, 1 RedhefferIntegrator, 2 DifferentialIntegrator
 * (src/core/RCWAIntegrator.h:25, RedhefferIntegrator.h:34, DifferentialIntegrator.h:25).
 * The sampler must outlive the integrator. */
int vrc_integrator_create(vrc_ctx* ctx, int kind, double z0, double z1, vrc_sampler* s, vrc_integrator** out);
int vrc_integrator_destroy(vrc_integrator* it);
/* RCWAIntegrator::compute(N) (src/core/RCWAIntegrator.cpp:20-87): N midpoint slices,
 * solved one after another. */
int vrc_rcwa_compute(vrc_integrator* it, int N);
/* Same result; slices sampled and solved in batches before the ordered cascade. */
int vrc_rcwa_compute_batched(vrc_integrator* it, int N);
/* Extension: explicit slice edges (midpoint sampling, per-slice width). */
int vrc_rcwa_compute_edges(vrc_integrator* it, int nedges, const double* z_edges);
/* RedhefferIntegrator::compute(p, maxN, pe) (src/core/RedhefferIntegrator.cpp:174-224);
 * p and maxN are accepted and unused, as in the reference. The reference recurses until
 * every segment meets pe; here trisection stops at depth 40 (segments of 3^-40 of the
 * interval; "redheffer.max_depth" of vrc_tune_set), accepts the segment and returns
 * status 5. */
int vrc_redheffer_compute(vrc_integrator* it, int p, int maxN, double pe);
/* RedhefferIntegrator::compute for several integrators at once (a wavelength or geometry
 * sweep: one sampler each, one context, equal N). Results per integrator are those of
 * vrc_redheffer_compute; the subdivision trees grow level by level, and every new
 * reference eigensystem of a level, across segments and integrators, is solved in one
 * batched launch (the reference solves them one at a time, depth first:
 * src/core/RedhefferIntegrator.cpp:279-313). Status as vrc_redheffer_compute. */
int vrc_redheffer_compute_many(int count, vrc_integrator* const* its, int p, int maxN, double pe);
/* DifferentialIntegrator::compute(N) (src/core/DifferentialIntegrator.cpp:21-134). */
int vrc_differential_compute(vrc_integrator* it, int N);
/* Getters Tdd/Tuu/Rud/Rdu/W0/V0/gam0/W1/V1/gam1 (src/core/RCWAIntegrator.h:28-38):
 * which = 0..9 in that order; host buffer of N*N (or N for gam) complex values. */
int vrc_integrator_get(vrc_integrator* it, int which, double* host_out);
/* N() and nEigVal() (src/core/RedhefferIntegrator.h:38-39). */
int vrc_integrator_counts(vrc_integrator* it, int* N, int* nEigVal);
/* Accepted varRCWA segments in cascade order as (z0, z1) pairs; returns the count
 * (writes at most cap pairs). No reference counterpart (subdivision replay). */
int vrc_integrator_tree(vrc_integrator* it, int cap, double* host_out);

/* ------------------------------------------------------ batched stacks */
/* No reference counterpart: many independent layer stacks (wavelength, geometry)
 * in one call. Every instance: periodic cell taken from its bottom layer, nlayers
 * slabs of given thickness sampled once each (RCWAIntegrator semantics per slab),
 * half-space media bottom/top as end bases. All instances of a call share nx, ny
 * and nlayers. S_out: per instance [Tdd|Tuu|Rud|Rdu] (4 N x N, device); the
 * optional W0/V0/gam0/W1/V1/gam1 outputs may be NULL. Instances are solved in chunks
 * of at most `chunk`, split evenly over the chunks; with single-valued half-spaces (and no
 * tape) the slab eigensolves of each cascade group go in even launches of at most
 * chunk * nlayers slabs, across instance boundaries. chunk <= 0: full eigensolver launches
 * (the QR kernel's slots: 1024, 512 or 256 matrices by N, see "zgeev.qr_variant") and then
 * one launch of the rest. ninst = 0 does nothing (S_out may then be NULL).
 * info_out (host, ninst ints, or NULL) receives per instance 0, 1 (an eigensolve of the
 * instance did not converge) or 4 (its S-matrix holds NaN or infinity, e.g. a zero
 * propagation constant or a singular system); the status is then the first flagged
 * instance's code, and vrc_last_error names that instance. */
typedef struct vrc_stack_instance {
  double lambda;
  int nlayers;
  const vrc_layer* layers;
  const double* thickness;
  const vrc_layer* bottom;
  const vrc_layer* top;
} vrc_stack_instance;
int vrc_batch_stack_rcwa(vrc_ctx* ctx, int nx, int ny, int ninst, const vrc_stack_instance* inst, int chunk,
                         void* S_out, void* W0, void* V0, void* gam0, void* W1, void* V1, void* gam1,
                         int* info_out);

/* vrc_batch_stack_rcwa with a harmonic count per instance (no reference counterpart; the
 * reference solves one instance at a time, src/gpu/RCWAIntegratorGPU.cpp:191-232):
 * instance i uses nx[i] x ny[i] harmonics, N_i = 2 nx[i] ny[i]. Outputs are packed in
 * instance order: instance i's [Tdd|Tuu|Rud|Rdu] starts at complex offset
 * 4 * sum_{j<i} N_j^2 of S_out, its W0/V0/W1/V1 at sum_{j<i} N_j^2 of those arrays and
 * its gam0/gam1 at sum_{j<i} N_j (optional outputs may be NULL). Instances of one
 * harmonic class are solved together (largest class first), with the launches of
 * vrc_batch_stack_rcwa; results per instance equal that call's. chunk, info_out and
 * the status as vrc_batch_stack_rcwa. */
int vrc_batch_stack_rcwa_mixed(vrc_ctx* ctx, int ninst, const int* nx, const int* ny, const vrc_stack_instance* inst,
                               int chunk, void* S_out, void* W0, void* V0, void* g0, void* W1, void* V1, void* g1,
                               int* info_out);

/* Recorded batches (no reference counterpart): vrc_batch_plan_create records the work of
 * vrc_batch_stack_rcwa(ctx, nx, ny, ninst, inst, chunk, S_out, W0, ..., g1) once into a
 * HIP graph (nothing runs yet); each vrc_batch_plan_launch replays it on the context's
 * stream, solving the same instances again into the same output buffers (which must stay
 * valid while the plan exists), then reads the per-instance codes back (synchronizes):
 * info_out and the status as vrc_batch_stack_rcwa. The plan owns the device work space
 * and host staging the recorded work uses, until vrc_batch_plan_destroy. Launches of one
 * plan, and other work on the context, are ordered by the context's stream. */
typedef struct vrc_batch_plan vrc_batch_plan;
int vrc_batch_plan_create(vrc_ctx* ctx, int nx, int ny, int ninst, const vrc_stack_instance* inst, int chunk,
                          void* S_out, void* W0, void* V0, void* g0, void* W1, void* V1, void* g1,
                          vrc_batch_plan** out);
int vrc_batch_plan_launch(vrc_batch_plan* plan, int* info_out);
int vrc_batch_plan_destroy(vrc_batch_plan* plan);
/* Plane-wave diffraction efficiencies of solved stacks; replaces the power bookkeeping of
 * RCWASolver::scatterPlaneWave (src/rcwa/RCWASolver.cpp:1636-1733) for the integrator's
 * S-matrix. S, W0, W1: device outputs of vrc_batch_stack_rcwa for the same instances;
 * e_inc: device, per instance the incident field harmonics [Ex; Ey] (2 nx ny complex) in
 * the bottom medium (modal coefficients W0^{-1} e_inc; the wave travels up). REF, TRN:
 * device, nx*ny doubles per instance: |Re Kz| / (sqrt(Re eps) k0) (|Ex|^2 + |Ey|^2 + |Ez|^2)
 * per harmonic of the reflected (bottom, eps of its first cell) and transmitted (top)
 * field, Ez from the plane-wave divergence condition; NaN for an instance with a
 * transversely moving component (Kz = 0), which the reference rejects. */
int vrc_stack_efficiencies(vrc_ctx* ctx, int nx, int ny, int ninst, const vrc_stack_instance* inst, const void* S,
                           const void* W0, const void* W1, const void* e_inc, double* REF, double* TRN);
/* Power carried by modes through a window; replaces evaluateEnergy
 * (src/tests/test_cpu.cpp:335-412) for nmodes modes in one call. W, V: device, column b
 * (at W + b*ldw, V + b*ldv, ld >= 2 nx ny) holds one mode's field harmonics [Ex; Ey] and
 * [Hx; Hy] (index j + i*ny, i the x-harmonic), as the columns of the integrator's W0/V0 or
 * W1/V1. energy: device, nmodes doubles: 0.5 Re(Ex Hy* - Ey Hx*) dx dy summed over the
 * midpoints of an nres_x x nres_y grid on [x0, x1] x [y0, y1], fields from the wavelets
 * exp(-i 2 pi m x / Lx), exp(-i 2 pi n y / Ly). The same sum is formed exactly through the
 * window Gram matrices of the wavelets, so the cost does not grow with the grid. */
int vrc_modal_power(vrc_ctx* ctx, int nx, int ny, int nmodes, const void* W, long long ldw, const void* V,
                    long long ldv, double Lx, double Ly, int nres_x, int nres_y, double x0, double x1, double y0,
                    double y1, double* energy);
/* findGuidedModes (src/tests/test_cpu.cpp:311-333), host only: the modes whose gamma/k0^2
 * has real part in [re_min, re_max] and imaginary part below imag_coeff, by decreasing real
 * part (ties keep index order). gamma: host, n complex; index_out (n ints) and neff_out
 * (n complex gamma/k0^2): host; *count: number found. Makes no device call; ctx may be
 * NULL (errors then carry no text). */
int vrc_find_guided_modes(vrc_ctx* ctx, int n, const void* gamma, double k0, double imag_coeff, double re_min,
                          double re_max, int* index_out, void* neff_out, int* count);

/* ------------------------------------------------------ adjoint (gradients) */
/* No reference counterpart: the reference differentiates nothing on this path.
 * Convention for every *_vjp: for a real objective L and complex Z, Z_bar is the
 * array with dL = Re sum(conj(Z_bar) * dZ) (PyTorch's convention; L = |z|^2 gives
 * z_bar = 2z). Cotangent inputs are device arrays in the layout of the forward
 * output, results are written in the layout of the forward input.
 *
 * vrc_batch_stack_forward: vrc_batch_stack_rcwa that also returns a tape (NULL tape
 * pointer = no tape) for the reverse pass: the geometry, the end media's eigen-systems,
 * the cascade before the end projection (4 N x N per instance), the cascade prefix at the
 * start of every block of K layers (checkpoints, 4 N x N each) and, by default, every
 * slab's eigen-system (2 N x N + N complex each). With recompute = 1 the slabs'
 * eigen-systems are not kept and the reverse pass solves them again, one block of K
 * layers (all instances of a pass) per launch: the tape then grows with L / K instead of
 * L, at the cost of one more eigensolve per slab. K <= 0: min(L, 16). The reverse pass
 * holds the intermediates of one block, not of the whole stack.
 * vrc_batch_stack_vjp: eps_bar (vrc_tape_cells complex values) receives, per
 * instance and slab in order, the cotangent of every eps cell of the slab in the
 * layout of vrc_layer.eps, for S_bar = the cotangent of S_out ([Tdd|Tuu|Rud|Rdu]
 * per instance). Slab coordinates must be non-decreasing (no painter's refill),
 * periodic cells only. chunk (<= 0: automatic) bounds the instances per pass.
 * vrc_batch_stack_vjp_shape: the same reverse pass giving (as well as, or instead of,
 * eps_bar: may be NULL) the gradient with respect to shape parameters of the slabs'
 * x coordinates (a Parameterization, src/rcwa/Parameterization.h:5-25, per slab;
 * chained through the closed-form coordinate tangents of src/rcwa/Layer.cpp:104-162):
 * slab s = instance * nlayers + layer has npara[s] parameters (0: none) and nvar[s]
 * varied interior x coordinates, listed in slab order in var_index (coordinate
 * indices) and coeff (nvar[s] x npara[s] per slab, row-major: coordinate var_index[k]
 * moves by coeff[k * npara[s] + p] per unit of parameter p). shape_bar_host receives
 * the sum over slabs of npara doubles, slab by slab (synchronizes).
 * vrc_tape_bytes: device memory the tape holds. */
typedef struct vrc_tape vrc_tape;
int vrc_batch_stack_forward(vrc_ctx* ctx, int nx, int ny, int ninst, const vrc_stack_instance* inst, int chunk,
                            void* S_out, void* W0, void* V0, void* gam0, void* W1, void* V1, void* gam1,
                            vrc_tape** tape, int* info_out);
int vrc_batch_stack_forward_ex(vrc_ctx* ctx, int nx, int ny, int ninst, const vrc_stack_instance* inst, int chunk,
                               void* S_out, void* W0, void* V0, void* gam0, void* W1, void* V1, void* gam1,
                               int recompute, int layer_block, vrc_tape** tape, int* info_out);
long long vrc_tape_cells(const vrc_tape* tape);
long long vrc_tape_bytes(const vrc_tape* tape);
int vrc_batch_stack_vjp(vrc_tape* tape, const void* S_bar, void* eps_bar, int chunk);
int vrc_batch_stack_vjp_shape(vrc_tape* tape, const void* S_bar, const int* npara, const int* nvar,
                              const int* var_index, const double* coeff, double* shape_bar_host, void* eps_bar,
                              int chunk);
int vrc_tape_destroy(vrc_tape* tape);

/* Stage adjoints (each the vector-Jacobian product of the forward entry point of
 * the same name above, with the same argument layout):
 *  - vrc_assemble_toeplitz_vjp: fe11/fe22/fe33 cotangents -> eps_bar (all layers'
 *    cells concatenated, layer order, vrc_layer.eps layout);
 *  - vrc_assemble_pq_vjp: periodic P/Q cotangents -> eps_bar (region gives the cell);
 *  - vrc_eig_pq_vjp: (W_bar, V_bar, gam_bar) -> (P_bar, Q_bar); W_bar and gam_bar are
 *    overwritten. Valid for objectives invariant under the eigenvector gauge
 *    (W -> W D with V -> V D), as every S-matrix objective is;
 *  - vrc_layer_smatrix_vjp: (Tuu_bar, Rud_bar) -> (W_bar, V_bar, gam_bar), reference
 *    basis constant;
 *  - vrc_star_product_vjp: S (Tdd..Rdu) is the cascade before the product; the S_bar
 *    arrays hold the cotangent after the product on entry and before it on return;
 *    slice cotangents (tdd_bar..rdu_bar) are written with stride N*N;
 *  - vrc_project_left_vjp: S is the cascade before the projection; S_bar as above. */
int vrc_assemble_toeplitz_vjp(vrc_ctx* ctx, int nx, int ny, int nlayers, const vrc_layer* layers,
                              const void* fe11_bar, const void* fe22_bar, const void* fe33_bar, long stride,
                              void* eps_bar);
int vrc_assemble_pq_vjp(vrc_ctx* ctx, int nx, int ny, int nlayers, const vrc_layer* layers, double lambda,
                        const vrc_layer* region, const void* P_bar, const void* Q_bar, long stride, void* eps_bar);
int vrc_eig_pq_vjp(vrc_ctx* ctx, int N, int batch, const void* P, const void* Q, long stride, const void* W,
                   const void* V, const void* gam, void* W_bar, const void* V_bar, void* gam_bar, void* P_bar,
                   void* Q_bar);
int vrc_layer_smatrix_vjp(vrc_ctx* ctx, int N, int batch, const double* lambda_host, const double* dz_host,
                          const void* W, const void* V, const void* gam, const void* W0, const void* V0,
                          long ref_stride, const void* Tuu_bar, const void* Rud_bar, void* W_bar, void* V_bar,
                          void* gam_bar);
int vrc_star_product_vjp(vrc_ctx* ctx, int N, int batch, const void* Tdd, const void* Tuu, const void* Rud,
                         const void* Rdu, const void* tdd, const void* tuu, const void* rud, const void* rdu,
                         long slice_stride, void* Tdd_bar, void* Tuu_bar, void* Rud_bar, void* Rdu_bar,
                         void* tdd_bar, void* tuu_bar, void* rud_bar, void* rdu_bar);
int vrc_project_left_vjp(vrc_ctx* ctx, int N, int batch, const void* Tdd, const void* Tuu, const void* Rud,
                         const void* Rdu, const void* W1, const void* V1, const void* W0, const void* V0,
                         long ref_stride, void* Tdd_bar, void* Tuu_bar, void* Rud_bar, void* Rdu_bar);

#ifdef __cplusplus
}
#endif

#endif /* VARRCWA_H */
