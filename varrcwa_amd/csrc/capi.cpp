// extern "C" entry points (see include/varrcwa.h for the contract of each).
#include "../../include/varrcwa.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>

#include "vrc_adjoint.hpp"
#include "vrc_context.hpp"

using namespace vrc;

namespace {
inline cd c2(const double* p) { return p ? mk(p[0], p[1]) : mk(0.0, 0.0); }
inline void need(bool ok, const char* what) {
  if (!ok) throw ArgFailure(what);
}

// guarded() for the calls that solve whole problems: the context's numerical-failure
// record is reset on entry and read back (one small synchronizing copy) on exit; a
// recorded failure returns its positive code (H_EIG .. H_DEPTH) with the text in
// vrc_last_error. The results are written as computed.
template <class F>
int checked(vrc_ctx* c, F&& f) {
  int num = 0;
  const int rc = guarded(c, [&] {
    health_reset(*c);
    f();
    std::string m;
    num = health_read(*c, m);
    if (num) c->last_error = m;
  });
  return rc != 0 ? rc : num;
}
}  // namespace

// Trisection depth limit of the varRCWA recursion (vrc_tune_set "redheffer.max_depth").
static int g_redheffer_depth = 40;

extern "C" {

const char* vrc_version(void) { return "varrcwa_amd 0.1 (gfx950)"; }

int vrc_ctx_create(int device, void* stream, vrc_ctx** out) {
  if (!out) return -1;
  *out = nullptr;
  vrc_ctx* c = new vrc_ctx();
  c->device = device;
  // NULL selects the device's default (null) stream, which orders the library's work
  // with every other default-stream user in the process (e.g. torch's default stream).
  int rc = guarded(c, [&] {
    c->stream = (hipStream_t)stream;
    noalloc_init_device(device);
    VRC_HIP(hipMalloc(&c->health, sizeof(int) * HEALTH_INTS));
    fill_int(c->stream, c->health, 0, HEALTH_INTS);
    VRC_HIP(hipStreamSynchronize(c->stream));
  });
  if (rc != 0) {
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int vrc_ctx_destroy(vrc_ctx* c) {
  if (!c) return 0;
  ctx_release(c);
  return 0;
}

const char* vrc_last_error(const vrc_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

int vrc_sync(vrc_ctx* c) {
  return guarded(c, [&] { VRC_HIP(hipStreamSynchronize(c->stream)); });
}

long long vrc_pool_bytes(const vrc_ctx* c) { return c ? (long long)c->ws.reserved() : 0; }

int vrc_pool_trim(vrc_ctx* c) {
  return guarded(c, [&] {
    VRC_HIP(hipStreamSynchronize(c->stream));
    c->ws.trim();
  });
}

int vrc_tune_set(const char* key, long long value) {
  if (!key) return -1;
  tune_manual_begin();
  if (std::strcmp(key, "zgemm.large_tile") == 0) {
    zgemm_set_big(value != 0);
    return 0;
  }
  if (std::strcmp(key, "zgemm.large_tile_min_k") == 0) {
    zgemm_set_big_min_k((int)value);
    return 0;
  }
  if (std::strcmp(key, "zgemm.big_waves") == 0) {
    zgemm_set_big_waves((int)value);
    return 0;
  }
  if (std::strcmp(key, "zgemm.small_waves") == 0) {
    zgemm_set_small_waves((int)value);
    return 0;
  }
  if (std::strcmp(key, "zgemm.big_n") == 0) {
    zgemm_set_big_n((int)value);
    return 0;
  }
  if (std::strcmp(key, "redheffer.max_depth") == 0) {
    if (value < 0 || value > 60) return -1;
    g_redheffer_depth = (int)value;
    return 0;
  }
  if (std::strcmp(key, "zgeev.max_sweeps") == 0) {
    zgeev_set_max_sweeps((int)value);
    return 0;
  }
  if (std::strcmp(key, "zgeev.qr_variant") == 0) {
    if (value < -1 || value > 4 || value == 3) return -1;
    zgeev_set_qr_variant((int)value);
    return 0;
  }
  if (std::strcmp(key, "zgeev.bulges_w48") == 0) {
    zgeev_set_bulges_w48((int)value);
    return 0;
  }
  if (std::strcmp(key, "zgeev.bulges_w32") == 0) {
    zgeev_set_bulges_w32((int)value);
    return 0;
  }
  if (std::strcmp(key, "zgeev.bulges") == 0) {
    zgeev_set_bulges((int)value);
    return 0;
  }
  if (std::strcmp(key, "zgeev.aed_window") == 0) {
    zgeev_set_aed((int)value);
    return 0;
  }
  if (std::strcmp(key, "rcwa.group") == 0) {
    rcwa_set_group((int)value);
    return 0;
  }
  if (std::strcmp(key, "stack.cascade_group") == 0) {
    stack_set_cascade_group((int)value);
    return 0;
  }
  return -1;
}

int vrc_timing_enable(vrc_ctx* c, int on) {
  return guarded(c, [&] { c->timing = on != 0; });
}

int vrc_timing_read(vrc_ctx* c, double* ms, int* counts, int nstages) {
  return guarded(c, [&] {
    need(ms && nstages > 0, "bad arguments");
    VRC_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < nstages; ++i) {
      ms[i] = 0.0;
      if (counts) counts[i] = 0;
    }
    for (auto& r : c->recs) {
      float t = 0.0f;
      VRC_HIP(hipEventElapsedTime(&t, r.a, r.b));
      if (r.stage >= 0 && r.stage < nstages) {
        ms[r.stage] += t;
        if (counts) counts[r.stage] += 1;
      }
      c->event_pool.push_back(r.a);
      c->event_pool.push_back(r.b);
    }
    c->recs.clear();
  });
}

int vrc_qr_stats(vrc_ctx* c, long long* out) {
  return guarded(c, [&] {
    need(out != nullptr, "null output");
    for (int k = 0; k < 6; ++k) {
      out[k] = c->qr_stats[k];
      c->qr_stats[k] = 0;
    }
  });
}

int vrc_zgemm(vrc_ctx* c, int opA, int opB, int m, int n, int k, const double* alpha, const void* A, long lda,
              long sA, const void* B, long ldb, long sB, const double* beta, void* C, long ldc, long sC,
              int batch) {
  return guarded(c, [&] {
    need(opA >= 0 && opA <= 2 && opB >= 0 && opB <= 2, "op must be 0 (N), 1 (T) or 2 (C)");
    need(m >= 0 && n >= 0 && k >= 0 && batch >= 0, "negative size");
    need(C != nullptr || m == 0 || n == 0, "null C");
    zgemm_strided(c->stream, opA, opB, m, n, k, c2(alpha), (const cd*)A, lda, sA, (const cd*)B, ldb, sB, c2(beta),
                  (cd*)C, ldc, sC, batch);
  });
}

int vrc_zgetrf(vrc_ctx* c, int n, void* A, long lda, long sA, int* ipiv, long sP, int* info, int batch) {
  return guarded(c, [&] {
    need(n >= 0 && batch >= 0 && lda >= n, "bad size");
    need(A && ipiv, "null pointer");
    if (info) fill_int(c->stream, info, 0, batch);  // a fill kernel: safe under graph capture
    zgetrf(c->stream, n, (cd*)A, lda, sA, ipiv, sP, info, batch, c->ws);
  });
}

int vrc_zgetrs(vrc_ctx* c, int n, int nrhs, const void* LU, long lda, long sA, const int* ipiv, long sP,
               const void* B, long ldb, long sB, void* X, long ldx, long sX, int batch) {
  return guarded(c, [&] {
    need(n >= 0 && nrhs >= 0 && batch >= 0 && lda >= n && ldb >= n && ldx >= n, "bad size");
    zgetrs(c->stream, n, nrhs, (const cd*)LU, lda, sA, ipiv, sP, (const cd*)B, ldb, sB, (cd*)X, ldx, sX, batch,
           c->ws);
  });
}

int vrc_zgeev(vrc_ctx* c, int n, void* A, long lda, long sA, void* w, long sw, void* VR, long ldv, long sV,
              int* info, int batch) {
  return guarded(c, [&] {
    need(n >= 0 && batch >= 0 && lda >= n && ldv >= n, "bad size");
    need(A && w && VR && info, "null pointer");
    zgeev_batched(c->stream, n, (cd*)A, lda, sA, (cd*)w, sw, (cd*)VR, ldv, sV, info, batch, c->ws);
  });
}

}  // extern "C"

// ============================================================ path-level entry points
#include "vrc_ops.hpp"

struct vrc_sampler {
  vrc_ctx* ctx;
  std::unique_ptr<vrc::Sampler> s;
};
struct vrc_rcwa_solver {
  vrc_ctx* ctx;
  std::unique_ptr<vrc::LegacySolver> s;
};
struct vrc_integrator {
  vrc_ctx* ctx;
  int kind;
  std::unique_ptr<vrc::Integrator> it;
};

namespace {
HostLayer to_host(const vrc_layer* l) {
  if (!l || !l->x || !l->y || !l->eps) throw ArgFailure("null layer field");
  if (l->ncoord_x < 2 || l->ncoord_y < 2) throw ArgFailure("layer needs >= 2 coordinates per axis");
  HostLayer h;
  h.x.assign(l->x, l->x + l->ncoord_x);
  h.y.assign(l->y, l->y + l->ncoord_y);
  const size_t ne = (size_t)(l->ncoord_x - 1) * (l->ncoord_y - 1);
  h.eps.resize(ne);
  for (size_t i = 0; i < ne; ++i) h.eps[i] = mk(l->eps[2 * i], l->eps[2 * i + 1]);
  return h;
}
std::vector<HostLayer> to_host(const vrc_layer* l, int n) {
  if (n < 0 || (n > 0 && !l)) throw ArgFailure("bad layer array");
  std::vector<HostLayer> v;
  v.reserve(n);
  for (int i = 0; i < n; ++i) v.push_back(to_host(l + i));
  return v;
}
void check_harmonics(int nx, int ny) {
  need(nx >= 1 && ny >= 1 && (nx & 1) && (ny & 1), "nx, ny must be odd and positive");
  need(nx <= 63 && ny <= 63, "nx, ny above 63 not supported");
}
}  // namespace

// Host-only entry points: no device call; ctx may be NULL (then no error text is kept).
template <class F>
static int host_guarded(vrc_ctx* c, F&& f) {
  try {
    f();
    return 0;
  } catch (const ArgFailure& e) {
    if (c) c->last_error = e.what;
    return -1;
  } catch (const std::exception& e) {
    if (c) c->last_error = e.what();
    return -3;
  }
}

extern "C" {

int vrc_assemble_toeplitz(vrc_ctx* c, int nx, int ny, int nlayers, const vrc_layer* layers, void* fe11, void* fe22,
                          void* fe33, long stride) {
  return guarded(c, [&] {
    check_harmonics(nx, ny);
    need(nlayers >= 0 && fe11 && fe22 && fe33, "bad arguments");
    const long n = (long)nx * ny;
    need(stride >= n * n || nlayers <= 1, "stride smaller than n*n");
    std::vector<HostLayer> hl = to_host(layers, nlayers);
    AssemblyBatch a;
    a.nx = nx;
    a.ny = ny;
    a.nlayers = nlayers;
    for (auto& l : hl) a.layers.push_back(&l);
    if (nlayers == 0) return;
    assemble_fe(*c, a, (cd*)fe11, (cd*)fe22, (cd*)fe33, stride);
  });
}

int vrc_assemble_pq(vrc_ctx* c, int nx, int ny, int nlayers, const vrc_layer* layers, double lambda, int enable_pml,
                    const vrc_layer* region, void* P, void* Q, long stride) {
  return guarded(c, [&] {
    check_harmonics(nx, ny);
    need(nlayers >= 0 && P && Q, "bad arguments");
    const long N = 2L * nx * ny;
    need(stride == N * N, "P/Q stride must be N*N");
    std::vector<HostLayer> hl = to_host(layers, nlayers);
    VarSampler s(c, lambda, nx, ny, to_host(region), enable_pml != 0);
    s.sample_layers(hl, (cd*)P, (cd*)Q);
  });
}

int vrc_wave_eqn_derivatives(vrc_ctx* c, int nx, int ny, const vrc_layer* layer, double lambda, int enable_pml,
                             const vrc_layer* region, int nvar, const int* var_index, const double* var_coeff,
                             int npara, void* dP, void* dQ, void* dfe11, void* dfe22, void* dfe33, void* dPl,
                             void* dQl) {
  return guarded(c, [&] {
    check_harmonics(nx, ny);
    need(layer && region && npara >= 0 && nvar >= 0, "bad arguments");
    need(npara == 0 || (dP && dQ), "dP and dQ are required when there are parameters");
    VarSampler s(c, lambda, nx, ny, to_host(region), enable_pml != 0);
    s.shape_derivatives(to_host(layer), nvar, var_index, var_coeff, npara, (cd*)dP, (cd*)dQ, (cd*)dfe11,
                        (cd*)dfe22, (cd*)dfe33, (cd*)dPl, (cd*)dQl);
  });
}

int vrc_eig_pq(vrc_ctx* c, int N, int batch, const void* P, const void* Q, long stride, void* W, void* V, void* gam,
               int* info) {
  return guarded(c, [&] {
    need(N > 0 && batch >= 0 && P && Q && W && V && gam && info, "bad arguments");
    need(stride == (long)N * N, "stride must be N*N");
    fill_int(c->stream, info, 0, batch);
    wave_eqn(*c, N, (const cd*)P, (const cd*)Q, stride, (cd*)W, (cd*)V, (cd*)gam, info, batch);
  });
}

int vrc_layer_smatrix(vrc_ctx* c, int N, int batch, const double* lam, const double* dz, const void* W, const void* V,
                      const void* gam, const void* W0, const void* V0, long ref_stride, void* Tuu, void* Rud) {
  return guarded(c, [&] {
    need(N > 0 && batch >= 0 && lam && dz && W && V && gam && W0 && V0 && Tuu && Rud, "bad arguments");
    std::vector<double> s(batch);
    for (int b = 0; b < batch; ++b) {
      need(lam[b] > 0, "wavelength must be positive");
      s[b] = dz[b] * lam[b] / (2.0 * 3.1415926535897932);
    }
    DevBuf ds = c->ws.get(sizeof(double) * (batch > 0 ? batch : 1));
    VRC_HIP(hipMemcpyAsync(ds.ptr, s.data(), sizeof(double) * batch, hipMemcpyHostToDevice, c->stream));
    VRC_HIP(hipStreamSynchronize(c->stream));
    layer_smatrix(*c, N, batch, ds.as<double>(), (const cd*)W, (const cd*)V, (const cd*)gam, (const cd*)W0,
                  (const cd*)V0, ref_stride, (cd*)Tuu, (cd*)Rud);
  });
}

int vrc_star_product(vrc_ctx* c, int N, int batch, void* Tdd, void* Tuu, void* Rud, void* Rdu, const void* tdd,
                     const void* tuu, const void* rud, const void* rdu, long ss) {
  return guarded(c, [&] {
    need(N > 0 && batch >= 0 && Tdd && Tuu && Rud && Rdu && tdd && tuu && rud && rdu, "bad arguments");
    SMat S;
    S.Tdd = (cd*)Tdd;
    S.Tuu = (cd*)Tuu;
    S.Rud = (cd*)Rud;
    S.Rdu = (cd*)Rdu;
    star_product(*c, N, batch, S, (const cd*)tdd, (const cd*)tuu, (const cd*)rud, (const cd*)rdu, ss);
  });
}

int vrc_project_left(vrc_ctx* c, int N, int batch, void* Tdd, void* Tuu, void* Rud, void* Rdu, const void* W1,
                     const void* V1, const void* W0, const void* V0, long ref_stride) {
  return guarded(c, [&] {
    need(N > 0 && batch >= 0 && Tdd && Tuu && Rud && Rdu && W1 && V1 && W0 && V0, "bad arguments");
    const long NN = (long)N * N;
    SMat S;
    S.Tdd = (cd*)Tdd;
    S.Tuu = (cd*)Tuu;
    S.Rud = (cd*)Rud;
    S.Rdu = (cd*)Rdu;
    DevBuf LW = c->ws.get(sizeof(cd) * NN * batch), LV = c->ws.get(sizeof(cd) * NN * batch);
    DevBuf pW = c->ws.get(sizeof(int) * (size_t)N * batch), pV = c->ws.get(sizeof(int) * (size_t)N * batch);
    DevBuf inf = c->ws.get(sizeof(int) * (size_t)batch);
    fill_int(c->stream, inf.as<int>(), 0, batch);
    zgeam(c->stream, N, N, mk(1, 0), (const cd*)W1, N, ref_stride, mk(0, 0), nullptr, 0, 0, LW.as<cd>(), N, NN,
          batch);
    zgeam(c->stream, N, N, mk(1, 0), (const cd*)V1, N, ref_stride, mk(0, 0), nullptr, 0, 0, LV.as<cd>(), N, NN,
          batch);
    zgetrf(c->stream, N, LW.as<cd>(), N, NN, pW.as<int>(), N, inf.as<int>(), batch, c->ws);
    zgetrf(c->stream, N, LV.as<cd>(), N, NN, pV.as<int>(), N, inf.as<int>(), batch, c->ws);
    LUPair lu;
    lu.W = LW.as<cd>();
    lu.V = LV.as<cd>();
    lu.pW = pW.as<int>();
    lu.pV = pV.as<int>();
    lu.sM = NN;
    project_left(*c, N, batch, S, lu, (const cd*)W0, (const cd*)V0, ref_stride);
  });
}

int vrc_redheffer_segment(vrc_ctx* c, int N, double lambda, double z0, double z1, const void* Pr, const void* Qr,
                          const void* W, const void* V, const void* gam, const void* P0, const void* Pm,
                          const void* P1, const void* Q0, const void* Qm, const void* Q1, void* tdd, void* tuu,
                          void* rud, void* rdu, double* err) {
  return guarded(c, [&] {
    need(N > 0 && lambda > 0 && Pr && Qr && W && V && gam && P0 && Pm && P1 && Q0 && Qm && Q1, "bad arguments");
    need(tdd && tuu && rud && rdu && err, "bad arguments");
    const long NN = (long)N * N;
    DevBuf LW = c->ws.get(sizeof(cd) * NN), LV = c->ws.get(sizeof(cd) * NN);
    DevBuf pW = c->ws.get(sizeof(int) * N), pV = c->ws.get(sizeof(int) * N), inf = c->ws.get(sizeof(int));
    fill_int(c->stream, inf.as<int>(), 0, 1);
    zgeam(c->stream, N, N, mk(1, 0), (const cd*)W, N, 0, mk(0, 0), nullptr, 0, 0, LW.as<cd>(), N, 0, 1);
    zgeam(c->stream, N, N, mk(1, 0), (const cd*)V, N, 0, mk(0, 0), nullptr, 0, 0, LV.as<cd>(), N, 0, 1);
    zgetrf(c->stream, N, LW.as<cd>(), N, 0, pW.as<int>(), N, inf.as<int>(), 1, c->ws);
    zgetrf(c->stream, N, LV.as<cd>(), N, 0, pV.as<int>(), N, inf.as<int>(), 1, c->ws);
    RefSys R;
    R.W = (const cd*)W;
    R.V = (const cd*)V;
    R.gam = (const cd*)gam;
    R.LU.W = LW.as<cd>();
    R.LU.V = LV.as<cd>();
    R.LU.pW = pW.as<int>();
    R.LU.pV = pV.as<int>();
    R.LU.sM = NN;
    const cd* P3[3] = {(const cd*)P0, (const cd*)Pm, (const cd*)P1};
    const cd* Q3[3] = {(const cd*)Q0, (const cd*)Qm, (const cd*)Q1};
    *err = redheffer_perturbation(*c, N, 2.0 * 3.1415926535897932 / lambda, z0, z1, (const cd*)Pr, (const cd*)Qr, R,
                                  P3, Q3, (cd*)tdd, (cd*)tuu, (cd*)rud, (cd*)rdu);
  });
}

int vrc_var_sampler_create(vrc_ctx* c, double lambda, int nx, int ny, const vrc_layer* base, int enable_pml,
                           vrc_sampler** out) {
  return guarded(c, [&] {
    need(out != nullptr, "null out");
    *out = nullptr;
    check_harmonics(nx, ny);
    auto s = std::make_unique<vrc_sampler>();
    s->ctx = c;
    s->s = std::make_unique<VarSampler>(c, lambda, nx, ny, to_host(base), enable_pml != 0);
    ctx_retain(c);
    *out = s.release();
  });
}

int vrc_sampler_add_rule(vrc_sampler* s, int axis, int layout, double (*rule)(double, void*), void* user) {
  if (!s) return -1;
  return guarded(s->ctx, [&] {
    auto* v = dynamic_cast<VarSampler*>(s->s.get());
    need(v != nullptr, "rules apply to a VarLayerSampler only");
    need(axis == 0 || axis == 1, "axis must be 0 (x) or 1 (y)");
    v->add_rule(axis, layout, rule, user);
  });
}

int vrc_callback_sampler_create(vrc_ctx* c, double lambda, int nx, int ny, const vrc_region* region,
                                int enable_pml, vrc_sample_host_fn host_fn, vrc_sample_device_fn device_fn,
                                void* user, vrc_sampler** out) {
  if (out) *out = nullptr;
  return guarded(c, [&] {
    need(out && region, "bad arguments");
    check_harmonics(nx, ny);
    Region r;
    r.Lx = region->Lx;
    r.Ly = region->Ly;
    r.b0 = region->b0;
    r.b1 = region->b1;
    r.u0 = region->u0;
    r.u1 = region->u1;
    r.l0 = region->l0;
    r.l1 = region->l1;
    r.r0 = region->r0;
    r.r1 = region->r1;
    auto h = std::make_unique<vrc_sampler>();
    h->ctx = c;
    h->s = std::make_unique<CallbackSampler>(c, lambda, nx, ny, r, enable_pml != 0, host_fn, device_fn, user);
    ctx_retain(c);
    *out = h.release();
  });
}

int vrc_sampler_region(const vrc_sampler* s, vrc_region* out) {
  if (!s || !out) return -1;
  const Region& r = s->s->region();
  *out = vrc_region{r.Lx, r.Ly, r.b0, r.b1, r.u0, r.u1, r.l0, r.l1, r.r0, r.r1};
  return 0;
}

int vrc_stack_sampler_create(vrc_ctx* c, double lambda, int nx, int ny, int nlayers, const vrc_layer* layers,
                             const double* z_edges, const vrc_layer* bottom, const vrc_layer* top, int enable_pml,
                             vrc_sampler** out) {
  return guarded(c, [&] {
    need(out != nullptr && z_edges != nullptr && nlayers >= 0, "bad arguments");
    *out = nullptr;
    check_harmonics(nx, ny);
    auto s = std::make_unique<vrc_sampler>();
    s->ctx = c;
    s->s = std::make_unique<StackSampler>(c, lambda, nx, ny, to_host(layers, nlayers),
                                          std::vector<double>(z_edges, z_edges + nlayers + 1), to_host(bottom),
                                          to_host(top), enable_pml != 0);
    ctx_retain(c);
    *out = s.release();
  });
}

int vrc_gds_sampler_create(vrc_ctx* c, double lambda, int nx, int ny, const char* filename, const double* shift,
                           const double* pml_space, const double* eps_background, const double* eps_device,
                           double thickness, vrc_sampler** out) {
  return guarded(c, [&] {
    need(out != nullptr && filename != nullptr && eps_background && eps_device, "bad arguments");
    *out = nullptr;
    check_harmonics(nx, ny);
    const double zero3[3] = {0.0, 0.0, 0.0};
    // the reference's defaults (src/core/GDSIISampler.h:27-35)
    const double dflt[8] = {0.5, 0.5, 1.5, 1.5, 0.5, 0.5, 1.08, 1.0};
    auto s = std::make_unique<vrc_sampler>();
    s->ctx = c;
    s->s = std::make_unique<GdsSampler>(c, lambda, nx, ny, std::string(filename), shift ? shift : zero3,
                                        pml_space ? pml_space : dflt, mk(eps_background[0], eps_background[1]),
                                        mk(eps_device[0], eps_device[1]), thickness);
    ctx_retain(c);
    *out = s.release();
  });
}

int vrc_gds_polygons(const char* filename, double unit, int cell, int max_polygons, int* npolygons, int* counts,
                     int max_points, double* xy) {
  return host_guarded(nullptr, [&] {
    need(filename && npolygons, "bad arguments");
    GdsLibrary lib = read_gds(filename, unit);
    need(cell >= 0 && cell < (int)lib.cells.size(), "no such cell");
    const auto& polys = lib.cells[cell].polygons;
    *npolygons = (int)polys.size();
    if ((int)polys.size() > max_polygons || !counts) return;  // count only
    long total = 0;
    for (const auto& p : polys) total += (long)p.x.size();
    for (size_t k = 0; k < polys.size(); ++k) counts[k] = (int)polys[k].x.size();
    if (total > max_points || !xy) return;
    long o = 0;
    for (const auto& p : polys)
      for (size_t i = 0; i < p.x.size(); ++i) {
        xy[2 * o] = p.x[i];
        xy[2 * o + 1] = p.y[i];
        ++o;
      }
  });
}

int vrc_gds_sampler_bounds(const vrc_sampler* s, double* out4) {
  if (!s || !out4) return -1;
  const GdsSampler* g = dynamic_cast<const GdsSampler*>(s->s.get());
  if (!g) {
    s->ctx->last_error = "not a GDSII sampler";
    return -1;
  }
  out4[0] = g->minX();
  out4[1] = g->maxX();
  out4[2] = g->minZ();
  out4[3] = g->maxZ();
  return 0;
}

int vrc_sampler_layout(vrc_sampler* s, double z, int max_coord_x, double* x, int* ncoord_x, int max_coord_y,
                       double* y, int* ncoord_y, double* eps) {
  if (!s) return -1;
  return host_guarded(s->ctx, [&] {
    need(ncoord_x && ncoord_y, "bad arguments");
    HostLayer L = s->s->layer_at(z);
    *ncoord_x = (int)L.x.size();
    *ncoord_y = (int)L.y.size();
    if ((int)L.x.size() > max_coord_x || (int)L.y.size() > max_coord_y || !x || !y || !eps)
      return;  // sizes only
    std::copy(L.x.begin(), L.x.end(), x);
    std::copy(L.y.begin(), L.y.end(), y);
    std::copy(L.eps.begin(), L.eps.end(), (cd*)eps);
  });
}

int vrc_rcwa_solver_create(vrc_ctx* c, int nX, int nY, vrc_rcwa_solver** out) {
  return guarded(c, [&] {
    need(out != nullptr, "bad arguments");
    *out = nullptr;
    check_harmonics(2 * nX + 1, 2 * nY + 1);
    auto s = std::make_unique<vrc_rcwa_solver>();
    s->ctx = c;
    s->s = std::make_unique<LegacySolver>(c, nX, nY);
    ctx_retain(c);
    *out = s.release();
  });
}

int vrc_rcwa_solver_add_layer(vrc_rcwa_solver* s, const vrc_layer* layer) {
  if (!s) return -1;
  return host_guarded(s->ctx, [&] {
    need(layer != nullptr, "bad arguments");
    s->s->add_layer(to_host(layer));
  });
}

int vrc_rcwa_solver_solve(vrc_rcwa_solver* s, double lambda, int nstack, const int* stack, const double* thickness,
                          const double* translation) {
  if (!s) return -1;
  return checked(s->ctx, [&] {
    need(nstack >= 0 && (nstack == 0 || (stack && thickness)), "bad arguments");
    s->s->solve(lambda, std::vector<int>(stack, stack + nstack), std::vector<double>(thickness, thickness + nstack),
                translation);
  });
}

int vrc_rcwa_solver_get(vrc_rcwa_solver* s, int which, int layer_type, double* host) {
  if (!s) return -1;
  return guarded(s->ctx, [&] {
    need(host != nullptr, "bad arguments");
    s->s->get(which, layer_type, (cd*)host);
  });
}

int vrc_rcwa_solver_intermediate_field(vrc_rcwa_solver* s, int ncols, const double* input_host, double* c_m_host,
                                       double* c_p_host) {
  if (!s) return -1;
  return guarded(s->ctx, [&] {
    need(input_host && c_m_host && c_p_host, "null buffer");
    s->s->intermediate_field(ncols, (const cd*)input_host, (cd*)c_m_host, (cd*)c_p_host);
  });
}

int vrc_rcwa_solver_dims(const vrc_rcwa_solver* s, int* N, int* ntypes) {
  if (!s || !N || !ntypes) return -1;
  *N = s->s->N();
  *ntypes = s->s->ntypes();
  return 0;
}

int vrc_rcwa_solver_destroy(vrc_rcwa_solver* s) {
  if (!s) return 0;
  vrc_ctx* c = s->ctx;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  delete s;
  ctx_release(c);
  return 0;
}

int vrc_sampler_destroy(vrc_sampler* s) {
  if (!s) return 0;
  vrc_ctx* c = s->ctx;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  delete s;
  ctx_release(c);
  return 0;
}

int vrc_sampler_ndim(const vrc_sampler* s) { return s ? s->s->nDim() : -1; }

int vrc_sampler_sample(vrc_sampler* s, double z, double* Ph, double* Qh) {
  if (!s) return -1;
  return guarded(s->ctx, [&] {
    need(Ph && Qh, "null output");
    s->s->sample_host(z, (cd*)Ph, (cd*)Qh);
  });
}

int vrc_sampler_sample_gpu(vrc_sampler* s, double z, void* P, void* Q) {
  if (!s) return -1;
  return guarded(s->ctx, [&] {
    need(P && Q, "null output");
    s->s->sample_gpu({z}, (cd*)P, (cd*)Q);
  });
}

int vrc_sampler_k(vrc_sampler* s, double* Kx, double* Ky) {
  if (!s) return -1;
  return guarded(s->ctx, [&] {
    need(Kx && Ky, "null output");
    std::memcpy(Kx, s->s->Kx().data(), sizeof(cd) * s->s->Kx().size());
    std::memcpy(Ky, s->s->Ky().data(), sizeof(cd) * s->s->Ky().size());
  });
}

int vrc_integrator_create(vrc_ctx* c, int kind, double z0, double z1, vrc_sampler* s, vrc_integrator** out) {
  return guarded(c, [&] {
    need(out && s && s->ctx == c, "bad arguments");
    need(kind >= 0 && kind <= 2, "kind must be 0, 1 or 2");
    *out = nullptr;
    auto it = std::make_unique<vrc_integrator>();
    it->ctx = c;
    it->kind = kind;
    it->it = std::make_unique<Integrator>(c, z0, z1, s->s.get());
    ctx_retain(c);
    *out = it.release();
  });
}

int vrc_integrator_destroy(vrc_integrator* it) {
  if (!it) return 0;
  vrc_ctx* c = it->ctx;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  delete it;
  ctx_release(c);
  return 0;
}

static std::vector<std::pair<double, double>> uniform_slices(double z0, double z1, int N) {
  std::vector<std::pair<double, double>> v;
  const double dz = (z1 - z0) / (double)N;
  for (int i = 0; i < N; ++i) v.push_back({z0 + (i + 0.5) * dz, dz});
  return v;
}

int vrc_rcwa_compute(vrc_integrator* it, int N) {
  if (!it) return -1;
  return checked(it->ctx, [&] {
    need(it->kind == 0 && N >= 1, "RCWA integrator with N >= 1 expected");
    it->it->rcwa(uniform_slices(it->it->it_z0(), it->it->it_z1(), N), false);
  });
}

int vrc_rcwa_compute_batched(vrc_integrator* it, int N) {
  if (!it) return -1;
  return checked(it->ctx, [&] {
    need(it->kind == 0 && N >= 1, "RCWA integrator with N >= 1 expected");
    it->it->rcwa(uniform_slices(it->it->it_z0(), it->it->it_z1(), N), true);
  });
}

int vrc_rcwa_compute_edges(vrc_integrator* it, int ne, const double* ze) {
  if (!it) return -1;
  return checked(it->ctx, [&] {
    need(it->kind == 0 && ne >= 2 && ze, "RCWA integrator with >= 2 edges expected");
    std::vector<std::pair<double, double>> v;
    for (int i = 0; i + 1 < ne; ++i) {
      need(ze[i + 1] > ze[i], "edges must increase");
      v.push_back({0.5 * (ze[i] + ze[i + 1]), ze[i + 1] - ze[i]});
    }
    it->it->rcwa(v, false);
  });
}

int vrc_redheffer_compute(vrc_integrator* it, int p, int maxN, double pe) {
  if (!it) return -1;
  (void)p;
  (void)maxN;
  return checked(it->ctx, [&] {
    need(it->kind == 1 && pe > 0, "Redheffer integrator with pe > 0 expected");
    it->it->redheffer(pe, g_redheffer_depth);
  });
}

int vrc_redheffer_compute_many(int count, vrc_integrator* const* its, int p, int maxN, double pe) {
  if (count < 1 || !its || !its[0]) return -1;
  (void)p;
  (void)maxN;
  vrc_ctx* c = its[0]->ctx;
  return checked(c, [&] {
    need(pe > 0, "pe > 0 expected");
    std::vector<Integrator*> v;
    for (int i = 0; i < count; ++i) {
      need(its[i] && its[i]->ctx == c, "every integrator of the batch must exist and share the first one's context");
      need(its[i]->kind == 1, "Redheffer integrators expected");
      v.push_back(its[i]->it.get());
    }
    Integrator::redheffer_many(v, pe, g_redheffer_depth);
  });
}

int vrc_differential_compute(vrc_integrator* it, int N) {
  if (!it) return -1;
  return checked(it->ctx, [&] {
    need(it->kind == 2 && N >= 1, "differential integrator with N >= 1 expected");
    it->it->differential(N);
  });
}

int vrc_integrator_get(vrc_integrator* it, int which, double* host) {
  if (!it) return -1;
  return guarded(it->ctx, [&] {
    need(host != nullptr, "null output");
    it->it->get(which, (cd*)host);
  });
}

int vrc_integrator_counts(vrc_integrator* it, int* N, int* nEig) {
  if (!it) return -1;
  if (N) *N = it->it->N_;
  if (nEig) *nEig = it->it->nEig_;
  return 0;
}

int vrc_integrator_tree(vrc_integrator* it, int cap, double* out) {
  if (!it) return -1;
  const auto& t = it->it->tree_;
  for (int i = 0; i < (int)t.size() && i < cap && out; ++i) {
    out[2 * i] = t[i].first;
    out[2 * i + 1] = t[i].second;
  }
  return (int)t.size();
}

namespace {
// Host copies of a batch of stack instances (the StackInstance pointers refer into it).
struct HostStack {
  std::vector<std::vector<HostLayer>> keep;
  std::vector<HostLayer> bottoms, tops;
  std::vector<StackInstance> si;
  HostStack(int ninst, const vrc_stack_instance* inst) : keep(ninst), si(ninst) {
    bottoms.reserve(ninst);
    tops.reserve(ninst);
    for (int i = 0; i < ninst; ++i) {
      const vrc_stack_instance& I = inst[i];
      need(I.nlayers >= 1 && I.thickness, "instance needs >= 1 layer and thicknesses");
      keep[i] = to_host(I.layers, I.nlayers);
      bottoms.push_back(to_host(I.bottom));
      tops.push_back(to_host(I.top));
    }
    for (int i = 0; i < ninst; ++i) {
      si[i].lambda = inst[i].lambda;
      for (auto& l : keep[i]) si[i].layers.push_back(&l);
      si[i].thickness.assign(inst[i].thickness, inst[i].thickness + inst[i].nlayers);
      si[i].bottom = &bottoms[i];
      si[i].top = &tops[i];
    }
  }
};
}  // namespace

struct vrc_tape {
  vrc_ctx* ctx;
  std::unique_ptr<StackTape> t;
};

int vrc_batch_stack_rcwa(vrc_ctx* c, int nx, int ny, int ninst, const vrc_stack_instance* inst, int chunk,
                         void* S_out, void* W0, void* V0, void* g0, void* W1, void* V1, void* g1, int* info_out) {
  return vrc_batch_stack_forward(c, nx, ny, ninst, inst, chunk, S_out, W0, V0, g0, W1, V1, g1, nullptr, info_out);
}

int vrc_batch_stack_rcwa_mixed(vrc_ctx* c, int ninst, const int* nx, const int* ny, const vrc_stack_instance* inst,
                               int chunk, void* S_out, void* W0, void* V0, void* g0, void* W1, void* V1, void* g1,
                               int* info_out) {
  int num = 0;
  const int rc = guarded(c, [&] {
    need(ninst >= 0 && (ninst == 0 || (nx && ny && inst && S_out)), "bad arguments");
    if (ninst == 0) return;
    for (int i = 0; i < ninst; ++i) check_harmonics(nx[i], ny[i]);
    // packed offsets (complex elements) of every instance's outputs
    std::vector<long> oS(ninst), oM(ninst), oG(ninst);
    long aS = 0, aM = 0, aG = 0;
    for (int i = 0; i < ninst; ++i) {
      const long N = 2L * nx[i] * ny[i];
      oS[i] = aS;
      oM[i] = aM;
      oG[i] = aG;
      aS += 4 * N * N;
      aM += N * N;
      aG += N;
    }
    // harmonic classes, largest first
    std::map<std::pair<int, int>, std::vector<int>> cls;
    for (int i = 0; i < ninst; ++i) cls[{nx[i], ny[i]}].push_back(i);
    std::vector<std::pair<std::pair<int, int>, std::vector<int>>> order(cls.begin(), cls.end());
    std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) {
      return a.first.first * a.first.second > b.first.first * b.first.second;
    });
    health_reset(*c);
    std::vector<int> fl(ninst, 0);
    std::string m;
    for (auto& [hk, idx] : order) {
      const int cx = hk.first, cy = hk.second, k = (int)idx.size();
      const long N = 2L * cx * cy, NN = N * N;
      std::vector<vrc_stack_instance> sub(k);
      for (int q = 0; q < k; ++q) sub[q] = inst[idx[q]];
      HostStack hs(k, sub.data());
      DevBuf So = c->ws.get(sizeof(cd) * 4 * NN * k);
      const bool wantM = W0 || V0 || W1 || V1, wantG = g0 || g1;
      DevBuf Mo = c->ws.get(wantM ? sizeof(cd) * 4 * NN * k : 256), Go = c->ws.get(wantG ? sizeof(cd) * 2 * N * k : 256);
      cd* M = Mo.as<cd>();
      cd* G = Go.as<cd>();
      DevBuf flags = c->ws.get(sizeof(int) * (size_t)k);
      fill_int(c->stream, flags.as<int>(), 0, k);
      batch_stack_rcwa(*c, cx, cy, hs.si, chunk, So.as<cd>(), W0 ? M : nullptr, V0 ? M + NN * k : nullptr,
                       g0 ? G : nullptr, W1 ? M + 2 * NN * k : nullptr, V1 ? M + 3 * NN * k : nullptr,
                       g1 ? G + N * k : nullptr, nullptr, flags.as<int>());
      // class outputs (instance-major) into the packed per-instance layout
      auto put = [&](void* dst, const cd* src, long per, const std::vector<long>& off) {
        for (int q = 0; q < k; ++q)
          VRC_HIP(hipMemcpyAsync((cd*)dst + off[idx[q]], src + per * q, sizeof(cd) * per, hipMemcpyDeviceToDevice,
                                 c->stream));
      };
      put(S_out, So.as<cd>(), 4 * NN, oS);
      if (W0) put(W0, M, NN, oM);
      if (V0) put(V0, M + NN * k, NN, oM);
      if (W1) put(W1, M + 2 * NN * k, NN, oM);
      if (V1) put(V1, M + 3 * NN * k, NN, oM);
      if (g0) put(g0, G, N, oG);
      if (g1) put(g1, G + N * k, N, oG);
      std::vector<int> f(k);
      VRC_HIP(hipMemcpyAsync(f.data(), flags.ptr, sizeof(int) * k, hipMemcpyDeviceToHost, c->stream));
      VRC_HIP(hipStreamSynchronize(c->stream));
      for (int q = 0; q < k; ++q) fl[idx[q]] = f[q];
    }
    num = health_read(*c, m);  // synchronizes
    if (info_out) std::copy(fl.begin(), fl.end(), info_out);
    for (int i = 0; i < ninst; ++i)
      if (fl[i]) {
        num = fl[i];
        m = "instance " + std::to_string(i) + ": " + health_text(fl[i]) + (m.empty() ? "" : "; " + m);
        break;
      }
    if (num) c->last_error = m;
  });
  return rc != 0 ? rc : num;
}

// ------------------------------------------------------------------ recorded batches
struct vrc_batch_plan {
  vrc_ctx* ctx = nullptr;
  int ninst = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<std::pair<void*, size_t>> dev;  // device blocks the graph reads and writes
  std::vector<void*> pinned;                  // host staging the graph's copies read
  int* flags = nullptr;                       // per-instance codes (one of the device blocks)
  ~vrc_batch_plan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& b : dev) (void)hipFree(b.first);
    for (void* p : pinned) (void)hipHostFree(p);
  }
};

int vrc_batch_plan_create(vrc_ctx* c, int nx, int ny, int ninst, const vrc_stack_instance* inst, int chunk,
                          void* S_out, void* W0, void* V0, void* g0, void* W1, void* V1, void* g1,
                          vrc_batch_plan** out) {
  if (out) *out = nullptr;
  return guarded(c, [&] {
    check_harmonics(nx, ny);
    need(out && ninst >= 1 && inst && S_out, "bad arguments");
    HostStack hs(ninst, inst);
    auto plan = std::make_unique<vrc_batch_plan>();
    plan->ctx = c;
    plan->ninst = ninst;
    VRC_HIP(hipStreamSynchronize(c->stream));
    hipStream_t cap = nullptr;
    VRC_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    // the context records on the capture stream, without stage timers, its pool noting
    // every block the recorded work touches and its uploads staged in blocks the plan keeps
    struct Recording {
      vrc_ctx* c;
      hipStream_t prev, cap;
      bool timing;
      bool capturing = false;
      Recording(vrc_ctx* cc, hipStream_t cs, std::vector<void*>* hold)
          : c(cc), prev(cc->stream), cap(cs), timing(cc->timing) {
        c->stream = cs;
        c->timing = false;
        c->pinned.hold = hold;
        c->ws.track_begin();
      }
      ~Recording() {
        if (capturing) {  // an error while recording: end it and drop the partial graph
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(cap, &g);
          if (g) (void)hipGraphDestroy(g);
        }
        if (c->ws.tracking()) {  // an error before the plan claimed its blocks: nothing ran, free them
          size_t leased = 0;
          for (auto& b : c->ws.track_end(&leased)) (void)hipFree(b.first);
        }
        c->stream = prev;
        c->timing = timing;
        c->pinned.hold = nullptr;
        (void)hipStreamSynchronize(cap);
        (void)hipStreamDestroy(cap);
      }
    } rec(c, cap, &plan->pinned);
    VRC_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed));
    rec.capturing = true;
    {
      DevBuf flags = c->ws.get(sizeof(int) * (size_t)ninst);
      plan->flags = flags.as<int>();
      fill_int(cap, flags.as<int>(), 0, ninst);
      fill_int(cap, c->health, 0, HEALTH_INTS);
      batch_stack_rcwa(*c, nx, ny, hs.si, chunk, (cd*)S_out, (cd*)W0, (cd*)V0, (cd*)g0, (cd*)W1, (cd*)V1, (cd*)g1,
                       nullptr, flags.as<int>());
    }
    rec.capturing = false;
    VRC_HIP(hipStreamEndCapture(cap, &plan->graph));
    size_t leased = 0;
    plan->dev = c->ws.track_end(&leased);
    need(leased == 0, "internal: a block of the recorded work is still leased");
    VRC_HIP(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
    ctx_retain(c);
    *out = plan.release();
  });
}

int vrc_batch_plan_launch(vrc_batch_plan* p, int* info_out) {
  if (!p) return -1;
  vrc_ctx* c = p->ctx;
  int num = 0;
  const int rc = guarded(c, [&] {
    c->host_code = 0;
    c->host_msg.clear();
    VRC_HIP(hipGraphLaunch(p->exec, c->stream));
    std::vector<int> fl(p->ninst);
    VRC_HIP(hipMemcpyAsync(fl.data(), p->flags, sizeof(int) * p->ninst, hipMemcpyDeviceToHost, c->stream));
    std::string m;
    num = health_read(*c, m);  // synchronizes
    if (info_out) std::copy(fl.begin(), fl.end(), info_out);
    for (int i = 0; i < p->ninst; ++i)
      if (fl[i]) {
        num = fl[i];
        m = "instance " + std::to_string(i) + ": " + health_text(fl[i]) + (m.empty() ? "" : "; " + m);
        break;
      }
    if (num) c->last_error = m;
  });
  return rc != 0 ? rc : num;
}

int vrc_batch_plan_destroy(vrc_batch_plan* p) {
  if (!p) return 0;
  vrc_ctx* c = p->ctx;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  delete p;
  ctx_release(c);
  return 0;
}

// ------------------------------------------------------------------ autotuner
int vrc_autotune_stack(vrc_ctx* c, int nx, int ny, int ninst, const vrc_stack_instance* inst, double budget_s,
                       const char* cache_path, char* report, int report_len) {
  return guarded(c, [&] {
    check_harmonics(nx, ny);
    need(ninst >= 1 && inst && budget_s > 0, "bad arguments");
    HostStack hs(ninst, inst);
    TuneKnobs best;
    const std::string rep = autotune_stack(*c, nx, ny, hs.si, budget_s, &best);
    tune_store(2 * nx * ny, best);
    if (cache_path) tune_save(cache_path, 2 * nx * ny, best, rep);
    if (report && report_len > 0) {
      std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
      report[report_len - 1] = 0;
    }
  });
}

int vrc_autotune_load(const char* path) {
  if (!path) return -1;
  try {
    return tune_load(path);
  } catch (...) {
    return -1;
  }
}

int vrc_autotune_clear(void) {
  tune_clear();
  return 0;
}

int vrc_modal_power(vrc_ctx* c, int nx, int ny, int nmodes, const void* W, long long ldw, const void* V,
                    long long ldv, double Lx, double Ly, int nres_x, int nres_y, double x0, double x1, double y0,
                    double y1, double* energy) {
  return guarded(c, [&] {
    check_harmonics(nx, ny);
    const long n2 = 2L * nx * ny;
    need(nmodes >= 0 && (nmodes == 0 || (W && V && energy)), "bad arguments");
    need(ldw >= n2 && ldv >= n2, "leading dimensions must be at least 2 nx ny");
    need(nres_x > 0 && nres_y > 0, "the sampling grid needs at least one point per axis");
    need(Lx > 0 && Ly > 0, "cell periods must be positive");
    modal_power(*c, nx, ny, nmodes, (const cd*)W, (long)ldw, (const cd*)V, (long)ldv, Lx, Ly, nres_x, nres_y, x0,
                x1, y0, y1, energy);
  });
}

int vrc_find_guided_modes(vrc_ctx* c, int n, const void* gamma, double k0, double imag_coeff, double re_min,
                          double re_max, int* index_out, void* neff_out, int* count) {
  return host_guarded(c, [&] {
    need(n >= 0 && (n == 0 || (gamma && index_out && neff_out)) && count, "bad arguments");
    need(k0 > 0, "k0 must be positive");
    const cd* g = (const cd*)gamma;
    std::vector<std::pair<int, cd>> rec;
    for (int i = 0; i < n; ++i) {
      const cd gi = mk(g[i].x / k0 / k0, g[i].y / k0 / k0);
      if (re_min <= gi.x && gi.x <= re_max && gi.y < imag_coeff) rec.emplace_back(i, gi);
    }
    std::stable_sort(rec.begin(), rec.end(), [](const std::pair<int, cd>& a, const std::pair<int, cd>& b) {
      return a.second.x > b.second.x;
    });
    for (size_t k = 0; k < rec.size(); ++k) {
      index_out[k] = rec[k].first;
      ((cd*)neff_out)[k] = rec[k].second;
    }
    *count = (int)rec.size();
  });
}

int vrc_stack_efficiencies(vrc_ctx* c, int nx, int ny, int ninst, const vrc_stack_instance* inst, const void* S,
                           const void* W0, const void* W1, const void* e_inc, double* REF, double* TRN) {
  return guarded(c, [&] {
    check_harmonics(nx, ny);
    need(ninst >= 0 && (ninst == 0 || (inst && S && W0 && W1 && e_inc && REF && TRN)), "bad arguments");
    HostStack hs(ninst, inst);
    stack_efficiencies(*c, nx, ny, hs.si, (const cd*)S, (const cd*)W0, (const cd*)W1, (const cd*)e_inc, REF, TRN);
  });
}

int vrc_batch_stack_forward(vrc_ctx* c, int nx, int ny, int ninst, const vrc_stack_instance* inst, int chunk,
                            void* S_out, void* W0, void* V0, void* g0, void* W1, void* V1, void* g1,
                            vrc_tape** tape, int* info_out) {
  return vrc_batch_stack_forward_ex(c, nx, ny, ninst, inst, chunk, S_out, W0, V0, g0, W1, V1, g1, 0, 0, tape,
                                    info_out);
}

int vrc_batch_stack_forward_ex(vrc_ctx* c, int nx, int ny, int ninst, const vrc_stack_instance* inst, int chunk,
                               void* S_out, void* W0, void* V0, void* g0, void* W1, void* V1, void* g1,
                               int recompute, int layer_block, vrc_tape** tape, int* info_out) {
  if (tape) *tape = nullptr;
  int num = 0;
  const int rc = guarded(c, [&] {
    check_harmonics(nx, ny);
    need(ninst >= 0 && (ninst == 0 || (inst && S_out)), "bad arguments");
    if (ninst == 0) {  // nothing to solve
      need(tape == nullptr, "a gradient tape needs at least one instance");
      return;
    }
    HostStack hs(ninst, inst);
    std::unique_ptr<vrc_tape> tp;
    if (tape) {
      tp = std::make_unique<vrc_tape>();
      tp->ctx = c;
      tp->t = make_stack_tape(*c, nx, ny, hs.si, recompute != 0, layer_block);
    }
    health_reset(*c);
    DevBuf flags = c->ws.get(sizeof(int) * (size_t)ninst);
    fill_int(c->stream, flags.as<int>(), 0, ninst);
    batch_stack_rcwa(*c, nx, ny, hs.si, chunk, (cd*)S_out, (cd*)W0, (cd*)V0, (cd*)g0, (cd*)W1, (cd*)V1, (cd*)g1,
                     tp ? tp->t.get() : nullptr, flags.as<int>());
    std::vector<int> fl(ninst);
    VRC_HIP(hipMemcpyAsync(fl.data(), flags.ptr, sizeof(int) * ninst, hipMemcpyDeviceToHost, c->stream));
    std::string m;
    num = health_read(*c, m);  // synchronizes
    if (info_out) std::copy(fl.begin(), fl.end(), info_out);
    for (int i = 0; i < ninst; ++i)
      if (fl[i]) {
        num = fl[i];
        m = "instance " + std::to_string(i) + ": " + health_text(fl[i]) + (m.empty() ? "" : "; " + m);
        break;
      }
    if (num) c->last_error = m;
    if (tape) {
      ctx_retain(c);
      *tape = tp.release();
    }
  });
  return rc != 0 ? rc : num;
}

long long vrc_tape_cells(const vrc_tape* t) { return (t && t->t) ? (long long)t->t->cells() : -1; }

long long vrc_tape_bytes(const vrc_tape* t) { return (t && t->t) ? (long long)t->t->device_bytes() : -1; }

int vrc_batch_stack_vjp(vrc_tape* t, const void* S_bar, void* eps_bar, int chunk) {
  if (!t) return -1;
  return guarded(t->ctx, [&] {
    need(S_bar && eps_bar, "bad arguments");
    stack_tape_vjp(*t->t, (const cd*)S_bar, (cd*)eps_bar, chunk);
  });
}

int vrc_batch_stack_vjp_shape(vrc_tape* t, const void* S_bar, const int* npara, const int* nvar,
                              const int* var_index, const double* coeff, double* shape_bar_host, void* eps_bar,
                              int chunk) {
  if (!t) return -1;
  return guarded(t->ctx, [&] {
    need(S_bar && npara && nvar && shape_bar_host, "bad arguments");
    const StackTape& tp = *t->t;
    const int S = tp.ninst * tp.L;
    ShapeParams sp;
    sp.npara.assign(npara, npara + S);
    sp.nvar.assign(nvar, nvar + S);
    sp.pofs.assign(S + 1, 0);
    sp.vofs.assign(S + 1, 0);
    sp.cofs.assign(S + 1, 0);
    for (int s = 0; s < S; ++s) {
      need(npara[s] >= 0 && nvar[s] >= 0, "negative parameter or coordinate count");
      need(nvar[s] == 0 || npara[s] > 0, "varied coordinates need at least one parameter");
      sp.pofs[s + 1] = sp.pofs[s] + npara[s];
      sp.vofs[s + 1] = sp.vofs[s] + nvar[s];
      sp.cofs[s + 1] = sp.cofs[s] + (long)nvar[s] * npara[s];
    }
    need(sp.vofs[S] == 0 || (var_index && coeff), "coordinate lists missing");
    sp.var_index.assign(var_index, var_index + sp.vofs[S]);
    sp.coeff.assign(coeff, coeff + sp.cofs[S]);
    for (int s = 0; s < S; ++s)
      for (long k = sp.vofs[s]; k < sp.vofs[s + 1]; ++k) {
        const HostLayer* L = tp.slab(s / tp.L, s % tp.L);
        need(sp.var_index[k] > 0 && sp.var_index[k] < (int)L->x.size() - 1,
             "only interior x coordinates can be varied");
      }
    const long P = sp.pofs[S];
    vrc_ctx* c = t->ctx;
    DevBuf sb = c->ws.get(sizeof(double) * (size_t)std::max<long>(1, P));
    if (P > 0) {
      const std::vector<double> zero((size_t)P, 0.0);
      c->pinned.upload(c->stream, sb.ptr, zero.data(), sizeof(double) * P);
    }
    stack_tape_vjp(*t->t, (const cd*)S_bar, (cd*)eps_bar, chunk, &sp, sb.as<double>());
    if (P > 0) {
      VRC_HIP(hipMemcpyAsync(shape_bar_host, sb.ptr, sizeof(double) * P, hipMemcpyDeviceToHost, c->stream));
      VRC_HIP(hipStreamSynchronize(c->stream));
    }
  });
}

int vrc_tape_destroy(vrc_tape* t) {
  if (!t) return 0;
  vrc_ctx* c = t->ctx;
  delete t;  // device buffers go back to the context's pool (stream-ordered reuse)
  ctx_release(c);
  return 0;
}

int vrc_assemble_toeplitz_vjp(vrc_ctx* c, int nx, int ny, int nlayers, const vrc_layer* layers, const void* fe11b,
                              const void* fe22b, const void* fe33b, long stride, void* eps_bar) {
  return guarded(c, [&] {
    check_harmonics(nx, ny);
    need(nlayers >= 0 && fe11b && fe22b && fe33b && eps_bar, "bad arguments");
    const long n = (long)nx * ny;
    need(stride >= n * n || nlayers <= 1, "stride smaller than n*n");
    std::vector<HostLayer> hl = to_host(layers, nlayers);
    std::vector<const HostLayer*> lp;
    std::vector<long> off;
    long o = 0;
    for (auto& l : hl) {
      lp.push_back(&l);
      off.push_back(o);
      o += (long)l.eps.size();
    }
    if (nlayers == 0) return;
    assemble_fe_vjp(*c, nx, ny, lp, (const cd*)fe11b, (const cd*)fe22b, (const cd*)fe33b, stride, (cd*)eps_bar, off);
  });
}

int vrc_assemble_pq_vjp(vrc_ctx* c, int nx, int ny, int nlayers, const vrc_layer* layers, double lambda,
                        const vrc_layer* region, const void* Pb, const void* Qb, long stride, void* eps_bar) {
  return guarded(c, [&] {
    check_harmonics(nx, ny);
    need(nlayers >= 0 && Pb && Qb && eps_bar && region, "bad arguments");
    need(lambda > 0, "wavelength must be positive");
    const long N = 2L * nx * ny;
    need(stride == N * N, "P/Q stride must be N*N");
    if (nlayers == 0) return;
    std::vector<HostLayer> hl = to_host(layers, nlayers);
    HostLayer reg = to_host(region);
    const int n = nx * ny;
    std::vector<cd> al, be;
    alpha_beta(nx, ny, reg.x.back() - reg.x.front(), reg.y.back() - reg.y.front(), al, be);
    std::vector<cd> ab(2 * (size_t)n);
    std::copy(al.begin(), al.end(), ab.begin());
    std::copy(be.begin(), be.end(), ab.begin() + n);
    const double k0 = 2.0 * 3.1415926535897932 / lambda;
    std::vector<double> k2(nlayers, k0 * k0);
    DevBuf dab = c->ws.get(sizeof(cd) * ab.size()), dk = c->ws.get(sizeof(double) * nlayers);
    VRC_HIP(hipMemcpyAsync(dab.ptr, ab.data(), sizeof(cd) * ab.size(), hipMemcpyHostToDevice, c->stream));
    VRC_HIP(hipMemcpyAsync(dk.ptr, k2.data(), sizeof(double) * nlayers, hipMemcpyHostToDevice, c->stream));
    AssemblyBatch a;
    a.nx = nx;
    a.ny = ny;
    a.nlayers = nlayers;
    a.periodic = true;
    std::vector<long> off;
    long o = 0;
    for (auto& l : hl) {
      a.layers.push_back(&l);
      off.push_back(o);
      o += (long)l.eps.size();
    }
    a.d_alpha = dab.as<cd>();
    a.d_beta = a.d_alpha + n;
    a.sab = 0;
    a.d_k0sq = dk.as<double>();
    assemble_pq_vjp(*c, a, (const cd*)Pb, (const cd*)Qb, stride, (cd*)eps_bar, off);
  });
}

int vrc_eig_pq_vjp(vrc_ctx* c, int N, int batch, const void* P, const void* Q, long stride, const void* W,
                   const void* V, const void* gam, void* W_bar, const void* V_bar, void* gam_bar, void* P_bar,
                   void* Q_bar) {
  return guarded(c, [&] {
    need(N > 0 && batch >= 0 && P && Q && W && V && gam && W_bar && V_bar && gam_bar && P_bar && Q_bar,
         "bad arguments");
    need(stride == (long)N * N, "stride must be N*N");
    if (batch == 0) return;
    wave_eqn_vjp(*c, N, batch, (const cd*)P, (const cd*)Q, stride, (const cd*)W, (const cd*)V, (const cd*)gam,
                 nullptr, (cd*)W_bar, (const cd*)V_bar, (cd*)gam_bar, (cd*)P_bar, (cd*)Q_bar);
  });
}

int vrc_layer_smatrix_vjp(vrc_ctx* c, int N, int batch, const double* lam, const double* dz, const void* W,
                          const void* V, const void* gam, const void* W0, const void* V0, long ref_stride,
                          const void* Tuu_bar, const void* Rud_bar, void* W_bar, void* V_bar, void* gam_bar) {
  return guarded(c, [&] {
    need(N > 0 && batch >= 0 && lam && dz && W && V && gam && W0 && V0, "bad arguments");
    need(Tuu_bar && Rud_bar && W_bar && V_bar && gam_bar, "bad arguments");
    if (batch == 0) return;
    std::vector<double> s(batch);
    for (int b = 0; b < batch; ++b) {
      need(lam[b] > 0, "wavelength must be positive");
      s[b] = dz[b] * lam[b] / (2.0 * 3.1415926535897932);
    }
    DevBuf ds = c->ws.get(sizeof(double) * batch);
    VRC_HIP(hipMemcpyAsync(ds.ptr, s.data(), sizeof(double) * batch, hipMemcpyHostToDevice, c->stream));
    VRC_HIP(hipStreamSynchronize(c->stream));
    layer_smatrix_vjp(*c, N, batch, ds.as<double>(), (const cd*)W, (const cd*)V, (const cd*)gam, (const cd*)W0,
                      (const cd*)V0, ref_stride, (const cd*)Tuu_bar, (const cd*)Rud_bar, (cd*)W_bar, (cd*)V_bar,
                      (cd*)gam_bar, nullptr);
  });
}

int vrc_star_product_vjp(vrc_ctx* c, int N, int batch, const void* Tdd, const void* Tuu, const void* Rud,
                         const void* Rdu, const void* tdd, const void* tuu, const void* rud, const void* rdu,
                         long slice_stride, void* Tdd_bar, void* Tuu_bar, void* Rud_bar, void* Rdu_bar, void* tdd_bar,
                         void* tuu_bar, void* rud_bar, void* rdu_bar) {
  return guarded(c, [&] {
    need(N > 0 && batch >= 0 && Tdd && Tuu && Rud && Rdu && tdd && tuu && rud && rdu, "bad arguments");
    need(Tdd_bar && Tuu_bar && Rud_bar && Rdu_bar && tdd_bar && tuu_bar && rud_bar && rdu_bar, "bad arguments");
    if (batch == 0) return;
    SMat S0, Sb;
    S0.Tdd = (cd*)Tdd;
    S0.Tuu = (cd*)Tuu;
    S0.Rud = (cd*)Rud;
    S0.Rdu = (cd*)Rdu;
    Sb.Tdd = (cd*)Tdd_bar;
    Sb.Tuu = (cd*)Tuu_bar;
    Sb.Rud = (cd*)Rud_bar;
    Sb.Rdu = (cd*)Rdu_bar;
    star_product_vjp(*c, N, batch, S0, (const cd*)tdd, (const cd*)tuu, (const cd*)rud, (const cd*)rdu, slice_stride,
                     Sb, (cd*)tdd_bar, (cd*)tuu_bar, (cd*)rud_bar, (cd*)rdu_bar, (long)N * N);
  });
}

int vrc_project_left_vjp(vrc_ctx* c, int N, int batch, const void* Tdd, const void* Tuu, const void* Rud,
                         const void* Rdu, const void* W1, const void* V1, const void* W0, const void* V0,
                         long ref_stride, void* Tdd_bar, void* Tuu_bar, void* Rud_bar, void* Rdu_bar) {
  return guarded(c, [&] {
    need(N > 0 && batch >= 0 && Tdd && Tuu && Rud && Rdu && W1 && V1 && W0 && V0, "bad arguments");
    need(Tdd_bar && Tuu_bar && Rud_bar && Rdu_bar, "bad arguments");
    if (batch == 0) return;
    const long NN = (long)N * N;
    SMat S, Sb;
    S.Tdd = (cd*)Tdd;
    S.Tuu = (cd*)Tuu;
    S.Rud = (cd*)Rud;
    S.Rdu = (cd*)Rdu;
    Sb.Tdd = (cd*)Tdd_bar;
    Sb.Tuu = (cd*)Tuu_bar;
    Sb.Rud = (cd*)Rud_bar;
    Sb.Rdu = (cd*)Rdu_bar;
    DevBuf LW = c->ws.get(sizeof(cd) * NN * batch), LV = c->ws.get(sizeof(cd) * NN * batch);
    DevBuf pW = c->ws.get(sizeof(int) * (size_t)N * batch), pV = c->ws.get(sizeof(int) * (size_t)N * batch);
    DevBuf inf = c->ws.get(sizeof(int) * (size_t)batch);
    fill_int(c->stream, inf.as<int>(), 0, batch);
    zgeam(c->stream, N, N, mk(1, 0), (const cd*)W1, N, ref_stride, mk(0, 0), nullptr, 0, 0, LW.as<cd>(), N, NN,
          batch);
    zgeam(c->stream, N, N, mk(1, 0), (const cd*)V1, N, ref_stride, mk(0, 0), nullptr, 0, 0, LV.as<cd>(), N, NN,
          batch);
    zgetrf(c->stream, N, LW.as<cd>(), N, NN, pW.as<int>(), N, inf.as<int>(), batch, c->ws);
    zgetrf(c->stream, N, LV.as<cd>(), N, NN, pV.as<int>(), N, inf.as<int>(), batch, c->ws);
    LUPair lu;
    lu.W = LW.as<cd>();
    lu.V = LV.as<cd>();
    lu.pW = pW.as<int>();
    lu.pV = pV.as<int>();
    lu.sM = NN;
    project_left_vjp(*c, N, batch, S, lu, (const cd*)W0, (const cd*)V0, ref_stride, Sb);
  });
}

}  // extern "C"
