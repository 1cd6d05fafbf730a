// The legacy stacked-layer solver (RCWASolver, src/rcwa/RCWASolver.cpp:73-580): every layer
// type is solved once in one batched eigensolve, interface matrices come from LU solves with
// the neighbouring types' eigenvectors, and the RTCM recursion runs from the last layer to
// the first with one inverse and four products per interface, all on the device.
#include <cmath>
#include <map>

#include "vrc_context.hpp"
#include "vrc_ops.hpp"

namespace vrc {

static constexpr double PI_L = 3.1415926535897932;  // src/rcwa/defns.h:29
static const cd ONE = {1.0, 0.0}, ZERO = {0.0, 0.0};

LegacySolver::LegacySolver(vrc_ctx* c, int nX, int nY) : c_(c), nx_(2 * nX + 1), ny_(2 * nY + 1) {
  if (nX < 0 || nY < 0) throw ArgFailure("harmonic orders must be non-negative");
}

// RCWASolver::addLayer (src/rcwa/RCWASolver.cpp:73-102): every layer shares the first one's
// cell and PML region (the reference prints an error and skips a mismatching layer).
void LegacySolver::add_layer(const HostLayer& L) {
  if (L.x.size() < 4 || L.y.size() < 4) throw ArgFailure("a layer needs at least 4 coordinates per axis (PML)");
  if ((long)L.eps.size() != (long)(L.x.size() - 1) * (long)(L.y.size() - 1))
    throw ArgFailure("eps size does not match coordinates");
  if (!layers_.empty()) {
    const HostLayer& f = layers_[0];
    auto region = [](const HostLayer& l) {
      const size_t nx = l.x.size(), ny = l.y.size();
      return std::vector<double>{l.x[nx - 1] - l.x[0], l.y[ny - 1] - l.y[0], l.y[0], l.y[1], l.y[ny - 2],
                                 l.y[ny - 1], l.x[0], l.x[1], l.x[nx - 2], l.x[nx - 1]};
    };
    if (region(f) != region(L)) throw ArgFailure("the new layer does not match the layers before");
  }
  layers_.push_back(L);
  solved_ = false;
}

void LegacySolver::solve(double lambda, const std::vector<int>& stack, const std::vector<double>& thickness,
                         const double* translation) {
  if (layers_.empty()) throw ArgFailure("the layers are empty");
  const int nL = (int)stack.size();
  if (nL < 2) throw ArgFailure("a stack needs at least two layers");
  if ((int)thickness.size() != nL) throw ArgFailure("thickness needs one value per stack entry");
  for (int t : stack)
    if (t < 0 || t >= (int)layers_.size()) throw ArgFailure("stack entry is not a layer type");
  hipStream_t st = c_->stream;
  const int n = nx_ * ny_, N = 2 * n, L = (int)layers_.size();
  const long NN = (long)N * N;
  // Layer::solve (src/rcwa/Layer.cpp:173-190) for every type: P, Q with the PML K of the
  // first layer's region, batched eigensolve
  VarSampler smp(c_, lambda, nx_, ny_, layers_[0], true);
  const double k0 = smp.k0();
  k0_ = k0;
  DevBuf P = c_->ws.get(sizeof(cd) * NN * L), Q = c_->ws.get(sizeof(cd) * NN * L);
  smp.sample_layers(layers_, P.as<cd>(), Q.as<cd>());
  W_ = c_->ws.get(sizeof(cd) * NN * L);
  V_ = c_->ws.get(sizeof(cd) * NN * L);
  gam_ = c_->ws.get(sizeof(cd) * (size_t)N * L);
  DevBuf info = c_->ws.get(sizeof(int) * L);
  fill_int(st, info.as<int>(), 0, L);
  wave_eqn(*c_, N, P.as<cd>(), Q.as<cd>(), NN, W_.as<cd>(), V_.as<cd>(), gam_.as<cd>(), info.as<int>(), L);
  P = DevBuf();
  Q = DevBuf();
  // LU of every type's E and H (types needed by an interface only)
  DevBuf LUw = c_->ws.get(sizeof(cd) * NN * L), LUv = c_->ws.get(sizeof(cd) * NN * L);
  DevBuf pw = c_->ws.get(sizeof(int) * (size_t)N * L), pv = c_->ws.get(sizeof(int) * (size_t)N * L),
         inf2 = c_->ws.get(sizeof(int) * 2 * L);
  fill_int(st, inf2.as<int>(), 0, 2 * L);
  zgeam(st, N, N, ONE, W_.as<cd>(), N, NN, ZERO, nullptr, 0, 0, LUw.as<cd>(), N, NN, L);
  zgeam(st, N, N, ONE, V_.as<cd>(), N, NN, ZERO, nullptr, 0, 0, LUv.as<cd>(), N, NN, L);
  zgetrf(st, N, LUw.as<cd>(), N, NN, pw.as<int>(), N, inf2.as<int>(), L, c_->ws);
  zgetrf(st, N, LUv.as<cd>(), N, NN, pv.as<int>(), N, inf2.as<int>() + L, L, c_->ws);
  // interfaces (src/rcwa/RCWASolver.cpp:460-580): unique pairs without translation, one per
  // stack step with it; translation scales the eigenvector rows of x-harmonic m by
  // exp(-i 2 pi m t / Lx) (Layer::permuteEigVecX), so E = E_a^-1 diag(d_b / d_a) E_b.
  std::vector<int> imap(nL - 1);
  std::vector<std::pair<int, int>> pairs;
  if (!translation) {
    std::map<std::pair<int, int>, int> uniq;
    for (int i = 0; i + 1 < nL; ++i) {
      auto key = std::make_pair(stack[i], stack[i + 1]);
      auto it = uniq.find(key);
      if (it == uniq.end()) {
        it = uniq.emplace(key, (int)pairs.size()).first;
        pairs.push_back(key);
      }
      imap[i] = it->second;
    }
  } else {
    for (int i = 0; i + 1 < nL; ++i) {
      pairs.emplace_back(stack[i], stack[i + 1]);
      imap[i] = i;
    }
  }
  const int NP = (int)pairs.size();
  DevBuf IE = c_->ws.get(sizeof(cd) * NN * NP), IH = c_->ws.get(sizeof(cd) * NN * NP);
  const double Lx = layers_[0].x.back() - layers_[0].x.front();
  DevBuf dscale = c_->ws.get(sizeof(cd) * N), tmp = c_->ws.get(sizeof(cd) * NN);
  for (int k = 0; k < NP; ++k) {
    const int a = pairs[k].first, b = pairs[k].second;
    cd* E = IE.as<cd>() + k * NN;
    cd* H = IH.as<cd>() + k * NN;
    if (!translation && a == b) {
      zset(st, N, N, ONE, ZERO, E, N, 0, 1);
      zset(st, N, N, ONE, ZERO, H, N, 0, 1);
      continue;
    }
    const cd* Eb = W_.as<cd>() + b * NN;
    const cd* Hb = V_.as<cd>() + b * NN;
    if (translation) {
      std::vector<cd> d(N);
      const double ta = translation[k], tb = translation[k + 1];
      for (int i = 0; i < nx_; ++i) {
        const int m = i - (nx_ - 1) / 2;
        const double pa = -2.0 * PI_L * m / Lx * ta, pb = -2.0 * PI_L * m / Lx * tb;
        const cd r = cdiv(mk(std::cos(pb), std::sin(pb)), mk(std::cos(pa), std::sin(pa)));
        for (int j = 0; j < ny_; ++j) d[(size_t)i * ny_ + j] = d[(size_t)i * ny_ + j + n] = r;
      }
      VRC_HIP(hipMemcpyAsync(dscale.ptr, d.data(), sizeof(cd) * N, hipMemcpyHostToDevice, st));
      zdscale(st, N, N, dscale.as<cd>(), 0, nullptr, 0, 0, Eb, N, 0, tmp.as<cd>(), N, 0, 1);
      zgetrs(st, N, N, LUw.as<cd>() + a * NN, N, 0, pw.as<int>() + a * N, 0, tmp.as<cd>(), N, 0, E, N, 0, 1,
             c_->ws);
      zdscale(st, N, N, dscale.as<cd>(), 0, nullptr, 0, 0, Hb, N, 0, tmp.as<cd>(), N, 0, 1);
      zgetrs(st, N, N, LUv.as<cd>() + a * NN, N, 0, pv.as<int>() + a * N, 0, tmp.as<cd>(), N, 0, H, N, 0, 1,
             c_->ws);
      VRC_HIP(hipStreamSynchronize(st));  // the host scale vector is reused
    } else {
      zgetrs(st, N, N, LUw.as<cd>() + a * NN, N, 0, pw.as<int>() + a * N, 0, Eb, N, 0, E, N, 0, 1, c_->ws);
      zgetrs(st, N, N, LUv.as<cd>() + a * NN, N, 0, pv.as<int>() + a * N, 0, Hb, N, 0, H, N, 0, 1, c_->ws);
    }
  }
  LUw = DevBuf();
  LUv = DevBuf();
  REF_ = c_->ws.get(sizeof(cd) * NN * (nL - 1));
  // RTCM (src/rcwa/RCWASolver.cpp:343-457), from the last layer to the first
  Tdd_ = c_->ws.get(sizeof(cd) * NN);
  Rud_ = c_->ws.get(sizeof(cd) * NN);
  cd* Tdd = Tdd_.as<cd>();
  cd* Rud = Rud_.as<cd>();
  zset(st, N, N, ONE, ZERO, Tdd, N, 0, 1);
  zset(st, N, N, ZERO, ZERO, Rud, N, 0, 1);
  std::vector<double> sv(nL, 0.0);
  for (int i = 1; i + 1 < nL; ++i) sv[i] = thickness[i] / k0;  // exp(i (gamma / k0) t)
  DevBuf ds = c_->ws.get(sizeof(double) * nL);
  VRC_HIP(hipMemcpyAsync(ds.ptr, sv.data(), sizeof(double) * nL, hipMemcpyHostToDevice, st));
  DevBuf phi = c_->ws.get(sizeof(cd) * N), wR = c_->ws.get(sizeof(cd) * NN), wT = c_->ws.get(sizeof(cd) * NN),
         F = c_->ws.get(sizeof(cd) * NN), G = c_->ws.get(sizeof(cd) * NN), S = c_->ws.get(sizeof(cd) * NN),
         tau = c_->ws.get(sizeof(cd) * NN), piv = c_->ws.get(sizeof(int) * N), inf = c_->ws.get(sizeof(int));
  fill_int(st, inf.as<int>(), 0, 1);
  for (int i = 0; i + 1 < nL; ++i) {
    const int index = nL - i - 1;
    const int type = stack[index];
    // phi = 1 on the last (semi-infinite) layer: s = 0 there
    phases(st, N, gam_.as<cd>() + (long)type * N, 0, ds.as<double>() + (i == 0 ? 0 : index), phi.as<cd>(), 0, 1);
    zdscale(st, N, N, phi.as<cd>(), 0, phi.as<cd>(), 0, 0, Rud, N, 0, wR.as<cd>(), N, 0, 1);
    zdscale(st, N, N, nullptr, 0, phi.as<cd>(), 0, 0, Tdd, N, 0, wT.as<cd>(), N, 0, 1);
    const cd* E = IE.as<cd>() + (long)imap[index - 1] * NN;
    const cd* H = IH.as<cd>() + (long)imap[index - 1] * NN;
    zgeam(st, N, N, ONE, E, N, 0, ZERO, nullptr, 0, 0, F.as<cd>(), N, 0, 1);
    zgemm_strided(st, OP_N, OP_N, N, N, N, ONE, E, N, 0, wR.as<cd>(), N, 0, ONE, F.as<cd>(), N, 0, 1);
    zgeam(st, N, N, ONE, H, N, 0, ZERO, nullptr, 0, 0, G.as<cd>(), N, 0, 1);
    zgemm_strided(st, OP_N, OP_N, N, N, N, mk(-1.0, 0.0), H, N, 0, wR.as<cd>(), N, 0, ONE, G.as<cd>(), N, 0, 1);
    zgeam(st, N, N, ONE, F.as<cd>(), N, 0, ONE, G.as<cd>(), N, 0, S.as<cd>(), N, 0, 1);
    zgetrf(st, N, S.as<cd>(), N, 0, piv.as<int>(), 0, inf.as<int>(), 1, c_->ws);
    zset(st, N, N, ONE, ZERO, tau.as<cd>(), N, 0, 1);
    zgetrs(st, N, N, S.as<cd>(), N, 0, piv.as<int>(), 0, tau.as<cd>(), N, 0, tau.as<cd>(), N, 0, 1, c_->ws);
    zgemm_strided(st, OP_N, OP_N, N, N, N, mk(-2.0, 0.0), G.as<cd>(), N, 0, tau.as<cd>(), N, 0, ZERO, Rud, N, 0, 1);
    zadd_diag(st, N, ONE, nullptr, 0, Rud, N, 0, 1);
    zgemm_strided(st, OP_N, OP_N, N, N, N, mk(2.0, 0.0), wT.as<cd>(), N, 0, tau.as<cd>(), N, 0, ZERO, Tdd, N, 0, 1);
    // Rud of this step: the reflection REFMatrices[i + 1] of the intermediate-field pass
    zgeam(st, N, N, ONE, Rud, N, 0, ZERO, nullptr, 0, 0, REF_.as<cd>() + (long)i * NN, N, 0, 1);
  }
  IE_ = std::move(IE);
  IH_ = std::move(IH);
  ds_ = std::move(ds);
  stack_ = stack;
  imap_ = imap;
  VRC_HIP(hipStreamSynchronize(st));
  solved_ = true;
}

// The amplitude pass of RCWASolver::evaluateIntermediateField (src/rcwa/RCWASolver.cpp:1478-1500)
// on all input columns at once. Interface i (between stack entries index - 1 and index,
// index = nL - 1 - i) in scattering form: t11 = (E + H) / 2, t12 = (E - H) / 2,
// S11 = t11 - t12 t11^-1 t12, S12 = t12 t11^-1, S21 = -t11^-1 t12, S22 = t11^-1, applied
// as solves with t11 and S11; REFMatrices[i + 1] is the RTCM step's Rud. The first value,
// S12[0] t11 T, is t12 T (the same product).
void LegacySolver::intermediate_field(int ncols, const cd* input, cd* c_m, cd* c_p) {
  if (!solved_) throw ArgFailure("the scatter matrix has not been solved");
  const int nL = (int)stack_.size();
  if (nL < 3) throw ArgFailure("the stack has no inner layer");
  if (ncols < 1) throw ArgFailure("at least one input column expected");
  hipStream_t st = c_->stream;
  const int N = 2 * nx_ * ny_, K = ncols, nin = nL - 2;
  const long NN = (long)N * N, NK = (long)N * K;
  DevBuf vb = c_->ws.get(sizeof(cd) * NK);
  c_->pinned.upload(st, vb.ptr, input, sizeof(cd) * NK);
  const cd* v = vb.as<cd>();
  DevBuf outm = c_->ws.get(sizeof(cd) * NK * nin), outp = c_->ws.get(sizeof(cd) * NK * nin);
  cd* cm = outm.as<cd>();
  cd* cp = outp.as<cd>();
  DevBuf t11b = c_->ws.get(sizeof(cd) * NN), t12b = c_->ws.get(sizeof(cd) * NN), Xb = c_->ws.get(sizeof(cd) * NN),
         S11b = c_->ws.get(sizeof(cd) * NN), piv1 = c_->ws.get(sizeof(int) * N), piv2 = c_->ws.get(sizeof(int) * N),
         inf = c_->ws.get(sizeof(int) * 2), phi = c_->ws.get(sizeof(cd) * N), y = c_->ws.get(sizeof(cd) * NK),
         z = c_->ws.get(sizeof(cd) * NK), d = c_->ws.get(sizeof(cd) * NK);
  cd *t11 = t11b.as<cd>(), *t12 = t12b.as<cd>(), *X = Xb.as<cd>(), *S11 = S11b.as<cd>();
  const cd half = mk(0.5, 0.0), mhalf = mk(-0.5, 0.0), mone = mk(-1.0, 0.0);
  // t11 (LU in place), t12, X = t11^-1 t12, S11 = t11 - t12 X (LU in place) of interface i
  auto interface = [&](int i) {
    const int index = nL - 1 - i;
    const cd* E = IE_.as<cd>() + (long)imap_[index - 1] * NN;
    const cd* H = IH_.as<cd>() + (long)imap_[index - 1] * NN;
    zgeam(st, N, N, half, E, N, 0, half, H, N, 0, t11, N, 0, 1);
    zgeam(st, N, N, half, E, N, 0, mhalf, H, N, 0, t12, N, 0, 1);
    zgeam(st, N, N, ONE, t11, N, 0, ZERO, nullptr, 0, 0, S11, N, 0, 1);
    fill_int(st, inf.as<int>(), 0, 2);
    zgetrf(st, N, t11, N, 0, piv1.as<int>(), 0, inf.as<int>(), 1, c_->ws);
    zgetrs(st, N, N, t11, N, 0, piv1.as<int>(), 0, t12, N, 0, X, N, 0, 1, c_->ws);
    zgemm_strided(st, OP_N, OP_N, N, N, N, mone, t12, N, 0, X, N, 0, ONE, S11, N, 0, 1);
    zgetrf(st, N, S11, N, 0, piv2.as<int>(), 0, inf.as<int>() + 1, 1, c_->ws);
  };
  auto mm = [&](cd a, const cd* A, const cd* B, cd b, cd* C) {
    zgemm_strided(st, OP_N, OP_N, N, K, N, a, A, N, 0, B, N, 0, b, C, N, 0, 1);
  };
  // phase of stack entry `index` (RCWASolver uses exp(i gamma t), gamma = sqrt(eig) / k0)
  auto phase = [&](int index) {
    phases(st, N, gam_.as<cd>() + (long)stack_[index] * N, 0, ds_.as<double>() + index, phi.as<cd>(), 0, 1);
  };
  // one interface of the pass: forward amplitude a (in), backward am (in): temp = S11^-1 (am -
  // S12 a), out = S21 temp + S22 a = t11^-1 (a - t12 temp)
  auto through = [&](int i, const cd* a, const cd* am, cd* out) {
    interface(i);
    zgetrs(st, N, K, t11, N, 0, piv1.as<int>(), 0, a, N, 0, y.as<cd>(), N, 0, 1, c_->ws);  // t11^-1 a
    zgeam(st, N, K, ONE, am, N, 0, ZERO, nullptr, 0, 0, z.as<cd>(), N, 0, 1);
    mm(mone, t12, y.as<cd>(), ONE, z.as<cd>());                                          // am - S12 a
    zgetrs(st, N, K, S11, N, 0, piv2.as<int>(), 0, z.as<cd>(), N, 0, y.as<cd>(), N, 0, 1, c_->ws);  // temp
    zgeam(st, N, K, ONE, a, N, 0, ZERO, nullptr, 0, 0, z.as<cd>(), N, 0, 1);
    mm(mone, t12, y.as<cd>(), ONE, z.as<cd>());                                          // a - t12 temp
    zgetrs(st, N, K, t11, N, 0, piv1.as<int>(), 0, z.as<cd>(), N, 0, out, N, 0, 1, c_->ws);
  };
  // c_m[nin - 1] = t12_0 Tdd v (interface 0: the last inner layer and the last layer)
  {
    const int index = nL - 1;
    const cd* E = IE_.as<cd>() + (long)imap_[index - 1] * NN;
    const cd* H = IH_.as<cd>() + (long)imap_[index - 1] * NN;
    zgeam(st, N, N, half, E, N, 0, mhalf, H, N, 0, t12, N, 0, 1);
    mm(ONE, Tdd_.as<cd>(), v, ZERO, z.as<cd>());
    mm(ONE, t12, z.as<cd>(), ZERO, cm + (long)(nin - 1) * NK);
  }
  // first inner layer: forward amplitude from the incident v and reflected Rud v
  DevBuf Rv = c_->ws.get(sizeof(cd) * NK);
  mm(ONE, Rud_.as<cd>(), v, ZERO, Rv.as<cd>());
  through(nL - 2, v, Rv.as<cd>(), cp);
  phase(1);
  zdscale(st, N, K, phi.as<cd>(), 0, nullptr, 0, 0, cp, N, 0, d.as<cd>(), N, 0, 1);
  for (int i = nin - 2; i >= 0; --i) {
    cd* cmi = cm + (long)(nin - i - 2) * NK;
    cd* cpi = cp + (long)(nin - i - 1) * NK;
    mm(ONE, REF_.as<cd>() + (long)(i + 1) * NN, d.as<cd>(), ZERO, cmi);  // REFMatrices[i + 2]
    through(i + 1, d.as<cd>(), cmi, cpi);
    phase(nL - 1 - (i + 1));
    zdscale(st, N, K, phi.as<cd>(), 0, nullptr, 0, 0, cpi, N, 0, d.as<cd>(), N, 0, 1);
  }
  VRC_HIP(hipMemcpyAsync(c_m, cm, sizeof(cd) * NK * nin, hipMemcpyDeviceToHost, st));
  VRC_HIP(hipMemcpyAsync(c_p, cp, sizeof(cd) * NK * nin, hipMemcpyDeviceToHost, st));
  VRC_HIP(hipStreamSynchronize(st));
}

void LegacySolver::get(int which, int k, cd* host) const {
  if (!solved_) throw ArgFailure("the scatter matrix has not been solved");
  const int N = 2 * nx_ * ny_;
  const long NN = (long)N * N;
  const cd* src = nullptr;
  size_t cnt = (size_t)NN;
  if (which == 0) {
    src = Tdd_.as<cd>();
  } else if (which == 1) {
    src = Rud_.as<cd>();
  } else if (which >= 2 && which <= 4) {
    if (k < 0 || k >= (int)layers_.size()) throw ArgFailure("no such layer type");
    if (which == 2) src = W_.as<cd>() + k * NN;
    if (which == 3) src = V_.as<cd>() + k * NN;
    if (which == 4) {
      src = gam_.as<cd>() + (long)k * N;
      cnt = N;
    }
  } else {
    throw ArgFailure("unknown output");
  }
  VRC_HIP(hipMemcpyAsync(host, src, sizeof(cd) * cnt, hipMemcpyDeviceToHost, c_->stream));
  VRC_HIP(hipStreamSynchronize(c_->stream));
  if (which == 4) {  // Layer::gamma() is sqrt(eig) / k0 (src/rcwa/Layer.cpp:187)
    const double s = 1.0 / k0_;
    for (size_t i = 0; i < cnt; ++i) host[i] = mk(host[i].x * s, host[i].y * s);
  }
}

}  // namespace vrc
