// Drives the header-only C++ API (include/varrcwa.hpp) the way a caller of the
// reference's C++ classes does, on the two scenario geometries of the reference
// drivers, and writes the ten results as raw complex128 for the Python parity test.
//
//   cpp_api_check <metasurface|var12|user|usergpu|gds:FILE> <rcwa|redheffer|redheffermany|differential> <nh> <N or pe> <out.bin>
//
// Scenarios (geometry as data): metasurface cell of src/tests/test_metasurface_cpu.cpp:129-164
// (periodic, x[2] follows the pillar rule; "user"/"usergpu": the same cell through a
// caller-written LayerSampler subclass overriding sample() / sampleOnGPU()) and the PML waveguide taper of
// src/tests/test_cpu.cpp:128-176 (rule "var12" on x[3]). Output: int32 N, nEigVal, nDim,
// then Tdd, Tuu, Rud, Rdu, W0, V0, gam0, W1, V1, gam1 (column-major complex128).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/varrcwa.hpp"

using namespace varrcwa;

static constexpr double kPi = 3.1415926535897932;

// A caller's own sampler, written the way a reference user subclasses LayerSampler
// (src/core/LayerSampler.h:5-77): it assigns the simulation region, calls
// postInitialization(), and produces P, Q itself -- here by moving the metasurface pillar
// edge and asking a private layer description for the operator at that edge position. With
// on_gpu it overrides sampleOnGPU (device buffers on the context stream), else sample().
class PillarSampler : public LayerSampler {
 public:
  PillarSampler(int nx, int ny, bool on_gpu)
      : LayerSampler(1.55, nx, ny, false), on_gpu_(on_gpu) {
    postInitialization();
  }
  void sample(scalar z, MatrixXcs& P, MatrixXcs& Q) override { layer(z)->sample(0.0, P, Q); }
  void sampleOnGPU(scalar z, void* P, void* Q) override {
    if (!on_gpu_) return LayerSampler::sampleOnGPU(z, P, Q);
    layer(z)->sampleOnGPU(0.0, P, Q);
  }

 protected:
  void assignSimulationRegion() override {
    Lx_ = Ly_ = 0.65;
    l0_ = b0_ = 0.0;
    l1_ = b1_ = 0.1;
    r0_ = u0_ = 0.5;
    r1_ = u1_ = 0.65;
  }

 private:
  // the pillar cell with its right edge at the metasurface rule's position for z
  std::shared_ptr<VarLayerSampler> layer(scalar z) {
    const double xe = z <= 1.0 ? 0.5 : 0.2 * std::cos(0.5 * 3.1415926535897932 * (z - 1.0)) + 0.3;
    MatrixXcs eps(3, 3);
    for (long k = 0; k < eps.size(); ++k) eps(k) = scalex(1.0, 0.0);
    eps(1, 1) = scalex(3.48 * 3.48, 0.0);
    return std::make_shared<VarLayerSampler>(lambda_, nx_, ny_, VectorXs{0.0, 0.1, xe, 0.65},
                                             VectorXs{0.0, 0.1, 0.5, 0.65}, eps, false, ctx_);
  }
  bool on_gpu_;
};

static MatrixXcs grid(int rows, int cols, double bg) {
  MatrixXcs m(rows, cols);
  for (long k = 0; k < m.size(); ++k) m(k) = scalex(bg, 0.0);
  return m;
}

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s <metasurface|var12|gds:FILE> <rcwa|redheffer|redheffermany|differential> <nh> <N|pe> <out>\n",
                 argv[0]);
    return 2;
  }
  const std::string scen = argv[1], method = argv[2];
  const int nh = std::atoi(argv[3]);
  const double param = std::atof(argv[4]);
  try {
    double z0, z1;
    std::shared_ptr<LayerSampler> sampler;
    if (scen == "metasurface") {
      z0 = 0.0;
      z1 = 2.0;
      const double co = 3.48 * 3.48;
      MatrixXcs eps = grid(3, 3, 1.0);
      eps(1, 1) = scalex(co, 0.0);
      auto vs = std::make_shared<VarLayerSampler>(1.55, 2 * nh + 1, 2 * nh + 1, VectorXs{0.0, 0.1, 0.5, 0.65},
                                                  VectorXs{0.0, 0.1, 0.5, 0.65}, eps, false);
      vs->addXRule(2, [](scalar z) { return z <= 1.0 ? 0.5 : 0.2 * std::cos(0.5 * kPi * (z - 1.0)) + 0.3; });
      sampler = vs;
    } else if (scen == "var12") {
      z0 = -0.1;
      z1 = 10.1;
      MatrixXcs eps = grid(6, 7, 1.445 * 1.445);
      eps(3, 2) = scalex(3.48 * 3.48, 0.0);
      auto vs = std::make_shared<VarLayerSampler>(1.55, 2 * nh + 1, 2 * nh + 1,
                                                  VectorXs{0.0, 0.5, 1.5, 2.6, 3.7, 3.9, 5.0, 5.5},
                                                  VectorXs{0.0, 0.5, 1.5, 1.58, 1.8, 2.8, 3.3}, eps);
      vs->addXRule(3, [](scalar z) {
        if (z < 0.0) return 2.6;
        if (z > 10.0) return 3.7;
        return (1.0 - 0.1 * z) * 2.6 + 0.1 * z * 3.7;
      });
      sampler = vs;
    } else if (scen == "user" || scen == "usergpu") {
      // the metasurface scenario through a caller-written LayerSampler subclass
      z0 = 0.0;
      z1 = 2.0;
      sampler = std::make_shared<PillarSampler>(2 * nh + 1, 2 * nh + 1, scen == "usergpu");
      if (sampler->Lx() != 0.65 || sampler->Ly() != 0.65) {
        std::fprintf(stderr, "Lx()/Ly() do not report the assigned region\n");
        return 1;
      }
    } else if (scen.rfind("gds:", 0) == 0) {
      // src/tests/test_gds_cpu.cpp:129-145: the contour layout, ny = 2 nh - 1
      z0 = -1.63;
      z1 = 1.36;
      sampler = std::make_shared<GDSIISampler>(1.55, 2 * nh + 1, 2 * nh - 1, scen.substr(4),
                                               std::vector<scalar>{2.3647, 0.0, 0.0}, 0.5, 0.5, 1.0, 1.3);
    } else {
      std::fprintf(stderr, "unknown scenario %s\n", scen.c_str());
      return 2;
    }

    std::unique_ptr<detail::IntegratorBase> it;
    int n_used = 0, n_eig = 0;
    if (method == "rcwa") {
      auto r = std::make_unique<RCWAIntegrator>(z0, z1, sampler);
      r->compute((int)param);
      n_used = (int)param;
      it = std::move(r);
    } else if (method == "redheffer") {
      auto r = std::make_unique<RedhefferIntegrator>(z0, z1, sampler);
      r->compute(0, 1000000, param);
      n_used = r->N();
      n_eig = r->nEigVal();
      it = std::move(r);
    } else if (method == "redheffermany") {
      // two integrators of the same layer at once (RedhefferIntegrator::computeMany); the
      // second must agree with the first
      auto r = std::make_unique<RedhefferIntegrator>(z0, z1, sampler);
      RedhefferIntegrator twin(z0, z1, sampler);
      RedhefferIntegrator::computeMany({r.get(), &twin}, 0, 1000000, param);
      n_used = r->N();
      n_eig = r->nEigVal();
      if (twin.N() != n_used) {
        std::fprintf(stderr, "computeMany: N differs between equal integrators\n");
        return 1;
      }
      const MatrixXcs& a = r->Tuu();
      const MatrixXcs& b = twin.Tuu();
      double dmax = 0.0, amax = 0.0;
      for (long k = 0; k < a.size(); ++k) {
        dmax = std::fmax(dmax, std::abs(a(k) - b(k)));
        amax = std::fmax(amax, std::abs(a(k)));
      }
      if (!(dmax <= 1e-12 * amax)) {
        std::fprintf(stderr, "computeMany: Tuu differs between equal integrators (%g)\n", dmax / amax);
        return 1;
      }
      it = std::move(r);
    } else if (method == "differential") {
      auto r = std::make_unique<DifferentialIntegrator>(z0, z1, sampler);
      r->compute((int)param);
      n_used = (int)param;
      it = std::move(r);
    } else {
      std::fprintf(stderr, "unknown method %s\n", method.c_str());
      return 2;
    }

    FILE* f = std::fopen(argv[5], "wb");
    if (!f) {
      std::perror("open output");
      return 1;
    }
    const int head[3] = {n_used, n_eig, sampler->nDim()};
    std::fwrite(head, sizeof(int), 3, f);
    const MatrixXcs* outs[10] = {&it->Tdd(), &it->Tuu(), &it->Rud(), &it->Rdu(), &it->W0(),
                                 &it->V0(),  &it->gam0(), &it->W1(), &it->V1(), &it->gam1()};
    for (auto* m : outs) std::fwrite(m->data(), sizeof(scalex), (size_t)m->size(), f);
    std::fclose(f);
    std::printf("ok %s %s nh=%d N=%d nEig=%d\n", scen.c_str(), method.c_str(), nh, n_used, n_eig);
  } catch (const Error& e) {
    std::fprintf(stderr, "error (%d): %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
