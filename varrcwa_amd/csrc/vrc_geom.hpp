// Layer geometry (host form) and the device view used by the Fourier assembly.
#pragma once

#include <vector>

#include "vrc_common.hpp"

namespace vrc {

// A layer in the reference's Layer(x, y, eps) form (src/rcwa/Layer.cpp:18-34):
// x: nCx coordinates, y: nCy coordinates, eps: (nCy-1) x (nCx-1) column-major
// (row i = y strip [y_i, y_{i+1}], column j = x cell [x_j, x_{j+1}]).
struct HostLayer {
  std::vector<double> x, y;
  std::vector<cd> eps;
  int ncx() const { return (int)x.size() - 1; }
  int ncy() const { return (int)y.size() - 1; }
  cd at(int i, int j) const { return eps[(size_t)i + (size_t)j * ncy()]; }
};

constexpr int FOURIER_FFT_MAX = 2048;  // longest grid axis transformed by FFT (1.5 x that in LDS: 48 KB)

// Device arrays for a batch of layers (capacities MX cells, MY strips per layer).
struct FourierDev {
  int nx, ny;
  long MX, MY;
  const int* ncx;
  const int* ncy;
  const double* Lx;
  const double* Ly;
  const double* xu;  // [l][MX+1] coordinates in the given order (Fourier modes X and XY)
  const double* xs;  // [l][MX+1] sorted coordinates (mode Y intervals)
  const double* ys;  // [l][MY+1]
  const cd* val;     // [l][MY][MX] painter-order (refilled) permittivity, sorted cell order
  // log2 of the cell count when the layer is a uniform grid of a power-of-two number of
  // cells along x (fftx) / y (ffty), else 0: those 1D tables come from an FFT
  const int* fftx;
  const int* ffty;
  int fft_max;  // largest FFT length of the batch (0: none)
  cd *FX, *FXi, *FY1, *FYi, *FX1, *tFFx, *tFFy, *F2, *G11, *G22;
};

// Coefficient tables (FX, FXi, FY1, FYi, FX1) and inverse Toeplitz blocks (tFFx, tFFy).
void fourier_tables(hipStream_t s, const FourierDev& g, int nlayers);
// fourier_tables + contraction + expansion into fe11, fe22, fe33 (n x n, stride sfe).
void fourier_assemble(hipStream_t s, const FourierDev& g, int nlayers, cd* fe11, cd* fe22, cd* fe33, long sfe);
// Workspace of the shape tangents of one layer (npara parameters; MX cells, MY strips).
struct ShapeWork {
  cd *dFX, *dFXi, *dFX1;  // [p][MY][2nx-1], [p][MY][2nx-1], [p][MX][2nx-1]
  cd *D, *Y, *dtFF;       // [p][MY][nx][nx]
  cd *F2, *G11, *G22;     // one layer's contraction tables
};
// Derivatives of fe11, fe22, fe33 (n x n each, stride sfe per parameter) of layer 0 of g
// (after fourier_tables) with respect to npara parameters of its x coordinates; cfs:
// device [(MX+1) x npara], the coefficient of each sorted coordinate position.
void fourier_tangents(hipStream_t s, const FourierDev& g, int npara, const double* cfs, const ShapeWork& w,
                      cd* dfe11, cd* dfe22, cd* dfe33, long sfe);

}  // namespace vrc
