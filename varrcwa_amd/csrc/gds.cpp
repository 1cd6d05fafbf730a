// GDSII front-end: the polygon subset of a GDSII stream file, as the reference reads it
// through gdstk::read_gds (src/gdstk/library.cpp:160-420), and the cross-section sampler
// GDSIISampler (src/core/GDSIISampler.cpp:17-146) built on it.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <limits>

#include "vrc_context.hpp"
#include "vrc_ops.hpp"

namespace vrc {

// GDSII 8-byte real: sign bit, 7-bit excess-64 base-16 exponent, 56-bit mantissa in [1/16, 1)
// (value = mantissa / 2^56 * 16^(exponent - 64)).
double gdsii_real(uint64_t r) {
  const int e = (int)((r >> 56) & 0x7F) - 64;
  const double m = (double)(r & 0x00FFFFFFFFFFFFFFULL) / 72057594037927936.0;  // 2^56
  const double v = std::ldexp(m, 4 * e);
  return (r >> 63) ? -v : v;
}

namespace {
uint16_t be16(const unsigned char* p) { return (uint16_t)((p[0] << 8) | p[1]); }
uint32_t be32(const unsigned char* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}
uint64_t be64(const unsigned char* p) { return ((uint64_t)be32(p) << 32) | be32(p + 4); }
std::string gds_string(const unsigned char* p, size_t n) {
  while (n > 0 && p[n - 1] == 0) --n;
  return std::string((const char*)p, n);
}
}  // namespace

GdsLibrary read_gds(const std::string& path, double unit) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw ArgFailure("cannot open GDSII file " + path);
  std::vector<unsigned char> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  GdsLibrary lib;
  double factor = 1.0;
  GdsCell* cell = nullptr;
  GdsPolygon* poly = nullptr;
  enum { NONE, POLY, OTHER } elem = NONE;
  size_t o = 0;
  while (o + 4 <= buf.size()) {
    const uint16_t len = be16(&buf[o]);
    if (len < 4) break;  // zero padding after ENDLIB
    if (o + len > buf.size()) throw ArgFailure("truncated GDSII record in " + path);
    const unsigned char rec = buf[o + 2];
    const unsigned char* d = &buf[o + 4];
    const size_t n = len - 4;
    switch (rec) {
      case 0x02:  // LIBNAME
        lib.name = gds_string(d, n);
        break;
      case 0x03: {  // UNITS: database unit in user units, in meters
        if (n < 16) throw ArgFailure("bad UNITS record");
        const double db_in_user = gdsii_real(be64(d)), db_in_m = gdsii_real(be64(d + 8));
        if (unit > 0) {
          factor = db_in_m / unit;
          lib.unit = unit;
        } else {
          factor = db_in_user;
          lib.unit = db_in_m / db_in_user;
        }
        lib.precision = db_in_m;
      } break;
      case 0x04:  // ENDLIB
        return lib;
      case 0x05:  // BGNSTR
        lib.cells.emplace_back();
        cell = &lib.cells.back();
        break;
      case 0x06:  // STRNAME
        if (cell) cell->name = gds_string(d, n);
        break;
      case 0x07:  // ENDSTR
        cell = nullptr;
        break;
      case 0x08:  // BOUNDARY
      case 0x2D:  // BOX
        elem = POLY;
        poly = nullptr;
        if (cell) {
          cell->polygons.emplace_back();
          poly = &cell->polygons.back();
        }
        break;
      case 0x09:  // PATH
      case 0x0A:  // SREF
      case 0x0B:  // AREF
      case 0x0C:  // TEXT
      case 0x15:  // NODE
        elem = OTHER;
        poly = nullptr;
        break;
      case 0x0D:  // LAYER
        if (poly && n >= 2) poly->layer = (int16_t)be16(d);
        break;
      case 0x0E:  // DATATYPE
      case 0x2E:  // BOXTYPE
        if (poly && n >= 2) poly->datatype = (int16_t)be16(d);
        break;
      case 0x10:  // XY
        if (elem == POLY && poly) {
          for (size_t k = 0; k + 8 <= n; k += 8) {
            poly->x.push_back(factor * (double)(int32_t)be32(d + k));
            poly->y.push_back(factor * (double)(int32_t)be32(d + k + 4));
          }
        }
        break;
      case 0x11:  // ENDEL: GDSII polygons repeat the first point at the end
        if (elem == POLY && poly && !poly->x.empty()) {
          poly->x.pop_back();
          poly->y.pop_back();
        }
        elem = NONE;
        poly = nullptr;
        break;
      default:
        break;
    }
    o += len;
  }
  return lib;
}

// ---------------------------------------------------------------- GDSIISampler
GdsSampler::GdsSampler(vrc_ctx* c, double lambda, int nx, int ny, const std::string& file, const double shift[3],
                       const double pml_space[8], cd eps_background, cd eps_device, double thickness)
    : Sampler(c, lambda, nx, ny, true), eps_bg_(eps_background), eps_dev_(eps_device), thickness_(thickness) {
  for (int i = 0; i < 3; ++i) shift_[i] = shift[i];
  const double lPML = pml_space[0], rPML = pml_space[1], lspace = pml_space[2], rspace = pml_space[3];
  const double bPML = pml_space[4], uPML = pml_space[5], bspace = pml_space[6], uspace = pml_space[7];
  GdsLibrary lib = read_gds(file, 1e-6);
  if (lib.cells.empty()) throw ArgFailure("GDSII file has no cell: " + file);
  polys_ = lib.cells[0].polygons;  // the reference reads the first cell only
  // GDSIISampler::assignSimulationRegion (src/core/GDSIISampler.cpp:49-92)
  minX_ = minZ_ = std::numeric_limits<double>::infinity();
  maxX_ = maxZ_ = -std::numeric_limits<double>::infinity();
  for (const auto& p : polys_)
    for (size_t i = 0; i < p.x.size(); ++i) {
      minX_ = std::min(minX_, p.x[i]);
      maxX_ = std::max(maxX_, p.x[i]);
      minZ_ = std::min(minZ_, p.y[i]);
      maxZ_ = std::max(maxZ_, p.y[i]);
    }
  minX_ += shift_[0];
  maxX_ += shift_[0];
  minZ_ += shift_[2];
  maxZ_ += shift_[2];
  const double b0 = shift_[1], b1 = b0 + bPML, u0 = b1 + bspace + thickness + uspace, u1 = u0 + uPML;
  const double l1 = minX_ - lspace, l0 = l1 - lPML, r0 = maxX_ + rspace, r1 = r0 + rPML;
  xr_ = {l0, l1, r0, r1};
  y_ = {b0, b1, b1 + bspace, b1 + bspace + thickness, u0, u1};
  post_init(xr_, {b0, b1, u0, u1});
}

// GDSIISampler::sampleLayout (src/core/GDSIISampler.cpp:94-146): the polygon edges crossing
// the line at z give the x coordinates (sorted); the device material fills every other
// x cell between them in the device strip of y.
HostLayer GdsSampler::layer_at(double z) const {
  constexpr double tol = 1e-12;
  std::vector<double> xs;
  for (const auto& p : polys_) {
    const size_t m = p.x.size();
    for (size_t i = 0; i < m; ++i) {
      const size_t j = (i + 1 == m) ? 0 : i + 1;
      const double y0 = p.y[i], y1 = p.y[j];
      if ((y0 <= z && z < y1) || (y1 <= z && z < y0)) {
        double h = y1 - y0;
        if (std::fabs(h) < tol) h += 2.0 * tol;  // bias to p0
        const double a = (z - y0) / h;
        xs.push_back((1 - a) * p.x[i] + a * p.x[j]);
      }
    }
  }
  std::sort(xs.begin(), xs.end());
  const int Nx = (int)xs.size() + 4;
  HostLayer L;
  L.x.resize(Nx);
  L.x[0] = xr_[0];
  L.x[1] = xr_[1];
  L.x[Nx - 2] = xr_[2];
  L.x[Nx - 1] = xr_[3];
  for (int i = 2; i <= Nx - 3; ++i) L.x[i] = xs[i - 2] + shift_[0];
  L.y = y_;
  const int ncy = (int)y_.size() - 1, ncx = Nx - 1;
  L.eps.assign((size_t)ncy * ncx, eps_bg_);
  const int device_strip = 2;  // yIndicator = {0, 0, 1, 0, 0}
  for (int j = 2; j <= Nx - 4; j += 2) L.eps[(size_t)device_strip + (size_t)j * ncy] = eps_dev_;
  return L;
}

}  // namespace vrc
