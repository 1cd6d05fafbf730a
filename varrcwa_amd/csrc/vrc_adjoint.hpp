// Vector-Jacobian products of the S-matrix chain and the batched stack tape.
#pragma once

#include <memory>
#include <vector>

#include "vrc_ops.hpp"

namespace vrc {

// eps cotangents of layers (periodic cell, non-decreasing coordinates) from the
// cotangents of fe11/fe22/fe33 (n x n, stride sfe). epsb: cell (i strip, j column) of
// layer l at offs[l] + i + j*ncy (the input layout of the layer's eps).
void assemble_fe_vjp(vrc_ctx& c, int nx, int ny, const std::vector<const HostLayer*>& layers, const cd* fe11b,
                     const cd* fe22b, const cd* fe33b, long sfe, cd* epsb, const std::vector<long>& offs);
// eps cotangents from the cotangents of periodic P and Q (N x N, stride sPQ).
void assemble_pq_vjp(vrc_ctx& c, const AssemblyBatch& a, const cd* Pb, const cd* Qb, long sPQ, cd* epsb,
                     const std::vector<long>& offs);
// P, Q cotangents of wave_eqn given those of W (in/out: the V path is added), V and
// gamma (in/out). Wi: W^{-1} (stride N*N) or null to compute it.
void wave_eqn_vjp(vrc_ctx& c, int N, int batch, const cd* P, const cd* Q, long sPQ, const cd* W, const cd* V,
                  const cd* gam, const cd* Wi, cd* Wb, const cd* Vb, cd* gb, cd* Pb, cd* Qb);
// W, V, gamma cotangents of layer_smatrix given those of Tuu (tb) and Rud (rb); the
// reference basis (W0, V0) is a constant. Wi_out (optional) receives W^{-1}.
void layer_smatrix_vjp(vrc_ctx& c, int N, int batch, const double* d_s, const cd* W, const cd* V, const cd* gam,
                       const cd* W0, const cd* V0, long sRef, const cd* tb, const cd* rb, cd* Wb, cd* Vb, cd* gb,
                       cd* Wi_out);
// Forward intermediates of the slab S-matrix kept for its reverse pass (N x N per slab;
// Li, E: the inverses of the two even/odd systems and their solutions, 2 per slab;
// x: the N phases exp(i gamma s)); Tuu and Rud are the slab S-matrix itself.
struct SlabSaved {
  DevBuf Wi, Vi, F0, G0, Li, E, Rud, Tuu, x;
};
void layer_smatrix_fwd_saved(vrc_ctx& c, int N, int batch, const double* d_s, const cd* W, const cd* V, const cd* gam,
                             const cd* W0, const cd* V0, long sRef, SlabSaved& sv);
void layer_smatrix_bwd_saved(vrc_ctx& c, int N, int batch, const double* d_s, SlabSaved& sv, const cd* tb,
                             const cd* rb, cd* Wb, cd* Vb, cd* gb);
// star_product: S0 is the cascade before the product, Sb the cotangent after it on
// entry and of S0 on return; slab cotangents are written with stride sb. saved
// (optional): the STAR_SAVE intermediates kept by star_product(..., save).
void star_product_vjp(vrc_ctx& c, int N, int batch, const SMat& S0, const cd* tdd, const cd* tuu, const cd* rud,
                      const cd* rdu, long ss, SMat& Sb, cd* tddb, cd* tuub, cd* rudb, cd* rdub, long sb,
                      const cd* saved = nullptr);
// project_left with constant end bases: S is the cascade before the projection.
void project_left_vjp(vrc_ctx& c, int N, int batch, const SMat& S, const LUPair& LU1, const cd* W0, const cd* V0,
                      long sRef, SMat& Sb);

// Everything the reverse pass of batch_stack_rcwa needs: copies of the geometry and
// the eigen-systems of every instance (bottom medium, top medium, slabs).
struct StackTape {
  vrc_ctx* c = nullptr;
  int nx = 0, ny = 0, L = 0, ninst = 0;
  std::vector<HostLayer> layers;  // [inst][L] slabs, then bottoms [inst], tops [inst]
  std::vector<double> lambda, thickness;  // [inst], [inst][L]
  DevBuf Wb, Vb, gb;  // bottoms: [inst] N x N, N x N, N
  DevBuf Wt, Vt, gt;  // tops
  DevBuf Ws, Vs, gs;  // slabs [inst][L]
  const HostLayer* slab(int i, int l) const { return &layers[(size_t)i * L + l]; }
  const HostLayer* bottom(int i) const { return &layers[(size_t)ninst * L + i]; }
  const HostLayer* top(int i) const { return &layers[(size_t)ninst * L + ninst + i]; }
  long cells() const;  // total eps entries of all slabs (length of the gradient)
};

std::unique_ptr<StackTape> make_stack_tape(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst);
// Gradient of a real objective of the stack S-matrices: Sb holds, per instance, the
// cotangents of [Tdd|Tuu|Rud|Rdu] (4 N x N, device, the layout of S_out); epsb
// receives, per instance and slab, the cotangent of every eps cell (input layout),
// concatenated in instance-then-slab order. chunk bounds the instances per pass.
void stack_tape_vjp(StackTape& t, const cd* Sb, cd* epsb, int chunk);

}  // namespace vrc
