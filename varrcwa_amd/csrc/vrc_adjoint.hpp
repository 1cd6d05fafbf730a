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

// Everything the reverse pass of batch_stack_rcwa needs: copies of the geometry, the
// end media's eigen-systems, the cascade before the end projection, the cascade prefixes
// at the start of every block of K layers (checkpoints), and, unless the slabs'
// eigen-systems are recomputed in the reverse pass, those eigen-systems too.
struct StackTape {
  vrc_ctx* c = nullptr;
  int nx = 0, ny = 0, L = 0, ninst = 0;
  bool recompute = false;  // slab eigen-systems recomputed block by block in the reverse pass
  int K = 0;               // layers per block of the reverse pass
  std::vector<HostLayer> layers;  // [inst][L] slabs, then bottoms [inst], tops [inst]
  std::vector<double> lambda, thickness;  // [inst], [inst][L]
  DevBuf Wb, Vb, gb;  // bottoms: [inst] N x N, N x N, N
  DevBuf Wt, Vt, gt;  // tops
  DevBuf Ws, Vs, gs;  // slabs [inst][L] (not kept with recompute)
  DevBuf ckpt;        // [inst][block 1 .. nblocks-1] prefix before the block's first layer, 4 N x N
  DevBuf Sfin;        // [inst] cascade of all slabs before the projection, 4 N x N
  int nblocks() const { return K > 0 ? (L + K - 1) / K : 0; }
  const HostLayer* slab(int i, int l) const { return &layers[(size_t)i * L + l]; }
  const HostLayer* bottom(int i) const { return &layers[(size_t)ninst * L + i]; }
  const HostLayer* top(int i) const { return &layers[(size_t)ninst * L + ninst + i]; }
  long cells() const;  // total eps entries of all slabs (length of the gradient)
  size_t device_bytes() const;  // device memory the tape holds
};

// recompute: keep no slab eigen-systems (they are recomputed in the reverse pass);
// K: layers per block of the reverse pass (<= 0: min(L, 16)).
// Shape parameters of the slabs (a Parameterization, src/rcwa/Parameterization.h:5-25, per
// slab): slab s = instance * L + layer has npara[s] parameters (0: none); entry k of its
// list (var_index[vofs[s] + k], k < nvar[s]) moves that interior x coordinate by
// coeff[cofs[s] + k * npara[s] + p] per unit of parameter p. The gradient of parameter p
// of slab s goes to position pofs[s] + p.
struct ShapeParams {
  std::vector<int> npara, nvar;
  std::vector<long> pofs, vofs, cofs;
  std::vector<int> var_index;
  std::vector<double> coeff;
};
// out[p] += Re sum_i conj(A[i]) B[p * sB + i] for p < npara (device; deterministic order).
void dot_re_accumulate(vrc_ctx& c, long count, const cd* A, const cd* B, long sB, int npara, double* out);

std::unique_ptr<StackTape> make_stack_tape(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& inst,
                                           bool recompute = false, int K = 0);
// Gradient of a real objective of the stack S-matrices: Sb holds, per instance, the
// cotangents of [Tdd|Tuu|Rud|Rdu] (4 N x N, device, the layout of S_out); epsb
// receives, per instance and slab, the cotangent of every eps cell (input layout),
// concatenated in instance-then-slab order. chunk bounds the instances per pass.
// shape (optional): with shape_bar (device, one double per parameter, accumulated into)
// the gradient with respect to the slabs' shape parameters is formed as well; epsb may
// then be null.
void stack_tape_vjp(StackTape& t, const cd* Sb, cd* epsb, int chunk, const ShapeParams* shape = nullptr,
                    double* shape_bar = nullptr);

}  // namespace vrc
