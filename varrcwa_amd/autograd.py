"""torch.autograd binding of the batched stack solve.

``stack_smatrix(eps, problem)`` returns the S-matrices of many layer stacks as a
differentiable function of the permittivity of every slab cell; the backward pass
runs the adjoint kernels of ``libvarrcwa.so`` (vrc_batch_stack_forward with a tape,
then vrc_batch_stack_vjp), which keep each instance's eigen-systems from the forward
pass and recompute the slab S-matrices and the cascade prefixes.

The reference computes no derivatives on this path; the objective helper
``zeroth_order_transmission`` is the 0th-order power objective used for
metasurface inverse design (TE+TM, field-basis transmission between the (0,0)
harmonics of the end media).

``stack_smatrix_shape(theta, problem)`` is the same for geometric parameters (e.g. one
pillar radius per layer): ``ShapeProblem`` builds the stacks from theta and says how the
slabs' x coordinates move with it, and the backward pass chains the closed-form
coordinate tangents of P and Q into the reverse pass (vrc_batch_stack_vjp_shape).
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np

from .api import Layer, batch_stack_rcwa
from ._lib import Context


class StackProblem:
    """Geometry of a batch of stacks whose slab permittivities are the variables.

    instances: dicts with lam, layers [(x, y, eps)], thickness, bottom, top (as for
    ``batch_stack_rcwa``). The slab coordinates stay fixed; the eps values given here
    are the starting point returned by ``eps_vector``.
    """

    def __init__(self, ctx: Context, nx: int, ny: int, instances: Sequence[dict], chunk: int = 0):
        self.ctx, self.nx, self.ny, self.chunk = ctx, int(nx), int(ny), int(chunk)
        self.instances = []
        self.shapes: List[List[tuple]] = []
        for inst in instances:
            layers = [l if isinstance(l, Layer) else Layer(*l) for l in inst["layers"]]
            self.instances.append(dict(inst, layers=layers))
            self.shapes.append([l.eps.shape for l in layers])
        sizes = [int(np.prod(s)) for shp in self.shapes for s in shp]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)

    @property
    def cells(self) -> int:
        return int(self.offsets[-1])

    def eps_vector(self):
        """All slab cells as one complex128 vector (instance, slab, column-major cells)."""
        v = np.concatenate([l.eps.reshape(-1, order="F") for inst in self.instances for l in inst["layers"]])
        return self.ctx.torch.from_numpy(v.astype(np.complex128))

    def instances_with(self, eps: np.ndarray):
        out, k = [], 0
        for inst, shp in zip(self.instances, self.shapes):
            layers = []
            for l, s in zip(inst["layers"], shp):
                e = eps[self.offsets[k]:self.offsets[k + 1]].reshape(s, order="F")
                layers.append(Layer(l.x, l.y, e))
                k += 1
            out.append(dict(inst, layers=layers))
        return out


def _autograd_function():
    import torch

    class StackSMatrix(torch.autograd.Function):
        @staticmethod
        def forward(fctx, eps, problem):
            insts = problem.instances_with(eps.detach().cpu().numpy())
            out = batch_stack_rcwa(problem.ctx, problem.nx, problem.ny, insts, chunk=problem.chunk, want_bases=True,
                                   tape=True)
            fctx.tape = out["tape"]
            fctx.chunk = problem.chunk
            fctx.real_input = not eps.is_complex()
            fctx.mark_non_differentiable(out["W0"], out["W1"])
            return out["S"], out["W0"], out["W1"]

        @staticmethod
        def backward(fctx, gS, gW0, gW1):
            if gS is None:
                return None, None
            # the tape is read, not consumed: it stays with the graph node (released with the
            # graph), so a second backward with retain_graph=True sees the same eigen-systems
            g = fctx.tape.vjp(gS, chunk=fctx.chunk)
            return (g.real if fctx.real_input else g), None

    return StackSMatrix


_FN = None


def stack_smatrix(eps, problem: StackProblem):
    """S (ninst, 4, N, N) [Tdd, Tuu, Rud, Rdu] and the end bases W0, W1 (ninst, N, N),
    all in column-major storage (``.transpose(-1, -2)`` views them as matrices).
    eps: (cells,) tensor, real or complex; gradients flow to it through S."""
    global _FN
    if _FN is None:
        _FN = _autograd_function()
    return _FN.apply(eps, problem)


def zeroth_order_transmission(S, W0, W1, nx: int, ny: int):
    """sum_{p,q in {x,y}} |T_pq|^2 with T the field-basis transmission W1 Tuu W0^{-1}
    restricted to the (0,0) harmonic of both end media; one value per instance."""
    import torch
    n = nx * ny
    N = 2 * n
    c = (nx // 2) * ny + ny // 2
    idx = [c, n + c]
    Tuu = S[:, 1].transpose(-1, -2).contiguous()
    W0m = W0.transpose(-1, -2)
    W1m = W1.transpose(-1, -2)
    E = torch.zeros((N, 2), dtype=S.dtype, device=S.device)
    E[idx[0], 0] = 1.0
    E[idx[1], 1] = 1.0
    with torch.no_grad():  # the end bases are constants of the objective
        X = torch.linalg.solve(W0m, E.expand(S.shape[0], N, 2).contiguous())
        R = W1m[:, idx, :].contiguous()
    T = R @ (Tuu @ X)
    return (T.abs() ** 2).sum(dim=(-2, -1))


# ------------------------------------------------------------------ shape parameters
def pillar_layer(r: float, cell: float, nstrips: int, eps_pillar: complex, eps_bg: complex):
    """A circular pillar of radius r centred in a square cell, staircased on nstrips fixed
    y strips (uniform over the cell; even, so the strips pair up about the centre): strip j
    holds the pillar over |x - c| < w_j = sqrt(r^2 - (y_j - c)^2) (y_j its midpoint), none
    when |y_j - c| >= r. Returns (Layer, var_index, coeff): the interior x coordinates that
    move with r and dx/dr for each (-+ r / w at c -+ w)."""
    if nstrips < 2 or nstrips % 2:
        raise ValueError("nstrips must be even and >= 2")
    c = 0.5 * cell
    if not 0.0 < r < c:
        raise ValueError("the radius must lie in (0, cell / 2)")
    y = np.linspace(0.0, cell, nstrips + 1)
    dy = np.abs(0.5 * (y[1:] + y[:-1]) - c)
    w = np.where(dy < r, np.sqrt(np.maximum(r * r - dy * dy, 0.0)), 0.0)
    ws = np.unique(w[w > 0])  # ascending
    x = np.concatenate([[0.0], c - ws[::-1], c + ws, [cell]])
    xm = 0.5 * (x[1:] + x[:-1])
    eps = np.where(np.abs(xm[None, :] - c) < w[:, None], eps_pillar, eps_bg).astype(complex)
    U = ws.size
    var_index = list(range(1, 2 * U + 1))
    coeff = np.concatenate([-r / ws[::-1], r / ws]).reshape(-1, 1)
    return Layer(x, y, eps), var_index, coeff


class ShapeProblem:
    """Stacks whose slab geometry is a function of parameters theta.

    build(theta) -> (instances, shape), with instances as for ``batch_stack_rcwa`` and
    shape one entry per slab (instance-major, then layer): None, or (var_index, coeff,
    theta_index): coordinate var_index[k] moves by coeff[k, p] per unit of theta[
    theta_index[p]]. recompute / layer_block: the tape options of batch_stack_rcwa."""

    def __init__(self, ctx: Context, nx: int, ny: int, build, chunk: int = 0, recompute: bool = False,
                 layer_block: int = 0):
        self.ctx, self.nx, self.ny, self.build = ctx, int(nx), int(ny), build
        self.chunk, self.recompute, self.layer_block = int(chunk), bool(recompute), int(layer_block)


def _shape_function():
    import torch

    class StackSMatrixShape(torch.autograd.Function):
        @staticmethod
        def forward(fctx, theta, problem):
            insts, shape = problem.build(theta.detach().cpu().numpy())
            out = batch_stack_rcwa(problem.ctx, problem.nx, problem.ny, insts, chunk=problem.chunk, want_bases=True,
                                   tape=True, recompute=problem.recompute, layer_block=problem.layer_block)
            fctx.tape = out["tape"]
            fctx.shape = shape
            fctx.chunk = problem.chunk
            fctx.ntheta = theta.numel()
            fctx.dtype, fctx.device = theta.dtype, theta.device
            fctx.mark_non_differentiable(out["W0"], out["W1"])
            return out["S"], out["W0"], out["W1"]

        @staticmethod
        def backward(fctx, gS, gW0, gW1):
            if gS is None:
                return None, None
            per, _ = fctx.tape.vjp_shape(gS, [None if e is None else (e[0], e[1]) for e in fctx.shape],
                                         chunk=fctx.chunk)
            g = np.zeros(fctx.ntheta)
            for e, gs in zip(fctx.shape, per):
                if e is not None:
                    np.add.at(g, np.asarray(e[2], dtype=np.int64), gs)
            return torch.from_numpy(g).to(dtype=fctx.dtype, device=fctx.device), None

    return StackSMatrixShape


_FN_SHAPE = None


def stack_smatrix_shape(theta, problem: ShapeProblem):
    """S, W0, W1 as ``stack_smatrix``, differentiable with respect to the real shape
    parameters theta (1-D tensor) through ``problem.build``."""
    global _FN_SHAPE
    if _FN_SHAPE is None:
        _FN_SHAPE = _shape_function()
    return _FN_SHAPE.apply(theta, problem)
