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
