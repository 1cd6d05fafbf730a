"""Reverse pass in blocks of layers, with or without recomputing the slab eigen-systems, and
gradients with respect to shape parameters (pillar radii) of the slabs.

* the permittivity cotangents of one random S-matrix cotangent agree (1e-10 relative)
  between the whole-stack pass, a blocked pass (cascade checkpoints every K layers) and a
  blocked pass that solves the slab eigen-systems again; the recomputing tape is smaller;
* d J / d r for one pillar radius per layer (J: zeroth-order transmission, TE+TM, summed
  over wavelengths) against central differences of the GPU forward (h = 1e-6, bar 1e-5 of
  the largest component) and of the CPU oracle on two radii;
* the deep-stack shape: 200 layers with r(z) = 0.2 + 0.08 sin(2 pi z / 8 + phi) um on
  eps = 2.1, two wavelengths in 1.50-1.60 um, recomputing tape, blocks of 16 layers.
"""
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as O
from tests import geometries as G
from varrcwa_amd import api
from varrcwa_amd.autograd import ShapeProblem, pillar_layer, stack_smatrix_shape, zeroth_order_transmission

pytestmark = pytest.mark.gpu

TOL_GRAD = 1e-5


def pixel_stack(rng, nlayers, lam):
    cell = 0.8
    x = np.linspace(0.0, cell, 5)
    layers = [(x, x, (1.0 + 11.25 * rng.random((4, 4))).astype(complex)) for _ in range(nlayers)]
    return dict(lam=lam, layers=[api.Layer(*l) for l in layers], thickness=np.full(nlayers, 0.12),
                bottom=api.Layer(*G.homogeneous(cell, 2.1)), top=api.Layer(*G.homogeneous(cell, 1.0)))


def test_blocked_and_recomputed_reverse_pass_agree(ctx):
    torch = ctx.torch
    rng = np.random.default_rng(5)
    insts = [pixel_stack(rng, 5, lam) for lam in (1.0, 1.3)]
    nh = 3
    grads, sizes = [], []
    Sbar = None
    for recompute, K in [(False, 5), (False, 2), (True, 2), (True, 1)]:
        out = api.batch_stack_rcwa(ctx, nh, nh, insts, tape=True, recompute=recompute, layer_block=K)
        if Sbar is None:
            g = torch.Generator().manual_seed(3)
            Sbar = torch.complex(torch.randn(out["S"].shape, generator=g, dtype=torch.float64),
                                 torch.randn(out["S"].shape, generator=g, dtype=torch.float64)).cuda()
        tape = out["tape"]
        sizes.append(tape.bytes())
        grads.append(tape.vjp(Sbar).cpu().numpy())
        tape.close()
    ref = grads[0]
    for g in grads[1:]:
        assert np.abs(g - ref).max() <= 1e-10 * np.abs(ref).max(), np.abs(g - ref).max() / np.abs(ref).max()
    print("tape bytes (keep K=5, keep K=2, recompute K=2, recompute K=1):", sizes)
    assert sizes[2] < sizes[1]


def pillar_problem(ctx, nh, radii, lams, th, nstrips=16, recompute=False, layer_block=0):
    cell = 1.0
    bottom = api.Layer(*G.homogeneous(cell, 2.1))
    top = api.Layer(*G.homogeneous(cell, 1.0))

    def build(theta):
        insts, shape = [], []
        for lam in lams:
            layers = []
            for l, r in enumerate(theta):
                L, vi, cf = pillar_layer(float(r), cell, nstrips, 12.25, 1.0)
                layers.append(L)
                shape.append((vi, cf, [l]) if vi else None)
            insts.append(dict(lam=lam, layers=layers, thickness=np.full(len(theta), th), bottom=bottom, top=top))
        return insts, shape

    return ShapeProblem(ctx, nh, nh, build, recompute=recompute, layer_block=layer_block), build


def objective(problem, theta):
    S, W0, W1 = stack_smatrix_shape(theta, problem)
    return zeroth_order_transmission(S, W0, W1, problem.nx, problem.ny).sum()


def away_from_kinks(radii, nstrips):
    """Radii at least 1e-4 from a strip midpoint distance (the staircase changes there)."""
    dy = np.abs((np.arange(nstrips) + 0.5) / nstrips - 0.5)
    return all(np.abs(dy - r).min() > 1e-4 for r in radii)


def fd_gpu(problem, theta0, k, h=1e-6):
    torch = problem.ctx.torch
    vals = []
    with torch.no_grad():
        for sgn in (1, -1):
            t = theta0.copy()
            t[k] += sgn * h
            vals.append(float(objective(problem, torch.from_numpy(t))))
    return (vals[0] - vals[1]) / (2 * h)


def test_pillar_radius_gradient(ctx):
    torch = ctx.torch
    rng = np.random.default_rng(8)
    radii = rng.uniform(0.13, 0.27, 6)
    assert away_from_kinks(radii, 16)
    lams = [1.52, 1.58]
    prob, build = pillar_problem(ctx, 5, radii, lams, 0.15)
    theta = torch.from_numpy(radii.copy()).requires_grad_(True)
    J = objective(prob, theta)
    J.backward()
    g = theta.grad.numpy()
    fd = np.array([fd_gpu(prob, radii, k) for k in range(radii.size)])
    err = np.abs(g - fd).max() / np.abs(fd).max()
    print(f"J={float(J.detach()):.10f} grad {g} fd {fd} rel {err:.2e}", flush=True)
    assert err < TOL_GRAD
    # the CPU oracle: J and central differences on two radii (3 x 3 harmonics)
    prob3, build3 = pillar_problem(ctx, 3, radii, lams, 0.15)
    t3 = torch.from_numpy(radii.copy()).requires_grad_(True)
    J3 = objective(prob3, t3)
    J3.backward()

    def oracle_J(rr):
        insts, _ = build3(rr)
        tot = 0.0
        for inst in insts:
            lay = [(l.x, l.y, l.eps) for l in inst["layers"]]
            ze = np.concatenate([[0.0], np.cumsum(inst["thickness"])])
            s = O.StackLayerSampler(inst["lam"], 3, 3, lay, ze, (inst["bottom"].x, inst["bottom"].y,
                                                                  inst["bottom"].eps),
                                    (inst["top"].x, inst["top"].y, inst["top"].eps))
            r = O.rcwa_integrate(s, ze[0], ze[-1], len(lay), z_edges=ze)
            c = 4
            idx = [c, 9 + c]
            T = r.W1 @ r.Tuu @ np.linalg.inv(r.W0)
            tot += float((np.abs(T[np.ix_(idx, idx)]) ** 2).sum())
        return tot

    assert abs(oracle_J(radii) - float(J3.detach())) < 1e-10 * abs(oracle_J(radii))
    for k in (1, 4):
        rp, rm = radii.copy(), radii.copy()
        rp[k] += 1e-6
        rm[k] -= 1e-6
        fdo = (oracle_J(rp) - oracle_J(rm)) / 2e-6
        assert abs(t3.grad.numpy()[k] - fdo) <= TOL_GRAD * np.abs(t3.grad.numpy()).max(), (k, t3.grad[k], fdo)


def test_deep_stack_radius_gradient_recomputed(ctx):
    """200 layers of 0.04 um, 3 x 3 harmonics, two wavelengths: radii gradient through a
    recomputing tape in blocks of 16 layers; central differences on four radii."""
    torch = ctx.torch
    L = 200
    phi = float(np.random.default_rng(5).uniform(0, 2 * math.pi))
    z = (np.arange(L) + 0.5) * 8.0 / L
    radii = 0.2 + 0.08 * np.sin(2 * math.pi * z / 8.0 + phi)
    nstrips = 16
    dy = np.abs((np.arange(nstrips) + 0.5) / nstrips - 0.5)
    check = [k for k in (10, 70, 130, 190) if np.abs(dy - radii[k]).min() > 1e-4]
    assert len(check) >= 3
    prob, _ = pillar_problem(ctx, 3, radii, [1.50, 1.60], 0.04, nstrips, recompute=True, layer_block=16)
    theta = torch.from_numpy(radii.copy()).requires_grad_(True)
    J = objective(prob, theta)
    J.backward()
    g = theta.grad.numpy()
    assert np.isfinite(g).all() and np.abs(g).max() > 0
    for k in check:
        fd = fd_gpu(prob, radii, k)
        print(f"radius {k}: adjoint {g[k]:.10e} fd {fd:.10e}", flush=True)
        assert abs(g[k] - fd) <= TOL_GRAD * np.abs(g).max(), (k, g[k], fd)
    # the same gradient from a tape that keeps every eigen-system
    prob_k, _ = pillar_problem(ctx, 3, radii, [1.50, 1.60], 0.04, nstrips, recompute=False, layer_block=16)
    tk = torch.from_numpy(radii.copy()).requires_grad_(True)
    objective(prob_k, tk).backward()
    assert np.abs(tk.grad.numpy() - g).max() <= 1e-10 * np.abs(g).max()
