"""Adjoint (reverse-mode) kernels against central finite differences of the forward
GPU path, stage by stage and end to end, plus the end-to-end gradient against finite
differences of the CPU oracle.

Every stage test checks a directional derivative: for a random real objective
phi(X) = Re<C, F(X)> and a random complex direction D, the adjoint gives
Re<F^*(C), D>, the forward gives (phi(X + hD) - phi(X - hD)) / 2h. Tolerances:
1e-6 relative on the stages (h chosen so truncation and rounding stay below 1e-8),
1e-5 relative on the permittivity gradients: central differences at h = 1e-6 on the
small stack (every cell), directional fourth-order differences on the 64x64-pixel
metasurface, whose rounding noise floor rules out h = 1e-6 there (see that test).
"""
import numpy as np
import pytest

from oracle import rcwa_oracle as O
from tests import geometries as G
from varrcwa_amd import api
from varrcwa_amd.autograd import StackProblem, stack_smatrix, zeroth_order_transmission

pytestmark = pytest.mark.gpu

TOL_STAGE = 1e-6
TOL_GRAD = 1e-5


def crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def inner(a, b):
    """Re sum(conj(a) * b) of two tensors/arrays."""
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    b = b.detach().cpu().numpy() if hasattr(b, "detach") else np.asarray(b)
    return float(np.real(np.vdot(a.ravel(), b.ravel())))


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def check(ad, fd, tol):
    err = abs(ad - fd) / max(abs(fd), 1e-300)
    print(f"adjoint {ad:.12e}  fd {fd:.12e}  rel {err:.2e}", flush=True)
    assert err < tol, (ad, fd, err)


def pixel_layer(rng, cell, nxp, nyp, lossy=True):
    x = np.linspace(0.0, cell, nxp + 1)
    y = np.linspace(0.0, cell, nyp + 1)
    eps = 1.0 + 11.25 * rng.random((nyp, nxp)) + (0.3j * rng.random((nyp, nxp)) if lossy else 0.0)
    return x, y, eps.astype(complex)


def flat(layers):
    return np.concatenate([np.asarray(l[2]).reshape(-1, order="F") for l in layers])


def unflat(layers, v):
    out, o = [], 0
    for x, y, e in layers:
        k = e.size
        out.append((x, y, v[o:o + k].reshape(e.shape, order="F")))
        o += k
    return out


def test_toeplitz_vjp(ctx):
    torch = ctx.torch
    rng = np.random.default_rng(11)
    layers = [pixel_layer(rng, 0.8, 6, 5), G.random_rect_layer(rng, 0.6), G.homogeneous(0.7, 2.1 + 0.1j)]
    nx, ny = 5, 3
    fe = api.assemble_toeplitz(ctx, nx, ny, layers)
    C = dev(torch, crandn(rng, *fe.shape))
    eb = api.assemble_toeplitz_vjp(ctx, nx, ny, layers, C)
    e0 = flat(layers)
    D = crandn(rng, e0.size)
    h = 1e-5

    def phi(t):
        return inner(C, api.assemble_toeplitz(ctx, nx, ny, unflat(layers, e0 + t * D)))

    check(inner(eb, D), (phi(h) - phi(-h)) / (2 * h), TOL_STAGE)


def test_pq_vjp(ctx):
    torch = ctx.torch
    rng = np.random.default_rng(12)
    layers = [pixel_layer(rng, 0.8, 5, 4), pixel_layer(rng, 0.8, 3, 7)]
    region = G.homogeneous(0.8, 1.0)
    nx, ny, lam = 3, 5, 1.1
    P, Q = api.assemble_pq(ctx, nx, ny, layers, lam, False, region)
    Cp, Cq = dev(torch, crandn(rng, *P.shape)), dev(torch, crandn(rng, *Q.shape))
    eb = api.assemble_pq_vjp(ctx, nx, ny, layers, lam, region, Cp, Cq)
    e0 = flat(layers)
    D = crandn(rng, e0.size)
    h = 1e-5

    def phi(t):
        P, Q = api.assemble_pq(ctx, nx, ny, unflat(layers, e0 + t * D), lam, False, region)
        return inner(Cp, P) + inner(Cq, Q)

    check(inner(eb, D), (phi(h) - phi(-h)) / (2 * h), TOL_STAGE)


def _well_conditioned(rng, B, N, scale=0.3):
    return np.eye(N)[None] + scale * crandn(rng, B, N, N) / np.sqrt(N)


def test_layer_smatrix_vjp(ctx):
    torch = ctx.torch
    rng = np.random.default_rng(13)
    B, N = 3, 12
    W = _well_conditioned(rng, B, N)
    V = _well_conditioned(rng, B, N)
    g = rng.uniform(0.5, 3.0, (B, N)) + 1j * rng.uniform(0.0, 1.0, (B, N))
    g[:, :3] = 0.2 + 1j * rng.uniform(3.0, 8.0, (B, 3))  # evanescent modes
    W0 = _well_conditioned(rng, 1, N)[0]
    V0 = _well_conditioned(rng, 1, N)[0]
    lam, dz = np.array([1.0, 1.3, 0.9]), np.array([0.3, 0.15, 0.5])
    tW, tV, tg, tW0, tV0 = (dev(torch, a) for a in (W, V, g, W0, V0))
    Ct, Cr = dev(torch, crandn(rng, B, N, N)), dev(torch, crandn(rng, B, N, N))
    Wb, Vb, gb = api.layer_smatrix_vjp(ctx, lam, dz, tW, tV, tg, tW0, tV0, Ct, Cr)
    dW, dV, dg = crandn(rng, B, N, N), crandn(rng, B, N, N), crandn(rng, B, N)
    h = 1e-6

    def phi(t):
        Tuu, Rud = api.layer_smatrix(ctx, lam, dz, dev(torch, W + t * dW), dev(torch, V + t * dV),
                                     dev(torch, g + t * dg), tW0, tV0)
        return inner(Ct, Tuu) + inner(Cr, Rud)

    ad = inner(Wb, dW) + inner(Vb, dV) + inner(gb, dg)
    check(ad, (phi(h) - phi(-h)) / (2 * h), TOL_STAGE)


def test_eig_and_slab_vjp(ctx):
    """P, Q -> eig -> slab S: gauge-invariant composite, so the eigen adjoint applies."""
    torch = ctx.torch
    rng = np.random.default_rng(14)
    nx = ny = 3
    lam, dz = 1.05, 0.12
    layers = [pixel_layer(rng, 0.8, 4, 4), pixel_layer(rng, 0.8, 5, 3)]
    region = G.homogeneous(0.8, 1.0)
    P, Q = api.assemble_pq(ctx, nx, ny, layers, lam, False, region)
    P0, Q0 = api.assemble_pq(ctx, nx, ny, [G.homogeneous(0.8, 2.1)], lam, False, region)
    W0, V0, _, _ = api.eig_pq(ctx, P0, Q0)
    B, N, _ = P.shape

    def slab(P, Q):
        W, V, g, info = api.eig_pq(ctx, P, Q)
        assert int(info.abs().sum()) == 0
        return (W, V, g) + api.layer_smatrix(ctx, [lam] * B, [dz] * B, W, V, g, W0[0], V0[0])

    W, V, g, Tuu, Rud = slab(P, Q)
    Ct, Cr = dev(torch, crandn(rng, B, N, N)), dev(torch, crandn(rng, B, N, N))
    Wb, Vb, gb = api.layer_smatrix_vjp(ctx, [lam] * B, [dz] * B, W, V, g, W0[0], V0[0], Ct, Cr)
    Pb, Qb = api.eig_pq_vjp(ctx, P, Q, W, V, g, Wb, Vb, gb)
    Pn, Qn = P.cpu().numpy(), Q.cpu().numpy()
    dP = crandn(rng, *Pn.shape) * np.abs(Pn).max()
    dQ = crandn(rng, *Qn.shape) * np.abs(Qn).max()
    h = 1e-7

    def phi(t):
        _, _, _, Tu, Ru = slab(dev(torch, Pn + t * dP), dev(torch, Qn + t * dQ))
        return inner(Ct, Tu) + inner(Cr, Ru)

    check(inner(Pb, dP) + inner(Qb, dQ), (phi(h) - phi(-h)) / (2 * h), TOL_STAGE)


def _random_S(rng, B, N):
    return np.stack([0.5 * _well_conditioned(rng, B, N, 0.5), 0.5 * _well_conditioned(rng, B, N, 0.5),
                     0.2 * crandn(rng, B, N, N) / np.sqrt(N), 0.2 * crandn(rng, B, N, N) / np.sqrt(N)])


def test_star_product_vjp(ctx):
    torch = ctx.torch
    rng = np.random.default_rng(15)
    B, N = 2, 10
    S0, s = _random_S(rng, B, N), _random_S(rng, B, N)
    C = crandn(rng, 4, B, N, N)
    S0b, sb = api.star_product_vjp(ctx, dev(torch, S0), dev(torch, s), dev(torch, C))
    dS, ds = crandn(rng, *S0.shape), crandn(rng, *s.shape)
    h = 1e-6

    def phi(t):
        S = dev(torch, S0 + t * dS)
        api.star_product(ctx, S, dev(torch, s + t * ds))
        return inner(C, S)

    check(inner(S0b, dS) + inner(sb, ds), (phi(h) - phi(-h)) / (2 * h), TOL_STAGE)


def test_project_left_vjp(ctx):
    torch = ctx.torch
    rng = np.random.default_rng(16)
    B, N = 2, 10
    S = _random_S(rng, B, N)
    W1, V1, W0, V0 = (dev(torch, _well_conditioned(rng, B, N)) for _ in range(4))
    C = crandn(rng, 4, B, N, N)
    Sb = api.project_left_vjp(ctx, dev(torch, S), W1, V1, W0, V0, dev(torch, C))
    dS = crandn(rng, *S.shape)
    h = 1e-6

    def phi(t):
        X = dev(torch, S + t * dS)
        api.project_left(ctx, X, W1, V1, W0, V0)
        return inner(C, X)

    check(inner(Sb, dS), (phi(h) - phi(-h)) / (2 * h), TOL_STAGE)


# ---------------------------------------------------------------------------------- end to end
def _pixel_stack(rng, nlayers, nxp, cell, lam, th):
    layers = []
    for _ in range(nlayers):
        x, y, e = pixel_layer(rng, cell, nxp, nxp, lossy=False)
        layers.append((x, y, e))
    air = G.homogeneous(cell, 1.0)
    return dict(lam=lam, layers=layers, thickness=np.full(nlayers, th), bottom=air, top=air)


def _objective(ctx, problem, eps):
    S, W0, W1 = stack_smatrix(eps, problem)
    return zeroth_order_transmission(S, W0, W1, problem.nx, problem.ny).sum()


def _oracle_objective(inst, nx, ny):
    ze = np.concatenate([[0.0], np.cumsum(inst["thickness"])])
    s = O.StackLayerSampler(inst["lam"], nx, ny, inst["layers"], ze, inst["bottom"], inst["top"])
    r = O.rcwa_integrate(s, ze[0], ze[-1], len(inst["layers"]), z_edges=ze)
    n = nx * ny
    c = (nx // 2) * ny + ny // 2
    idx = [c, n + c]
    T = r.W1 @ r.Tuu @ np.linalg.inv(r.W0)
    return float((np.abs(T[np.ix_(idx, idx)]) ** 2).sum())


def test_stack_gradient_small(ctx):
    """Two instances x three 4x4-pixel slabs, 3x3 harmonics: autograd gradient vs
    finite differences of the GPU forward (all cells) and of the CPU oracle (4 cells)."""
    torch = ctx.torch
    rng = np.random.default_rng(21)
    nx = ny = 3
    insts = [_pixel_stack(rng, 3, 4, 0.8, lam, 0.15) for lam in (1.0, 1.2)]
    prob = StackProblem(ctx, nx, ny, insts)
    eps = prob.eps_vector().real.clone().to("cuda").requires_grad_(True)
    J = _objective(ctx, prob, eps)
    J.backward()
    g = eps.grad.cpu().numpy()
    e0 = eps.detach().cpu().numpy()
    h = 1e-6
    fd = np.empty_like(g)
    with torch.no_grad():
        for k in range(e0.size):
            ep, em = e0.copy(), e0.copy()
            ep[k] += h
            em[k] -= h
            jp = float(_objective(ctx, prob, torch.from_numpy(ep).to("cuda")))
            jm = float(_objective(ctx, prob, torch.from_numpy(em).to("cuda")))
            fd[k] = (jp - jm) / (2 * h)
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    print(f"J={float(J.detach()):.10f} |g|={np.linalg.norm(g):.3e} rel err vs GPU FD {err:.2e}", flush=True)
    assert err < TOL_GRAD
    # CPU oracle: J itself and finite differences on 4 cells (instance 0, slab 0)
    j_or = sum(_oracle_objective(i, nx, ny) for i in insts)
    assert abs(j_or - float(J.detach())) < 1e-10 * abs(j_or)
    for k in rng.choice(16, 4, replace=False):
        vals = []
        for sgn in (1, -1):
            e = e0.copy()
            e[k] += sgn * h
            ii = prob.instances_with(e.astype(complex))
            vals.append(sum(_oracle_objective(dict(i, layers=[(l.x, l.y, l.eps) for l in i["layers"]]), nx, ny)
                            for i in ii))
        fdo = (vals[0] - vals[1]) / (2 * h)
        assert abs(g[k] - fdo) <= TOL_GRAD * np.abs(fd).max(), (k, g[k], fdo)


def test_metasurface_inverse_design_gradient(ctx):
    """Freeform metasurface: 5 layers of 64x64 pixels, eps = 1 + 11.25 sigmoid(z),
    z ~ N(0,1) (seed 42), period 0.8 um, lambda 1.0 um, 11x11 harmonics, 0.15 um layers;
    objective: 0th-order transmitted power (TE+TM). Gradient w.r.t. all 20480 pixels by
    the adjoint.

    This objective carries rounding noise on any implementation: over six pixels,
    J(eps + 1e-12 e_k) - J(eps) reaches 1.1e-10 with LAPACK zgeev + the reference formulas
    (the oracle), 2.0e-10 with these kernels' P, Q fed to the oracle, and 6.7e-12 with these
    kernels throughout (MI355X); P and Q agree with the oracle's to 1.4e-15. A central
    difference at h = 1e-6 is therefore noise-dominated (relative error ~1e-3 on the 32
    pixels with these kernels, several times more with LAPACK's noise), and is only
    printed. Checked instead:
      - the derivative along the unit gradient direction (= |g| if g is right) by the
        fourth-order central difference at h = 3e-3: relative error <= 1e-5 (measured
        5.3e-8 on MI355X);
      - a random unit direction (signal 1.7e-3), h = 1e-2: <= 1e-5 (measured 5.1e-8);
      - 32 random pixels (seed 7), fourth-order difference at h = 5e-2 and its sixth-order
        Richardson extrapolation with h = 2.5e-2: relative error of the 32-vector <= 1e-5
        each (measured 7.2e-8 and 6.7e-8)."""
    torch = ctx.torch
    rng = np.random.default_rng(42)
    z = rng.standard_normal((5, 64, 64))
    epsv = 1.0 + 11.25 / (1.0 + np.exp(-z))
    x = np.linspace(0.0, 0.8, 65)
    layers = [(x, x, epsv[k].astype(complex)) for k in range(5)]
    air = G.homogeneous(0.8, 1.0)
    inst = dict(lam=1.0, layers=layers, thickness=np.full(5, 0.15), bottom=air, top=air)
    nx = ny = 11
    prob = StackProblem(ctx, nx, ny, [inst])
    assert prob.cells == 20480
    eps = prob.eps_vector().real.clone().to("cuda").requires_grad_(True)
    J = _objective(ctx, prob, eps)
    J.backward()
    g = eps.grad.cpu().numpy()
    e0 = eps.detach().cpu().numpy()

    def Jdir(d, t):
        return float(_objective(ctx, prob, torch.from_numpy(e0 + t * d).to("cuda")))

    def fd4(d, h):
        return (8.0 * (Jdir(d, h) - Jdir(d, -h)) - (Jdir(d, 2 * h) - Jdir(d, -2 * h))) / (12.0 * h)

    with torch.no_grad():
        r = np.random.default_rng(3).standard_normal(e0.size)
        for name, d, h, tol in (("gradient", g / np.linalg.norm(g), 3e-3, TOL_GRAD),
                                ("random", r / np.linalg.norm(r), 1e-2, TOL_GRAD)):
            ad, fd = float(g @ d), fd4(d, h)
            rel = abs(ad - fd) / abs(fd)
            print(f"direction {name}: adjoint {ad:.12e} fd4(h={h:.0e}) {fd:.12e} rel {rel:.2e}", flush=True)
            assert rel < tol
        pick = np.random.default_rng(7).choice(e0.size, 32, replace=False)
        fdp = np.empty(32)
        fdr = np.empty(32)
        fd6 = np.empty(32)
        for t, k in enumerate(pick):
            ek = np.zeros(e0.size)
            ek[k] = 1.0
            fdp[t] = fd4(ek, 5e-2)
            fdr[t] = (16.0 * fd4(ek, 2.5e-2) - fdp[t]) / 15.0  # Richardson: sixth order
            fd6[t] = (Jdir(ek, 1e-6) - Jdir(ek, -1e-6)) / 2e-6
    err = np.linalg.norm(g[pick] - fdp) / np.linalg.norm(fdp)
    errr = np.linalg.norm(g[pick] - fdr) / np.linalg.norm(fdr)
    err6 = np.linalg.norm(g[pick] - fd6) / np.linalg.norm(fd6)
    print(f"J={float(J.detach()):.12f} |g|={np.linalg.norm(g):.3e}; 32 pixels: rel err vs 4th-order FD (h=5e-2) {err:.2e}, "
          f"vs 6th-order (Richardson of h=5e-2, 2.5e-2) {errr:.2e}, "
          f"vs central FD (h=1e-6, noise-dominated) {err6:.2e}", flush=True)
    assert err < TOL_GRAD and errr < TOL_GRAD, (err, errr)
