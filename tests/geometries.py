"""Scenario geometries of the reference drivers, restated as data (shared by tests)."""
import math

import numpy as np

PI = 3.1415926535897932


def metasurface(nh=2):
    """src/tests/test_metasurface_cpu.cpp:129-164 (periodic, pillar edge moves with z)."""
    x = np.array([0.0, 0.1, 0.5, 0.65])
    y = x.copy()
    cl, co = 1.0, 3.48 * 3.48
    eps = np.array([[cl, cl, cl], [cl, co, cl], [cl, cl, cl]], dtype=complex)

    def rule(z):
        return 0.5 if z <= 1.0 else 0.2 * math.cos(0.5 * PI * (z - 1.0)) + 0.3

    return dict(lam=1.55, nx=2 * nh + 1, ny=2 * nh + 1, x=x, y=y, eps=eps, pml=False, rules={2: rule},
                z0=0.0, z1=2.0)


def var12(nhx=2, nhy=2):
    """src/tests/test_cpu.cpp:128-176 (PML waveguide taper, var12 rule)."""
    x = np.array([0.0, 0.5, 1.5, 2.6, 3.7, 3.9, 5.0, 5.5])
    y = np.array([0.0, 0.5, 1.5, 1.58, 1.8, 2.8, 3.3])
    cl, co = 1.445 ** 2, 3.48 ** 2
    eps = np.full((6, 7), cl, dtype=complex)
    eps[3, 2] = co

    def rule(z):
        if z < 0.0:
            return 2.6
        if z > 10.0:
            return 3.7
        return (1.0 - 0.1 * z) * 2.6 + 0.1 * z * 3.7

    return dict(lam=1.55, nx=2 * nhx + 1, ny=2 * nhy + 1, x=x, y=y, eps=eps, pml=True, rules={3: rule},
                z0=-0.1, z1=10.1)


def guided_benchmark(nhx=2, nhy=2):
    """src/tests/test_cpu.cpp:416-451 (PML waveguide whose core edge moves 3.6 -> 4.7 over
    z in [0, 10]; the sampler's enablePML defaults to true, src/core/VarLayerSampler.h:19-37)."""
    x = np.array([0.0, 0.5, 2.5, 3.6, 4.7, 6.7, 7.2])
    y = np.array([0.0, 0.5, 2.0, 2.08, 2.3, 3.8, 4.3])
    cl, co = 1.445 * 1.445, 3.48 * 3.48
    eps = np.full((6, 6), cl, dtype=complex)
    eps[3, 2] = co

    def rule(z):
        if z < 0.0:
            return 3.6
        if z > 10.0:
            return 4.7
        return (1.0 - 0.1 * z) * 3.6 + 0.1 * z * 4.7

    return dict(lam=1.55, nx=2 * nhx + 1, ny=2 * nhy + 1, x=x, y=y, eps=eps, pml=True, rules={3: rule},
                z0=-0.1, z1=10.1)


def oracle_sampler(g):
    from oracle import rcwa_oracle as O
    s = O.VarLayerSampler(g["lam"], g["nx"], g["ny"], g["x"], g["y"], g["eps"], g["pml"])
    for k, f in g["rules"].items():
        s.addXRule(k, f)
    return s


def gpu_sampler(g, ctx=None):
    from varrcwa_amd.api import VarLayerSampler
    s = VarLayerSampler(g["lam"], g["nx"], g["ny"], g["x"], g["y"], g["eps"], g["pml"], ctx=ctx)
    for k, f in g["rules"].items():
        s.addXRule(k, f)
    return s


def random_rect_layer(rng, cell=0.6, eps_choices=(4.0, 12.25), nmax=4, side=(0.05, 0.3)):
    """A union of 1..nmax axis-aligned rectangles (periodically wrapped) in a square
    cell, on an eps=1 background, as an exact Layer(x, y, eps) grid."""
    nrect = int(rng.integers(1, nmax + 1))
    rects = []
    for _ in range(nrect):
        w, h = rng.uniform(*side), rng.uniform(*side)
        x0, y0 = rng.uniform(0, cell), rng.uniform(0, cell)
        e = float(eps_choices[int(rng.integers(0, len(eps_choices)))])
        rects.append((x0, y0, w, h, e))
    return rects_to_layer(rects, cell)


def rects_to_layer(rects, cell):
    """Rasterise rectangles (x0, y0, w, h, eps) painted in order onto an exact grid."""
    xs, ys = {0.0, cell}, {0.0, cell}
    pieces = []
    for (x0, y0, w, h, e) in rects:
        for dxs in (0.0, -cell):
            for dys in (0.0, -cell):
                a, b = x0 + dxs, x0 + w + dxs
                c, d = y0 + dys, y0 + h + dys
                a, b = max(a, 0.0), min(b, cell)
                c, d = max(c, 0.0), min(d, cell)
                if b > a and d > c:
                    pieces.append((a, b, c, d, e))
                    xs.update([a, b])
                    ys.update([c, d])
    x = np.array(sorted(xs))
    y = np.array(sorted(ys))
    eps = np.ones((y.size - 1, x.size - 1), dtype=complex)
    xm = 0.5 * (x[1:] + x[:-1])
    ym = 0.5 * (y[1:] + y[:-1])
    for (a, b, c, d, e) in pieces:
        mx = (xm > a) & (xm < b)
        my = (ym > c) & (ym < d)
        eps[np.ix_(my, mx)] = e
    return x, y, eps


def homogeneous(cell, eps):
    return np.array([0.0, cell]), np.array([0.0, cell]), np.array([[eps]], dtype=complex)


def pillar_taper_c1(grid=256, nh=11, nlayers=20):
    """Staircased tapered pillar on a substrate (plain RCWA, CPU-runnable case): square cell
    of period 0.5 um, lambda = 1.0 um, circular pillar of eps = 12.25 in air whose radius
    tapers linearly from 0.10 to 0.20 um over 1.0 um height, staircased into `nlayers`
    layers (radius at each layer's mid-height), each a grid x grid pixel map (pixel centres
    inside the circle); substrate eps = 2.1 below, air above; nh x nh harmonics."""
    cell, h = 0.5, 1.0
    x = np.linspace(0.0, cell, grid + 1)
    c = 0.5 * (x[1:] + x[:-1]) - 0.5 * cell
    r2 = c[None, :] ** 2 + c[:, None] ** 2
    layers, thick = [], []
    for k in range(nlayers):
        r = 0.10 + 0.10 * (k + 0.5) / nlayers
        eps = np.where(r2 < r * r, 12.25, 1.0).astype(complex)
        layers.append((x, x.copy(), eps))
        thick.append(h / nlayers)
    return dict(lam=1.0, nx=nh, ny=nh, cell=cell, layers=layers, thickness=thick,
                bottom=homogeneous(cell, 2.1), top=homogeneous(cell, 1.0))


def mode_converter_c2(nh=15):
    """PML waveguide mode converter (varRCWA case): a rectangular core of eps = 12.1 and
    height 0.22 um in eps = 2.1 cladding, computational window 4 x 2 um with 0.5 um
    complex-stretched PML on every side, lambda = 1.55 um, nh x nh harmonics; over
    z in [0, 10] um the core width w(z) = 0.4 + 0.8 (3 s^2 - 2 s^3), s = z / 10 (cubic
    profile, centred at x = 2 um), constant before and after."""
    x = np.array([0.0, 0.5, 1.8, 2.2, 3.5, 4.0])
    y = np.array([0.0, 0.5, 0.89, 1.11, 1.5, 2.0])
    cl, co = 2.1, 12.1
    eps = np.full((5, 5), cl, dtype=complex)
    eps[2, 2] = co

    def width(z):
        s = min(max(z / 10.0, 0.0), 1.0)
        return 0.4 + 0.8 * (3 * s * s - 2 * s * s * s)

    def left(z):
        return 2.0 - 0.5 * width(z)

    def right(z):
        return 2.0 + 0.5 * width(z)

    return dict(lam=1.55, nx=nh, ny=nh, x=x, y=y, eps=eps, pml=True, rules={2: left, 3: right},
                z0=0.0, z1=10.0)


def stack_oracle_sampler(g):
    from oracle import rcwa_oracle as O
    ze = np.concatenate([[0.0], np.cumsum(g["thickness"])])
    return O.StackLayerSampler(g["lam"], g["nx"], g["ny"], g["layers"], ze, g["bottom"], g["top"]), ze


def stack_gpu_sampler(g, ctx=None):
    from varrcwa_amd.api import Layer, StackLayerSampler
    ze = np.concatenate([[0.0], np.cumsum(g["thickness"])])
    layers = [Layer(*l) for l in g["layers"]]
    return StackLayerSampler(g["lam"], g["nx"], g["ny"], layers, ze, Layer(*g["bottom"]), Layer(*g["top"]),
                             ctx=ctx), ze
