"""Generate the seeded rectangle-stack geometries of the S-matrix throughput workload.

Per seed (numpy.random.default_rng(seed), seeds 0..7), 10 layers are drawn in this
order: thickness ~ U(0.05, 0.2) um; rectangle count ~ integers(1, 5); then per
rectangle: width, height ~ U(0.05, 0.3) um, corner x0, y0 ~ U(0, 0.6) um, eps index
~ integers(0, 2) into {4.0, 12.25}. Rectangles are wrapped periodically into the
0.6 um cell, painted in order on an eps = 1 background and stored as exact
Layer(x, y, eps) grids. Wavelengths: linspace(0.8, 1.6, 64) um. Half-space media:
eps = 1 above and below. Instance i of the 512-instance sweep is geometry i % 8 at
wavelength i // 8.

The mixed-order variant of the sweep gives instance i the harmonics c3_orders[i] x
c3_orders[i], drawn with numpy.random.default_rng(100).choice([7, 11, 15, 21], 512).

Writes tests/golden/metric_geometry.json (run: python tests/golden/make_metric_geometry.py).
"""
import json
from pathlib import Path

import numpy as np

CELL = 0.6
EPS = (4.0, 12.25)
SEEDS = range(8)
NLAYERS = 10


def rects_to_layer(rects, cell):
    xs, ys = {0.0, cell}, {0.0, cell}
    pieces = []
    for (x0, y0, w, h, e) in rects:
        for dxs in (0.0, -cell):
            for dys in (0.0, -cell):
                a, b = max(x0 + dxs, 0.0), min(x0 + w + dxs, cell)
                c, d = max(y0 + dys, 0.0), min(y0 + h + dys, cell)
                if b > a and d > c:
                    pieces.append((a, b, c, d, e))
                    xs.update([a, b])
                    ys.update([c, d])
    x = np.array(sorted(xs))
    y = np.array(sorted(ys))
    eps = np.ones((y.size - 1, x.size - 1))
    xm = 0.5 * (x[1:] + x[:-1])
    ym = 0.5 * (y[1:] + y[:-1])
    for (a, b, c, d, e) in pieces:
        eps[np.ix_((ym > c) & (ym < d), (xm > a) & (xm < b))] = e
    return x, y, eps


def geometry(seed):
    rng = np.random.default_rng(seed)
    layers, thick = [], []
    for _ in range(NLAYERS):
        thick.append(float(rng.uniform(0.05, 0.2)))
        nrect = int(rng.integers(1, 5))
        rects = []
        for _ in range(nrect):
            w, h = rng.uniform(0.05, 0.3), rng.uniform(0.05, 0.3)
            x0, y0 = rng.uniform(0, CELL), rng.uniform(0, CELL)
            e = EPS[int(rng.integers(0, 2))]
            rects.append((x0, y0, w, h, e))
        x, y, eps = rects_to_layer(rects, CELL)
        layers.append(dict(x=x.tolist(), y=y.tolist(), eps=eps.tolist()))
    return dict(seed=seed, thickness=thick, layers=layers)


def main():
    out = dict(cell=CELL, wavelengths=np.linspace(0.8, 1.6, 64).tolist(), eps_background=1.0,
               geometries=[geometry(s) for s in SEEDS],
               c3_orders=[int(v) for v in np.random.default_rng(100).choice([7, 11, 15, 21], 512)])
    path = Path(__file__).with_name("metric_geometry.json")
    path.write_text(json.dumps(out))
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
