"""Write the oracle fixtures of the full-size GPU parity tests (test infrastructure).

Each fixture holds what the CPU oracle (the restated reference CPU path) computes for one
scenario, reduced to a few hundred KB: for each basis-invariant transform
(W1 Tuu W0^-1, W0 Tdd W1^-1, W0 Rdu W0^-1, W1 Rud W1^-1) its largest modulus, 512 seeded
entries and the block of the zeroth harmonic; both end media's propagation constants; and,
for varRCWA, the accepted segments in cascade order. Scenarios:

  metric_inst{0,511}  the throughput workload at 21 x 21 harmonics (N = 882): geometry 0 at
                      0.8 um and geometry 7 at 1.6 um of tests/golden/metric_geometry.json
  c1_pillar           staircased tapered pillar, 256 x 256 pixel layers, 11 x 11 harmonics,
                      20 slices through RCWAIntegrator
  var12_25x21         the reference's PML waveguide taper at its own 25 x 21 harmonics
                      (src/tests/test_cpu.cpp:128-176), RCWAIntegrator with 4 slices
  c2_converter        PML mode converter, 15 x 15 harmonics, RedhefferIntegrator at pe = 1e-2
  c3_mixed            the mixed-order sweep's instances 0, 2, 3 and 5 (21 x 21, 7 x 7, 15 x 15
                      and 11 x 11 harmonics: one of each class), keys prefixed "i<instance>_"
  metasurface_25x25   the reference's metasurface scenario at its own 25 x 25 harmonics
                      (src/tests/test_metasurface_cpu.cpp:129-164), N = 1250, RCWAIntegrator
                      with 4 slices

Run in the build container: python tests/golden/make_parity_fixtures.py [name ...]
(minutes of CPU; the GPU tests only read the .npz files).
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import rcwa_oracle as O  # noqa: E402
from tests import geometries as G  # noqa: E402

OUT = Path(__file__).resolve().parent
NSAMPLE = 512


def reduce_result(r, nx, ny, seed=0):
    """The fixture arrays of one SResult."""
    inv = O.invariant_transforms(r)
    N = r.W0.shape[0]
    n = nx * ny
    z = (nx // 2) * ny + ny // 2
    idx0 = np.array([z, n + z])
    rng = np.random.default_rng(seed)
    ij = rng.integers(0, N, size=(NSAMPLE, 2))
    out = dict(N=np.array(N), nx=np.array(nx), ny=np.array(ny), sample_ij=ij,
               gam0=np.asarray(r.gam0), gam1=np.asarray(r.gam1))
    for k, M in inv.items():
        out[f"{k}_max"] = np.array(np.abs(M).max())
        out[f"{k}_samples"] = M[ij[:, 0], ij[:, 1]]
        out[f"{k}_zeroth"] = M[np.ix_(idx0, idx0)]
        out[f"{k}_fro"] = np.array(np.linalg.norm(M))
    if r.tree:
        out["tree"] = np.asarray(r.tree, dtype=np.float64)
    return out


def metric(i, nh=21):
    d = json.loads((OUT / "metric_geometry.json").read_text())
    cell = d["cell"]
    geoms, lams = d["geometries"], d["wavelengths"]
    g = geoms[i % len(geoms)]
    lam = lams[(i // len(geoms)) % len(lams)]
    air = G.homogeneous(cell, 1.0)
    layers = [(np.array(l["x"]), np.array(l["y"]), np.array(l["eps"], dtype=complex)) for l in g["layers"]]
    ze = np.concatenate([[0.0], np.cumsum(g["thickness"])])
    s = O.StackLayerSampler(lam, nh, nh, layers, ze, air, air)
    return reduce_result(O.rcwa_integrate(s, ze[0], ze[-1], len(layers), z_edges=ze), nh, nh)


C3_INSTANCES = (0, 2, 3, 5)


def c3():
    orders = json.loads((OUT / "metric_geometry.json").read_text())["c3_orders"]
    out = {}
    for i in C3_INSTANCES:
        for k, v in metric(i, orders[i]).items():
            out[f"i{i}_{k}"] = v
    out["N"] = np.array(max(int(out[f"i{i}_N"]) for i in C3_INSTANCES))
    return out


def c1():
    g = G.pillar_taper_c1()
    s, ze = G.stack_oracle_sampler(g)
    return reduce_result(O.rcwa_integrate(s, ze[0], ze[-1], len(ze) - 1, z_edges=ze), g["nx"], g["ny"])


def var12():
    g = G.var12(12, 10)
    return reduce_result(O.rcwa_integrate(G.oracle_sampler(g), g["z0"], g["z1"], 4), g["nx"], g["ny"])


def c2():
    g = G.mode_converter_c2()
    r = O.redheffer_integrate(G.oracle_sampler(g), g["z0"], g["z1"], 1e-2)
    return reduce_result(r, g["nx"], g["ny"])


def metasurface_full():
    g = G.metasurface(12)
    return reduce_result(O.rcwa_integrate(G.oracle_sampler(g), g["z0"], g["z1"], 4), g["nx"], g["ny"])


SCENARIOS = {"metric_inst0": lambda: metric(0), "metric_inst511": lambda: metric(511), "c1_pillar": c1,
             "var12_25x21": var12, "c2_converter": c2, "c3_mixed": c3,
             "metasurface_25x25": metasurface_full}


def main(names):
    for name in names or SCENARIOS:
        t0 = time.perf_counter()
        arrs = SCENARIOS[name]()
        np.savez_compressed(OUT / f"parity_{name}.npz", **arrs)
        print(f"{name}: N={int(arrs['N'])} in {time.perf_counter() - t0:.1f} s"
              + (f", {len(arrs['tree'])} segments" if "tree" in arrs else ""), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
