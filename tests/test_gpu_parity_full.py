"""Parity with the CPU oracle at the full sizes of the scenarios the reference and the
workload define, against fixtures the oracle wrote in the build container
(tests/golden/make_parity_fixtures.py):

* the throughput workload's instances 0 and 511 at 21 x 21 harmonics (N = 882), inside
  batches of four through vrc_batch_stack_rcwa;
* the staircased tapered pillar (256 x 256 pixel layers, 11 x 11 harmonics, 20 slices)
  through the drop-in RCWAIntegrator;
* the reference's PML waveguide taper var12 at its own 25 x 21 harmonics
  (src/tests/test_cpu.cpp:128-176), N = 1050, through RCWAIntegrator;
* the PML mode converter (15 x 15 harmonics, N = 450) through RedhefferIntegrator, which
  must take the oracle's subdivision tree segment for segment;
* the mixed-order sweep (harmonics per instance from {7, 11, 15, 21} squared) in one
  vrc_batch_stack_rcwa_mixed call.

Bar: 512 seeded entries of each basis-invariant transform (W1 Tuu W0^-1, W0 Tdd W1^-1,
W0 Rdu W0^-1, W1 Rud W1^-1) within 1e-9 of the transform's largest modulus, the zeroth
harmonic's 2 x 2 block (the propagating orders) within 1e-9 relative entrywise on the
moduli, Frobenius norms within 1e-9, and both end media's propagation constants within
1e-10 relative under the optimal pairing.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import rcwa_oracle as O
from tests import geometries as G
from varrcwa_amd import api
from varrcwa_amd._lib import from_dev_colmajor

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
TOL = 1e-9


def fixture(name):
    return dict(np.load(GOLDEN / f"parity_{name}.npz", allow_pickle=False))


def spectrum_rel(a, b):
    from scipy.optimize import linear_sum_assignment
    d = np.abs(a[:, None] - b[None, :])
    r, c = linear_sum_assignment(d)
    return d[r, c].max() / np.abs(b).max()


def compare(got: O.SResult, fx, label):
    inv = O.invariant_transforms(got)
    ij = fx["sample_ij"]
    nx, ny = int(fx["nx"]), int(fx["ny"])
    n = nx * ny
    z = (nx // 2) * ny + ny // 2
    idx0 = np.array([z, n + z])
    worst = {}
    for k, M in inv.items():
        e_s = np.abs(M[ij[:, 0], ij[:, 1]] - fx[f"{k}_samples"]).max() / float(fx[f"{k}_max"])
        e_f = abs(np.linalg.norm(M) - float(fx[f"{k}_fro"])) / float(fx[f"{k}_fro"])
        ga, gb = np.abs(M[np.ix_(idx0, idx0)]), np.abs(fx[f"{k}_zeroth"])
        big = gb > 1e-6 * float(fx[f"{k}_max"])
        e_z = float((np.abs(ga - gb)[big] / gb[big]).max()) if big.any() else 0.0
        worst[k] = max(e_s, e_f, e_z)
    g0 = spectrum_rel(np.asarray(got.gam0), fx["gam0"])
    g1 = spectrum_rel(np.asarray(got.gam1), fx["gam1"])
    print(f"{label}: transforms {worst}, gamma0 {g0:.1e}, gamma1 {g1:.1e}")
    assert max(worst.values()) < TOL, (label, worst)
    assert g0 < 1e-10 and g1 < 1e-10, (label, g0, g1)


def integrator_result(it):
    return O.SResult(it.Tdd(), it.Tuu(), it.Rud(), it.Rdu(), it.W0(), it.V0(), it.gam0(), it.W1(), it.V1(),
                     it.gam1())


def metric_instances(first, count):
    d = json.loads((GOLDEN / "metric_geometry.json").read_text())
    cell = d["cell"]
    geoms, lams = d["geometries"], d["wavelengths"]
    air = api.Layer([0.0, cell], [0.0, cell], [[1.0]])
    out = []
    for i in range(first, first + count):
        g = geoms[i % len(geoms)]
        lam = lams[(i // len(geoms)) % len(lams)]
        layers = [api.Layer(l["x"], l["y"], np.asarray(l["eps"], dtype=complex)) for l in g["layers"]]
        out.append(dict(lam=lam, layers=layers, thickness=g["thickness"], bottom=air, top=air))
    return out


@pytest.mark.parametrize("inst,first", [(0, 0), (511, 508)])
def test_metric_workload_N882(ctx, inst, first):
    fx = fixture(f"metric_inst{inst}")
    out = api.batch_stack_rcwa(ctx, 21, 21, metric_instances(first, 4), want_bases=True)
    k = inst - first
    S = from_dev_colmajor(out["S"][k])
    got = O.SResult(S[0], S[1], S[2], S[3], from_dev_colmajor(out["W0"][k]), from_dev_colmajor(out["V0"][k]),
                    out["gam0"][k].cpu().numpy(), from_dev_colmajor(out["W1"][k]), from_dev_colmajor(out["V1"][k]),
                    out["gam1"][k].cpu().numpy())
    compare(got, fx, f"metric instance {inst}")


def test_c1_pillar_plain_rcwa(ctx):
    fx = fixture("c1_pillar")
    g = G.pillar_taper_c1()
    s, ze = G.stack_gpu_sampler(g, ctx)
    it = api.RCWAIntegrator(ze[0], ze[-1], s)
    it.compute_edges(ze)
    assert it.N() == 20
    compare(integrator_result(it), fx, "C1 pillar")


def test_var12_reference_harmonics(ctx):
    fx = fixture("var12_25x21")
    g = G.var12(12, 10)
    s = G.gpu_sampler(g, ctx)
    assert s.nDim() == 1050
    it = api.RCWAIntegrator(g["z0"], g["z1"], s)
    it.compute(4)
    compare(integrator_result(it), fx, "var12 25x21")


def test_c2_converter_varrcwa_same_tree(ctx):
    fx = fixture("c2_converter")
    g = G.mode_converter_c2()
    s = G.gpu_sampler(g, ctx)
    it = api.RedhefferIntegrator(g["z0"], g["z1"], s)
    it.compute(0, 10, 1e-2)
    tree = np.asarray(it.tree())
    assert tree.shape == fx["tree"].shape, (tree.shape, fx["tree"].shape)
    assert np.abs(tree - fx["tree"]).max() < 1e-12
    compare(integrator_result(it), fx, "C2 converter")


def test_c3_mixed_orders_one_call(ctx):
    """The mixed-order sweep (orders per instance from {7, 11, 15, 21}): instances 0..5 in
    one vrc_batch_stack_rcwa_mixed call, the four with fixtures (one per harmonic class)
    against the oracle, each also equal to the single-class call on its own."""
    fx = fixture("c3_mixed")
    orders = json.loads((GOLDEN / "metric_geometry.json").read_text())["c3_orders"]
    insts = metric_instances(0, 6)
    harm = [(orders[i], orders[i]) for i in range(6)]
    out = api.batch_stack_rcwa_mixed(ctx, harm, insts, want_bases=True)
    for i in (0, 2, 3, 5):
        sub = {k[len(f"i{i}_"):]: v for k, v in fx.items() if k.startswith(f"i{i}_")}
        S = from_dev_colmajor(out["S"][i])
        got = O.SResult(S[0], S[1], S[2], S[3], from_dev_colmajor(out["W0"][i]), from_dev_colmajor(out["V0"][i]),
                        out["gam0"][i].cpu().numpy(), from_dev_colmajor(out["W1"][i]),
                        from_dev_colmajor(out["V1"][i]), out["gam1"][i].cpu().numpy())
        assert got.Tdd.shape[0] == 2 * orders[i] ** 2
        compare(got, sub, f"C3 instance {i} ({orders[i]}x{orders[i]})")
        one = from_dev_colmajor(api.batch_stack_rcwa(ctx, orders[i], orders[i], [insts[i]])["S"][0])
        assert np.abs(one - S).max() <= 1e-12 * np.abs(one).max()


def test_metasurface_reference_harmonics(ctx):
    """The reference's metasurface scenario at its own 25 x 25 harmonics (N = 1250), four
    slices through the drop-in RCWAIntegrator."""
    fx = fixture("metasurface_25x25")
    g = G.metasurface(12)
    s = G.gpu_sampler(g, ctx)
    assert s.nDim() == 1250
    it = api.RCWAIntegrator(g["z0"], g["z1"], s)
    it.compute(4)
    compare(integrator_result(it), fx, "metasurface 25x25")


def test_c2_converter_pe_1e6_converges_with_rcwa(ctx):
    """The converter through varRCWA at pe = 1e-6 (a tree too large for the oracle to replay):
    midpoint-slice RCWA converges to it at second order, |S_RCWA(n) - S_var| shrinking about
    4x per doubling of n (n = 250, 500, 1000), the last difference small."""
    g = G.mode_converter_c2()
    s = G.gpu_sampler(g, ctx)
    var = api.RedhefferIntegrator(g["z0"], g["z1"], s)
    var.compute(0, 10, 1e-6)
    ref = O.invariant_transforms(integrator_result(var))
    errs = []
    for n in (250, 500, 1000):
        it = api.RCWAIntegrator(g["z0"], g["z1"], s)
        it.compute(n)
        inv = O.invariant_transforms(integrator_result(it))
        errs.append(max(np.abs(inv[k] - ref[k]).max() / np.abs(ref[k]).max() for k in ref))
    print(f"varRCWA pe=1e-6: {var.N()} segments; RCWA differences {errs}")
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0, errs
    assert errs[2] < 1e-3, errs
