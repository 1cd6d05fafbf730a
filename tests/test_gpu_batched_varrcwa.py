"""varRCWA over several integrators in one call (vrc_redheffer_compute_many): a wavelength
sweep gives, per wavelength, the subdivision tree and S-matrix of RedhefferIntegrator.compute
on that wavelength alone, and the oracle's tree and S-matrix
(RedhefferIntegrator::compute, src/core/RedhefferIntegrator.cpp:174-314)."""
import numpy as np
import pytest

from oracle import rcwa_oracle as O
from tests import geometries as G
from varrcwa_amd import api

pytestmark = pytest.mark.gpu


def result(it):
    return O.SResult(it.Tdd(), it.Tuu(), it.Rud(), it.Rdu(), it.W0(), it.V0(), it.gam0(), it.W1(), it.V1(),
                     it.gam1())


def max_rel(a, b):
    return max(np.abs(x - y).max() / np.abs(y).max() for x, y in zip(a, b))


def test_sweep_equals_single_calls_and_oracle(ctx):
    lams = [1.55, 1.4, 1.7]
    pe = 5e-2
    geos = []
    for lam in lams:
        g = G.metasurface(2)
        g["lam"] = lam
        geos.append(g)
    samplers = [G.gpu_sampler(g, ctx) for g in geos]
    batch = [api.RedhefferIntegrator(g["z0"], g["z1"], s) for g, s in zip(geos, samplers)]
    api.RedhefferIntegrator.compute_many(batch, 0, 10, pe)
    for g, s, itb in zip(geos, samplers, batch):
        one = api.RedhefferIntegrator(g["z0"], g["z1"], s)
        one.compute(0, 10, pe)
        assert one.tree() == itb.tree()
        assert one.N() == itb.N()
        rb, r1 = result(itb), result(one)
        # the same kernels on the same matrices: equal to rounding of the batched launches
        assert max_rel([rb.Tdd, rb.Tuu, rb.Rud, rb.Rdu], [r1.Tdd, r1.Tuu, r1.Rud, r1.Rdu]) < 1e-12
        ref = O.redheffer_integrate(G.oracle_sampler(g), g["z0"], g["z1"], pe)
        assert len(itb.tree()) == len(ref.tree)
        assert np.allclose(itb.tree(), ref.tree, rtol=0, atol=1e-14)
        a, b = O.invariant_transforms(rb), O.invariant_transforms(ref)
        err = max(np.abs(a[k] - b[k]).max() / np.abs(b[k]).max() for k in a)
        print(f"lambda {g['lam']}: {itb.N()} segments, deviation from the oracle {err:.2e}")
        assert err < 1e-9, (g["lam"], err)


def test_sweep_rejects_mixed_sizes(ctx):
    g1, g2 = G.metasurface(1), G.metasurface(2)
    its = [api.RedhefferIntegrator(g["z0"], g["z1"], G.gpu_sampler(g, ctx)) for g in (g1, g2)]
    with pytest.raises(api.VrcError, match="equal N"):
        api.RedhefferIntegrator.compute_many(its, 0, 10, 5e-2)
