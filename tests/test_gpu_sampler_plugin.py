"""The sampler plugin boundary from Python: a caller-written subclass of
varrcwa_amd.api.LayerSampler (the reference's abstract LayerSampler,
src/core/LayerSampler.h:5-77) drives the three integrators; its P and Q reach the library
through vrc_callback_sampler_create. The subclass here moves the metasurface pillar edge
itself and asks a private library layer description for P and Q at that edge position,
on the host (sample) or into the device buffers (sampleOnGPU). Bar: 1e-9 relative on the
basis-invariant transforms against the oracle.
"""
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as O
from tests import geometries as G
from varrcwa_amd import api

pytestmark = pytest.mark.gpu


class PillarSampler(api.LayerSampler):
    def __init__(self, nh, on_gpu, ctx):
        super().__init__(1.55, 2 * nh + 1, 2 * nh + 1, enablePML=False, ctx=ctx)
        self.on_gpu = on_gpu
        self.calls = 0
        self.postInitialization()

    def assignSimulationRegion(self):
        self.Lx_ = self.Ly_ = 0.65
        self.l0_ = self.b0_ = 0.0
        self.l1_ = self.b1_ = 0.1
        self.r0_ = self.u0_ = 0.5
        self.r1_ = self.u1_ = 0.65

    def _layer(self, z):
        xe = 0.5 if z <= 1.0 else 0.2 * math.cos(0.5 * math.pi * (z - 1.0)) + 0.3
        eps = np.ones((3, 3), dtype=complex)
        eps[1, 1] = 3.48 ** 2
        return api.VarLayerSampler(self.lam, self.nx, self.ny, [0.0, 0.1, xe, 0.65], [0.0, 0.1, 0.5, 0.65], eps,
                                   enablePML=False, ctx=self.ctx)

    def sample(self, z):
        self.calls += 1
        return self._layer(z).sample(0.0)

    def sampleOnGPU(self, z, P, Q):
        if not self.on_gpu:
            return NotImplemented
        self.calls += 1
        lay = self._layer(z)
        self.ctx.check(self.ctx.lib.vrc_sampler_sample_gpu(lay.h, 0.0, P, Q), "sampleOnGPU")


def close(got, ref, tol=1e-9):
    a, b = O.invariant_transforms(got), O.invariant_transforms(ref)
    for k in a:
        err = np.abs(a[k] - b[k]).max() / np.abs(b[k]).max()
        assert err < tol, (k, err)


def result(it):
    return O.SResult(it.Tdd(), it.Tuu(), it.Rud(), it.Rdu(), it.W0(), it.V0(), it.gam0(), it.W1(), it.V1(),
                     it.gam1())


@pytest.mark.parametrize("on_gpu", [False, True])
def test_python_user_sampler_drives_all_integrators(ctx, on_gpu):
    g = G.metasurface(2)
    s = PillarSampler(2, on_gpu, ctx)
    assert s.nDim() == 50 and s.Lx() == 0.65
    assert np.allclose(np.diag(s.Kx()), O.alpha_beta(5, 5, 0.65, 0.65)[0])
    it = api.RCWAIntegrator(g["z0"], g["z1"], s)
    it.compute(6)
    assert s.calls >= 8
    close(result(it), O.rcwa_integrate(G.oracle_sampler(g), g["z0"], g["z1"], 6))
    pe = 5e-2
    it = api.RedhefferIntegrator(g["z0"], g["z1"], s)
    it.compute(0, 10, pe)
    ref = O.redheffer_integrate(G.oracle_sampler(g), g["z0"], g["z1"], pe)
    assert it.N() == ref.N
    close(result(it), ref)
    # the differential method in its well-conditioned regime (3 x 3 harmonics, 32 steps): at
    # 5 x 5 harmonics and 8 steps the oracle itself moves by 1e-3 under 1e-15 relative
    # perturbations of P and Q (evanescent growth over one step amplifies rounding)
    g1 = G.metasurface(1)
    s1 = PillarSampler(1, on_gpu, ctx)
    it = api.DifferentialIntegrator(g1["z0"], g1["z1"], s1)
    it.compute(32)
    close(result(it), O.differential_integrate(G.oracle_sampler(g1), g1["z0"], g1["z1"], 32))


def test_python_user_sampler_errors(ctx):
    class Broken(PillarSampler):
        def sample(self, z):
            raise RuntimeError("no P for you")

    s = Broken(1, False, ctx)
    it = api.RCWAIntegrator(0.0, 1.0, s)
    with pytest.raises(api.VrcError, match="callback"):
        it.compute(2)

    class NoRegion(api.LayerSampler):
        pass

    with pytest.raises(NotImplementedError):
        NoRegion(1.0, 3, 3, ctx=ctx).postInitialization()
