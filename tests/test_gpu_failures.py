"""Numerical failures surface as statuses, not as silent NaN results.

The reference checks nothing (cuSOLVER/MAGMA info ignored, the division by gamma at
src/rcwa/WaveEquationSolver.cpp:36 unguarded). Here the calls that solve whole problems
read the device failure record back and return a positive status naming the cause, and
batched stacks flag the failing instances only. Failures are forced from the host: an
exact Rayleigh/Wood anomaly wavelength, a zero or NaN permittivity, the eigensolver's own
sweep budget set to one sweep, the varRCWA depth limit set to one level.
"""
import ctypes as C

import numpy as np
import pytest

from tests import geometries as G
from varrcwa_amd import api

pytestmark = pytest.mark.gpu


def rect_instance(lam, eps_rect=12.25, eps_bg=1.0, cell=0.6, nl=2):
    x = [0.0, 0.1, 0.35, cell]
    y = [0.0, 0.2, 0.45, cell]
    e = np.full((3, 3), eps_bg, dtype=complex)
    e[1, 1] = eps_rect
    air = api.Layer([0.0, cell], [0.0, cell], [[1.0]])
    layers = [api.Layer(x, y, e) for _ in range(nl)]
    return dict(lam=lam, layers=layers, thickness=[0.1] * nl, bottom=air, top=air)


def raw_batch(ctx, nh, insts):
    """vrc_batch_stack_rcwa through ctypes: (status, info per instance, S tensor)."""
    import torch
    from varrcwa_amd._lib import ptr
    arr, keep = api._stack_array(insts)
    N = 2 * nh * nh
    S = torch.empty((len(insts), 4, N, N), dtype=torch.complex128, device="cuda")
    info = np.full(len(insts), -7, dtype=np.int32)
    rc = ctx.lib.vrc_batch_stack_rcwa(ctx.h, nh, nh, len(insts), arr, 0, ptr(S), None, None, None, None, None,
                                      None, info.ctypes.data_as(C.POINTER(C.c_int)))
    del keep
    return rc, info, S


def tune(ctx, key, value):
    assert ctx.lib.vrc_tune_set(key.encode(), int(value)) == 0


def test_healthy_batch_reports_zero(ctx):
    rc, info, S = raw_batch(ctx, 3, [rect_instance(0.9), rect_instance(1.3)])
    assert rc == 0 and (info == 0).all(), (rc, info)
    assert bool(S.isfinite().all())


def test_wood_anomaly_flags_only_that_instance(ctx):
    """lambda = period: the (+-1, 0) orders of the air half-spaces graze (k0^2 = alpha^2 exactly),
    gamma = 0 and V = Q W / gamma is undefined."""
    rc, info, S = raw_batch(ctx, 3, [rect_instance(0.9), rect_instance(0.6), rect_instance(1.3)])
    msg = ctx.lib.vrc_last_error(ctx.h).decode()
    assert rc == 4, (rc, msg)
    assert list(info) == [0, 4, 0], info
    assert "instance 1" in msg and "Wood" in msg, msg
    assert bool(S[0].isfinite().all()) and bool(S[2].isfinite().all())
    assert not bool(S[1].isfinite().all())
    # the Python binding raises, or returns the codes on request
    with pytest.raises(api.VrcError, match="instance 1"):
        api.batch_stack_rcwa(ctx, 3, 3, [rect_instance(0.9), rect_instance(0.6)])
    out = api.batch_stack_rcwa(ctx, 3, 3, [rect_instance(0.9), rect_instance(0.6)], return_info=True)
    assert list(out["info"]) == [0, 4]


@pytest.mark.parametrize("bad", [0.0, float("nan"), float("inf")])
def test_zero_or_nonfinite_permittivity_is_a_bad_argument(ctx, bad):
    inst = rect_instance(1.0, eps_rect=bad)
    rc, info, _ = raw_batch(ctx, 3, [rect_instance(0.9), inst])
    msg = ctx.lib.vrc_last_error(ctx.h).decode()
    assert rc == -1 and "permittivity" in msg, (rc, msg)
    # a bad half-space medium as well
    inst = rect_instance(1.0)
    inst["top"] = api.Layer([0.0, 0.6], [0.0, 0.6], [[bad]])
    rc, _, _ = raw_batch(ctx, 3, [inst])
    assert rc == -1
    # and the sampler of the drop-in integrators
    g = G.metasurface(1)
    eps = g["eps"].copy()
    eps[1, 1] = bad
    with pytest.raises(api.VrcError, match="permittivity"):
        api.VarLayerSampler(g["lam"], g["nx"], g["ny"], g["x"], g["y"], eps, enablePML=False, ctx=ctx)


def test_unconverged_eigensolver(ctx):
    """One QR sweep is not enough for a 162 x 162 patterned-layer eigenproblem: every slab's
    eigensolve reports info > 0, every instance is flagged 1, and the drop-in integrator
    returns status 1."""
    tune(ctx, "zgeev.max_sweeps", 1)
    try:
        rc, info, _ = raw_batch(ctx, 9, [rect_instance(0.9), rect_instance(1.2)])
        msg = ctx.lib.vrc_last_error(ctx.h).decode()
        assert rc == 1 and list(info) == [1, 1], (rc, info, msg)
        assert "did not converge" in msg
        g = G.metasurface(4)  # 9 x 9 harmonics
        s = G.gpu_sampler(g, ctx)
        it = api.RCWAIntegrator(g["z0"], g["z1"], s)
        rc = ctx.lib.vrc_rcwa_compute(it.h, 2)
        assert rc == 1, (rc, ctx.lib.vrc_last_error(ctx.h).decode())
    finally:
        tune(ctx, "zgeev.max_sweeps", 0)
    # the budget restored, the same problems solve cleanly
    rc, info, _ = raw_batch(ctx, 9, [rect_instance(0.9), rect_instance(1.2)])
    assert rc == 0 and (info == 0).all()
    assert ctx.lib.vrc_rcwa_compute(it.h, 2) == 0


def test_varrcwa_depth_limit(ctx):
    """varRCWA that cannot meet pe within the depth limit accepts the segments and says so."""
    g = G.metasurface(1)
    s = G.gpu_sampler(g, ctx)
    it = api.RedhefferIntegrator(g["z0"], g["z1"], s)
    tune(ctx, "redheffer.max_depth", 1)
    try:
        rc = ctx.lib.vrc_redheffer_compute(it.h, 0, 10, 1e-14)
        msg = ctx.lib.vrc_last_error(ctx.h).decode()
        assert rc == 5 and "depth limit" in msg, (rc, msg)
        assert it.N() == 3  # one level of trisection, every segment accepted
    finally:
        tune(ctx, "redheffer.max_depth", 40)
    assert ctx.lib.vrc_redheffer_compute(it.h, 0, 10, 1e-2) == 0
