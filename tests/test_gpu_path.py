"""Parity of the HIP S-matrix path with the CPU oracle (restated reference CPU path).

Tolerances: relative 1e-9 on basis-invariant S transforms and on
|S| of non-degenerate modes; Fourier/P/Q stages compared elementwise at 1e-12.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from oracle import rcwa_oracle as O
from tests import geometries as G
from varrcwa_amd import api
from varrcwa_amd._lib import from_dev_colmajor, to_dev_colmajor

pytestmark = pytest.mark.gpu

TOL_S = 1e-9
ROOT = Path(__file__).resolve().parent.parent


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def spectrum_rel(a, b):
    """Max distance between two eigenvalue multisets under the optimal pairing
    (order-free; robust to near-degenerate pairs), relative to max |b|."""
    from scipy.optimize import linear_sum_assignment
    d = np.abs(a[:, None] - b[None, :])
    r, c = linear_sum_assignment(d)
    return d[r, c].max() / np.abs(b).max()


def test_eig_pq_large_N(ctx):
    """Deep-stack harmonic count (25 x 25 harmonics, N = 1250): the Hessenberg panel's rows no
    longer fit one row per thread, the LU panel leaves the register path; eigen residual,
    branch rule, V = Q W / gamma and the spectrum against LAPACK."""
    g = G.metasurface(12)
    s_or = G.oracle_sampler(g)
    P, Q = s_or.sample(0.7)
    assert P.shape == (1250, 1250)
    W, V, gam = _eig_system(ctx, P, Q)
    PQ = P @ Q
    lam = gam * gam
    res = np.abs(PQ @ W - W * lam[None, :]).max() / np.abs(PQ).max()
    assert res < 1e-12 * PQ.shape[0], res
    assert np.allclose(gam, O.gamma_sqrt(lam), rtol=1e-12, atol=0)
    assert rel(V, (Q @ W) / gam[None, :]) < 1e-11
    assert spectrum_rel(lam, np.linalg.eigvals(PQ)) < 1e-11


def layer_set():
    rng = np.random.default_rng(3)
    g = G.metasurface(2)
    crossing = (np.array([0.0, 0.4, 0.2, 0.65]), np.array([0.0, 0.1, 0.5, 0.65]), g["eps"])
    pix = (np.linspace(0, 0.8, 9), np.linspace(0, 0.8, 9), 1 + 11.25 * rng.random((8, 8)) + 0j)
    rect = G.random_rect_layer(rng)
    v = G.var12(2, 2)
    return [(g["x"], g["y"], g["eps"]), crossing, pix, rect, (v["x"], v["y"], v["eps"])]


@pytest.mark.parametrize("nh", [(2, 2), (3, 1), (5, 5)])
def test_fourier_assembly_matches_oracle(ctx, nh):
    nx, ny = 2 * nh[0] + 1, 2 * nh[1] + 1
    layers = layer_set()
    fe = api.assemble_toeplitz(ctx, nx, ny, layers).cpu().numpy()
    for i, (x, y, eps) in enumerate(layers):
        for k, opt in enumerate([O.DISCONTINUOUS_X, O.DISCONTINUOUS_Y, O.CONTINUOUS_XY]):
            ref = O.fourier2d(x, y, eps, nx, ny, opt)
            got = fe[k, i].T  # column-major storage
            assert rel(got, ref) < 1e-12, (i, k, rel(got, ref))


def test_fourier_pixel_grids_fft_path(ctx):
    """Uniform power-of-two pixel grids take the FFT path (1D tables from a radix-2 FFT in LDS
    with the sinc/phase correction); other grids the direct sum. Against the oracle's direct
    sums at 1e-12: 64 x 64 (freeform layer), 256 x 256 (staircased pillar), 128 x 32 with a
    shifted origin, 48 x 16 (not a power of two along x: direct along x, FFT along y)."""
    rng = np.random.default_rng(11)

    def grid(gx, gy, x0=0.0, y0=0.0, Lx=0.8, Ly=0.8):
        return (x0 + Lx * np.arange(gx + 1) / gx, y0 + Ly * np.arange(gy + 1) / gy,
                1 + 11.25 * rng.random((gy, gx)) + 0.3j * rng.random((gy, gx)))

    layers = [grid(64, 64), grid(256, 256, Lx=1.0, Ly=1.0), grid(128, 32, x0=0.13, y0=-0.2, Lx=0.6, Ly=0.9),
              grid(48, 16)]
    for nx, ny in [(11, 11), (5, 9)]:
        fe = api.assemble_toeplitz(ctx, nx, ny, layers).cpu().numpy()
        for i, (x, y, eps) in enumerate(layers):
            for k, opt in enumerate([O.DISCONTINUOUS_X, O.DISCONTINUOUS_Y, O.CONTINUOUS_XY]):
                ref = O.fourier2d(x, y, eps, nx, ny, opt)
                got = fe[k, i].T
                assert rel(got, ref) < 1e-12, (nx, ny, i, k, rel(got, ref))


@pytest.mark.parametrize("which", ["metasurface", "var12"])
def test_pq_matches_oracle(ctx, which):
    g = G.metasurface(2) if which == "metasurface" else G.var12(2, 2)
    s_or = G.oracle_sampler(g)
    zs = [0.3, 1.5] if which == "metasurface" else [-0.1, 4.0]
    P, Q = [], []
    for z in zs:
        p, q = s_or.sample(z)
        P.append(p)
        Q.append(q)
    s_gpu = G.gpu_sampler(g, ctx)
    for z, p, q in zip(zs, P, Q):
        pg, qg = s_gpu.sample(z)
        assert rel(pg, p) < 1e-11, rel(pg, p)
        assert rel(qg, q) < 1e-12, rel(qg, q)
    Kx, Ky = s_gpu.Kx(), s_gpu.Ky()
    assert rel(Kx, s_or.Kx) < 1e-13 and rel(Ky, s_or.Ky) < 1e-13


def _eig_system(ctx, P, Q):
    dP, dQ = to_dev_colmajor(P[None]), to_dev_colmajor(Q[None])
    W, V, g, info = api.eig_pq(ctx, dP, dQ)
    ctx.sync()
    assert int(info.cpu()[0]) == 0
    return from_dev_colmajor(W)[0], from_dev_colmajor(V)[0], g.cpu().numpy()[0]


@pytest.mark.parametrize("which", ["metasurface", "var12"])
def test_eig_pq_properties(ctx, which):
    g = G.metasurface(5) if which == "metasurface" else G.var12(3, 3)
    s_or = G.oracle_sampler(g)
    P, Q = s_or.sample(0.7 if which == "metasurface" else 3.0)
    W, V, gam = _eig_system(ctx, P, Q)
    PQ = P @ Q
    lam = gam * gam
    res = np.abs(PQ @ W - W * lam[None, :]).max() / np.abs(PQ).max()
    assert res < 1e-12 * PQ.shape[0], res
    # branch rule (src/rcwa/WaveEquationSolver.cpp:12-21)
    assert np.all(gam.imag >= 0) or np.all((gam.imag >= 0) | ((np.abs(gam.imag) < np.abs(gam.real)) & (gam.real > 0)))
    assert np.allclose(gam, O.gamma_sqrt(lam), rtol=1e-12, atol=0)
    # V = Q W diag(1/gamma)
    assert rel(V, (Q @ W) / gam[None, :]) < 1e-12
    # eigenvalue multiset vs LAPACK
    ref = np.linalg.eigvals(PQ)
    assert spectrum_rel(lam, ref) < 1e-12


def test_layer_smatrix_and_star_product(ctx):
    g = G.metasurface(2)
    s_or = G.oracle_sampler(g)
    P1, Q1 = s_or.sample(2.0)
    P, Q = s_or.sample(1.3)
    W1, V1, g1 = _eig_system(ctx, P1, Q1)
    W, V, gm = _eig_system(ctx, P, Q)
    lam, dz = g["lam"], 0.25
    Tuu, Rud = api.layer_smatrix(ctx, lam, dz, to_dev_colmajor(W[None]), to_dev_colmajor(V[None]),
                                 ctx.torch.from_numpy(gm[None].copy()).cuda(), to_dev_colmajor(W1[None]),
                                 to_dev_colmajor(V1[None]))
    rt, rr = O.rcwa_scatter_matrix_from_eig(lam, dz, W, V, gm, W1, V1)
    assert rel(from_dev_colmajor(Tuu)[0], rt) < 1e-11
    assert rel(from_dev_colmajor(Rud)[0], rr) < 1e-11
    # star product of two generic S-matrices (batched, 2 entries)
    rng = np.random.default_rng(7)
    n = 20

    def rs(scale):
        return scale * (rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))) / np.sqrt(n)

    S = [np.eye(n) + rs(0.2), np.eye(n) + rs(0.2), rs(0.3), rs(0.3)]
    s = [np.eye(n) + rs(0.2), np.eye(n) + rs(0.2), rs(0.3), rs(0.3)]
    dS = ctx.torch.stack([to_dev_colmajor(m) for m in S])
    ds = ctx.torch.stack([to_dev_colmajor(m) for m in s])
    api.star_product(ctx, dS, ds)
    got = [from_dev_colmajor(dS[k]) for k in range(4)]
    for b in range(2):
        ref = O.star_product(tuple(m[b] for m in S), s[0][b], s[1][b], s[2][b], s[3][b])
        for k in range(4):
            assert rel(got[k][b], ref[k]) < 1e-12, (b, k)


def invariants_close(r_gpu, r_or, tol=TOL_S):
    a = O.invariant_transforms(r_gpu)
    b = O.invariant_transforms(r_or)
    for k in a:
        assert rel(a[k], b[k]) < tol, (k, rel(a[k], b[k]))


def gpu_result(it):
    return O.SResult(it.Tdd(), it.Tuu(), it.Rud(), it.Rdu(), it.W0(), it.V0(), it.gam0(), it.W1(), it.V1(),
                     it.gam1(), N=it.N(), nEigVal=it.nEigVal())


@pytest.mark.parametrize("which", ["metasurface", "var12"])
def test_rcwa_integrator_parity(ctx, which):
    g = G.metasurface(2) if which == "metasurface" else G.var12(2, 2)
    N = 8
    ref = O.rcwa_integrate(G.oracle_sampler(g), g["z0"], g["z1"], N)
    s = G.gpu_sampler(g, ctx)
    it = api.RCWAIntegrator(g["z0"], g["z1"], s)
    it.compute(N)
    r1 = gpu_result(it)
    invariants_close(r1, ref)
    it.compute_batched(N)
    r2 = gpu_result(it)
    invariants_close(r2, ref)
    # gamma sets agree with the oracle's (order-free)
    assert spectrum_rel(r1.gam1, ref.gam1) < 1e-11


def test_redheffer_parity_metasurface(ctx):
    g = G.metasurface(2)
    pe = 5e-2
    ref = O.redheffer_integrate(G.oracle_sampler(g), g["z0"], g["z1"], pe)
    s = G.gpu_sampler(g, ctx)
    it = api.RedhefferIntegrator(g["z0"], g["z1"], s)
    it.compute(0, 1000000, pe)
    tree = it.tree()
    assert len(tree) == len(ref.tree) and np.allclose(tree, ref.tree, rtol=0, atol=1e-14), (len(tree), len(ref.tree))
    assert it.N() == ref.N
    invariants_close(gpu_result(it), ref)


def test_differential_parity_metasurface(ctx):
    # well-conditioned regime of the 4th-order step (evanescent growth over one step
    # stays O(1)); at coarse steps the method itself amplifies rounding
    g = G.metasurface(1)
    N = 32
    ref = O.differential_integrate(G.oracle_sampler(g), g["z0"], g["z1"], N)
    s = G.gpu_sampler(g, ctx)
    it = api.DifferentialIntegrator(g["z0"], g["z1"], s)
    it.compute(N)
    invariants_close(gpu_result(it), ref, tol=1e-9)


def _stack_case(seed, nlayers=2):
    rng = np.random.default_rng(seed)
    cell = 0.6
    layers = [G.random_rect_layer(rng, cell) for _ in range(nlayers)]
    th = rng.uniform(0.05, 0.2, nlayers)
    return dict(lam=float(rng.uniform(0.8, 1.6)), layers=layers, thickness=th, bottom=G.homogeneous(cell, 2.1),
                top=G.homogeneous(cell, 1.0))


def test_batch_stack_parity(ctx):
    nx = ny = 3
    insts = [_stack_case(s) for s in range(3)]
    out = api.batch_stack_rcwa(ctx, nx, ny, insts, want_bases=True)
    S = from_dev_colmajor(out["S"])
    bases = {k: (from_dev_colmajor(out[k]) if not k.startswith("gam") else out[k].cpu().numpy())
             for k in ["W0", "V0", "gam0", "W1", "V1", "gam1"]}
    for i, inst in enumerate(insts):
        ze = np.concatenate([[0.0], np.cumsum(inst["thickness"])])
        so = O.StackLayerSampler(inst["lam"], nx, ny, inst["layers"], ze, inst["bottom"], inst["top"])
        ref = O.rcwa_integrate(so, ze[0], ze[-1], len(inst["layers"]), z_edges=ze)
        got = O.SResult(S[i, 0], S[i, 1], S[i, 2], S[i, 3], bases["W0"][i], bases["V0"][i], bases["gam0"][i],
                        bases["W1"][i], bases["V1"][i], bases["gam1"][i])
        invariants_close(got, ref)
    # the plain drop-in path over the same instance agrees with the batch
    inst = insts[0]
    ze = np.concatenate([[0.0], np.cumsum(inst["thickness"])])
    s = api.StackLayerSampler(inst["lam"], nx, ny, inst["layers"], ze, inst["bottom"], inst["top"], ctx=ctx)
    it = api.RCWAIntegrator(ze[0], ze[-1], s)
    it.compute_edges(ze)
    got = O.SResult(S[0, 0], S[0, 1], S[0, 2], S[0, 3], bases["W0"][0], bases["V0"][0], bases["gam0"][0],
                    bases["W1"][0], bases["V1"][0], bases["gam1"][0])
    invariants_close(gpu_result(it), got, tol=1e-11)


def test_batch_stack_large_N(ctx):
    """One stack at 25 x 25 harmonics (N = 1250, the deep-stack configuration's size): LU panels
    beyond the register path in the slab and star-product solves, against the oracle."""
    nx = ny = 25
    inst = _stack_case(7)
    out = api.batch_stack_rcwa(ctx, nx, ny, [inst], want_bases=True)
    S = from_dev_colmajor(out["S"])[0]
    b = {k: (from_dev_colmajor(out[k])[0] if not k.startswith("gam") else out[k].cpu().numpy()[0])
         for k in ["W0", "V0", "gam0", "W1", "V1", "gam1"]}
    ze = np.concatenate([[0.0], np.cumsum(inst["thickness"])])
    so = O.StackLayerSampler(inst["lam"], nx, ny, inst["layers"], ze, inst["bottom"], inst["top"])
    ref = O.rcwa_integrate(so, ze[0], ze[-1], len(inst["layers"]), z_edges=ze)
    got = O.SResult(S[0], S[1], S[2], S[3], b["W0"], b["V0"], b["gam0"], b["W1"], b["V1"], b["gam1"])
    invariants_close(got, ref)


def test_batch_stack_chunks_and_groups(ctx):
    """Launch chunks and cascade groups only regroup the batches: every instance's S is the
    same (to rounding) for one chunk, chunks of 2 in one group, and chunks of 2 in groups
    of 2 instances."""
    nx = ny = 3
    insts = [_stack_case(s) for s in range(7)]
    ref = from_dev_colmajor(api.batch_stack_rcwa(ctx, nx, ny, insts)["S"])
    try:
        for group in (128, 2):
            assert ctx.lib.vrc_tune_set(b"stack.cascade_group", group) == 0
            got = from_dev_colmajor(api.batch_stack_rcwa(ctx, nx, ny, insts, chunk=2)["S"])
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), group
    finally:
        ctx.lib.vrc_tune_set(b"stack.cascade_group", 512)  # the library's default


def test_batch_stack_full_launch_and_remainder(ctx):
    """A batch larger than one full eigensolver launch: at 7 x 7 harmonics (N = 98) 400
    instances of 3 slabs run one launch of 1024 slabs on the four-per-CU QR kernel and one of
    the remaining 176 on the one-per-CU kernel. Every instance equals the run with the
    one-per-CU kernel everywhere (to rounding), and the first and last instances (one in each
    launch) equal the oracle."""
    nx = ny = 7
    insts = [_stack_case(s, nlayers=3) for s in range(400)]
    lib = ctx.lib
    nst = 10  # stage timers of vrc_timing_read; the QR stage is the fourth
    ms, cnt, qs = (C.c_double * nst)(), (C.c_int * nst)(), (C.c_longlong * 6)()
    assert lib.vrc_timing_enable(ctx.h, 1) == 0
    try:
        lib.vrc_timing_read(ctx.h, ms, cnt, nst)  # reset
        lib.vrc_qr_stats(ctx.h, qs)
        out = api.batch_stack_rcwa(ctx, nx, ny, insts, want_bases=True)
        assert lib.vrc_timing_read(ctx.h, ms, cnt, nst) == 0 and lib.vrc_qr_stats(ctx.h, qs) == 0
    finally:
        lib.vrc_timing_enable(ctx.h, 0)
    assert (cnt[3], qs[0]) == (2, 1200), (cnt[3], qs[0])  # two QR launches, 1200 dense eigensolves
    S = from_dev_colmajor(out["S"])
    try:
        assert ctx.lib.vrc_tune_set(b"zgeev.qr_variant", 0) == 0
        S0 = from_dev_colmajor(api.batch_stack_rcwa(ctx, nx, ny, insts)["S"])
    finally:
        ctx.lib.vrc_tune_set(b"zgeev.qr_variant", -1)
    scale = np.abs(S0).reshape(len(insts), -1).max(axis=1)
    dev = np.abs(S - S0).reshape(len(insts), -1).max(axis=1) / scale
    assert dev.max() <= 1e-11, (int(dev.argmax()), float(dev.max()))
    bases = {k: (from_dev_colmajor(out[k]) if not k.startswith("gam") else out[k].cpu().numpy())
             for k in ["W0", "V0", "gam0", "W1", "V1", "gam1"]}
    for i in (0, len(insts) - 1):
        inst = insts[i]
        ze = np.concatenate([[0.0], np.cumsum(inst["thickness"])])
        so = O.StackLayerSampler(inst["lam"], nx, ny, inst["layers"], ze, inst["bottom"], inst["top"])
        ref = O.rcwa_integrate(so, ze[0], ze[-1], len(inst["layers"]), z_edges=ze)
        got = O.SResult(S[i, 0], S[i, 1], S[i, 2], S[i, 3], bases["W0"][i], bases["V0"][i], bases["gam0"][i],
                        bases["W1"][i], bases["V1"][i], bases["gam1"][i])
        invariants_close(got, ref)


def test_batch_stack_patterned_ends(ctx):
    """End media with several cells take the dense eigensolver path of the batch."""
    nx = ny = 3
    rng = np.random.default_rng(9)
    insts = []
    for s in range(2):
        c = _stack_case(s)
        c["bottom"] = G.random_rect_layer(rng, 0.6)
        insts.append(c)
    out = api.batch_stack_rcwa(ctx, nx, ny, insts, want_bases=True)
    S = from_dev_colmajor(out["S"])
    bases = {k: (from_dev_colmajor(out[k]) if not k.startswith("gam") else out[k].cpu().numpy())
             for k in ["W0", "V0", "gam0", "W1", "V1", "gam1"]}
    for i, inst in enumerate(insts):
        ze = np.concatenate([[0.0], np.cumsum(inst["thickness"])])
        so = O.StackLayerSampler(inst["lam"], nx, ny, inst["layers"], ze, inst["bottom"], inst["top"])
        ref = O.rcwa_integrate(so, ze[0], ze[-1], len(inst["layers"]), z_edges=ze)
        got = O.SResult(S[i, 0], S[i, 1], S[i, 2], S[i, 3], bases["W0"][i], bases["V0"][i], bases["gam0"][i],
                        bases["W1"][i], bases["V1"][i], bases["gam1"][i])
        invariants_close(got, ref)


def _efficiency_case(seed, cell=1.4, lam=0.9):
    rng = np.random.default_rng(seed)
    layers = [G.random_rect_layer(rng, cell) for _ in range(2)]
    return dict(lam=lam, layers=layers, thickness=np.array([0.13, 0.18]), bottom=G.homogeneous(cell, 2.1),
                top=G.homogeneous(cell, 1.0))


def _efficiency_check(ctx, nx, ny, insts, e_inc, tol=1e-10):
    import torch
    n = nx * ny
    out = api.batch_stack_rcwa(ctx, nx, ny, insts, want_bases=True)
    REF, TRN = api.stack_efficiencies(ctx, nx, ny, insts, out,
                                      torch.from_numpy(e_inc).to(f"cuda:{ctx.device}").contiguous())
    REF, TRN = REF.cpu().numpy(), TRN.cpu().numpy()
    for i, inst in enumerate(insts):
        ze = np.concatenate([[0.0], np.cumsum(inst["thickness"])])
        so = O.StackLayerSampler(inst["lam"], nx, ny, inst["layers"], ze, inst["bottom"], inst["top"])
        ref = O.rcwa_integrate(so, ze[0], ze[-1], len(inst["layers"]), z_edges=ze)
        cell = inst["bottom"][0][-1] - inst["bottom"][0][0]
        al, be = O.alpha_beta(nx, ny, cell, cell)
        r0, t0 = O.scatter_plane_wave(ref, al, be, 2 * O.PI / inst["lam"], inst["bottom"][2][0, 0],
                                      inst["top"][2][0, 0], e_inc[i])
        scale = max(np.abs(r0).max(), np.abs(t0).max())
        assert np.abs(REF[i] - r0).max() <= tol * scale, (i, np.abs(REF[i] - r0).max())
        assert np.abs(TRN[i] - t0).max() <= tol * scale, (i, np.abs(TRN[i] - t0).max())
    return REF, TRN


def test_stack_efficiencies(ctx):
    """Plane-wave diffraction efficiencies of batched stacks (RCWASolver::scatterPlaneWave)
    against the oracle's bookkeeping on the oracle's S-matrix, to 1e-10 of the largest
    efficiency: x- and y-polarised normally incident waves and a random field over all
    harmonics. The lossless stacks with several propagating orders keep the power balance
    for the plane waves; each side is normalised by its own medium's sqrt(Re eps) k0
    (src/rcwa/RCWASolver.cpp:1732-1733), so sum(REF) + sum(TRN) sqrt(eps_trn / eps_ref) = 1."""
    nx, ny = 5, 3
    n = nx * ny
    insts = [_efficiency_case(s) for s in range(3)]
    e = np.zeros((3, 2 * n), dtype=complex)
    c0 = (nx // 2) * ny + ny // 2
    e[0, c0] = 1.0
    e[1, n + c0] = 1.0
    rng = np.random.default_rng(5)
    e[2] = rng.standard_normal(2 * n) + 1j * rng.standard_normal(2 * n)
    REF, TRN = _efficiency_check(ctx, nx, ny, insts, e)
    for i in (0, 1):
        assert (TRN[i] > 1e-6).sum() >= 3  # several transmitted orders carry power
        bal = REF[i].sum() + TRN[i].sum() * np.sqrt(1.0 / 2.1)
        assert abs(bal - 1.0) < 1e-10, (i, bal)


def test_stack_efficiencies_patterned_bottom(ctx):
    """A patterned bottom medium gives a dense end basis W0: the incident coefficients then
    come from the LU solve W0 c = e_inc."""
    nx = ny = 3
    rng = np.random.default_rng(21)
    insts = []
    for s in range(2):
        c = _efficiency_case(s, cell=0.6, lam=1.1)
        c["bottom"] = G.random_rect_layer(rng, 0.6)
        insts.append(c)
    e = rng.standard_normal((2, 2 * nx * ny)) + 1j * rng.standard_normal((2, 2 * nx * ny))
    _efficiency_check(ctx, nx, ny, insts, e, tol=1e-9)


def test_modal_power_guided_modes(ctx):
    """The guided-mode benchmark of src/tests/test_cpu.cpp:416-507 at 7 x 5 harmonics: the
    guided modes of both ends (findGuidedModes), the power of every end mode on the driver's
    500 x 500 grid over [0.5, 6.7] x [0.5, 3.8] against the oracle's direct field sum
    (1e-10 of the largest |power|, 1e-9 relative for guided modes), the driver's normalised
    transmissions |Tuu(i,j)|^2 e1_i / e0_j, Parseval on the full cell, and a padded leading
    dimension."""
    import torch
    from varrcwa_amd._lib import ptr
    g = G.guided_benchmark(3, 2)
    nx, ny = g["nx"], g["ny"]
    n, N = nx * ny, 2 * nx * ny
    it = api.RCWAIntegrator(g["z0"], g["z1"], G.gpu_sampler(g, ctx))
    it.compute(6)
    k0 = 2 * O.PI / g["lam"]
    dev = f"cuda:{ctx.device}"
    win = (7.2, 4.3, 500, 500, 0.5, 6.7, 0.5, 3.8)
    energies, guided = [], []
    for W, V, gam in ((it.W0(), it.V0(), it.gam0()), (it.W1(), it.V1(), it.gam1())):
        rec = api.find_guided_modes(ctx, gam, k0, 0.2, 1.445, 3.48)
        assert rec == O.find_guided_modes(gam, k0, 0.2, 1.445, 3.48)
        assert len(rec) >= 1 and 1.445 < rec[0][1].real < 3.48
        Wd = torch.from_numpy(np.ascontiguousarray(W.T)).to(dev)
        Vd = torch.from_numpy(np.ascontiguousarray(V.T)).to(dev)
        e_gpu = api.modal_power(ctx, nx, ny, Wd, Vd, *win).cpu().numpy()
        e_ref = np.array([O.evaluate_energy(W[:, i], V[:, i], nx, ny, *win) for i in range(N)])
        assert np.abs(e_gpu - e_ref).max() <= 1e-10 * np.abs(e_ref).max()
        for i, _ in rec:
            assert abs(e_gpu[i] - e_ref[i]) <= 1e-9 * abs(e_ref[i]), (i, e_gpu[i], e_ref[i])
        # Parseval on the full cell (the sampled sum is exact there)
        full = api.modal_power(ctx, nx, ny, Wd, Vd, 7.2, 4.3, 2 * nx, 2 * ny, 0.0, 7.2, 0.0, 4.3).cpu().numpy()
        pars = 0.5 * 7.2 * 4.3 * np.real(np.sum(W[:n] * np.conj(V[n:]) - W[n:] * np.conj(V[:n]), axis=0))
        assert np.abs(full - pars).max() <= 1e-11 * np.abs(pars).max()
        # leading dimension larger than 2 nx ny
        ld = N + 3
        Wp = torch.zeros((N, ld), dtype=torch.complex128, device=dev)
        Vp = torch.zeros((N, ld), dtype=torch.complex128, device=dev)
        Wp[:, :N] = Wd
        Vp[:, :N] = Vd
        ep = torch.empty(N, dtype=torch.float64, device=dev)
        ctx.check(ctx.lib.vrc_modal_power(ctx.h, nx, ny, N, ptr(Wp), ld, ptr(Vp), ld, *[float(v) for v in win[:2]],
                                          win[2], win[3], *[float(v) for v in win[4:]], ptr(ep)), "modal_power")
        assert np.abs(ep.cpu().numpy() - e_gpu).max() <= 1e-14 * np.abs(e_gpu).max()
        energies.append(e_gpu)
        guided.append(rec)
    # the driver's normalised transmission table |Tuu(i,j)|^2 e1 / e0 over guided modes
    Tuu = it.Tuu()
    tab = np.array([[abs(Tuu[i, j]) ** 2 * energies[1][i] / energies[0][j] for j, _ in guided[0]]
                    for i, _ in guided[1]])
    e0 = {j: O.evaluate_energy(it.W0()[:, j], it.V0()[:, j], nx, ny, *win) for j, _ in guided[0]}
    e1 = {i: O.evaluate_energy(it.W1()[:, i], it.V1()[:, i], nx, ny, *win) for i, _ in guided[1]}
    ref = np.array([[abs(Tuu[i, j]) ** 2 * e1[i] / e0[j] for j, _ in guided[0]] for i, _ in guided[1]])
    assert np.abs(tab - ref).max() <= 1e-9 * np.abs(ref).max()
    assert np.all(np.isfinite(tab))


def _shape_parity(ctx, x, y, eps, nx, ny, lam, pml, para, npara, Kx, Ky):
    layer = (x, y, eps)
    out = api.wave_eqn_derivatives(ctx, nx, ny, layer, lam, pml, layer, para, npara)
    dP, dQ, dPl, dQl, dfe = O.wave_eqn_derivatives(x, y, eps, Kx, Ky, lam, nx, ny, para, npara)
    for name, ref in (("dP", dP), ("dQ", dQ), ("dfe11", dfe[0]), ("dfe22", dfe[1]), ("dfe33", dfe[2])):
        got = from_dev_colmajor(out[name])
        for p in range(npara):
            assert np.abs(got[p] - ref[p]).max() <= 1e-11 * np.abs(ref[p]).max(), (name, p)
    for name, ref in (("dPl", dPl), ("dQl", dQl)):
        got = from_dev_colmajor(out[name])
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), name


def test_wave_eqn_derivatives(ctx):
    """Shape (two-parameter Parameterization of x coordinates) and wavelength derivatives of
    fe11/fe22/fe33, P and Q against the oracle's restatement of the reference's tangent path
    (itself pinned to finite differences): periodic metasurface with complex permittivity
    at 5 x 3 harmonics, and the PML waveguide taper (dense K held fixed) at 5 x 5."""
    g = G.metasurface(1)
    x, y = g["x"].copy(), g["y"].copy()
    eps = g["eps"].copy()
    eps[0, 2], eps[2, 0] = 2.0 + 0.1j, 5.0
    a, b = O.alpha_beta(5, 3, x[-1] - x[0], y[-1] - y[0])
    _shape_parity(ctx, x, y, eps, 5, 3, 1.55, False, {2: [1.0], 1: [0.3, -0.7]}, 2, np.diag(a), np.diag(b))
    v = G.var12(2, 2)
    s = O.VarLayerSampler(v["lam"], v["nx"], v["ny"], v["x"], v["y"], v["eps"], True)
    _shape_parity(ctx, v["x"].copy(), v["y"], v["eps"], v["nx"], v["ny"], v["lam"], True,
                  {3: [1.0, 0.5], 4: [0.0, 1.0], 2: [0.25]}, 3, s.Kx, s.Ky)


def test_wave_eqn_derivatives_rejects(ctx):
    """Crossing coordinates (the reference's own derivative there is marked unfinished and is
    not a derivative of its forward map), boundary coordinates and repeated coordinates are
    refused with an error, before any kernel runs."""
    g = G.metasurface(1)
    y, eps = g["y"], g["eps"]
    xc = np.array([0.0, 0.4, 0.2, 0.65])
    for x, para in ((xc, {1: [1.0]}), (g["x"], {0: [1.0]}), (g["x"], {3: [1.0]})):
        with pytest.raises(api.VrcError):
            api.wave_eqn_derivatives(ctx, 3, 3, (x, y, eps), 1.55, False, (x, y, eps), para)


def test_gds_sampler(ctx, tmp_path):
    """GDSIISampler from a written file and from the reference's layout file: bounds and the
    cross-section layouts equal the oracle's exactly, P/Q at a few z and the RCWA S-matrix of
    the contour device (src/tests/test_gds_cpu.cpp:129-145 at 5 x 3 harmonics, 12 slices)
    against the oracle at the usual 1e-9 bar."""
    import os
    from tests.gds_files import synthetic_cells, write_gds
    f = tmp_path / "t.gds"
    write_gds(f, synthetic_cells())
    contour = os.path.join(os.path.dirname(__file__), "golden", "contour_extract.gds")
    kw = dict(shift=(2.3647, 0.0, 0.0), lPML=0.5, rPML=0.5, lspace=1.0, rspace=1.3)
    fac = 1e-9 / 1e-6
    for path, args, zs in ((str(f), dict(shift=(0.25, 0.1, 0.0)), (fac * -500, 0.1, 0.6999, fac * 700)),
                           (contour, kw, (-1.63, -1.0, 0.0, 0.7, 1.35))):
        g = api.GDSIISampler(1.55, 5, 3, path, ctx=ctx, **args)
        o = O.GDSIISampler(1.55, 5, 3, path, **args)
        assert (g.minX(), g.maxX(), g.minZ(), g.maxZ()) == (o.minX, o.maxX, o.minZ, o.maxZ)
        for z in zs:
            x, y, e = g.sampleLayout(z)
            xo, yo, eo = o.layer_at(z)
            assert np.array_equal(x, xo) and np.array_equal(y, yo) and np.array_equal(e, eo), z
            P, Q = g.sample(z)
            Po, Qo = o.sample(z)
            assert np.abs(P - Po).max() <= 1e-11 * np.abs(Po).max()
            assert np.abs(Q - Qo).max() <= 1e-11 * np.abs(Qo).max()
    g = api.GDSIISampler(1.55, 5, 3, contour, ctx=ctx, **kw)
    o = O.GDSIISampler(1.55, 5, 3, contour, **kw)
    it = api.RCWAIntegrator(-1.63, 1.36, g)
    it.compute(12)
    invariants_close(gpu_result(it), O.rcwa_integrate(o, -1.63, 1.36, 12))


def _legacy_layer(eps_core, shift=0.0, pml=0.1):
    x = np.array([0.0, pml, 0.3 + shift, 0.55 + shift, 1.0 - pml, 1.0])
    y = np.array([0.0, pml, 0.45, 0.55, 1.0 - pml, 1.0])
    e = np.full((5, 5), 2.1, dtype=complex)
    e[2, 2] = eps_core
    return (x, y, e)


def test_legacy_rcwa_solver(ctx):
    """The legacy stacked-layer solver against the oracle's RTCM restatement: a stack with
    repeated types (shared interface matrices) and a lossy layer, and the translation
    option; compared through W_last Tdd W_first^-1 and W_first Rud W_first^-1 with each
    side's own eigenvectors (1e-9), the layer spectra gamma (sorted), the transmittance
    helpers on the device result, and the addLayer region check."""
    nX, nY = 2, 1
    types = [_legacy_layer(2.1), _legacy_layer(12.0), _legacy_layer(4.0 + 0.2j), _legacy_layer(12.0, 0.05)]
    g = api.RCWASolver(nX, nY, ctx=ctx)
    o = O.RCWASolver(nX, nY)
    for L in types:
        g.addLayer(*L)
        o.addLayer(*L)
    cases = (([0, 1, 2, 1, 2, 0], [0.0, 0.37, 0.21, 0.37, 0.21, 0.0], None),
             ([0, 1, 2, 3, 0], [0.0, 0.3, 0.2, 0.25, 0.0], [0.0, 0.11, -0.07, 0.05, 0.0]))
    for stack, th, tr in cases:
        g.solve(1.3, stack, th, tr)
        To, Ro = o.solve(1.3, stack, th, tr)
        first, last = stack[0], stack[-1]
        Wg0, Wg1 = g.eigvecE(first), g.eigvecE(last)
        Xg = Wg1 @ g.Tdd() @ np.linalg.inv(Wg0)
        Yg = Wg0 @ g.Rud() @ np.linalg.inv(Wg0)
        Xo = o.E[last] @ To @ np.linalg.inv(o.E[first])
        Yo = o.E[first] @ Ro @ np.linalg.inv(o.E[first])
        assert np.abs(Xg - Xo).max() <= 1e-9 * np.abs(Xo).max(), stack
        assert np.abs(Yg - Yo).max() <= 1e-9 * np.abs(Yo).max(), stack
        for k in range(len(types)):
            assert spectrum_rel(g.gamma(k), o.gamma[k]) <= 1e-10
        T = g.Tdd()
        assert abs(g.transmittance(3) - np.sum(np.abs(T[:, 3]) ** 2)) < 1e-12
        assert abs(g.transmittance(3, 5) - abs(T[5, 3]) ** 2) < 1e-15
        assert g.tdd(5, 3) == T[5, 3]
    x, y, e = _legacy_layer(2.1)
    with pytest.raises(api.VrcError):
        g.addLayer(x * 1.1, y, e)


def test_legacy_intermediate_field(ctx):
    """RCWASolver::evaluateIntermediateField on the device against the oracle (pinned by
    interface continuity): the amplitudes depend on each side's eigenvector basis, so the
    comparison is on the fields they describe, E_type(j) c_m[j] and E_type(j) c_p[j] of every
    inner layer, for one incident harmonic field per column (input = E_first^-1 e in each
    side's own basis), 1e-9 of the largest; a unit input mode through the C++-style call."""
    nX, nY = 2, 1
    types = [_legacy_layer(2.1), _legacy_layer(12.0), _legacy_layer(4.0 + 0.2j)]
    g = api.RCWASolver(nX, nY, ctx=ctx)
    o = O.RCWASolver(nX, nY)
    for L in types:
        g.addLayer(*L)
        o.addLayer(*L)
    stack, th = [0, 1, 2, 1, 2, 0], [0.0, 0.37, 0.21, 0.15, 0.3, 0.0]
    g.solve(1.3, stack, th)
    o.solve(1.3, stack, th)
    N = 2 * (2 * nX + 1) * (2 * nY + 1)
    rng = np.random.default_rng(4)
    e = rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))
    cmg, cpg = g.evaluateIntermediateField(np.linalg.solve(g.eigvecE(stack[0]), e), len(stack))
    cmo, cpo = o.evaluateIntermediateField(np.linalg.solve(o.E[stack[0]], e))
    assert len(cmg) == len(cmo) == len(stack) - 2
    for j in range(len(stack) - 2):
        t = stack[j + 1]
        for a, b in ((g.eigvecE(t) @ cmg[j], o.E[t] @ cmo[j]), (g.eigvecE(t) @ cpg[j], o.E[t] @ cpo[j])):
            assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max(), (j, np.abs(a - b).max() / np.abs(b).max())
    # inputMode variant: the unit vector of mode 3 in the device's own basis
    cm3, cp3 = g.evaluateIntermediateField(3, len(stack))
    e3 = np.zeros(N, complex)
    e3[3] = 1.0
    cmv, cpv = g.evaluateIntermediateField(e3, len(stack))
    assert all(np.array_equal(a, b) for a, b in zip(cm3, cmv))
    with pytest.raises(api.VrcError):
        g.solve(1.3, [0, 1], [0.0, 0.0])
        g.evaluateIntermediateField(0, 2)


def test_edge_cases_and_argument_errors(ctx, tmp_path):
    """Empty batches and zero-sized requests return empty results; a zero-thickness slab is
    transparent (Tuu = Tdd = I, Rud = Rdu = 0 between equal media); malformed requests are
    refused with an error before any kernel runs."""
    import torch
    from varrcwa_amd._lib import ptr
    nx = ny = 3
    n, N = nx * ny, 2 * nx * ny
    dev = f"cuda:{ctx.device}"
    out = api.batch_stack_rcwa(ctx, nx, ny, [], want_bases=True)
    assert tuple(out["S"].shape) == (0, 4, N, N)
    REF, TRN = api.stack_efficiencies(ctx, nx, ny, [], out, torch.zeros((0, 2 * n), dtype=torch.complex128,
                                                                        device=dev))
    assert REF.numel() == 0 and TRN.numel() == 0
    # zero-thickness patterned slab between equal media
    inst = _stack_case(4)
    inst["thickness"] = np.array([0.0, 0.0])
    inst["bottom"] = inst["top"] = G.homogeneous(0.6, 1.0)
    S = from_dev_colmajor(api.batch_stack_rcwa(ctx, nx, ny, [inst])["S"])[0]
    eye = np.eye(N)
    assert np.abs(S[0] - eye).max() < 1e-11 and np.abs(S[1] - eye).max() < 1e-11
    assert np.abs(S[2]).max() < 1e-11 and np.abs(S[3]).max() < 1e-11
    # instances with different slab counts in one batch, even / too many harmonics
    bad = [_stack_case(0, 2), _stack_case(1, 3)]
    with pytest.raises(api.VrcError):
        api.batch_stack_rcwa(ctx, nx, ny, bad)
    for hx in (4, 65):
        with pytest.raises(api.VrcError):
            api.batch_stack_rcwa(ctx, hx, 3, [_stack_case(0)])
    # modal power: no modes; an empty sampling grid
    W = torch.zeros((0, N), dtype=torch.complex128, device=dev)
    assert api.modal_power(ctx, nx, ny, W, W, 1.0, 1.0, 10, 10, 0.0, 1.0, 0.0, 1.0).numel() == 0
    W1 = torch.eye(N, dtype=torch.complex128, device=dev)
    e = torch.empty(N, dtype=torch.float64, device=dev)
    assert ctx.lib.vrc_modal_power(ctx.h, nx, ny, N, ptr(W1), N, ptr(W1), N, 1.0, 1.0, 0, 10, 0.0, 1.0, 0.0, 1.0,
                                   ptr(e)) != 0
    assert ctx.lib.vrc_modal_power(ctx.h, nx, ny, N, ptr(W1), N - 1, ptr(W1), N, 1.0, 1.0, 4, 4, 0.0, 1.0, 0.0, 1.0,
                                   ptr(e)) != 0
    # no shape parameters: only the wavelength derivatives
    g = G.metasurface(1)
    lay = (g["x"], g["y"], g["eps"])
    d = api.wave_eqn_derivatives(ctx, 5, 3, lay, 1.55, False, lay, {})
    a, b = O.alpha_beta(5, 3, g["x"][-1] - g["x"][0], g["y"][-1] - g["y"][0])
    _, _, dPl, dQl, _ = O.wave_eqn_derivatives(g["x"], g["y"], g["eps"], np.diag(a), np.diag(b), 1.55, 5, 3, {})
    assert d["dP"].shape[0] == 0
    assert np.abs(from_dev_colmajor(d["dQl"]) - dQl).max() <= 1e-12 * np.abs(dQl).max()
    assert np.abs(from_dev_colmajor(d["dPl"]) - dPl).max() <= 1e-12 * np.abs(dPl).max()
    # GDSII: missing file; legacy solver: stack entry out of range, a one-entry stack
    with pytest.raises(api.VrcError):
        api.GDSIISampler(1.55, 3, 3, str(tmp_path / "missing.gds"), ctx=ctx)
    s = api.RCWASolver(1, 1, ctx=ctx)
    s.addLayer(*_legacy_layer(2.1))
    for stack, th in (([0, 1], [0.0, 0.0]), ([0], [0.0])):
        with pytest.raises(api.VrcError):
            s.solve(1.3, stack, th)


def test_homogeneous_closed_form(ctx):
    """Single-value end media: W = I, PQ W = W Gamma^2 and V = Q W / Gamma against the
    oracle's P and Q (the closed form of a homogeneous medium)."""
    nx, ny, lam = 3, 5, 1.3
    rng = np.random.default_rng(4)
    bottom = G.homogeneous(0.7, 2.1 + 0.05j)
    inst = dict(lam=lam, layers=[G.random_rect_layer(rng, 0.7)], thickness=[0.1], bottom=bottom,
                top=G.homogeneous(0.7, 1.0))
    out = api.batch_stack_rcwa(ctx, nx, ny, [inst], want_bases=True)
    W0 = from_dev_colmajor(out["W0"])[0]
    V0 = from_dev_colmajor(out["V0"])[0]
    g0 = out["gam0"].cpu().numpy()[0]
    ze = np.array([0.0, 0.1])
    P, Q = O.StackLayerSampler(lam, nx, ny, inst["layers"], ze, bottom, inst["top"]).sample(-1.0)
    assert np.array_equal(W0, np.eye(W0.shape[0]))
    PQ = P @ Q
    assert np.abs(PQ @ W0 - W0 * (g0 ** 2)[None, :]).max() <= 1e-13 * np.abs(PQ).max()
    assert np.abs(V0 - (Q @ W0) / g0[None, :]).max() <= 1e-13 * np.abs(V0).max()
    assert np.all(g0.imag >= 0)


def test_context_in_a_fresh_process():
    """A process whose first GPU use is the library's context (torch.cuda not touched
    before) creates it and runs a small product: the loader brings torch's HIP runtime up
    before the library registers its kernels."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import numpy as np\n"
            "from varrcwa_amd import api\n"
            "ctx = api.default_context()\n"
            "import torch\n"
            "a = torch.eye(3, dtype=torch.complex128, device='cuda')\n"
            "print('ok', ctx.h.value is not None)\n") % str(ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "ok True" in r.stdout
