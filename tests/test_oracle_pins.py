"""Pins of the CPU oracle against what mathematics fixes independently of it.

The reference ships no golden vectors or assertions (its drivers only print
residuals), so each oracle function is pinned here by a closed form, a brute-force
integral, a structural identity or a physical invariant chosen so that a plausible
mistake (dropped term, wrong sign or index, transposed operand) fails.

Pinned: fourier1d/fourier1d_vec (quadrature + DFT closed form), refill (painter's
order), toeplitz_dense/block_toeplitz_dense (convolution identities), fourier2d
(homogeneous, single-strip inverse rules, Laurent = separable product), k_matrices
periodic (closed form) and PML (zero-width limit; non-zero widths against quadrature of the
stretch profile), pml_solve (quadrature of the stretch profile), svd_solve (minimum-norm
least squares with the rank threshold), wave_eqn_coeff (homogeneous closed form), wave_equation_solve (eigen residual,
homogeneous eigenvalues), gamma_sqrt (branch table), rcwa_scatter_matrix
(homogeneous slab), star_product (identity, associativity), rcwa_integrate
(z-invariance, energy conservation), redheffer_integrate (single segment for a
z-invariant structure, convergence to RCWA), differential_integrate (convergence, and
fourth order: successive-difference ratio 16),
gdsii_real (known answers), read_gds_polygons / GDSIISampler (written files with known
coordinates, the reference's layout file),
scatter_plane_wave (Airy slab, power balance in equal and unequal media), evaluate_energy
(Parseval, closed-form partial window), find_guided_modes (known answer), RCWASolver (the
legacy RTCM recursion against the star-product cascade, homogeneous closed form,
translation = shifted geometry), dsolve_1d /
dsolve_v2_1d / dsolve_para / fourier2d_tangents / wave_eqn_derivatives (central differences
of the forward maps, monotone coordinates).
Each pin added for PML, the differential method and svd_solve is also run on planted
mistakes (the oracle's own source with one fragment changed), which it must reject.
Parity unpinned: the derivative formulas for crossing coordinates (the reference's own is
unfinished there, and the library refuses that case).
"""
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as O
from tests import geometries as G


def quad_coeff(xs, fs, d, m, npts=20000):
    """(1/d) * integral of the piecewise constant f(x) exp(-2 pi i m x / d) over each cell (midpoint rule)."""
    tot = 0j
    for j in range(len(fs)):
        a, b = xs[j], xs[j + 1]
        t = a + (np.arange(npts) + 0.5) * (b - a) / npts
        tot += fs[j] * np.sum(np.exp(-2j * math.pi * m * t / d)) * (b - a) / npts
    return tot / d


def test_fourier1d_matches_quadrature():
    xs = np.array([0.0, 0.13, 0.4, 0.41, 0.9])
    fs = np.array([1.0, 3.5 - 0.2j, 7.0, 2.0 + 1j])
    d = xs[-1] - xs[0]
    M = 9
    got = O.fourier1d(xs, d, fs, M)
    vec = O.fourier1d_vec(xs, d, fs, M)
    for i in range(M):
        m = i - (M - 1) // 2
        ref = quad_coeff(xs, fs, d, m)
        assert abs(got[i] - ref) < 1e-7, (m, got[i], ref)
    assert np.abs(got - vec).max() < 1e-14


def test_fourier1d_pixel_grid_is_scaled_dft():
    # uniform grid of Gp pixels on [x0, x0+d]: F_m = sinc(m/Gp) exp(-i pi m/Gp) exp(-2 pi i m x0/d) DFT(f)_m / Gp
    rng = np.random.default_rng(0)
    Gp, d, x0 = 16, 0.8, 0.1
    f = rng.random(Gp) + 1j * rng.random(Gp)
    xs = x0 + d * np.arange(Gp + 1) / Gp
    M = 11
    got = O.fourier1d(xs, d, f, M)
    dft = np.fft.fft(f)
    for i in range(M):
        m = i - (M - 1) // 2
        s = 1.0 if m == 0 else math.sin(math.pi * m / Gp) / (math.pi * m / Gp)
        ref = s * np.exp(-1j * math.pi * m / Gp) * np.exp(-2j * math.pi * m * x0 / d) * dft[m % Gp] / Gp
        assert abs(got[i] - ref) < 1e-13


def test_refill_painters_order():
    # coordinates cross: cell 1 (x=0.4..0.2) is inverted; sorted intervals take the
    # material of the last non-inverted original cell covering them
    x = np.array([0.0, 0.4, 0.2, 0.65])
    index, rindex = O._rindex(x)
    assert list(index) == [0, 2, 1, 3]
    f = np.array([1.0, 2.0, 3.0], dtype=complex)
    newf = O.refill(x, f, rindex)
    # sorted intervals: [0,0.2] (cell 0 covers 0..0.4), [0.2,0.4] (cell 0, then cell 2 covers 0.2..0.65), [0.4,0.65] (cell 2)
    assert list(newf) == [1.0, 3.0, 3.0]
    # no crossing: identity
    assert list(O.refill(np.array([0.0, 0.1, 0.5, 0.65]), f, np.arange(4))) == [1.0, 2.0, 3.0]


def test_toeplitz_is_convolution():
    # Laurent rule: the coefficients of f*g equal T_f applied to the coefficients of g
    rng = np.random.default_rng(1)
    K = 3
    n = 2 * K + 1
    fc = rng.standard_normal(2 * n - 1) + 1j * rng.standard_normal(2 * n - 1)  # orders -2K..2K
    gc = rng.standard_normal(n) + 1j * rng.standard_normal(n)                    # orders -K..K
    T = O.toeplitz_dense(fc)
    prod = np.convolve(fc, gc)                                                   # orders -3K..3K
    centre = (prod.size - 1) // 2
    ref = prod[centre - K:centre + K + 1]
    assert np.abs(T @ gc - ref).max() < 1e-12


def test_block_toeplitz_index_map():
    rng = np.random.default_rng(2)
    nx, ny = 3, 2
    mat = rng.standard_normal((2 * nx - 1, 2 * ny - 1))
    B = O.block_toeplitz_dense(mat)
    for r in range(nx * ny):
        for c in range(nx * ny):
            m, n = divmod(r, ny)
            j, l = divmod(c, ny)
            assert B[r, c] == mat[nx - 1 + m - j, ny - 1 + n - l]


def test_fourier2d_homogeneous_and_inverse_rules():
    nx, ny = 5, 3
    x, y, eps = G.homogeneous(0.7, 4.0 + 0.1j)
    for opt in (O.DISCONTINUOUS_X, O.DISCONTINUOUS_Y, O.CONTINUOUS_XY):
        F = O.fourier2d(x, y, eps, nx, ny, opt)
        assert np.abs(F - (4.0 + 0.1j) * np.eye(nx * ny)).max() < 1e-13
    # one strip in y: inverse rule along x is inv(Toep(coeff(1/eps))) (x) I_y
    x = np.array([0.0, 0.2, 0.5, 1.0])
    y = np.array([0.0, 0.8])
    e = np.array([[2.0, 9.0, 1.0]], dtype=complex)
    F11 = O.fourier2d(x, y, e, nx, ny, O.DISCONTINUOUS_X)
    inv = np.linalg.inv(O.toeplitz_dense(O.fourier1d(x, 1.0, 1.0 / e[0], 2 * nx - 1)))
    assert np.abs(F11 - np.kron(inv, np.eye(ny))).max() < 1e-12
    # one cell in x: inverse rule along y is I_x (x) inv(Toep(coeff_y(1/eps)))
    x1 = np.array([0.0, 1.0])
    y1 = np.array([0.0, 0.3, 0.8])
    e1 = np.array([[2.0], [6.0]], dtype=complex)
    F22 = O.fourier2d(x1, y1, e1, nx, ny, O.DISCONTINUOUS_Y)
    invy = np.linalg.inv(O.toeplitz_dense(O.fourier1d(y1, 0.8, 1.0 / e1[:, 0], 2 * ny - 1)))
    assert np.abs(F22 - np.kron(np.eye(nx), invy)).max() < 1e-12
    # Laurent rule of a separable (rectangle) layer equals the product of 1D Laurent matrices
    x2 = np.array([0.0, 0.3, 0.6, 1.0])
    y2 = np.array([0.0, 0.2, 0.5, 0.8])
    ex = np.array([1.0, 5.0, 1.0])
    ey = np.array([1.0, 3.0, 1.0])
    # eps = 1 + (ex-1)(ey-1)/... use indicator product: eps = 1 + 8 * chi_x * chi_y
    chi = np.outer([0, 1, 0], [0, 1, 0]).astype(complex)
    e2 = 1.0 + 8.0 * chi
    F33 = O.fourier2d(x2, y2, e2, nx, ny, O.CONTINUOUS_XY)
    cx = O.toeplitz_dense(O.fourier1d(x2, 1.0, np.array([0, 1, 0], complex), 2 * nx - 1))
    cy = O.toeplitz_dense(O.fourier1d(y2, 0.8, np.array([0, 1, 0], complex), 2 * ny - 1))
    ref = np.eye(nx * ny) + 8.0 * np.kron(cx, cy)
    assert np.abs(F33 - ref).max() < 1e-12
    assert ex.size == ey.size


def test_k_matrices_periodic_and_pml_limits():
    nx, ny = 3, 5
    x = np.array([0.0, 0.5, 2.0, 2.5])
    y = np.array([0.0, 0.4, 1.6, 2.0])
    reg = dict(Lx=2.5, Ly=2.0, b0=0.0, b1=0.4, u0=1.6, u1=2.0, l0=0.0, l1=0.5, r0=2.0, r1=2.5)
    Kx, Ky = O.k_matrices(nx, ny, reg, False)
    a, b = O.alpha_beta(nx, ny, 2.5, 2.0)
    assert np.allclose(np.diag(Kx), a) and np.allclose(np.diag(Ky), b)
    assert a[0] == -2 * O.PI / 2.5 and b[1] == -2 * O.PI / 2.0 * 1
    # PML with zero-width absorbing regions is the unstretched identity: K = diag(alpha)
    reg0 = dict(Lx=2.5, Ly=2.0, b0=0.0, b1=0.0, u0=2.0, u1=2.0, l0=0.0, l1=0.0, r0=2.5, r1=2.5)
    Kx0, Ky0 = O.k_matrices(nx, ny, reg0, True)
    assert np.abs(Kx0 - np.diag(a)).max() < 1e-12 and np.abs(Ky0 - np.diag(b)).max() < 1e-12


def test_homogeneous_operator_eigensystem_and_slab():
    nx = ny = 3
    eps = 2.25
    lam = 1.3
    x, y, e = G.homogeneous(0.9, eps)
    s = O.StackLayerSampler(lam, nx, ny, [(x, y, e)], [0.0, 1.0], (x, y, e), (x, y, e))
    P, Q = s.sample(0.5)
    k0 = 2 * O.PI / lam
    a, b = O.alpha_beta(nx, ny, 0.9, 0.9)
    n = nx * ny
    Pref = np.block([[np.diag(a * b / eps), np.diag(k0 ** 2 - a * a / eps)],
                     [np.diag(b * b / eps - k0 ** 2), np.diag(-b * a / eps)]])
    Qref = np.block([[np.diag(-a * b), np.diag(a * a - k0 ** 2 * eps)],
                     [np.diag(-b * b + k0 ** 2 * eps), np.diag(b * a)]])
    assert np.abs(P - Pref).max() < 1e-11 and np.abs(Q - Qref).max() < 1e-11
    W, V, gam, lamv = O.wave_equation_solve(P, Q)
    kz2 = k0 ** 2 * (k0 ** 2 * eps - a * a - b * b)
    assert np.abs(np.sort_complex(lamv) - np.sort_complex(np.concatenate([kz2, kz2]))).max() < 1e-9
    # slab of the same medium in its own basis: T = X, R = 0
    dz = 0.37
    Tuu, Rud = O.rcwa_scatter_matrix(lam, dz, P, Q, W, V)
    X = np.exp(1j * gam * dz / k0)
    assert np.abs(Tuu - np.diag(X)).max() < 1e-12 and np.abs(Rud).max() < 1e-12
    assert n == 9


def test_gamma_sqrt_branch_table():
    vals = np.array([4.0, -4.0, 3 - 4j, -3 + 4j, -3 - 4j, 1e-3 - 1j])
    got = O.gamma_sqrt(vals)
    ref = np.array([2.0, 2j, 2 - 1j, 1 + 2j, -1 + 2j, None], dtype=object)
    for g, r in zip(got[:5], ref[:5]):
        assert abs(g - r) < 1e-15
    assert got[5].imag >= 0 or (abs(got[5].imag) < abs(got[5].real) and got[5].real > 0)
    assert np.allclose(got * got, vals)


def _rand_S(rng, n, scale=0.3):
    def r():
        return scale * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    return (np.eye(n) + r(), np.eye(n) + r(), r(), r())


def test_star_product_identity_and_associativity():
    rng = np.random.default_rng(4)
    n = 6
    I = (np.eye(n, dtype=complex), np.eye(n, dtype=complex), np.zeros((n, n), complex), np.zeros((n, n), complex))
    A, B, C = _rand_S(rng, n), _rand_S(rng, n), _rand_S(rng, n)
    AI = O.star_product(A, *I)
    assert all(np.abs(p - q).max() < 1e-13 for p, q in zip(AI, A))
    IA = O.star_product(I, *A)
    assert all(np.abs(p - q).max() < 1e-13 for p, q in zip(IA, A))
    ABC1 = O.star_product(O.star_product(A, *B), *C)
    BC = O.star_product(B, *C)
    ABC2 = O.star_product(A, *BC)
    assert all(np.abs(p - q).max() < 1e-12 for p, q in zip(ABC1, ABC2))


def _invariant_sampler():
    g = G.metasurface(1)
    s = O.VarLayerSampler(g["lam"], g["nx"], g["ny"], g["x"], g["y"], g["eps"], False)
    return s


def test_rcwa_z_invariant_slicing_and_single_varrcwa_segment():
    s = _invariant_sampler()
    r1 = O.rcwa_integrate(s, 0.0, 0.7, 1)
    r4 = O.rcwa_integrate(s, 0.0, 0.7, 4)
    a, b = O.invariant_transforms(r1), O.invariant_transforms(r4)
    for k in a:
        assert np.abs(a[k] - b[k]).max() < 1e-10 * max(1.0, np.abs(a[k]).max()), k
    rv = O.redheffer_integrate(s, 0.0, 0.7, 1e-3)
    assert rv.N == 1 and len(rv.tree) == 1
    c = O.invariant_transforms(rv)
    for k in a:
        assert np.abs(a[k] - c[k]).max() < 1e-10 * max(1.0, np.abs(a[k]).max()), k


def flux_gram(W, V):
    """Hermitian form G with flux(a) = a^H G a for the field (E, H) = (W a, V a):
    flux = 1/2 Re sum_h (Ex conj(Hy) - Ey conj(Hx)) (evaluateEnergy, src/tests/test_cpu.cpp:337-412,
    integrated over the cell by Parseval; the cell area is a common factor and dropped)."""
    n = W.shape[0] // 2
    M = V[n:].conj().T @ W[:n] - V[:n].conj().T @ W[n:]
    return 0.25 * (M + M.conj().T)


def energy_balance(r, k0, tol):
    """For each propagating incident mode j from below: incident - reflected = transmitted flux."""
    G0 = flux_gram(r.W0, r.V0)
    G1 = flux_gram(r.W1, r.V1)
    p0 = np.abs(r.gam0.imag) < 1e-9 * k0 ** 2
    p1 = np.abs(r.gam1.imag) < 1e-9 * k0 ** 2
    assert p0.sum() >= 2 and p1.sum() >= 2
    G0p, G1p = G0[np.ix_(p0, p0)], G1[np.ix_(p1, p1)]
    for j in np.nonzero(p0)[0]:
        rr = r.Rdu[p0, j]
        tt = r.Tuu[p1, j]
        inc = G0[j, j].real
        out = (tt.conj() @ G1p @ tt).real + (rr.conj() @ G0p @ rr).real
        assert abs(out - inc) < tol * abs(inc), (j, out, inc)


def test_energy_conservation_lossless_stack():
    # lossless periodic stack between glass and air: for each propagating incident mode,
    # transmitted + reflected propagating flux = incident flux
    rng = np.random.default_rng(5)
    nx = ny = 3
    cell = 0.6
    layers = [G.random_rect_layer(rng, cell) for _ in range(2)]
    ze = np.array([0.0, 0.11, 0.23])
    lam = 1.1
    s = O.StackLayerSampler(lam, nx, ny, layers, ze, G.homogeneous(cell, 2.1), G.homogeneous(cell, 1.0))
    r = O.rcwa_integrate(s, ze[0], ze[-1], 2, z_edges=ze)
    energy_balance(r, 2 * O.PI / lam, 1e-10)
    # a planted mistake (dropped reflection) breaks the balance
    bad = O.SResult(r.Tdd, r.Tuu, r.Rud, 0 * r.Rdu, r.W0, r.V0, r.gam0, r.W1, r.V1, r.gam1)
    with pytest.raises(AssertionError):
        energy_balance(bad, 2 * O.PI / lam, 1e-10)


def _plane_wave_case(nx, ny, layers, ze, lam, bottom, top, pol):
    cell = bottom[0][-1] - bottom[0][0]
    s = O.StackLayerSampler(lam, nx, ny, layers, ze, bottom, top)
    r = O.rcwa_integrate(s, ze[0], ze[-1], len(layers), z_edges=ze)
    al, be = O.alpha_beta(nx, ny, cell, cell)
    n = nx * ny
    e = np.zeros(2 * n, dtype=complex)
    e[(nx // 2) * ny + ny // 2 + (0 if pol == "x" else n)] = 1.0
    return O.scatter_plane_wave(r, al, be, 2 * O.PI / lam, bottom[2][0, 0], top[2][0, 0], e)


def test_plane_wave_homogeneous_slab_airy():
    """A homogeneous slab (index ns, thickness d) in air at normal incidence: only the
    (0,0) order carries power, T = 1 / (1 + F sin^2(k0 ns d)) with F = ((ns^2 - 1) / (2 ns))^2
    (Airy), and R = 1 - T."""
    nx = ny = 3
    cell, lam, d, eps = 0.7, 1.3, 0.37, 5.2
    slab = G.homogeneous(cell, eps)
    for pol in ("x", "y"):
        REF, TRN = _plane_wave_case(nx, ny, [slab], np.array([0.0, d]), lam, G.homogeneous(cell, 1.0),
                                    G.homogeneous(cell, 1.0), pol)
        c = (nx // 2) * ny + ny // 2
        ns, k0 = np.sqrt(eps), 2 * O.PI / lam
        F = ((ns * ns - 1) / (2 * ns)) ** 2
        T = 1.0 / (1.0 + F * np.sin(k0 * ns * d) ** 2)
        assert abs(TRN[c] - T) < 1e-12 and abs(REF[c] - (1 - T)) < 1e-12, (pol, TRN[c], T)
        others = np.ones(nx * ny, bool)
        others[c] = False
        assert np.abs(TRN[others]).max() < 1e-20 and np.abs(REF[others]).max() < 1e-20


def test_plane_wave_energy_conservation():
    """Lossless patterned stack between equal media: the efficiencies of all orders
    (propagating ones carry power, evanescent ones have Re Kz = 0) sum to 1; dropping Ez
    from the bookkeeping breaks the balance (oblique diffraction orders carry it)."""
    rng = np.random.default_rng(11)
    nx = ny = 5
    cell, lam = 1.4, 0.9
    layers = [G.random_rect_layer(rng, cell) for _ in range(2)]
    ze = np.array([0.0, 0.13, 0.31])
    air = G.homogeneous(cell, 1.0)
    for pol in ("x", "y"):
        REF, TRN = _plane_wave_case(nx, ny, layers, ze, lam, air, air, pol)
        assert (TRN > 1e-6).sum() >= 3  # several propagating diffraction orders
        assert abs(REF.sum() + TRN.sum() - 1.0) < 1e-10, (pol, REF.sum() + TRN.sum())
    # planted mistake: Ez left out
    s = O.StackLayerSampler(lam, nx, ny, layers, ze, air, air)
    r = O.rcwa_integrate(s, ze[0], ze[-1], 2, z_edges=ze)
    al, be = O.alpha_beta(nx, ny, cell, cell)
    n = nx * ny
    e = np.zeros(2 * n, dtype=complex)
    e[(nx // 2) * ny + ny // 2] = 1.0
    k0 = 2 * O.PI / lam
    txy = r.W1 @ (r.Tuu @ np.linalg.solve(r.W0, e))
    rxy = r.W0 @ (r.Rdu @ np.linalg.solve(r.W0, e))
    kz = np.sqrt(k0 * k0 - al * al - be * be)
    noz = sum((np.abs(kz.real) / k0 * (np.abs(v[:n]) ** 2 + np.abs(v[n:]) ** 2)).sum() for v in (rxy, txy))
    assert abs(noz - 1.0) > 1e-5  # far outside the 1e-10 balance above


def test_plane_wave_unequal_media_normalisation():
    """Incident from a substrate (eps 2.1) into air: the reference divides each side by its
    own medium's sqrt(Re eps) k0 (src/rcwa/RCWASolver.cpp:1732-1733), so the physical balance
    is sum(REF) + sum(TRN) sqrt(eps_trn / eps_ref) = 1; dividing TRN by the incident medium's
    index instead (the plausible slip) misses it by the index ratio."""
    rng = np.random.default_rng(12)
    nx, ny = 5, 3
    cell, lam = 1.4, 0.9
    layers = [G.random_rect_layer(rng, cell) for _ in range(2)]
    ze = np.array([0.0, 0.13, 0.31])
    for pol in ("x", "y"):
        REF, TRN = _plane_wave_case(nx, ny, layers, ze, lam, G.homogeneous(cell, 2.1), G.homogeneous(cell, 1.0),
                                    pol)
        assert abs(REF.sum() + TRN.sum() * np.sqrt(1.0 / 2.1) - 1.0) < 1e-10
        assert abs(REF.sum() + TRN.sum() / np.sqrt(2.1) * np.sqrt(1.0 / 2.1) - 1.0) > 1e-2


def test_evaluate_energy_parseval():
    """Full-cell window with more sample points than harmonic differences: the midpoint sums
    of the wavelet products are exactly orthogonal, so the sampled power is
    0.5 Lx Ly Re sum(Ex Hy* - Ey Hx*) over the harmonics (Parseval); a dropped 1/2, dx dy or
    a swapped component fails it."""
    rng = np.random.default_rng(3)
    nx, ny, Lx, Ly = 5, 3, 1.3, 0.7
    n = nx * ny
    for nres in ((5, 3), (7, 4), (16, 9)):
        e = rng.standard_normal(2 * n) + 1j * rng.standard_normal(2 * n)
        h = rng.standard_normal(2 * n) + 1j * rng.standard_normal(2 * n)
        got = O.evaluate_energy(e, h, nx, ny, Lx, Ly, nres[0], nres[1], 0.0, Lx, 0.0, Ly)
        want = 0.5 * Lx * Ly * np.real(np.sum(e[:n] * np.conj(h[n:]) - e[n:] * np.conj(h[:n])))
        assert abs(got - want) < 1e-12 * abs(want), (nres, got, want)
        swapped = 0.5 * Lx * Ly * np.real(np.sum(e[:n] * np.conj(h[:n]) - e[n:] * np.conj(h[n:])))
        assert abs(swapped - want) > 1e-3 * abs(want)


def test_evaluate_energy_partial_window_closed_form():
    """Ex only in the harmonic (m, n) = (1, 0), Hy only in (-1, 0), window partial in x and
    full in y: the power is 0.5 dx Ly Re(a b* sum_p exp(-i 4 pi x_p / Lx)), a geometric sum.
    This fixes the harmonic layout (index j + i ny, i the x-harmonic): the transposed
    layout moves both harmonics onto the y axis, where the full window cancels the term."""
    nx, ny, Lx, Ly = 5, 3, 1.3, 0.7
    n = nx * ny
    a, b = 0.7 - 0.4j, -0.3 + 1.1j
    x0, x1, P = 0.1 * Lx, 0.45 * Lx, 37

    def case(idx_e, idx_h):
        e = np.zeros(2 * n, complex)
        h = np.zeros(2 * n, complex)
        e[idx_e] = a
        h[n + idx_h] = b
        return O.evaluate_energy(e, h, nx, ny, Lx, Ly, P, 11, x0, x1, 0.0, Ly)

    i_p1, i_m1, j0 = (nx - 1) // 2 + 1, (nx - 1) // 2 - 1, (ny - 1) // 2
    got = case(j0 + i_p1 * ny, j0 + i_m1 * ny)
    s, th = (x1 - x0) / P, 4 * O.PI / Lx
    geo = np.exp(-1j * th * (x0 + s / 2)) * (1 - np.exp(-1j * th * s * P)) / (1 - np.exp(-1j * th * s))
    want = 0.5 * s * Ly * np.real(a * np.conj(b) * geo)
    assert abs(want) > 1e-2 and abs(got - want) < 1e-13, (got, want)
    transposed = case(i_p1 + j0 * nx, i_m1 + j0 * nx)
    assert abs(transposed - want) > 1e-2


def test_find_guided_modes_known_answer():
    """Hand-worked case: gamma / k0^2 = 2 + 0.1i (kept), 3 + 0.3i (imaginary part too large),
    1.2 (below the range), 2.5 - 0.01i (kept), 3.48 (range end, kept): sorted by real part."""
    k0 = 2 * O.PI / 1.55
    g = k0 * k0 * np.array([2 + 0.1j, 3 + 0.3j, 1.2, 2.5 - 0.01j, 3.48])
    rec = O.find_guided_modes(g, k0, 0.2, 1.445, 3.48)
    assert [r[0] for r in rec] == [4, 3, 0]
    assert abs(rec[1][1] - (2.5 - 0.01j)) < 1e-14


def _fd_shape_case(x, y, eps, nx, ny, lam, Kx, Ky, para, npara):
    dP, dQ, dPl, dQl, dfe = O.wave_eqn_derivatives(x, y, eps, Kx, Ky, lam, nx, ny, para, npara)
    h = 1e-6
    for p in range(npara):
        xp, xm = x.copy(), x.copy()
        for k, v in para.items():
            c = v[p] if p < len(v) else 0.0
            xp[k] += h * c
            xm[k] -= h * c
        Pp, Qp, fp = O.wave_eqn_coeff(xp, y, eps, Kx, Ky, lam, nx, ny, "lu", return_fe=True)
        Pm, Qm, fm = O.wave_eqn_coeff(xm, y, eps, Kx, Ky, lam, nx, ny, "lu", return_fe=True)
        for r in range(3):
            fd = (fp[r] - fm[r]) / (2 * h)
            assert np.abs(dfe[r][p] - fd).max() < 1e-7 * np.abs(fd).max(), (p, r)
        for got, a, b in ((dP[p], Pp, Pm), (dQ[p], Qp, Qm)):
            fd = (a - b) / (2 * h)
            assert np.abs(got - fd).max() < 1e-7 * np.abs(fd).max(), p
    Pp, Qp = O.wave_eqn_coeff(x, y, eps, Kx, Ky, lam + h, nx, ny, "lu")
    Pm, Qm = O.wave_eqn_coeff(x, y, eps, Kx, Ky, lam - h, nx, ny, "lu")
    assert np.abs(dPl - (Pp - Pm) / (2 * h)).max() < 1e-7 * np.abs(dPl).max()
    assert np.abs(dQl - (Qp - Qm) / (2 * h)).max() < 1e-7 * np.abs(dQl).max()
    return dP, dfe


def test_shape_derivatives_match_finite_differences():
    """The tangents of FourierSolver1D::dSolve/dSolve_v2, FourierSolver2D::solve's dFfs and
    Layer::waveEqnCoeff's dP/dQ/dPl/dQl are derivatives of the (separately pinned) forward
    maps: central differences (h = 1e-6) of fe11/fe22/fe33, P and Q in a two-parameter
    Parameterization (one coordinate in both parameters, coefficients missing = 0) and in
    lambda agree to 1e-7 relative, periodic and with dense PML K held fixed; a sign slip
    in d inv(T) = -inv(T) dT inv(T) would fail (planted below)."""
    g = G.metasurface(1)
    x, y = g["x"].copy(), g["y"].copy()
    eps = g["eps"].copy()
    eps[0, 2], eps[2, 0] = 2.0 + 0.1j, 5.0
    nx, ny = 5, 3
    a, b = O.alpha_beta(nx, ny, x[-1] - x[0], y[-1] - y[0])
    para = {2: [1.0], 1: [0.3, -0.7]}
    dP, dfe = _fd_shape_case(x, y, eps, nx, ny, 1.55, np.diag(a), np.diag(b), para, 2)
    v = G.var12(2, 1)
    s = O.VarLayerSampler(v["lam"], v["nx"], v["ny"], v["x"], v["y"], v["eps"], True)
    _fd_shape_case(v["x"].copy(), v["y"], v["eps"], v["nx"], v["ny"], v["lam"], s.Kx, s.Ky, {3: [1.0, 0.5], 4: [0.0, 1.0]},
                   2)
    # planted mistake: the inverse-rule tangent with the sign of d inv(T) flipped
    fe11_p, fe11_m = (O.fourier2d(x + dx, y, eps, nx, ny, O.DISCONTINUOUS_X) for dx in
                      (np.array([0, 0, 1e-6, 0]), np.array([0, 0, -1e-6, 0])))
    fd = (fe11_p - fe11_m) / 2e-6
    assert np.abs(-dfe[0][0] - fd).max() > 1e-2 * np.abs(fd).max()


def test_gdsii_real_known_answers():
    """The GDSII 8-byte real: 1.0 = 0x4110000000000000 and -2.5 = 0xC128000000000000 exactly
    (mantissa 1/16 and 5/32 at exponent 65), and the database units most layout tools write,
    0x3E4189374BC6A7EF = 1e-3 and 0x3944B82FA09B5A53 = 1e-9."""
    assert O.gdsii_real(0x4110000000000000) == 1.0
    assert O.gdsii_real(0xC128000000000000) == -2.5
    assert abs(O.gdsii_real(0x3E4189374BC6A7EF) - 1e-3) < 1e-18
    assert abs(O.gdsii_real(0x3944B82FA09B5A53) - 1e-9) < 1e-24


def test_gds_reader_and_cross_sections(tmp_path):
    """A written file with known integer coordinates (1 nm database unit, read in um):
    exact polygon points with the closing point dropped, boxes read as polygons, paths and
    the second cell left out; cross-sections: the rectangle's sides, the trapezoid's sloped
    edge by linear interpolation, the half-open rule y0 <= z < y1 at the top edge, and the
    device material in every other x cell of the device strip."""
    from tests.gds_files import synthetic_cells, write_gds
    f = tmp_path / "t.gds"
    write_gds(f, synthetic_cells())
    cells = O.read_gds_polygons(str(f))
    assert len(cells) == 2 and len(cells[0]) == 3
    fac = 1e-9 / 1e-6  # db_in_meters / unit, the reference's scale factor
    assert cells[0][0] == [(fac * 1000, fac * -500), (fac * 1600, fac * -500), (fac * 1600, fac * 700),
                           (fac * 1000, fac * 700)]
    assert cells[0][2] == [(fac * 3000, fac * -200), (fac * 3300, fac * -200), (fac * 3300, fac * 200),
                           (fac * 3000, fac * 200)]
    s = O.GDSIISampler(1.55, 3, 3, str(f), shift=(0.25, 0.1, 0.0))
    assert np.allclose([s.minX, s.maxX, s.minZ, s.maxZ], [1.25, 3.55, -0.5, 0.7], atol=1e-14, rtol=0)
    x, y, eps = s.layer_at(0.1)
    # sloped edge (2.6, -0.4) -> (2.3, 0.6) at z = 0.1: a = 0.5, x = 2.45
    want = [1.0, 1.6, 2.0, 2.45, 3.0, 3.3]
    assert np.allclose(x[2:-2], np.array(want) + 0.25, atol=1e-12, rtol=0)
    assert np.allclose(x[:2], [1.25 - 1.5 - 0.5, 1.25 - 1.5]) and np.allclose(x[-2:], [3.55 + 1.5, 3.55 + 2.0])
    assert np.allclose(y, [0.1, 0.6, 1.68, 1.9, 2.9, 3.4])
    dev = np.zeros(eps.shape, bool)
    dev[2, [2, 4, 6]] = True
    assert np.all(eps[dev] == 3.48 ** 2) and np.all(eps[~dev] == 1.445 ** 2)
    x_top, _, _ = s.layer_at(fac * 700)  # the rectangle's top edge is excluded: nothing crosses
    assert x_top.size == 4
    x_bot, _, _ = s.layer_at(fac * -500)  # its bottom edge is included (rectangle only)
    assert np.allclose(x_bot[2:-2], np.array([1.0, 1.6]) + 0.25)


def test_gds_fixture_cross_sections():
    """The reference's layout file (tests/golden/contour_extract.gds, the input of
    src/tests/test_gds_cpu.cpp): 8 closed polygons, so every generic cross-section crosses an
    even number of edges, inside the polygons' x range."""
    import os
    f = os.path.join(os.path.dirname(__file__), "golden", "contour_extract.gds")
    s = O.GDSIISampler(1.55, 3, 3, f, shift=(2.3647, 0.0, 0.0), lPML=0.5, rPML=0.5, lspace=1.0, rspace=1.3)
    assert len(s.polys) == 8
    for z in np.linspace(-1.6, 1.35, 23):
        x, _, _ = s.layer_at(z)
        inner = x[2:-2]
        assert inner.size % 2 == 0 and inner.size >= 2
        assert np.all(np.diff(inner) >= 0) and s.minX - 1e-12 <= inner[0] and inner[-1] <= s.maxX + 1e-12


def _legacy_layer(eps_core, shift=0.0, pml=0.1):
    x = np.array([0.0, pml, 0.3 + shift, 0.55 + shift, 1.0 - pml, 1.0])
    y = np.array([0.0, pml, 0.45, 0.55, 1.0 - pml, 1.0])
    e = np.full((5, 5), 2.1, dtype=complex)
    e[2, 2] = eps_core
    return (x, y, e)


def test_legacy_rtcm_matches_star_product_cascade():
    """The RTCM recursion of the legacy solver and the star-product cascade of the integrator
    (pinned separately) are different algorithms for the same stack: with the first stack
    entry as the bottom medium, W_last Tdd W_first^-1 equals W1 Tuu W0^-1 and
    W_first Rud W_first^-1 equals W0 Rdu W0^-1 (PML K of the first layer, two slabs)."""
    A, B, C = _legacy_layer(2.1), _legacy_layer(12.0), _legacy_layer(4.0 + 0.2j)
    s = O.RCWASolver(1, 1)
    for L in (A, B, C):
        s.addLayer(*L)
    Tdd, Rud = s.solve(1.3, [0, 1, 2, 0], [0.0, 0.37, 0.21, 0.0])
    so = O.StackLayerSampler(1.3, 3, 3, [B, C], [0.0, 0.37, 0.58], A, A, enable_pml=True)
    r = O.rcwa_integrate(so, 0.0, 0.58, 2, z_edges=[0.0, 0.37, 0.58])
    WA = s.E[0]
    X = WA @ Tdd @ np.linalg.inv(WA)
    Y = WA @ Rud @ np.linalg.inv(WA)
    inv = O.invariant_transforms(r)
    tk = [k for k in inv if "Tuu" in k][0]
    rk = [k for k in inv if "Rdu" in k][0]
    assert np.abs(X - inv[tk]).max() < 1e-11 * np.abs(inv[tk]).max()
    assert np.abs(Y - inv[rk]).max() < 1e-11 * np.abs(inv[rk]).max()
    # planted mistake: Rud without the 2 (tau taken as (F + G)^-1 / 2) breaks the match
    Tdd2, Rud2 = s.solve(1.3, [0, 1, 2, 0], [0.0, 0.37, 0.21, 0.0])
    bad = WA @ (Rud2 - np.eye(Rud2.shape[0])) @ np.linalg.inv(WA) * 0.5 + np.eye(Rud2.shape[0])
    assert np.abs(bad - inv[rk]).max() > 1e-3


def test_legacy_homogeneous_and_translation():
    """One homogeneous type repeated: Rud = 0 and Tdd = exp(i gamma t) over the inner
    thicknesses (the interfaces are identities). With zero-width PML (K periodic), shifting
    a layer's geometry by t equals the translation option with t for that stack entry
    (Layer::permuteEigVecX's phase exp(-i 2 pi m t / Lx)); the opposite sign does not."""
    s = O.RCWASolver(1, 1)
    s.addLayer(*_legacy_layer(2.1, pml=0.0))
    T, R = s.solve(1.3, [0, 0, 0, 0], [0.0, 0.2, 0.3, 0.0])
    assert np.abs(R).max() == 0.0
    assert np.abs(T - np.diag(np.exp(1j * s.gamma[0] * 0.5))).max() < 1e-14
    t = 0.17
    s = O.RCWASolver(1, 1)
    for L in (_legacy_layer(2.1, pml=0.0), _legacy_layer(12.0, pml=0.0), _legacy_layer(4.0, pml=0.0),
              _legacy_layer(12.0, t, pml=0.0)):
        s.addLayer(*L)

    def inv_t(T, R):
        WA, WC = s.E[0], s.E[2]
        return WC @ T @ np.linalg.inv(WA), WA @ R @ np.linalg.inv(WA)

    X1, Y1 = inv_t(*s.solve(1.3, [0, 3, 2], [0, 0.37, 0]))
    X2, Y2 = inv_t(*s.solve(1.3, [0, 1, 2], [0, 0.37, 0], translation=[0.0, t, 0.0]))
    X3, _ = inv_t(*s.solve(1.3, [0, 1, 2], [0, 0.37, 0], translation=[0.0, -t, 0.0]))
    assert np.abs(X1 - X2).max() < 1e-12 * np.abs(X1).max() and np.abs(Y1 - Y2).max() < 1e-12 * np.abs(Y1).max()
    assert np.abs(X1 - X3).max() > 1e-2 * np.abs(X1).max()


def _interface_residuals(s, stack, th, v, cm, cp):
    """Largest relative mismatch of the field continuity at every interface of the stack:
    with t11 = (E + H) / 2, t12 = (E - H) / 2 of the interface between entries j and j + 1,
    the left side's backward / forward amplitudes a+, a- and the right side's b+, b- must
    satisfy a+ = t11 b+ + t12 b-, a- = t12 b+ + t11 b- (E a (a+ + a-) = E b (b+ + b-),
    H a (a+ - a-) = H b (b+ - b-)). Layer 0: a- = v (incident), a+ = Rud v; inner layer
    j >= 1: forward c_p[j-1] at its first interface, backward c_m[j-1] at its second,
    carried across the layer by exp(i gamma t); last entry: b- = Tdd v, b+ = 0."""
    nL = len(stack)
    nin = nL - 2
    N = s.Tdd.shape[0]

    def phi(j):
        return np.exp(1j * s.gamma[stack[j]] * th[j])

    def res(ap, am, bp, bm, j):
        E, H = s.ifE[j], s.ifH[j]
        t11, t12 = 0.5 * (E + H), 0.5 * (E - H)
        return max(np.abs(ap - (t11 @ bp + t12 @ bm)).max() / np.abs(ap).max(),
                   np.abs(am - (t12 @ bp + t11 @ bm)).max() / np.abs(am).max())

    out = [res(s.Rud @ v, v, phi(1) * cm[0], cp[0], 0)]
    for j in range(1, nin):
        out.append(res(cm[j - 1], phi(j) * cp[j - 1], phi(j + 1) * cm[j], cp[j], j))
    out.append(res(cm[nin - 1], phi(nin) * cp[nin - 1], np.zeros(N), s.Tdd @ v, nin))
    return max(out)


def test_legacy_intermediate_field_interface_continuity():
    """RCWASolver::evaluateIntermediateField's amplitudes satisfy the tangential-field
    continuity at every interface of a 5-entry stack with three layer types (PML, lossy
    layer), for a random incident vector and for an input mode; planted mistake: one RTCM
    tau scaled by 1.01 breaks continuity at the inner interfaces."""
    A, B, C = _legacy_layer(2.1), _legacy_layer(12.0), _legacy_layer(4.0 + 0.2j)
    s = O.RCWASolver(1, 1)
    for L in (A, B, C):
        s.addLayer(*L)
    stack, th = [0, 1, 2, 1, 0], [0.0, 0.37, 0.21, 0.15, 0.0]
    s.solve(1.3, stack, th)
    N = s.Tdd.shape[0]
    rng = np.random.default_rng(0)
    v = rng.normal(size=N) + 1j * rng.normal(size=N)
    cm, cp = s.evaluateIntermediateField(v)
    assert len(cm) == len(cp) == 3
    assert _interface_residuals(s, stack, th, v, cm, cp) < 1e-12
    e4 = np.zeros(N, dtype=complex)
    e4[4] = 1.0
    cm4, cp4 = s.evaluateIntermediateField(e4)
    assert _interface_residuals(s, stack, th, e4, cm4, cp4) < 1e-12
    # columns at once equal one column at a time
    cmb, cpb = s.evaluateIntermediateField(np.stack([v, e4], axis=1))
    assert np.abs(cmb[1][:, 0] - cm[1]).max() < 1e-13 * np.abs(cm[1]).max()
    s.taus[1] = s.taus[1] * 1.01
    cmx, cpx = s.evaluateIntermediateField(v)
    assert _interface_residuals(s, stack, th, v, cmx, cpx) > 1e-4


def test_convergence_rcwa_varrcwa_differential():
    g = G.metasurface(1)
    s = O.VarLayerSampler(g["lam"], g["nx"], g["ny"], g["x"], g["y"], g["eps"], False)
    s.addXRule(2, g["rules"][2])
    ref = O.invariant_transforms(O.rcwa_integrate(s, 0.0, 2.0, 256))["Tuu"]
    errs = [np.abs(O.invariant_transforms(O.rcwa_integrate(s, 0.0, 2.0, N))["Tuu"] - ref).max() for N in (8, 32)]
    assert errs[1] < errs[0] / 4
    ev = [np.abs(O.invariant_transforms(O.redheffer_integrate(s, 0.0, 2.0, pe))["Tuu"] - ref).max()
          for pe in (1e-1, 1e-3)]
    assert ev[1] < ev[0] and ev[1] < 1e-3
    ed = [np.abs(O.invariant_transforms(O.differential_integrate(s, 0.0, 2.0, N))["Tuu"] - ref).max()
          for N in (32, 128)]
    assert ed[1] < ed[0]


# ---------------------------------------------------------------------------
# Planted mistakes: a pin is only as good as the slips it catches.  planted()
# rewrites one fragment of an oracle function's own source text and returns the
# mutated function, so every pin below is shown to fail on a plausible mistake.
# ---------------------------------------------------------------------------

def planted(func, old, new):
    import inspect
    import textwrap
    src = textwrap.dedent(inspect.getsource(func))
    assert old in src, f"planted fragment {old!r} not found in {func.__name__}"
    ns = dict(vars(O))
    exec(compile(src.replace(old, new), f"<planted {func.__name__}>", "exec"), ns)
    return ns[func.__name__]


def pml_profile(u, coeff, gamma):
    """Stretch profile whose Fourier coefficients PMLSolver1D::solve (src/rcwa/PMLSolver1D.cpp:15-30)
    returns: sin^2(pi u / 2) - (gamma / 4) sin^2(pi u) on the side whose outer end is u = 0
    (coeff = +1) and cos^2(pi u / 2) - (gamma / 4) sin^2(pi u) on the mirrored side
    (coeff = -1), u the position across the absorbing region.  Both equal 1 where they meet the
    interior and vanish at the outer end."""
    base = np.sin(0.5 * np.pi * u) ** 2 if coeff > 0 else np.cos(0.5 * np.pi * u) ** 2
    return base - 0.25 * gamma * np.sin(np.pi * u) ** 2


def quad_profile(b0, b1, L, n, fn, npts=200000):
    """(1/L) * integral over [b0, b1] of fn(u(x)) exp(-2 pi i n x / L), midpoint rule."""
    t = (np.arange(npts) + 0.5) / npts
    x = b0 + t * (b1 - b0)
    return np.sum(fn(t) * np.exp(-2j * np.pi * n * x / L)) * (b1 - b0) / npts / L


PML_GAMMA = 1.0 / complex(1.0, -1.0)  # src/core/LayerSampler.cpp:121-122


def pml_solve_error(fn):
    """Largest deviation of fn (pml_solve's signature) from quadrature of the stretch profile,
    over both sides, several region placements and orders."""
    worst = 0.0
    for (b0, b1, L) in [(0.0, 0.5, 2.5), (0.3, 0.9, 3.1), (2.0, 2.5, 2.5)]:
        for coeff in (1.0, -1.0):
            for n in range(-6, 7):
                ref = quad_profile(b0, b1, L, n, lambda u: pml_profile(u, coeff, PML_GAMMA))
                worst = max(worst, abs(fn(b0, b1, L, PML_GAMMA, n, coeff) - ref))
    return worst


def test_pml_solve_matches_stretch_profile_quadrature():
    assert pml_solve_error(O.pml_solve) < 1e-10
    # planted mistakes: conjugated stretch, flipped side, negated result, dropped term
    for old, new in [("r += gamma * sinc(n * bm - 1.0)", "r += np.conj(gamma) * sinc(n * bm - 1.0)"),
                     ("np.exp(1j * PI * (-(n * bp + coeff * 0.5)))", "np.exp(1j * PI * (-(n * bp - coeff * 0.5)))"),
                     ("return r * (bm / 16.0)", "return -r * (bm / 16.0)"),
                     ("r += -2.0 * gamma", "r += -0.0 * gamma"),
                     ("bp = (b1 + b0) / L", "bp = (b1 - b0) / L")]:
        assert pml_solve_error(planted(O.pml_solve, old, new)) > 1e-4, (old, new)


def kmat_from_quadrature(nx, ny, x, y):
    """Kx, Ky of a PML cell from the stretched-coordinate profile: s(x) = 1 inside, the
    stretch profiles over [x0, x1] and [x_-2, x_-1]; Kx[(m,n),(j,l)] = s_hat(m-j) delta(n,l)
    alpha_(j,l) (harmonic index m*ny+n), Ky likewise along y (src/core/LayerSampler.cpp:34-132)."""
    def coeffs(c, nh):
        L = c[-1] - c[0]
        out = {}
        for m in range(-(nh - 1), nh):
            v = quad_profile(c[0], c[1], L, m, lambda u: pml_profile(u, 1.0, PML_GAMMA))
            v += quad_profile(c[1], c[-2], L, m, lambda u: np.ones_like(u))
            v += quad_profile(c[-2], c[-1], L, m, lambda u: pml_profile(u, -1.0, PML_GAMMA))
            out[m] = v
        return out
    fx, fy = coeffs(x, nx), coeffs(y, ny)
    Tx = np.array([[fx[i - j] for j in range(nx)] for i in range(nx)])
    Ty = np.array([[fy[i - j] for j in range(ny)] for i in range(ny)])
    a, b = O.alpha_beta(nx, ny, x[-1] - x[0], y[-1] - y[0])
    return np.kron(Tx, np.eye(ny)) * a[None, :], np.kron(np.eye(nx), Ty) * b[None, :]


def test_pml_k_matrices_match_quadrature():
    nx, ny = 5, 3
    x = np.array([0.3, 0.8, 2.0, 2.4, 2.8])  # PML widths 0.5 and 0.4, cell offset from 0
    y = np.array([-0.2, 0.4, 1.5, 2.0])
    reg = dict(Lx=x[-1] - x[0], Ly=y[-1] - y[0], b0=y[0], b1=y[1], u0=y[-2], u1=y[-1],
               l0=x[0], l1=x[1], r0=x[-2], r1=x[-1])
    Kxr, Kyr = kmat_from_quadrature(nx, ny, x, y)
    scale = max(np.abs(Kxr).max(), np.abs(Kyr).max())

    def err(fn):
        Kx, Ky = fn(nx, ny, reg, True)
        return max(np.abs(Kx - Kxr).max(), np.abs(Ky - Kyr).max()) / scale

    assert err(O.k_matrices) < 1e-9
    # the absorbing regions really act: far from the unstretched diag(alpha)
    a, _ = O.alpha_beta(nx, ny, reg["Lx"], reg["Ly"])
    assert np.abs(Kxr - np.diag(a)).max() > 0.05 * scale
    for old, new in [("gamma = 1.0 / complex(1.0, -1.0)", "gamma = 1.0 / complex(1.0, 1.0)"),
                     ("fxx[i] += pml_solve(l0, l1, Lx, gamma, m, 1.0)", "fxx[i] += pml_solve(l0, l1, Lx, gamma, m, -1.0)"),
                     ("return Fx * alpha[None, :], Fy * beta[None, :]", "return alpha[:, None] * Fx, beta[:, None] * Fy"),
                     ("fyy[i] += pml_solve(u0, u1, Ly, gamma, n, -1.0)", "fyy[i] += pml_solve(u0, u1, Lx, gamma, n, -1.0)"),
                     ("/ np.exp(1j * (PI * m * (l1 + r0) / Lx)))", "* np.exp(1j * (PI * m * (l1 + r0) / Lx)))")]:
        assert err(planted(O.k_matrices, old, new)) > 1e-4, (old, new)


def differential_order_ratios(fn):
    """Successive-difference ratios |D(N) - D(2N)| / |D(2N) - D(4N)| of the differential method's
    W1 Tuu W0^-1 on a smoothly tapering pillar (z in [1, 2] of the metasurface rule); a method of
    order p gives 2^p."""
    g = G.metasurface(1)
    s = O.VarLayerSampler(g["lam"], g["nx"], g["ny"], g["x"], g["y"], g["eps"], False)
    s.addXRule(2, g["rules"][2])
    D = {N: O.invariant_transforms(fn(s, 1.0, 2.0, N))["Tuu"] for N in (8, 16, 32, 64)}
    return [np.abs(D[a] - D[2 * a]).max() / np.abs(D[2 * a] - D[4 * a]).max() for a in (8, 16)]


def test_differential_integrate_is_fourth_order():
    """DifferentialIntegrator's step (src/core/DifferentialIntegrator.cpp:58-65) is a 4th-order
    Magnus-type update: halving dz divides the error by 16.  Dropping its correction terms
    leaves a 2nd-order method (ratio 4)."""
    r = differential_order_ratios(O.differential_integrate)
    assert all(12.0 < x < 20.0 for x in r), r
    for old, new in [(" - b * (P2 @ Q1 @ P1 @ Q0))", ")"),
                     (" - c * (P1 @ Q1 @ P0 + P2 @ Q1 @ P1))", ")"),
                     ("(P0 + 4.0 * P1 + P2", "(P0 + 2.0 * P1 + P2"),
                     ("rdu = -np.linalg.solve(T22, T21)", "rdu = np.linalg.solve(T22, T21)")]:
        bad = differential_order_ratios(planted(O.differential_integrate, old, new))
        assert not all(12.0 < x < 20.0 for x in bad), (old, new, bad)


def test_svd_solve_least_squares_minimum_norm():
    """Eigen BDCSVD(A).solve(B) (src/rcwa/Layer.cpp:84-86): the minimum-norm least-squares
    solution over the singular values at or above max(rows, cols) * eps * s_max
    (Eigen SVDBase::rank and _solve_impl, Eigen >= 3.3)."""
    rng = np.random.default_rng(3)
    n = 12
    U, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    V, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    B = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
    # non-singular: equals the LU solve
    s = np.linspace(1.0, 3.0, n)
    A = (U * s) @ V.conj().T
    assert np.abs(O.svd_solve(A, B) - np.linalg.solve(A, B)).max() < 1e-12
    # rank-deficient, with one singular value below the threshold (dropped) and one above (kept)
    s = np.concatenate([np.linspace(1.0, 3.0, n - 3), [1e-9, 1e-20, 0.0]])[::-1]
    s = np.sort(s)[::-1]
    A = (U * s) @ V.conj().T

    def check(fn):
        x = fn(A, B)
        keep = s >= n * np.finfo(float).eps * s[0]
        # x lies in the span of the kept right singular vectors, with the exact coefficients
        c = V.conj().T @ x
        want = (U.conj().T @ B)[keep] / s[keep, None]
        return max(np.abs(c[~keep]).max() / np.abs(want).max(), np.abs(c[keep] - want).max() / np.abs(want).max())

    assert check(O.svd_solve) < 1e-5  # the kept 1e-9 direction amplifies rounding by 1e9
    for old, new in [("r = int(np.sum(s > thr))", "r = int(np.sum(s > 0))"),
                     ("return Vh[:r, :].conj().T @ tmp", "return Vh[:r, :].T @ tmp"),
                     ("thr = max(A.shape) * np.finfo(np.float64).eps * (s[0] if s.size else 0.0)",
                      "thr = 1e-8 * (s[0] if s.size else 0.0)")]:
        assert not check(planted(O.svd_solve, old, new)) < 1e-5, (old, new)
