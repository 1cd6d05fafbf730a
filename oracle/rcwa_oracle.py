"""CPU oracle for the varRCWA scattering-matrix hot path (TEST INFRASTRUCTURE ONLY).

This module restates, in numpy, the reference's CPU algorithm for the path
permittivity -> Fourier/Toeplitz convolution matrices -> P/Q -> eig(PQ) ->
layer S-matrix -> Redheffer star-product cascade -> end projections.
It exists only to check the HIP library.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker
or the timed CPU baseline; the product path (``varrcwa_amd``) never calls it.

Every function cites the reference file:line it follows (paths relative to the
reference repository root).  Linear algebra maps onto the same LAPACK kernels the
reference reaches through MKL:

* ``LAPACKE_zgeev(ROW_MAJOR,'N','V')`` (src/rcwa/MKLEigenSolver.hpp:45-55)
  -> ``numpy.linalg.eig`` (LAPACK ``zgeev`` jobvl='N', jobvr='V').
* Eigen ``PartialPivLU`` (MKL ``zgetrf``/``zgetrs`` under ``EIGEN_USE_MKL_ALL``)
  -> ``numpy.linalg.solve`` / ``inv`` (LAPACK ``zgesv``/``zgetri``).
* Eigen ``BDCSVD(...).solve`` (src/rcwa/Layer.cpp:84-86) -> explicit SVD solve
  with Eigen's rank threshold (``fe33_solve='svd'``, default) or LU (``'lu'``).

Parity pins (what fixes each function independently of this file) live in
``tests/test_oracle_pins.py``; functions without such a check are listed as
"parity unpinned" there.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

# src/rcwa/defns.h:29 -- the reference's pi literal (double precision).
PI = 3.1415926535897932

CONTINUOUS_XY = 0   # src/rcwa/FourierSolver2D.h:7-12
DISCONTINUOUS_X = 1
DISCONTINUOUS_Y = 2


# ---------------------------------------------------------------------------
# 1D Fourier coefficients of a piecewise-constant function
# ---------------------------------------------------------------------------

def sinc(x: float) -> float:
    """src/rcwa/MathFunction.h:6-9 (normalised sinc, exact 1 at 0)."""
    return 1.0 if x == 0.0 else math.sin(PI * x) / (PI * x)


def fourier1d(x: np.ndarray, d: float, f: np.ndarray, M: int) -> np.ndarray:
    """FourierSolver1D::solve (src/rcwa/FourierSolver1D.cpp:11-41).

    Ff[i] = sum_j f[j-1]/d * exp(-i m pi (x_j + x_{j-1})/d) * sinc(m dx/d) * dx,
    m = i - (M-1)/2, cells taken in the given (possibly unsorted) coordinate order.
    """
    x = np.asarray(x, dtype=np.float64)
    f = np.asarray(f, dtype=np.complex128)
    assert f.size == x.size - 1
    out = np.zeros(M, dtype=np.complex128)
    for i in range(M):
        m = i - (M - 1) // 2
        acc = 0j
        for j in range(1, x.size):
            x1, x0 = x[j], x[j - 1]
            if m == 0:
                acc += f[j - 1] / d * (x1 - x0)
            else:
                acc += (f[j - 1] / d * np.exp(-1j * (m * PI * (x1 + x0) / d))
                        * sinc(m * (x1 - x0) / d) * (x1 - x0))
        out[i] = acc
    return out


def fourier1d_vec(x: np.ndarray, d: float, f: np.ndarray, M: int) -> np.ndarray:
    """Vectorised form of :func:`fourier1d` (same formula, numpy broadcasting).

    ``f`` may be 2-D (rows, cells): every row is transformed with the same
    coordinates.  Used for pixel grids where the scalar loop is too slow.
    """
    x = np.asarray(x, dtype=np.float64)
    f = np.asarray(f, dtype=np.complex128)
    m = (np.arange(M) - (M - 1) // 2).astype(np.float64)
    x0 = x[:-1]
    x1 = x[1:]
    dx = x1 - x0
    arg = np.outer(m, dx) / d                      # (M, cells)
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(arg == 0.0, 1.0, np.sin(PI * arg) / (PI * arg))
    ph = np.exp(-1j * (np.outer(m, x1 + x0) * PI / d))
    ph[m == 0, :] = 1.0
    basis = ph * s * dx / d                         # (M, cells)
    return f @ basis.T if f.ndim == 2 else basis @ f


def reorder(x: np.ndarray) -> np.ndarray:
    """FourierSolver1D::reorder (src/rcwa/FourierSolver1D.cpp:145-154).

    Indices sorted by coordinate.  std::sort is not stable; for equal coordinates
    the oracle uses a stable order (documented divergence, affects only
    zero-width cells whose Fourier contribution is zero).
    """
    return np.argsort(np.asarray(x), kind="stable")


def refill(x: np.ndarray, f: np.ndarray, rindex: np.ndarray) -> np.ndarray:
    """FourierSolver1D::refill (src/rcwa/FourierSolver1D.cpp:156-193), painter's order."""
    f = np.asarray(f, dtype=np.complex128)
    need = any(rindex[i] != i for i in range(rindex.size - 1))
    newf = f.copy()
    if need:
        for i in range(f.size):
            if x[i] >= x[i + 1]:
                continue
            for j in range(newf.size):
                if rindex[i] <= j and rindex[i + 1] >= j + 1:
                    newf[j] = f[i]
    return newf


def _rindex(x: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    index = reorder(x)
    rindex = np.empty_like(index)
    rindex[index] = np.arange(index.size)
    return index, rindex


# ---------------------------------------------------------------------------
# Toeplitz helpers
# ---------------------------------------------------------------------------

def toeplitz_dense(vec: np.ndarray) -> np.ndarray:
    """ToeplitzMatrixXcs(vec).toDense() (src/rcwa/ToeplitzMatrixXcs.cpp:8-31).

    T(r, c) = vec[n - 1 + r - c] with n = (len(vec)+1)/2.
    """
    vec = np.asarray(vec)
    n = (vec.size + 1) // 2
    r = np.arange(n)
    return vec[n - 1 + r[:, None] - r[None, :]]


def block_toeplitz_dense(mat: np.ndarray) -> np.ndarray:
    """BlockToeplitzMatrixXcs(mat).toDense() (src/rcwa/BlockToeplitzMatrixXcs.cpp:5-52).

    Row/col index = m*ny + n; entry = mat[nx-1 + (m-j), ny-1 + (n-l)].
    """
    nx = (mat.shape[0] + 1) // 2
    ny = (mat.shape[1] + 1) // 2
    m = np.repeat(np.arange(nx), ny)
    n = np.tile(np.arange(ny), nx)
    return mat[nx - 1 + m[:, None] - m[None, :], ny - 1 + n[:, None] - n[None, :]]


# ---------------------------------------------------------------------------
# 2D Fourier solver (three convolution rules)
# ---------------------------------------------------------------------------

def fourier2d(coordX: np.ndarray, coordY: np.ndarray, f: np.ndarray,
              nx: int, ny: int, opt: int, fast: bool = True) -> np.ndarray:
    """FourierSolver2D::solve without derivatives (src/rcwa/FourierSolver2D.cpp:12-351).

    ``f`` has shape (nCy-1, nCx-1): row i is the y-strip [y_i, y_{i+1}].
    ``fast`` switches the 1D transforms to the vectorised formula (identical
    arithmetic up to summation order).
    """
    coordX = np.asarray(coordX, dtype=np.float64)
    coordY = np.asarray(coordY, dtype=np.float64)
    f = np.asarray(f, dtype=np.complex128)
    dx = coordX[-1] - coordX[0]
    dy = coordY[-1] - coordY[0]
    nCy, nCx = coordY.size, coordX.size
    f1 = fourier1d_vec if fast else fourier1d
    n = nx * ny
    Ff = np.zeros((n, n), dtype=np.complex128)
    index, rindex = _rindex(coordX)
    ii = np.repeat(np.arange(nx), ny)
    jj = np.tile(np.arange(ny), nx)

    if opt == DISCONTINUOUS_X:
        # src/rcwa/FourierSolver2D.cpp:43-144
        for i in range(nCy - 1):
            newf = refill(coordX, f[i, :], rindex)
            fInv = f1(coordX, dx, 1.0 / newf, 2 * nx - 1)
            tFF = np.linalg.inv(toeplitz_dense(fInv))
            fY1 = f1(np.array([coordY[i], coordY[i + 1]]), dy, np.array([1.0 + 0j]), 2 * ny - 1)
            # Ff[(a,b),(c,d)] += fY1[b - d + ny - 1] * tFF[a, c]
            Ff += fY1[jj[:, None] - jj[None, :] + ny - 1] * tFF[ii[:, None], ii[None, :]]
    elif opt == DISCONTINUOUS_Y:
        # src/rcwa/FourierSolver2D.cpp:145-249
        newf = np.stack([refill(coordX, f[i, :], rindex) for i in range(f.shape[0])])
        for i in range(nCx - 1):
            fInv = f1(coordY, dy, 1.0 / newf[:, i], 2 * ny - 1)
            tFF = np.linalg.inv(toeplitz_dense(fInv))
            xs = np.array([coordX[index[i]], coordX[index[i + 1]]])
            fX1 = f1(xs, dx, np.array([1.0 + 0j]), 2 * nx - 1)
            Ff += fX1[ii[:, None] - ii[None, :] + nx - 1] * tFF[jj[:, None], jj[None, :]]
    elif opt == CONTINUOUS_XY:
        # src/rcwa/FourierSolver2D.cpp:250-349
        four = np.zeros((2 * nx - 1, 2 * ny - 1), dtype=np.complex128)
        for i in range(nCy - 1):
            newf = refill(coordX, f[i, :], rindex)
            fF = f1(coordX, dx, newf, 2 * nx - 1)
            fY1 = f1(np.array([coordY[i], coordY[i + 1]]), dy, np.array([1.0 + 0j]), 2 * ny - 1)
            four += np.outer(fF, fY1)
        Ff = block_toeplitz_dense(four)
    else:
        raise ValueError("bad Fourier option")
    return Ff


# ---------------------------------------------------------------------------
# Shape and wavelength derivatives (tangents per parameter)
# ---------------------------------------------------------------------------

def dsolve_1d(x: np.ndarray, d: float, f: np.ndarray, M: int, layout: int) -> np.ndarray:
    """FourierSolver1D::dSolve (src/rcwa/FourierSolver1D.cpp:93-143): derivative of the M
    coefficients of a piecewise constant f with respect to the coordinate x[layout]:
    (f[layout-1] - f[layout]) / d * exp(-i 2 m pi x[layout] / d), a side absent at the ends."""
    out = np.zeros(M, dtype=np.complex128)
    for i in range(M):
        m = i - (M - 1) // 2
        e = 1.0 if m == 0 else np.exp(-1j * (2 * m * PI * x[layout] / d))
        if layout > 0:
            out[i] += f[layout - 1] / d * e
        if layout < x.size - 1:
            out[i] += -f[layout] / d * e
    return out


def dsolve_v2_1d(x: np.ndarray, d: float, f: np.ndarray, M: int, layout: int) -> np.ndarray:
    """FourierSolver1D::dSolve_v2 (src/rcwa/FourierSolver1D.cpp:195-249): dSolve on the
    painter's-order (refilled) values, neighbours taken at the coordinate's sorted position;
    only interior coordinates (0 < layout < len(x) - 1, sorted position likewise)."""
    _, rindex = _rindex(x)
    r = int(rindex[layout])
    if not (0 < layout < x.size - 1) or not (0 < r < x.size - 1):
        raise ValueError("only interior coordinates can be varied")
    newf = refill(x, np.asarray(f, dtype=np.complex128), rindex)
    out = np.zeros(M, dtype=np.complex128)
    for i in range(M):
        m = i - (M - 1) // 2
        e = 1.0 if m == 0 else np.exp(-1j * (2 * m * PI * x[layout] / d))
        out[i] += newf[r - 1] / d * e
        out[i] += -newf[r] / d * e
    return out


def _para_jacobian(para: Dict[int, Sequence[float]], npara: int) -> Tuple[List[int], np.ndarray]:
    """Parameterization (src/rcwa/Parameterization.h:5-25): x_k = sum_p a_kp p_p + x_k0.
    Coefficients beyond a coordinate's vector are 0 (the header's stated default)."""
    keys = list(para.keys())
    J = np.zeros((len(keys), npara))
    for r, k in enumerate(keys):
        v = np.asarray(para[k], dtype=np.float64)
        J[r, :v.size] = v
    return keys, J


def dsolve_para(x, d, f, M, para, npara, v2: bool) -> np.ndarray:
    """FourierSolver1D::dSolve / dSolve_v2 over a Parameterization
    (src/rcwa/FourierSolver1D.cpp:252-321): (M x nvar gradient) @ (nvar x npara jacobian)."""
    keys, J = _para_jacobian(para, npara)
    g = np.zeros((M, len(keys)), dtype=np.complex128)
    for c, k in enumerate(keys):
        g[:, c] = (dsolve_v2_1d if v2 else dsolve_1d)(x, d, f, M, k)
    return g @ J


def fourier2d_tangents(coordX, coordY, f, nx, ny, opt, para_x: Dict[int, Sequence[float]],
                       npara: int) -> List[np.ndarray]:
    """The dFfs outputs of FourierSolver2D::solve (src/rcwa/FourierSolver2D.cpp:12-351):
    per parameter the derivative of the convolution matrix, x coordinates only (the
    reference rejects y parameterizations, :19-23).

    DiscontinuousX: d inv(T) = -inv(T) dT inv(T) per strip with dT from dSolve_v2 of 1/f
    (:61-80); DiscontinuousY: the x-cell coefficient of each column differentiated at its
    two ends with dSolve (:188-212); ContinuousXY: dSolve_v2 of f (:295-321)."""
    coordX = np.asarray(coordX, dtype=np.float64)
    coordY = np.asarray(coordY, dtype=np.float64)
    f = np.asarray(f, dtype=np.complex128)
    npara = max(npara, max((len(v) for v in para_x.values()), default=0))
    dx = coordX[-1] - coordX[0]
    dy = coordY[-1] - coordY[0]
    nCy, nCx = coordY.size, coordX.size
    n = nx * ny
    out = [np.zeros((n, n), dtype=np.complex128) for _ in range(npara)]
    if npara == 0:
        return out
    index, rindex = _rindex(coordX)
    ii = np.repeat(np.arange(nx), ny)
    jj = np.tile(np.arange(ny), nx)
    if opt == DISCONTINUOUS_X:
        for i in range(nCy - 1):
            newf = refill(coordX, f[i, :], rindex)
            tFF = np.linalg.inv(toeplitz_dense(fourier1d_vec(coordX, dx, 1.0 / newf, 2 * nx - 1)))
            dfi = dsolve_para(coordX, dx, 1.0 / f[i, :], 2 * nx - 1, para_x, npara, True)
            fY1 = fourier1d_vec(np.array([coordY[i], coordY[i + 1]]), dy, np.array([1.0 + 0j]), 2 * ny - 1)
            for p in range(npara):
                dtFF = -tFF @ toeplitz_dense(dfi[:, p]) @ tFF
                out[p] += fY1[jj[:, None] - jj[None, :] + ny - 1] * dtFF[ii[:, None], ii[None, :]]
    elif opt == DISCONTINUOUS_Y:
        newf = np.stack([refill(coordX, f[i, :], rindex) for i in range(f.shape[0])])
        keys, J = _para_jacobian(para_x, npara)
        row = {k: J[r] for r, k in enumerate(keys)}
        for i in range(nCx - 1):
            tFF = np.linalg.inv(toeplitz_dense(fourier1d_vec(coordY, dy, 1.0 / newf[:, i], 2 * ny - 1)))
            xs = np.array([coordX[index[i]], coordX[index[i + 1]]])
            one = np.array([1.0 + 0j])
            dfX1 = np.zeros((2 * nx - 1, npara), dtype=np.complex128)
            if int(index[i]) in row:
                dfX1 += np.outer(dsolve_1d(xs, dx, one, 2 * nx - 1, 0), row[int(index[i])])
            if int(index[i + 1]) in row:
                dfX1 += np.outer(dsolve_1d(xs, dx, one, 2 * nx - 1, 1), row[int(index[i + 1])])
            for p in range(npara):
                out[p] += dfX1[ii[:, None] - ii[None, :] + nx - 1, p] * tFF[jj[:, None], jj[None, :]]
    elif opt == CONTINUOUS_XY:
        four = [np.zeros((2 * nx - 1, 2 * ny - 1), dtype=np.complex128) for _ in range(npara)]
        for i in range(nCy - 1):
            dfi = dsolve_para(coordX, dx, f[i, :], 2 * nx - 1, para_x, npara, True)
            fY1 = fourier1d_vec(np.array([coordY[i], coordY[i + 1]]), dy, np.array([1.0 + 0j]), 2 * ny - 1)
            for p in range(npara):
                four[p] += np.outer(dfi[:, p], fY1)
        out = [block_toeplitz_dense(fp) for fp in four]
    else:
        raise ValueError("bad Fourier option")
    return out


def wave_eqn_derivatives(coordX, coordY, eps, Kx, Ky, lam, nx, ny, para_x: Dict[int, Sequence[float]],
                         npara: int = 0):
    """The derivative outputs of Layer::waveEqnCoeff (src/rcwa/Layer.cpp:36-162): per
    parameter dP = [[-Kx T F2h, Kx T F1h], [-Ky T F2h, Ky T F1h]] with T = fe33^-1 dfe33,
    dQ = [[0, -k0^2 dfe22], [k0^2 dfe11, 0]] (K held fixed); and the wavelength derivatives
    dPl, dQl with d(k0^2)/d lambda = -8 pi^2 / lambda^3. Returns (dP, dQ, dPl, dQl, dfe)
    with dfe = (dfe11, dfe22, dfe33), lists over parameters."""
    npara = max(npara, max((len(v) for v in para_x.values()), default=0))
    fe11 = fourier2d(coordX, coordY, eps, nx, ny, DISCONTINUOUS_X)
    fe22 = fourier2d(coordX, coordY, eps, nx, ny, DISCONTINUOUS_Y)
    fe33 = fourier2d(coordX, coordY, eps, nx, ny, CONTINUOUS_XY)
    d11 = fourier2d_tangents(coordX, coordY, eps, nx, ny, DISCONTINUOUS_X, para_x, npara)
    d22 = fourier2d_tangents(coordX, coordY, eps, nx, ny, DISCONTINUOUS_Y, para_x, npara)
    d33 = fourier2d_tangents(coordX, coordY, eps, nx, ny, CONTINUOUS_XY, para_x, npara)
    n = nx * ny
    F1h = np.linalg.solve(fe33, Kx)
    F2h = np.linalg.solve(fe33, Ky)
    k02 = (2 * PI / lam) ** 2
    dk02 = -8.0 * PI * PI / (lam * lam * lam)
    dP, dQ = [], []
    for p in range(npara):
        T = np.linalg.solve(fe33, d33[p])
        m = np.zeros((2 * n, 2 * n), dtype=np.complex128)
        m[:n, :n] = Kx @ (-T @ F2h)
        m[:n, n:] = -Kx @ (-T @ F1h)
        m[n:, :n] = Ky @ (-T @ F2h)
        m[n:, n:] = -Ky @ (-T @ F1h)
        dP.append(m)
        q = np.zeros((2 * n, 2 * n), dtype=np.complex128)
        q[:n, n:] = -k02 * d22[p]
        q[n:, :n] = k02 * d11[p]
        dQ.append(q)
    I = np.eye(n)
    dPl = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    dPl[:n, n:] = dk02 * I
    dPl[n:, :n] = -dk02 * I
    dQl = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    dQl[:n, n:] = -dk02 * fe22
    dQl[n:, :n] = dk02 * fe11
    return dP, dQ, dPl, dQl, (d11, d22, d33)


# ---------------------------------------------------------------------------
# PML stretch and K matrices
# ---------------------------------------------------------------------------

def _expipi(x: float) -> complex:
    return complex(math.cos(PI * x), math.sin(PI * x))


def pml_solve(b0: float, b1: float, L: float, gamma: complex, n: int, coeff: float) -> complex:
    """PMLSolver1D::solve (src/rcwa/PMLSolver1D.cpp:15-30)."""
    bm = (b1 - b0) / L
    bp = (b1 + b0) / L
    r = 4.0 * sinc(n * bm - 0.5) * np.exp(1j * PI * (-(n * bp + coeff * 0.5)))
    r += 8.0 * np.exp(1j * PI * (-(n * bp))) * sinc(n * bm)
    r += -2.0 * gamma * np.exp(1j * PI * (-(n * bp))) * sinc(n * bm)
    r += 4.0 * sinc(n * bm + 0.5) * np.exp(1j * PI * (-(n * bp - coeff * 0.5)))
    r += gamma * sinc(n * bm - 1.0) * np.exp(1j * PI * (-(n * bp + coeff * 1.0)))
    r += gamma * sinc(n * bm + 1.0) * np.exp(1j * PI * (-(n * bp - coeff * 1.0)))
    return r * (bm / 16.0)


def alpha_beta(nx: int, ny: int, Lx: float, Ly: float) -> Tuple[np.ndarray, np.ndarray]:
    """LayerSampler::evaluateAlphaBeta (src/core/LayerSampler.cpp:9-32)."""
    Nx, Ny = (nx - 1) // 2, (ny - 1) // 2
    Kx, Ky = 2 * PI / Lx, 2 * PI / Ly
    m = np.repeat(np.arange(nx) - Nx, ny).astype(np.float64)
    n = np.tile(np.arange(ny) - Ny, nx).astype(np.float64)
    return Kx * m + 0j, Ky * n + 0j


def k_matrices(nx: int, ny: int, region: Dict[str, float], enable_pml: bool
               ) -> Tuple[np.ndarray, np.ndarray]:
    """LayerSampler::evaluateKMatrices + evaluatePML (src/core/LayerSampler.cpp:34-143)."""
    Lx, Ly = region["Lx"], region["Ly"]
    alpha, beta = alpha_beta(nx, ny, Lx, Ly)
    if not enable_pml:
        return np.diag(alpha), np.diag(beta)
    b0, b1, u0, u1 = region["b0"], region["b1"], region["u0"], region["u1"]
    l0, l1, r0, r1 = region["l0"], region["l1"], region["r0"], region["r1"]
    gamma = 1.0 / complex(1.0, -1.0)
    nPx, nPy = 2 * nx - 1, 2 * ny - 1
    maxPx, maxPy = (nPx - 1) // 2, (nPy - 1) // 2
    fyy = np.zeros(nPy, dtype=np.complex128)
    for i in range(nPy):
        n = i - maxPy
        fyy[i] = ((u0 - b1) / Ly * sinc(n * (u0 - b1) / Ly)
                  / np.exp(1j * PI * (n * (b1 + u0) / Ly)))
        fyy[i] += pml_solve(b0, b1, Ly, gamma, n, 1.0)
        fyy[i] += pml_solve(u0, u1, Ly, gamma, n, -1.0)
    fxx = np.zeros(nPx, dtype=np.complex128)
    for i in range(nPx):
        m = i - maxPx
        fxx[i] = ((r0 - l1) * sinc(m * (r0 - l1) / Lx) / Lx
                  / np.exp(1j * (PI * m * (l1 + r0) / Lx)))
        fxx[i] += pml_solve(l0, l1, Lx, gamma, m, 1.0)
        fxx[i] += pml_solve(r0, r1, Lx, gamma, m, -1.0)
    fyx = np.zeros(nPx, dtype=np.complex128)
    for i in range(nPx):
        m = i - maxPx
        if m == 0:
            fyx[i] = r1 - l0
        else:
            fyx[i] = ((np.exp(-1j * (2 * PI / Lx * m * r1)) - np.exp(-1j * (2 * PI / Lx * m * l0)))
                      * (1j * (Lx * Lx / (m * 2 * PI))))
    fxy = np.zeros(nPy, dtype=np.complex128)
    for i in range(nPy):
        n = i - maxPy
        if n == 0:
            fxy[i] = u1 - b0
        else:
            fxy[i] = ((np.exp(-1j * (2 * PI / Ly * n * u1)) - np.exp(-1j * (2 * PI / Ly * n * b0)))
                      * (1j * (Ly * Ly / (2 * n * PI))))
    fy = 1.0 / Lx * np.outer(fyx, fyy)
    fx = 1.0 / Ly * np.outer(fxx, fxy)
    Fx = block_toeplitz_dense(fx)
    Fy = block_toeplitz_dense(fy)
    return Fx * alpha[None, :], Fy * beta[None, :]


# ---------------------------------------------------------------------------
# Layer operator (P, Q)
# ---------------------------------------------------------------------------

def svd_solve(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """Eigen BDCSVD(A, ThinU|ThinV).solve(B) (src/rcwa/Layer.cpp:84-86).

    x = V diag(1/s) U^H b over the singular values above Eigen's default
    threshold (diagSize * eps relative to the largest).
    """
    U, s, Vh = np.linalg.svd(A)
    thr = max(A.shape) * np.finfo(np.float64).eps * (s[0] if s.size else 0.0)
    r = int(np.sum(s > thr))
    tmp = (U[:, :r].conj().T @ B) / s[:r, None]
    return Vh[:r, :].conj().T @ tmp


def wave_eqn_coeff(coordX, coordY, eps, Kx, Ky, lam, nx, ny, fe33_solve: str = "svd",
                   return_fe: bool = False):
    """Layer::waveEqnCoeff without derivatives (src/rcwa/Layer.cpp:36-100,134-138)."""
    fe11 = fourier2d(coordX, coordY, eps, nx, ny, DISCONTINUOUS_X)
    fe22 = fourier2d(coordX, coordY, eps, nx, ny, DISCONTINUOUS_Y)
    fe33 = fourier2d(coordX, coordY, eps, nx, ny, CONTINUOUS_XY)
    n = nx * ny
    if fe33_solve == "svd":
        F1h = svd_solve(fe33, Kx)
        F2h = svd_solve(fe33, Ky)
    else:
        F1h = np.linalg.solve(fe33, Kx)
        F2h = np.linalg.solve(fe33, Ky)
    k0 = 2 * PI / lam
    k02 = k0 * k0
    I = np.eye(n)
    P = np.empty((2 * n, 2 * n), dtype=np.complex128)
    P[:n, :n] = Kx @ F2h
    P[:n, n:] = -Kx @ F1h + k02 * I
    P[n:, :n] = Ky @ F2h - k02 * I
    P[n:, n:] = -Ky @ F1h
    Q = np.empty((2 * n, 2 * n), dtype=np.complex128)
    Q[:n, :n] = -Kx @ Ky
    Q[:n, n:] = Kx @ Kx - k02 * fe22
    Q[n:, :n] = -Ky @ Ky + k02 * fe11
    Q[n:, n:] = Ky @ Kx
    if return_fe:
        return P, Q, (fe11, fe22, fe33)
    return P, Q


# ---------------------------------------------------------------------------
# Samplers (the plugin boundary)
# ---------------------------------------------------------------------------

class LayerSampler:
    """LayerSampler (src/core/LayerSampler.h:5-77 / .cpp:9-153)."""

    def __init__(self, lam: float, nx: int, ny: int, enable_pml: bool = True):
        self.lam, self.nx, self.ny, self.enable_pml = lam, nx, ny, enable_pml
        self.region: Dict[str, float] = {}
        self.Kx = self.Ky = None
        self.fe33_solve = "svd"

    def k0(self) -> float:
        return 2 * PI / self.lam

    def nDim(self) -> int:
        return 2 * self.nx * self.ny

    def _post_init(self, x: np.ndarray, y: np.ndarray) -> None:
        # VarLayerSampler::assignSimulationRegion (src/core/VarLayerSampler.cpp:23-39)
        self.region = dict(Lx=x[-1] - x[0], Ly=y[-1] - y[0], b0=y[0], b1=y[1], u0=y[-2],
                           u1=y[-1], l0=x[0], l1=x[1], r0=x[-2], r1=x[-1])
        self.Kx, self.Ky = k_matrices(self.nx, self.ny, self.region, self.enable_pml)

    def layer_at(self, z: float):
        raise NotImplementedError

    def sample(self, z: float) -> Tuple[np.ndarray, np.ndarray]:
        x, y, eps = self.layer_at(z)
        return wave_eqn_coeff(x, y, eps, self.Kx, self.Ky, self.lam, self.nx, self.ny,
                              self.fe33_solve)


class VarLayerSampler(LayerSampler):
    """VarLayerSampler (src/core/VarLayerSampler.cpp:9-69): coordinates move with z."""

    def __init__(self, lam, nx, ny, x, y, eps, enable_pml=True):
        super().__init__(lam, nx, ny, enable_pml)
        self.x = np.asarray(x, dtype=np.float64).copy()
        self.y = np.asarray(y, dtype=np.float64).copy()
        self.eps = np.asarray(eps, dtype=np.complex128).copy()
        self.xrules: Dict[int, Callable[[float], float]] = {}
        self.yrules: Dict[int, Callable[[float], float]] = {}
        self._post_init(self.x, self.y)

    def addXRule(self, layout: int, rule: Callable[[float], float]) -> None:
        self.xrules[layout] = rule

    def addYRule(self, layout: int, rule: Callable[[float], float]) -> None:
        self.yrules[layout] = rule

    def layer_at(self, z):
        x = self.x.copy()
        y = self.y.copy()
        for i, f in self.xrules.items():
            x[i] = f(z)
        for i, f in self.yrules.items():
            y[i] = f(z)
        return x, y, self.eps


def gdsii_real(r: int) -> float:
    """GDSII 8-byte real (gdsii_real_to_double, src/gdstk/utils.cpp:46-51): sign bit,
    excess-64 base-16 exponent, 56-bit mantissa."""
    e = ((r >> 56) & 0x7F) - 64
    m = (r & 0x00FFFFFFFFFFFFFF) / 72057594037927936.0
    v = math.ldexp(m, 4 * e)
    return -v if r >> 63 else v


def read_gds_polygons(path: str, unit: float = 1e-6) -> List[List[List[Tuple[float, float]]]]:
    """gdstk::read_gds (src/gdstk/library.cpp:160-420) restricted to what GDSIISampler reads:
    per cell (in file order) its BOUNDARY/BOX polygons as point lists, coordinates
    factor * int32 with factor = db_in_meters / unit (UNITS record), the closing point of
    each polygon dropped at ENDEL (:386-391); PATH, SREF, AREF and TEXT elements are not
    polygons."""
    import struct
    data = open(path, "rb").read()
    cells: List[List[List[Tuple[float, float]]]] = []
    factor = 1.0
    cur = None
    elem = None
    o = 0
    while o + 4 <= len(data):
        ln, rec = struct.unpack(">HB", data[o:o + 3])
        if ln < 4:
            break
        body = data[o + 4:o + ln]
        if rec == 0x03:
            db_in_m = gdsii_real(struct.unpack(">Q", body[8:16])[0])
            factor = db_in_m / unit
        elif rec == 0x04:
            break
        elif rec == 0x05:
            cells.append([])
        elif rec in (0x08, 0x2D):
            elem = []
            cells[-1].append(elem)
        elif rec in (0x09, 0x0A, 0x0B, 0x0C, 0x15):
            elem = None
        elif rec == 0x10 and elem is not None:
            v = struct.unpack(">%di" % (len(body) // 4), body)
            elem.extend((factor * v[k], factor * v[k + 1]) for k in range(0, len(v) - 1, 2))
        elif rec == 0x11:
            if elem:
                elem.pop()
            elem = None
        o += ln
    return cells


class GDSIISampler(LayerSampler):
    """GDSIISampler (src/core/GDSIISampler.cpp:17-146): x-z cross-sections of the first cell's
    polygons (GDSII y is the propagation axis z), PML always on."""

    def __init__(self, lam, nx, ny, filename, shift=(0.0, 0.0, 0.0), lPML=0.5, rPML=0.5, lspace=1.5,
                 rspace=1.5, bPML=0.5, uPML=0.5, bspace=1.08, uspace=1.0, permBackground=1.445 * 1.445,
                 permDevice=3.48 * 3.48, thickness=0.22):
        super().__init__(lam, nx, ny, True)
        self.polys = read_gds_polygons(filename, 1e-6)[0]
        self.shift = [float(v) for v in shift]
        self.eps_bg, self.eps_dev = complex(permBackground), complex(permDevice)
        # assignSimulationRegion (src/core/GDSIISampler.cpp:49-92)
        pts = [p for poly in self.polys for p in poly]
        self.minX = min(p[0] for p in pts) + self.shift[0]
        self.maxX = max(p[0] for p in pts) + self.shift[0]
        self.minZ = min(p[1] for p in pts) + self.shift[2]
        self.maxZ = max(p[1] for p in pts) + self.shift[2]
        b0 = self.shift[1]
        b1 = b0 + bPML
        u0 = b1 + bspace + thickness + uspace
        u1 = u0 + uPML
        l1 = self.minX - lspace
        l0 = l1 - lPML
        r0 = self.maxX + rspace
        r1 = r0 + rPML
        self.xr = (l0, l1, r0, r1)
        self.y = np.array([b0, b1, b1 + bspace, b1 + bspace + thickness, u0, u1])
        self._post_init(np.array(self.xr), np.array([b0, b1, u0, u1]))

    def layer_at(self, z):
        """sampleLayout (src/core/GDSIISampler.cpp:94-146)."""
        xs = []
        for poly in self.polys:
            m = len(poly)
            for i in range(m):
                (x0, y0), (x1, y1) = poly[i], poly[(i + 1) % m]
                if (y0 <= z < y1) or (y1 <= z < y0):
                    h = y1 - y0
                    if abs(h) < 1e-12:
                        h += 2e-12
                    a = (z - y0) / h
                    xs.append((1 - a) * x0 + a * x1)
        xs.sort()
        Nx = len(xs) + 4
        x = np.empty(Nx)
        x[0], x[1], x[-2], x[-1] = self.xr
        x[2:Nx - 2] = np.array(xs) + self.shift[0]
        eps = np.full((self.y.size - 1, Nx - 1), self.eps_bg, dtype=np.complex128)
        for j in range(2, Nx - 3, 2):
            eps[2, j] = self.eps_dev
        return x, self.y.copy(), eps


class StackLayerSampler(LayerSampler):
    """Piecewise-constant-in-z stack of Layer(x, y, eps) slabs (extension).

    Layer k occupies the open interval (z_edges[k], z_edges[k+1]).  At or below
    z_edges[0] the sampler returns ``bottom``; at or above z_edges[-1] it returns
    ``top`` (half-space media).  With equal slab thicknesses,
    RCWAIntegrator(z_edges[0], z_edges[-1], stack).compute(len(layers)) samples
    each slab exactly once at its midpoint (src/core/RCWAIntegrator.cpp:41-45).
    All layers must share the outer cell [x0, x1] x [y0, y1].
    """

    def __init__(self, lam, nx, ny, layers, z_edges, bottom, top, enable_pml=False):
        super().__init__(lam, nx, ny, enable_pml)
        self.layers = layers
        self.z_edges = np.asarray(z_edges, dtype=np.float64)
        self.bottom, self.top = bottom, top
        x, y, _ = bottom
        self._post_init(np.asarray(x, np.float64), np.asarray(y, np.float64))

    def layer_at(self, z):
        if z <= self.z_edges[0]:
            return self.bottom
        if z >= self.z_edges[-1]:
            return self.top
        k = int(np.searchsorted(self.z_edges, z, side="right") - 1)
        return self.layers[k]


# ---------------------------------------------------------------------------
# Eigen-system and slab S-matrix
# ---------------------------------------------------------------------------

def gamma_sqrt(g: np.ndarray) -> np.ndarray:
    """GammaSqrtOp (src/rcwa/WaveEquationSolver.cpp:8-22)."""
    gv = np.sqrt(np.asarray(g, dtype=np.complex128))
    gv = np.where(gv.imag < 0, -gv, gv)
    flip = (np.abs(gv.imag) < np.abs(gv.real)) & (gv.real < 0)
    return np.where(flip, -gv, gv)


def wave_equation_solve(P: np.ndarray, Q: np.ndarray):
    """WaveEquationSolver::solve (src/rcwa/WaveEquationSolver.cpp:24-37).

    PQ = P Q; zgeev(jobvl='N', jobvr='V'); gamma = GammaSqrt(lambda); V = Q W diag(1/gamma).
    Returns (W, V, gamma, lambda).
    """
    PQ = P @ Q
    lam, W = np.linalg.eig(PQ)
    gam = gamma_sqrt(lam)
    V = (Q @ W) * (1.0 / gam)[None, :]
    return W, V, gam, lam


def rcwa_scatter_matrix(lam: float, wz: float, P, Q, W0, V0):
    """dtmm::RCWAScatterMatrix::compute (src/rcwa/RCWAScatterMatrix.cpp:26-85).

    Returns (Tuu, Rud); the reference aliases Tdd := Tuu and Rdu := Rud
    (src/rcwa/RCWAScatterMatrix.h:44-47).
    """
    W, V, gam, _ = wave_equation_solve(P, Q)
    return rcwa_scatter_matrix_from_eig(lam, wz, W, V, gam, W0, V0)


def rcwa_scatter_matrix_from_eig(lam: float, wz: float, W, V, gam, W0, V0):
    """The slab S-matrix from a given eigen-system: the arithmetic after the eigen
    step of RCWAScatterMatrix::compute, identical to RCWAScatterMatrix::compute_v2
    (src/rcwa/RCWAScatterMatrix.cpp:88-147). Returns (Tuu, Rud)."""
    invk0 = lam / (2.0 * PI)
    ss = invk0 * wz
    X = np.exp(1j * gam * ss)
    F0 = np.linalg.solve(W, W0)
    G0 = np.linalg.solve(V, V0)
    A = F0 + G0
    B = F0 - G0
    lu_rhs = np.linalg.solve(A, np.concatenate([B, X[:, None] * A, X[:, None] * B], axis=1))
    n = A.shape[0]
    invAB, invAXA, invAXB = lu_rhs[:, :n], lu_rhs[:, n:2 * n], lu_rhs[:, 2 * n:]
    I = np.eye(n)
    M = I - invAXB @ invAXB
    R1 = invAXB @ invAXA - invAB
    T1 = invAXA - invAXB @ invAB
    sol = np.linalg.solve(M, np.concatenate([R1, T1], axis=1))
    Rud = sol[:, :n]
    Tuu = sol[:, n:]
    return Tuu, Rud


def star_product(S, tdd, tuu, rud, rdu):
    """Redheffer star product as written in src/core/RCWAIntegrator.cpp:57-67.

    S = (Tdd, Tuu, Rud, Rdu) accumulated so far; returns the updated tuple.
    """
    Tdd, Tuu, Rud, Rdu = S
    n = Tdd.shape[0]
    I = np.eye(n)
    invC1 = I - rdu @ Rud
    C1 = np.linalg.inv(invC1)
    C2 = Rud @ C1 @ rdu + I
    Rdu_n = Rdu + Tdd @ rdu @ C2 @ Tuu
    Tdd_n = Tdd @ C1 @ tdd
    Rud_n = rud + tuu @ Rud @ C1 @ tdd
    Tuu_n = tuu @ C2 @ Tuu
    return Tdd_n, Tuu_n, Rud_n, Rdu_n


def left_end_projection(S, W1, V1, W0, V0):
    """End projection of src/core/RCWAIntegrator.cpp:73-86 (basis W1,V1 -> W0,V0)."""
    Tdd, Tuu, Rud, Rdu = S
    invW1_W0 = np.linalg.solve(W1, W0)
    invV1_V0 = np.linalg.solve(V1, V0)
    A = 0.5 * (invW1_W0 + invV1_V0)
    B = 0.5 * (invW1_W0 - invV1_V0)
    K = A - Rdu @ B
    sol = np.linalg.solve(K, np.concatenate([Rdu @ A - B, Tdd], axis=1))
    n = A.shape[0]
    Rdu_n = sol[:, :n]
    Tdd_n = sol[:, n:]
    Rud_n = Tuu @ B @ Tdd_n + Rud
    Tuu_n = Tuu @ (B @ Rdu_n + A)
    return Tdd_n, Tuu_n, Rud_n, Rdu_n


class RCWASolver:
    """The legacy stacked-layer solver RCWASolver (src/rcwa/RCWASolver.h:60-80; solve,
    src/rcwa/RCWASolver.cpp:104-113): layer types solved once (Layer::solve,
    src/rcwa/Layer.cpp:173-190: P, Q with PML K from the first layer's region, eigenvectors
    E = W, H = V, gamma = sqrt(eig) / k0), interface matrices E = W1^-1 W2, H = V1^-1 V2 per
    pair of neighbouring types (:460-580; with a translation per stack entry the eigenvector
    rows of x-harmonic m are scaled by exp(-i 2 pi m t / Lx) first, Layer::permuteEigVecX,
    src/rcwa/Layer.cpp:193-238), and the RTCM recursion from the last layer to the first
    (:343-457). Complete orthogonal decomposition solves are LU solves here (the same for the
    non-singular eigenvector matrices)."""

    def __init__(self, nX: int, nY: int):
        self.nx, self.ny = 2 * nX + 1, 2 * nY + 1
        self.layers = []

    def addLayer(self, x, y, eps):
        self.layers.append((np.asarray(x, np.float64), np.asarray(y, np.float64),
                            np.asarray(eps, np.complex128)))

    def solve(self, lam: float, stack: Sequence[int], thickness: Sequence[float],
              translation: Optional[Sequence[float]] = None):
        nx, ny = self.nx, self.ny
        n, N = nx * ny, 2 * nx * ny
        x0, y0, _ = self.layers[0]
        region = dict(Lx=x0[-1] - x0[0], Ly=y0[-1] - y0[0], b0=y0[0], b1=y0[1], u0=y0[-2], u1=y0[-1],
                      l0=x0[0], l1=x0[1], r0=x0[-2], r1=x0[-1])
        Kx, Ky = k_matrices(nx, ny, region, True)
        k0 = 2 * PI / lam
        self.E, self.H, self.gamma = [], [], []
        for (x, y, eps) in self.layers:
            P, Q = wave_eqn_coeff(x, y, eps, Kx, Ky, lam, nx, ny, "lu")
            W, V, gam, _ = wave_equation_solve(P, Q)
            self.E.append(W)
            self.H.append(V)
            self.gamma.append(gam / k0)
        nL = len(stack)

        def rows(t):
            m = np.repeat(np.arange(nx) - (nx - 1) // 2, ny)
            d = np.exp(-1j * (2.0 * PI * m / region["Lx"] * t))
            return np.concatenate([d, d])

        ifE, ifH = [], []
        for i in range(nL - 1):
            a, b = stack[i], stack[i + 1]
            if translation is None and a == b:
                ifE.append(np.eye(N, dtype=complex))
                ifH.append(np.eye(N, dtype=complex))
                continue
            Ea, Ha, Eb, Hb = self.E[a], self.H[a], self.E[b], self.H[b]
            if translation is not None:
                da, db = rows(translation[i])[:, None], rows(translation[i + 1])[:, None]
                Ea, Ha, Eb, Hb = da * Ea, da * Ha, db * Eb, db * Hb
            ifE.append(np.linalg.solve(Ea, Eb))
            ifH.append(np.linalg.solve(Ha, Hb))
        Tdd = np.eye(N, dtype=complex)
        Rud = np.zeros((N, N), dtype=complex)
        self.taus = []
        for i in range(nL - 1):
            index = nL - i - 1
            if i == 0:
                phi = np.ones(N, dtype=complex)
            else:
                phi = np.exp(1j * self.gamma[stack[index]] * thickness[index])
            waveRud = phi[:, None] * Rud * phi[None, :]
            waveTdd = Tdd * phi[None, :]
            E, H = ifE[index - 1], ifH[index - 1]
            F = E + E @ waveRud
            G = H - H @ waveRud
            tau = np.linalg.inv(F + G)
            self.taus.append(tau)
            Rud = -2.0 * G @ tau + np.eye(N)
            Tdd = 2.0 * waveTdd @ tau
        self.Tdd, self.Rud = Tdd, Rud
        self.ifE, self.ifH = ifE, ifH
        self.stack, self.thickness = list(stack), list(thickness)
        return Tdd, Rud

    def evaluateIntermediateField(self, input_vec: np.ndarray):
        """RCWASolver::evaluateIntermediateField (src/rcwa/RCWASolver.cpp:1373-1500; the
        inputMode variant at :1503-1621 is input_vec = e_mode): modal amplitudes c_m, c_p
        of every inner layer (nL - 2 of each) for the incident amplitudes input_vec of the
        first layer, from the interface matrices (t11 = (E + H) / 2, t12 = (E - H) / 2 in
        the scattering form S11 = t11 - t12 t11^-1 t12, S12 = t12 t11^-1, S21 = -t11^-1 t12,
        S22 = t11^-1), the RTCM taus and the stack's Rud, Tdd. input_vec may hold several
        columns (one solve per column). Complete orthogonal decomposition solves are LU
        solves here (the same for the non-singular S11)."""
        stack, thickness = self.stack, self.thickness
        nL = len(stack)
        N = self.Tdd.shape[0]
        v = np.asarray(input_vec, dtype=complex)
        S11, S12, S21, S22, phases, REF = [], [], [], [], [], [np.zeros((N, N), dtype=complex)]
        Rud = REF[0]
        for_inv_s22 = None
        for i in range(nL - 1):
            index = nL - i - 1
            if i == 0:
                phi = np.ones(N, dtype=complex)
            else:
                phi = np.exp(1j * self.gamma[stack[index]] * thickness[index])
            phases.append(phi)
            E, H = self.ifE[index - 1], self.ifH[index - 1]
            t11 = 0.5 * (E + H)
            t12 = 0.5 * (E - H)
            inv_t22 = np.linalg.inv(t11)
            inv_t22_t12 = np.linalg.solve(t11, t12)
            S11.append(t11 - t12 @ inv_t22_t12)
            S12.append(t12 @ inv_t22)
            S21.append(-inv_t22_t12)
            S22.append(inv_t22)
            if i == 0:
                for_inv_s22 = t11
            waveRud = phi[:, None] * Rud * phi[None, :]
            G = H - H @ waveRud
            Rud = -2.0 * G @ self.taus[i]
            Rud = Rud + np.eye(N)
            REF.append(Rud)
        R = self.Rud @ v
        T = self.Tdd @ v
        nin = nL - 2
        c_m = [None] * nin
        c_p = [None] * nin
        d = [None] * nin
        c_m[nin - 1] = S12[0] @ for_inv_s22 @ T
        temp = np.linalg.solve(S11[nL - 2], R - S12[nL - 2] @ v)
        c_p[0] = S21[nL - 2] @ temp + S22[nL - 2] @ v
        ph = phases[nL - 2] if v.ndim == 1 else phases[nL - 2][:, None]
        d[nin - 1] = ph * c_p[0]
        for i in range(nin - 2, -1, -1):
            c_m[nin - i - 2] = REF[i + 2] @ d[i + 1]
            tempi = np.linalg.solve(S11[i + 1], c_m[nin - i - 2] - S12[i + 1] @ d[i + 1])
            c_p[nin - i - 1] = S21[i + 1] @ tempi + S22[i + 1] @ d[i + 1]
            ph = phases[i + 1] if v.ndim == 1 else phases[i + 1][:, None]
            d[i] = ph * c_p[nin - i - 1]
        return c_m, c_p


@dataclass
class SResult:
    """Outputs of an integrator (the getters of src/core/RCWAIntegrator.h:28-38)."""
    Tdd: np.ndarray
    Tuu: np.ndarray
    Rud: np.ndarray
    Rdu: np.ndarray
    W0: np.ndarray
    V0: np.ndarray
    gam0: np.ndarray
    W1: np.ndarray
    V1: np.ndarray
    gam1: np.ndarray
    N: int = 0
    nEigVal: int = 0
    tree: List[Tuple[float, float]] = field(default_factory=list)


def scatter_plane_wave(r: "SResult", alpha: np.ndarray, beta: np.ndarray, k0: float, eps_ref: complex,
                       eps_trn: complex, e_inc: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Diffraction efficiencies per harmonic, the power bookkeeping of
    RCWASolver::scatterPlaneWave (src/rcwa/RCWASolver.cpp:1636-1733) applied to the
    integrator's S-matrix with the wave incident from the bottom medium.

    e_inc: incident field harmonics [Ex; Ey] (2n) in the bottom medium; its modal
    coefficients are c = W0^{-1} e_inc. Reflected coefficients Rdu c (bottom basis),
    transmitted Tuu c (top basis); fields E = W c per medium, Ez from the plane-wave
    divergence condition with Kz = -/+ sqrt(k0^2 eps - kx^2 - ky^2) (reflected / transmitted,
    src/rcwa/RCWASolver.cpp:1671-1700), and per harmonic
    REF_i = |Re Kz_i| / (sqrt(Re eps_ref) k0) * (|Ex|^2 + |Ey|^2 + |Ez|^2), TRN likewise with
    the top medium (src/rcwa/RCWASolver.cpp:1706-1732)."""
    n = alpha.shape[0]
    c = np.linalg.solve(r.W0, e_inc)
    rxy = r.W0 @ (r.Rdu @ c)
    txy = r.W1 @ (r.Tuu @ c)

    def power(xy, eps, sign):
        kz = sign * np.sqrt(k0 * k0 * eps - alpha * alpha - beta * beta)
        if np.any(kz == 0):
            raise ValueError("a plane-wave component moves transversely")
        ex, ey = xy[:n], xy[n:]
        ez = (alpha * ex + beta * ey) / kz
        return np.abs(kz.real) / (np.sqrt(np.real(eps)) * k0) * (np.abs(ex) ** 2 + np.abs(ey) ** 2 + np.abs(ez) ** 2)

    return power(rxy, eps_ref, -1.0), power(txy, eps_trn, 1.0)


def evaluate_energy(e: np.ndarray, h: np.ndarray, nx: int, ny: int, Lx: float, Ly: float, nres_x: int,
                    nres_y: int, x0: float, x1: float, y0: float, y1: float) -> float:
    """Power of one mode through a window, evaluateEnergy (src/tests/test_cpu.cpp:335-412):
    the field harmonics e = [Ex; Ey], h = [Hx; Hy] (index j + i ny, i the x-harmonic,
    :352-361) are summed to fields at the midpoints of an nres_x x nres_y grid over
    [x0, x1] x [y0, y1] with wavelets exp(-i 2 pi m x / Lx), exp(-i 2 pi n y / Ly)
    (:379-399), and 0.5 Re(Ex Hy* - Ey Hx*) dx dy is added up over the grid (:401-408)."""
    n = nx * ny
    kex, key = e[:n].reshape(nx, ny), e[n:].reshape(nx, ny)
    khx, khy = h[:n].reshape(nx, ny), h[n:].reshape(nx, ny)
    xstep, ystep = (x1 - x0) / nres_x, (y1 - y0) / nres_y
    xs = x0 + xstep * (np.arange(nres_x) + 0.5)
    ys = y0 + ystep * (np.arange(nres_y) + 0.5)
    mx = np.arange(nx) - (nx - 1) // 2
    my = np.arange(ny) - (ny - 1) // 2
    wx = np.exp(-2j * PI * np.outer(xs, mx) / Lx)  # (nres_x, nx)
    wy = np.exp(-2j * PI * np.outer(ys, my) / Ly)  # (nres_y, ny)

    def field(k):
        return wx @ k @ wy.T  # (nres_x, nres_y)

    ex, ey, hx, hy = field(kex), field(key), field(khx), field(khy)
    return float(0.5 * np.real(ex * np.conj(hy) - ey * np.conj(hx)).sum() * xstep * ystep)


def find_guided_modes(gamma: np.ndarray, k0: float, imag_coeff: float, re_min: float,
                      re_max: float) -> List[Tuple[int, complex]]:
    """findGuidedModes (src/tests/test_cpu.cpp:311-333): the modes whose gamma / k0^2 has real
    part in [re_min, re_max] and imaginary part below imag_coeff, as (index, gamma / k0^2),
    sorted by decreasing real part (ties keep index order; the reference's std::sort leaves
    their order unspecified)."""
    rec = []
    for i, g in enumerate(np.asarray(gamma)):
        gi = complex(g.real / k0 / k0, g.imag / k0 / k0)  # std::complex / double: per component
        if re_min <= gi.real <= re_max and gi.imag < imag_coeff:
            rec.append((i, complex(gi)))
    rec.sort(key=lambda r: -r[1].real)
    return rec


def rcwa_integrate(sampler: LayerSampler, z0: float, z1: float, N: int,
                   z_edges: Optional[Sequence[float]] = None) -> SResult:
    """RCWAIntegrator::compute(N) (src/core/RCWAIntegrator.cpp:20-87).

    ``z_edges`` (extension) replaces the uniform division by explicit slice
    edges; each slice is still sampled at its midpoint and uses its own width.
    """
    lam = sampler.lam
    P1, Q1 = sampler.sample(z1)
    W1, V1, gam1, _ = wave_equation_solve(P1, Q1)
    nD = W1.shape[0]
    S = (np.eye(nD, dtype=np.complex128), np.eye(nD, dtype=np.complex128),
         np.zeros((nD, nD), np.complex128), np.zeros((nD, nD), np.complex128))
    if z_edges is None:
        dz = (z1 - z0) / float(N)
        slices = [(z0 + (i + 0.5) * dz, dz) for i in range(N)]
    else:
        ze = list(z_edges)
        slices = [(0.5 * (ze[i] + ze[i + 1]), ze[i + 1] - ze[i]) for i in range(len(ze) - 1)]
    for zm, dz in slices:
        P, Q = sampler.sample(zm)
        tuu, rud = rcwa_scatter_matrix(lam, dz, P, Q, W1, V1)
        S = star_product(S, tuu, tuu, rud, rud)
    P0, Q0 = sampler.sample(z0)
    W0, V0, gam0, _ = wave_equation_solve(P0, Q0)
    S = left_end_projection(S, W1, V1, W0, V0)
    return SResult(*S, W0, V0, gam0, W1, V1, gam1, N=len(slices))


# ---------------------------------------------------------------------------
# varRCWA (RedhefferIntegrator, trisection branch)
# ---------------------------------------------------------------------------

def perturbation_and_difference(k0, Pr, Qr, W, V, gam, z0, z1, P0, Pm, P1, Q0, Qm, Q1):
    """RedhefferIntegrator::evaluatePerturbationAndDifference
    (src/core/RedhefferIntegrator.cpp:317-399).  Returns (tdd, tuu, rud, rdu, diffError)."""
    def dAB(Ps, Qs):
        a = np.linalg.solve(W, (Ps - Pr) @ V)
        b = np.linalg.solve(V, (Qs - Qr) @ W)
        return a + b, a - b
    dA0, dB0 = dAB(P0, Q0)
    dA1, dB1 = dAB(Pm, Qm)
    dA2, dB2 = dAB(P1, Q1)
    ph = np.exp(1j * (0.5 * (z1 - z0) / k0) * gam)
    pf = np.exp(1j * ((z1 - z0) / k0) * gam)
    c = 1j * ((z1 - z0) / (6 * 2 * k0))
    Dh = ph[:, None] * ph[None, :]
    err = 0.0
    tuu = c * (pf[:, None] * dA0 + 4.0 * Dh * dA1 + dA2 * pf[None, :])
    err = max(err, np.abs(tuu).max())
    tuu = tuu + np.diag(pf)
    rud = -c * (pf[:, None] * dB0 * pf[None, :] + 4.0 * Dh * dB1 + dB2)
    err = max(err, np.abs(rud).max())
    rdu = -c * (dB0 + 4.0 * Dh * dB1 + pf[:, None] * dB2 * pf[None, :])
    err = max(err, np.abs(rdu).max())
    tdd = c * (dA0 * pf[None, :] + 4.0 * Dh * dA1 + pf[:, None] * dA2)
    err = max(err, np.abs(tdd).max())
    tdd = tdd + np.diag(pf)
    return tdd, tuu, rud, rdu, err


def redheffer_integrate(sampler: LayerSampler, z0: float, z1: float, pe: float = 1e-2,
                        max_depth: int = 60) -> SResult:
    """RedhefferIntegrator::compute, active trisection branch
    (src/core/RedhefferIntegrator.cpp:174-224 and recursive_solve :226-314).

    ``tree`` records the accepted segments (z0, z1) in cascade order so that a
    device implementation can be checked on the same subdivision.
    """
    lam = sampler.lam
    k0 = 2.0 * PI / lam
    P0_, Q0_ = sampler.sample(z0)
    W0, V0, gam0, _ = wave_equation_solve(P0_, Q0_)
    P1_, Q1_ = sampler.sample(z1)
    W1, V1, gam1, _ = wave_equation_solve(P1_, Q1_)
    zm = 0.5 * (z0 + z1)
    Pm, Qm = sampler.sample(zm)
    Wm, Vm, gamm, _ = wave_equation_solve(Pm, Qm)
    nD = W0.shape[0]
    st = dict(S=(np.eye(nD, dtype=np.complex128), np.eye(nD, dtype=np.complex128),
                 np.zeros((nD, nD), np.complex128), np.zeros((nD, nD), np.complex128)),
              leftW=W0, leftV=V0, N=0, tree=[])

    def rec(Pr, Qr, W, V, gam, za, zb, Pa, Pmid, Pb, Qa, Qmid, Qb, depth):
        tdd, tuu, rud, rdu, err = perturbation_and_difference(
            k0, Pr, Qr, W, V, gam, za, zb, Pa, Pmid, Pb, Qa, Qmid, Qb)
        if err <= pe or depth >= max_depth:
            st["N"] += 1
            st["tree"].append((za, zb))
            invW_W0 = np.linalg.solve(W, st["leftW"])
            invV_V0 = np.linalg.solve(V, st["leftV"])
            A = 0.5 * (invW_W0 + invV_V0)
            B = 0.5 * (invW_W0 - invV_V0)
            K = A - rdu @ B
            rdu_n = np.linalg.solve(K, rdu @ A - B)
            tdd_n = np.linalg.solve(K, tdd)
            rud_n = tuu @ B @ tdd_n + rud
            tuu_n = tuu @ A + tuu @ B @ rdu_n
            st["leftW"], st["leftV"] = W, V
            st["S"] = star_product(st["S"], tdd_n, tuu_n, rud_n, rdu_n)
            return
        zm0 = 2.0 / 3.0 * za + 1.0 / 3.0 * zb
        zm1 = 1.0 / 3.0 * za + 2.0 / 3.0 * zb
        Pt, Qt = sampler.sample(0.5 * (za + zm0))
        Pk, Qk = sampler.sample(zm0)
        Wt, Vt, gamt, _ = wave_equation_solve(Pt, Qt)
        rec(Pt, Qt, Wt, Vt, gamt, za, zm0, Pa, Pt, Pk, Qa, Qt, Qk, depth + 1)
        Pt2, Qt2 = sampler.sample(zm1)
        rec(Pr, Qr, W, V, gam, zm0, zm1, Pk, Pmid, Pt2, Qk, Qmid, Qt2, depth + 1)
        Pk2, Qk2 = sampler.sample(0.5 * (zm1 + zb))
        Wt2, Vt2, gamt2, _ = wave_equation_solve(Pk2, Qk2)
        rec(Pk2, Qk2, Wt2, Vt2, gamt2, zm1, zb, Pt2, Pk2, Pb, Qt2, Qk2, Qb, depth + 1)

    rec(Pm, Qm, Wm, Vm, gamm, z0, z1, P0_, Pm, P1_, Q0_, Qm, Q1_, 0)
    Tdd, Tuu, Rud, Rdu = st["S"]
    # right projection (src/core/RedhefferIntegrator.cpp:211-223)
    invW0_W = np.linalg.solve(st["leftW"], W1)
    invV0_V = np.linalg.solve(st["leftV"], V1)
    A = 0.5 * (invW0_W + invV0_V)
    B = 0.5 * (invW0_W - invV0_V)
    K = A - Rud @ B
    Tuu = np.linalg.solve(K, Tuu)
    Rud = np.linalg.solve(K, Rud @ A - B)
    Rdu = Rdu + Tdd @ B @ Tuu
    Tdd = Tdd @ (B @ Rud + A)
    return SResult(Tdd, Tuu, Rud, Rdu, W0, V0, gam0, W1, V1, gam1, N=st["N"], nEigVal=2,
                   tree=st["tree"])


# ---------------------------------------------------------------------------
# Differential method
# ---------------------------------------------------------------------------

def differential_integrate(sampler: LayerSampler, z0: float, z1: float, N: int) -> SResult:
    """DifferentialIntegrator::compute (src/core/DifferentialIntegrator.cpp:21-134)."""
    lam = sampler.lam
    k0 = 2.0 * PI / lam
    P1_, Q1_ = sampler.sample(z1)
    W1, V1, gam1, _ = wave_equation_solve(P1_, Q1_)
    nD = W1.shape[0]
    S = (np.eye(nD, dtype=np.complex128), np.eye(nD, dtype=np.complex128),
         np.zeros((nD, nD), np.complex128), np.zeros((nD, nD), np.complex128))
    dz = (z1 - z0) / float(N)
    I = np.eye(nD)
    P0, Q0 = sampler.sample(z0)
    for i in range(N):
        zl = z0 + i * dz
        zr = z0 + (i + 1) * dz
        zm = 0.5 * (zl + zr)
        P1, Q1 = sampler.sample(zm)
        P2, Q2 = sampler.sample(zr)
        a = dz * dz / (6.0 * k0 * k0)
        b = dz * dz / (4.0 * k0 * k0)
        c = dz * dz / (2.0 * k0 * k0)
        L11 = I - a * (P1 @ Q0 + P1 @ Q1 + P2 @ Q1 - b * (P2 @ Q1 @ P1 @ Q0))
        L12 = (1j * dz / (6.0 * k0)) * (P0 + 4.0 * P1 + P2 - c * (P1 @ Q1 @ P0 + P2 @ Q1 @ P1))
        L21 = (1j * dz / (6.0 * k0)) * (Q0 + 4.0 * Q1 + Q2 - c * (Q1 @ P1 @ Q0 + Q2 @ P1 @ Q1))
        L22 = I - a * (Q1 @ P0 + Q1 @ P1 + Q2 @ P1 - b * (Q2 @ P1 @ Q1 @ P0))
        iW11 = np.linalg.solve(W1, L11)
        iW12 = np.linalg.solve(W1, L12)
        iV21 = np.linalg.solve(V1, L21)
        iV22 = np.linalg.solve(V1, L22)
        A11 = 0.5 * (iW11 + iV21)
        A12 = 0.5 * (iW12 + iV22)
        A21 = 0.5 * (iW11 - iV21)
        A22 = 0.5 * (iW12 - iV22)
        T11 = A11 @ W1 + A12 @ V1
        T12 = A11 @ W1 - A12 @ V1
        T21 = A21 @ W1 + A22 @ V1
        T22 = A21 @ W1 - A22 @ V1
        tdd = np.linalg.inv(T22)
        rud = T12 @ tdd
        rdu = -np.linalg.solve(T22, T21)
        tuu = T11 + T12 @ rdu
        S = star_product(S, tdd, tuu, rud, rdu)
        P0, Q0 = P2, Q2
    P0_, Q0_ = sampler.sample(z0)
    W0, V0, gam0, _ = wave_equation_solve(P0_, Q0_)
    S = left_end_projection(S, W1, V1, W0, V0)
    return SResult(*S, W0, V0, gam0, W1, V1, gam1, N=N)


# ---------------------------------------------------------------------------
# Basis-invariant views used by the parity tests
# ---------------------------------------------------------------------------

def invariant_transforms(r: SResult) -> Dict[str, np.ndarray]:
    """Harmonic-space maps that do not depend on the eigensolver's choice of
    eigenvector order, phase or rotation inside degenerate eigenspaces:
    W1 Tuu W0^-1, W0 Tdd W1^-1, W0 Rdu W0^-1, W1 Rud W1^-1."""
    iW0 = np.linalg.inv(r.W0)
    iW1 = np.linalg.inv(r.W1)
    return dict(Tuu=r.W1 @ r.Tuu @ iW0, Tdd=r.W0 @ r.Tdd @ iW1,
                Rdu=r.W0 @ r.Rdu @ iW0, Rud=r.W1 @ r.Rud @ iW1)
