"""Dense complex128 kernels (ZGEMM, LU, solve, eigensolver) against numpy/LAPACK."""
import ctypes as C

import numpy as np
import pytest

from varrcwa_amd._lib import cplx, from_dev_colmajor, ptr, to_dev_colmajor

pytestmark = pytest.mark.gpu


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def opm(a, op):
    return a if op == 0 else (np.swapaxes(a, -1, -2) if op == 1 else np.conj(np.swapaxes(a, -1, -2)))


@pytest.mark.parametrize("opA", [0, 1, 2])
@pytest.mark.parametrize("opB", [0, 1, 2])
def test_zgemm_ops(ctx, opA, opB):
    rng = np.random.default_rng(10 * opA + opB)
    m, n, k, batch = 70, 45, 37, 3
    A = crand(rng, batch, *((m, k) if opA == 0 else (k, m)))
    B = crand(rng, batch, *((k, n) if opB == 0 else (n, k)))
    Cm = crand(rng, batch, m, n)
    alpha, beta = 0.7 - 0.2j, -0.3 + 1.1j
    dA, dB, dC = to_dev_colmajor(A), to_dev_colmajor(B), to_dev_colmajor(Cm)
    lda, ldb = A.shape[1], B.shape[1]
    ctx.check(ctx.lib.vrc_zgemm(ctx.h, opA, opB, m, n, k, cplx(alpha), ptr(dA), lda, A.shape[1] * A.shape[2],
                                ptr(dB), ldb, B.shape[1] * B.shape[2], cplx(beta), ptr(dC), m, m * n, batch),
              "zgemm")
    ref = alpha * (opm(A, opA) @ opm(B, opB)) + beta * Cm
    got = from_dev_colmajor(dC)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_zgemm_large_and_beta0(ctx):
    rng = np.random.default_rng(1)
    m = n = k = 300
    A, B = crand(rng, 2, m, k), crand(rng, 2, k, n)
    dA, dB = to_dev_colmajor(A), to_dev_colmajor(B)
    import torch
    dC = torch.full((2, n, m), float("nan"), dtype=torch.complex128, device="cuda")
    ctx.check(ctx.lib.vrc_zgemm(ctx.h, 0, 0, m, n, k, cplx(1), ptr(dA), m, m * k, ptr(dB), k, k * n, cplx(0),
                                ptr(dC), m, m * n, 2), "zgemm")
    ref = A @ B
    assert np.abs(from_dev_colmajor(dC) - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("n", [5, 33, 64, 130, 242])
def test_getrf_getrs(ctx, n):
    import torch
    rng = np.random.default_rng(n)
    batch, nrhs = 3, 2 * n + 3
    A = crand(rng, batch, n, n)
    B = crand(rng, batch, n, nrhs)
    dA, dB = to_dev_colmajor(A), to_dev_colmajor(B)
    ipiv = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    ctx.check(ctx.lib.vrc_zgetrf(ctx.h, n, ptr(dA), n, n * n, ptr(ipiv), n, ptr(info), batch), "getrf")
    dX = torch.zeros_like(dB)
    ctx.check(ctx.lib.vrc_zgetrs(ctx.h, n, nrhs, ptr(dA), n, n * n, ptr(ipiv), n, ptr(dB), n, n * nrhs,
                                 ptr(dX), n, n * nrhs, batch), "getrs")
    X = from_dev_colmajor(dX)
    ref = np.linalg.solve(A, B)
    assert info.cpu().numpy().tolist() == [0] * batch
    err = np.abs(X - ref).max() / np.abs(ref).max()
    assert err < 1e-11, err
    # in-place solve (X aliases B)
    ctx.check(ctx.lib.vrc_zgetrs(ctx.h, n, nrhs, ptr(dA), n, n * n, ptr(ipiv), n, ptr(dB), n, n * nrhs,
                                 ptr(dB), n, n * nrhs, batch), "getrs in place")
    assert np.abs(from_dev_colmajor(dB) - ref).max() / np.abs(ref).max() < 1e-11


@pytest.mark.parametrize("n", [130, 1030])
def test_getrf_matches_lapack(ctx, n):
    """Pivot sequence and factors equal LAPACK zgetrf's (scipy lu_factor). n = 130 runs the
    register-resident panel (n <= 1024), n = 1030 the workgroup-per-panel kernel."""
    import torch
    from scipy.linalg import lu_factor
    rng = np.random.default_rng(7 + n)
    batch = 2
    A = crand(rng, batch, n, n)
    dA = to_dev_colmajor(A)
    ipiv = torch.zeros((batch, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    ctx.check(ctx.lib.vrc_zgetrf(ctx.h, n, ptr(dA), n, n * n, ptr(ipiv), n, ptr(info), batch), "getrf")
    LU = from_dev_colmajor(dA)
    piv = ipiv.cpu().numpy()
    assert info.cpu().numpy().tolist() == [0] * batch
    for b in range(batch):
        lu_ref, piv_ref = lu_factor(A[b])
        assert np.array_equal(piv[b], piv_ref), np.nonzero(piv[b] != piv_ref)[0][:5]
        assert np.abs(LU[b] - lu_ref).max() <= 1e-11 * np.abs(lu_ref).max()


def test_getrf_zero_column_info(ctx):
    """A zero column inside a panel: info names it (1-based), as LAPACK's."""
    import torch
    from scipy.linalg import lapack
    rng = np.random.default_rng(3)
    n = 70
    A = crand(rng, n, n)
    A[:, 37] = 0.0
    dA = to_dev_colmajor(A[None])
    ipiv = torch.zeros((1, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.check(ctx.lib.vrc_zgetrf(ctx.h, n, ptr(dA), n, n * n, ptr(ipiv), n, ptr(info), 1), "getrf")
    _, _, info_ref = lapack.zgetrf(A)
    assert int(info.cpu()[0]) == info_ref == 38


def test_getrf_singular_info(ctx):
    import torch
    n = 6
    A = np.eye(n, dtype=complex)
    A[3, 3] = 0.0
    dA = to_dev_colmajor(A[None])
    ipiv = torch.zeros((1, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.check(ctx.lib.vrc_zgetrf(ctx.h, n, ptr(dA), n, n * n, ptr(ipiv), n, ptr(info), 1), "getrf")
    assert int(info.cpu()[0]) == 4


def _eig_check(ctx, A, tol_res=1e-12):
    import torch
    batch, n, _ = A.shape
    dA = to_dev_colmajor(A)
    w = torch.zeros((batch, n), dtype=torch.complex128, device="cuda")
    VR = torch.zeros_like(dA)
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    ctx.check(ctx.lib.vrc_zgeev(ctx.h, n, ptr(dA), n, n * n, ptr(w), n, ptr(VR), n, n * n, ptr(info), batch),
              "zgeev")
    ctx.sync()
    assert info.cpu().numpy().tolist() == [0] * batch
    W = w.cpu().numpy()
    V = from_dev_colmajor(VR)
    for b in range(batch):
        nrm = np.linalg.norm(A[b], 2)
        res = np.abs(A[b] @ V[b] - V[b] * W[b][None, :]).max() / nrm
        assert res < tol_res * n, (b, res)
        # unit norm, largest component real and positive
        assert np.allclose(np.linalg.norm(V[b], axis=0), 1.0, atol=1e-12)
        k = np.argmax(np.abs(V[b]) ** 2, axis=0)
        big = V[b][k, np.arange(n)]
        assert np.all(np.abs(big.imag) == 0) and np.all(big.real > 0)
        # eigenvalues vs LAPACK (sorted, well separated for random matrices)
        ref = np.linalg.eigvals(A[b])
        key = lambda z: np.lexsort((np.round(z.imag, 6), np.round(z.real, 6)))
        a, r = np.sort_complex(W[b]), np.sort_complex(ref)
        assert np.abs(a - r).max() / np.abs(r).max() < 1e-11
    return W, V


@pytest.mark.parametrize("n", [3, 20, 48, 49, 97, 200])
def test_zgeev_random(ctx, n):
    rng = np.random.default_rng(1000 + n)
    _eig_check(ctx, crand(rng, 2, n, n))


def test_zgeev_triangular_and_diag(ctx):
    rng = np.random.default_rng(5)
    n = 80
    A = np.triu(crand(rng, n, n))
    D = np.diag(crand(rng, n))
    _eig_check(ctx, np.stack([A, D]))


@pytest.mark.parametrize("variant", [1, 2, 4])
@pytest.mark.parametrize("n", [20, 49, 97, 200, 441])
def test_zgeev_qr_kernel_variants(ctx, variant, n):
    """The other QR kernel configurations of "zgeev.qr_variant": 1 (64-wide window, window
    products staged through LDS), 2 (48-wide window on 256 threads, two matrices per CU),
    4 (32-wide window on 128 threads, four per CU) against LAPACK eigenvalues and
    eigen-residuals, like the default kernel."""
    rng = np.random.default_rng(2000 + n)
    A = crand(rng, 3, n, n)
    assert ctx.lib.vrc_tune_set(b"zgeev.qr_variant", variant) == 0
    try:
        _eig_check(ctx, A)
    finally:
        assert ctx.lib.vrc_tune_set(b"zgeev.qr_variant", -1) == 0
