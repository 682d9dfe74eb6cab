"""The device eigensolver (syev) and LU solve (gesv) against numpy/scipy LAPACK.

syev: eigenvalues against numpy.linalg.eigvalsh; eigenvectors through the invariants that
fix them (A V = V diag(W), V^T V = I), since their signs and, for repeated eigenvalues, the
basis are free. gesv: the solution against numpy.linalg.solve, and the pivots and packed LU
factors against scipy.linalg.lu_factor (LAPACK getrf), which the same pivot rule fixes.
Tolerances are relative to the matrix norm: FP64 1e-12, FP32 1e-4 (eigenvalues 1e-5)."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 1e-4}


def _sym(n, seed, dtype):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n))
    return ((B + B.T) / 2).astype(dtype)


def _check_eig(A, vals, vecs, tol):
    A = A.astype(np.float64)
    vals = vals.astype(np.float64)
    vecs = vecs.astype(np.float64)
    n = A.shape[0]
    scale = max(1.0, np.linalg.norm(A))
    assert np.all(np.diff(vals) >= 0)
    assert np.abs(vals - np.linalg.eigvalsh(A)).max() <= tol * scale
    assert np.linalg.norm(A @ vecs - vecs * vals) <= tol * scale * max(1, n) ** 0.5
    assert np.abs(vecs.T @ vecs - np.eye(n)).max() <= tol * max(1, n) ** 0.5


@pytest.mark.parametrize("n", [1, 2, 3, 7, 48, 100, 257])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_syev_random_symmetric(core, n, dtype):
    A = _sym(n, n, dtype)
    dA = dev(A)
    W = torch.empty(n, dtype=dA.dtype, device="cuda")
    core.syev(dA, W)
    _check_eig(A, host(W), host(dA), TOL[dtype])


def test_syev_repeated_eigenvalues_and_diagonal_input(core):
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((9, 9)))
    lam = np.array([1.0, 1.0, 1.0, 2.0, 2.0, -3.0, 0.0, 0.0, 5.0])
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    dA = dev(A)
    W = torch.empty(9, dtype=torch.float64, device="cuda")
    core.syev(dA, W)
    _check_eig(A, host(W), host(dA), 1e-12)
    # already diagonal: no rotation; eigenvalues sorted, eigenvectors a permutation
    D = np.diag([3.0, -1.0, 2.0, -1.0])
    dD = dev(D)
    W4 = torch.empty(4, dtype=torch.float64, device="cuda")
    core.syev(dD, W4)
    assert host(W4).tolist() == [-1.0, -1.0, 2.0, 3.0]
    assert np.array_equal(np.abs(host(dD)), np.array([[0, 0, 0, 1], [1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0]], float))


@pytest.mark.parametrize("uplo", ["U", "L"])
def test_syev_reads_one_triangle_only(uplo):
    """Through the C ABI on column-major storage, with NaN in the triangle not named."""
    from einsums_amd import _lib as L

    n = 33
    A = _sym(n, 5, np.float64)
    poison = A.copy()
    mask = np.tril(np.ones((n, n), bool), -1) if uplo == "U" else np.triu(np.ones((n, n), bool), 1)
    poison[mask] = np.nan
    # column-major storage with lda = n + 3
    lda = n + 3
    buf = np.zeros((n, lda))
    buf[:, :n] = poison.T  # row j of buf = column j of the matrix
    dA = dev(buf)
    W = torch.empty(n, dtype=torch.float64, device="cuda")
    info = torch.full((), 7, dtype=torch.int32, device="cuda")
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    L.check(L.lib.ein_syev_f64(b"V", uplo.encode(), n, vp(dA), lda, vp(W), vp(info), None))
    assert int(info.item()) == 0
    V = host(dA)[:, :n].T
    _check_eig(A, host(W), V, 1e-12)
    # eigenvalues only: W matches, A is not touched
    dA2 = dev(buf)
    W2 = torch.empty(n, dtype=torch.float64, device="cuda")
    L.check(L.lib.ein_syev_f64(b"N", uplo.encode(), n, vp(dA2), lda, vp(W2), vp(info), None))
    assert int(info.item()) == 0
    assert np.abs(host(W2) - np.linalg.eigvalsh(A)).max() <= 1e-12 * np.linalg.norm(A)
    assert np.array_equal(np.isnan(host(dA2)), np.isnan(buf))


def test_syev_argument_errors():
    from einsums_amd import _lib as L

    info = torch.zeros((), dtype=torch.int32, device="cuda")
    p = ctypes.c_void_p(info.data_ptr())
    assert L.lib.ein_syev_f64(b"X", b"U", 4, p, 4, p, p, None) == -1
    assert L.lib.ein_syev_f64(b"V", b"X", 4, p, 4, p, p, None) == -2
    assert L.lib.ein_syev_f64(b"V", b"U", -1, p, 4, p, p, None) == -3
    assert L.lib.ein_syev_f64(b"V", b"U", 4, p, 3, p, p, None) == -5
    assert L.lib.ein_syev_f64(b"V", b"U", 2000, p, 2000, p, p, None) == L.EIN_ERROR_UNSUPPORTED


@pytest.mark.parametrize("n,nrhs", [(1, 1), (2, 3), (9, 1), (100, 4), (513, 2)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gesv_vs_lapack(core, n, nrhs, dtype):
    import scipy.linalg as sl

    rng = np.random.default_rng(n * 7 + nrhs)
    A = rng.standard_normal((n, n)).astype(dtype)
    B = rng.standard_normal((n, nrhs)).astype(dtype)
    dA, dB = dev(A), dev(B)
    info, ipiv = core.gesv(dA, dB)
    assert info == 0
    A64 = A.astype(np.float64)
    X = np.linalg.solve(A64, B.astype(np.float64))
    cond = np.linalg.cond(A64)
    assert np.linalg.norm(host(dB) - X) <= TOL[dtype] * cond * max(1.0, np.linalg.norm(X))
    lu, piv = sl.lu_factor(A64)
    assert np.array_equal(host(ipiv), piv + 1)
    assert np.linalg.norm(host(dA) - lu) <= TOL[dtype] * cond * max(1.0, np.linalg.norm(lu))


def test_gesv_vector_rhs_and_singular(core):
    rng = np.random.default_rng(11)
    A = rng.standard_normal((6, 6))
    b = rng.standard_normal(6)
    dA, db = dev(A), dev(b)
    info, _ = core.gesv(dA, db)
    assert info == 0
    assert np.allclose(host(db), np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)
    # column 3 zero: U(4,4) is exactly zero (1-based info 4) and B stays as it was
    S = rng.standard_normal((6, 6))
    S[:, 3] = 0.0
    dS, db2 = dev(S), dev(b)
    info, _ = core.gesv(dS, db2)
    assert info == 4
    assert np.array_equal(host(db2), b)


def test_diis_system_like_reference(core):
    """The DIIS system of the reference HF test (compute_diis_coefs, QChemTests.cpp:136-166):
    error-vector overlaps bordered by ones, right-hand side e_last."""
    rng = np.random.default_rng(4)
    errs = [rng.standard_normal(49) * 10.0 ** -k for k in range(6)]
    m = len(errs)
    Bm = np.zeros((m + 1, m + 1))
    for i in range(m):
        for j in range(m):
            Bm[i, j] = errs[i] @ errs[j]
        Bm[i, m] = Bm[m, i] = 1.0
    rhs = np.zeros(m + 1)
    rhs[m] = 1.0
    dB, dr = dev(Bm), dev(rhs)
    info, _ = core.gesv(dB, dr)
    assert info == 0
    c = host(dr)
    assert abs(c[:m].sum() - 1.0) < 1e-10
    assert np.allclose(c, np.linalg.solve(Bm, rhs), rtol=1e-8, atol=1e-12)
