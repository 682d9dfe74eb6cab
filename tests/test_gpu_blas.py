"""GPU parity of the BLAS-shaped leaves (C ABI, column-major semantics) against the
CPU oracle. Tolerances: relative Frobenius 1e-13 (FP64), 1e-5 (FP32); integer-valued
known answers exact."""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
KA = json.loads((GOLD / "reference_known_answers.json").read_text())


def _lib():
    from einsums_amd import _lib

    return _lib


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


_alive = []


@pytest.fixture(autouse=True)
def _release_device_temporaries():
    yield
    torch.cuda.synchronize()
    _alive.clear()


def vd(a):
    """Device copy of a host array kept alive until the test ends, as a void*."""
    t = dev(a)
    _alive.append(t)
    return vp(t)


@pytest.mark.parametrize("trans", ["NN", "TN", "NT", "TT"])
def test_gemm_3x3_known_answers(trans):
    L = _lib()
    g = KA["gemm_3x3"]
    A = dev(np.array(g["A"], dtype=np.float64))
    B = dev(np.array(g["B"], dtype=np.float64))
    C = dev(np.full(9, np.nan))
    L.check(L.lib.ein_gemm_f64(trans[1].encode(), trans[0].encode(), 3, 3, 3, 1.0, vp(B), 3, vp(A), 3, 0.0, vp(C), 3,
                               None))
    assert host(C).tolist() == g["expected"][trans]


GEMM_SHAPES = [(1, 1, 1), (3, 5, 7), (17, 33, 65), (64, 64, 16), (128, 128, 16), (129, 257, 100), (300, 7, 2000),
               (70, 90, 5000), (256, 512, 48), (576, 576, 2304), (1000, 1000, 1000)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", GEMM_SHAPES)
def test_gemm_vs_oracle(oracle, dtype, ta, tb, m, n, k):
    L = _lib()
    rng = np.random.default_rng(m * 7919 + n * 31 + k)
    lda = (m if ta == "N" else k) + 3
    ldb = (k if tb == "N" else n) + 1
    ldc = m + 2
    A = rng.uniform(-1, 1, (k if ta == "N" else m) * lda).astype(dtype)
    B = rng.uniform(-1, 1, (n if tb == "N" else k) * ldb).astype(dtype)
    C0 = rng.uniform(-1, 1, n * ldc).astype(dtype)
    for alpha, beta in ((1.0, 0.0), (-0.75, 0.5)):
        ref = C0.astype(np.float64).copy()
        oracle.gemm_colmajor(ta, tb, m, n, k, alpha, A.astype(np.float64), lda, B.astype(np.float64), ldb, beta, ref, ldc)
        dA, dB, dC = dev(A), dev(B), dev(C0)
        fn = L.lib.ein_gemm_f64 if dtype == np.float64 else L.lib.ein_gemm_f32
        L.check(fn(ta.encode(), tb.encode(), m, n, k, alpha, vp(dA), lda, vp(dB), ldb, beta, vp(dC), ldc, None))
        got = host(dC)
        # padding rows of C (i >= m) must be untouched
        got2, ref2 = got.reshape(n, ldc), ref.reshape(n, ldc)
        assert np.array_equal(got2[:, m:], C0.reshape(n, ldc)[:, m:])
        tol = 1e-13 if dtype == np.float64 else 1e-5
        assert rel_fro(got2[:, :m], ref2[:, :m]) < tol


# Leading dimensions equal to the extents (no padding), so both operands take the 16-byte
# DMA staging. Edge tiles (m, n not multiples of 128), k not a multiple of 16, a stream-K
# tail (576 x 576 x 2304) and odd m (dword staging for A) are covered.
CONTIG_SHAPES = [(128, 128, 16), (256, 384, 100), (130, 260, 77), (576, 576, 2304), (1000, 1000, 1000),
                 (2304, 256, 48), (129, 200, 64)]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("tb", ["N", "T"])
@pytest.mark.parametrize("m,n,k", CONTIG_SHAPES)
def test_gemm_contiguous_vs_oracle(oracle, dtype, tb, m, n, k):
    L = _lib()
    rng = np.random.default_rng(m * 13 + n * 7 + k)
    ldb = k if tb == "N" else n
    A = rng.uniform(-1, 1, m * k).astype(dtype)
    B = rng.uniform(-1, 1, n * k).astype(dtype)
    C0 = rng.uniform(-1, 1, m * n).astype(dtype)
    fn = L.lib.ein_gemm_f64 if dtype == np.float64 else L.lib.ein_gemm_f32
    for alpha, beta in ((1.0, 0.0), (0.25, -1.5)):
        ref = C0.astype(np.float64).copy()
        oracle.gemm_colmajor("N", tb, m, n, k, alpha, A.astype(np.float64), m, B.astype(np.float64), ldb, beta, ref, m)
        dC = dev(C0)
        L.check(fn(b"N", tb.encode(), m, n, k, alpha, vd(A), m, vd(B), ldb, beta, vp(dC), m, None))
        assert rel_fro(host(dC), ref) < (1e-13 if dtype == np.float64 else 1e-5)


def test_gemm_beta_zero_clears_nan():
    L = _lib()
    rng = np.random.default_rng(5)
    A, B = rng.uniform(-1, 1, 40 * 30), rng.uniform(-1, 1, 30 * 20)
    dC = dev(np.full(40 * 20, np.nan))
    L.check(L.lib.ein_gemm_f64(b"N", b"N", 40, 20, 30, 1.0, vd((A)), 40, vd((B)), 30, 0.0, vp(dC), 40, None))
    assert np.isfinite(host(dC)).all()


def test_gemm_strided_batched(oracle):
    L = _lib()
    rng = np.random.default_rng(11)
    m, n, k, nb = 65, 47, 130, 9
    A = rng.uniform(-1, 1, (nb, k, m))  # each batch column-major m x k: A[b][p][i]
    B = rng.uniform(-1, 1, (nb, n, k))
    C = rng.uniform(-1, 1, (nb, n, m))
    dA, dB, dC = dev(A), dev(B), dev(C)
    L.check(L.lib.ein_gemm_strided_batched_f64(b"N", b"N", m, n, k, 1.5, vp(dA), m, k * m, vp(dB), k, n * k, 0.25, vp(dC),
                                               m, n * m, nb, None))
    got = host(dC)
    for b in range(nb):
        ref = C[b].reshape(-1).copy()
        oracle.gemm_colmajor("N", "N", m, n, k, 1.5, A[b].reshape(-1).copy(), m, B[b].reshape(-1).copy(), k, 0.25, ref, m)
        assert rel_fro(got[b].reshape(-1), ref) < 1e-13


@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("m,n", [(1, 1), (7, 300), (2304, 2304), (5000, 33), (33, 5000)])
@pytest.mark.parametrize("incx,incy", [(1, 1), (2, 3)])
def test_gemv(trans, m, n, incx, incy):
    L = _lib()
    rng = np.random.default_rng(m + 17 * n + incx)
    lda = m + 1
    A = rng.uniform(-1, 1, (n, lda))  # column-major storage: column j is A[j]
    lx, ly = (n, m) if trans == "N" else (m, n)
    x = rng.uniform(-1, 1, lx * incx)
    y0 = rng.uniform(-1, 1, ly * incy)
    Am = A[:, :m].T  # m x n matrix
    xv = x[::incx][:lx]
    ref = y0.copy()
    yv = (Am @ xv) if trans == "N" else (Am.T @ xv)
    ref[::incy][:ly] = 0.5 * yv - 2.0 * y0[::incy][:ly]
    dy = dev(y0)
    L.check(L.lib.ein_gemv_f64(trans.encode(), m, n, 0.5, vd((A)), lda, vd((x)), incx, -2.0, vp(dy), incy, None))
    assert rel_fro(host(dy), ref) < 1e-13


@pytest.mark.parametrize("n", [1, 100, 65536, 5308416, 16778453])
@pytest.mark.parametrize("inc", [1, 3])
def test_dot(n, inc):
    L = _lib()
    rng = np.random.default_rng(n + inc)
    x, y = rng.uniform(-1, 1, n * inc), rng.uniform(-1, 1, n * inc)
    r = dev(np.array([0.3]))
    L.check(L.lib.ein_dot_f64(n, 0.25, vd((x)), inc, vd((y)), inc, 2.0, vp(r), None))
    ref = 0.25 * float(np.dot(x[::inc][:n].astype(np.longdouble), y[::inc][:n].astype(np.longdouble))) + 0.6
    assert abs(host(r)[0] - ref) <= 1e-13 * max(1.0, abs(ref))
    # deterministic: a second call gives the identical bits
    r2 = dev(np.array([0.3]))
    L.check(L.lib.ein_dot_f64(n, 0.25, vd((x)), inc, vd((y)), inc, 2.0, vp(r2), None))
    assert host(r2)[0] == host(r)[0]


def test_ger_dirprod_scale_axpby():
    L = _lib()
    rng = np.random.default_rng(21)
    m, n = 300, 77
    x, y = rng.uniform(-1, 1, m), rng.uniform(-1, 1, n)
    A = rng.uniform(-1, 1, (n, m + 5))  # column-major lda = m + 5
    dA = dev(A)
    L.check(L.lib.ein_ger_f64(m, n, 1.5, vd((x)), 1, vd((y)), 1, vp(dA), m + 5, None))
    ref = A.copy()
    ref[:, :m] += 1.5 * np.outer(y, x)
    assert rel_fro(host(dA), ref) < 1e-14
    z0 = rng.uniform(-1, 1, 1000)
    xx, yy = rng.uniform(-1, 1, 1000), rng.uniform(-1, 1, 1000)
    dz = dev(z0)
    L.check(L.lib.ein_dirprod_f64(1000, 2.0, vd((xx)), 1, vd((yy)), 1, -1.0, vp(dz), 1, None))
    assert rel_fro(host(dz), 2.0 * xx * yy - z0) < 1e-15
    dz = dev(z0)
    L.check(L.lib.ein_scale_f64(500, 3.0, vp(dz), 2, None))
    ref = z0.copy()
    ref[::2] *= 3.0
    assert np.array_equal(host(dz), ref)
    dz = dev(z0)
    L.check(L.lib.ein_axpby_f64(1000, 2.0, vd((xx)), 1, 0.5, vp(dz), 1, None))
    assert rel_fro(host(dz), 2.0 * xx + 0.5 * z0) < 1e-15
