"""GPU parity of the reduction regime of the generic contraction (no MFMA: one operand
carries only link indices) and of the transposed GEMV at shapes that take the cut-K paths:
four waves per target adding in LDS, and in addition partials summed by the finish pass.
Tolerances (relative Frobenius): FP64 1e-13, FP32 1e-5."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro
from tests.test_gpu_einsum import run_both

pytestmark = pytest.mark.gpu

# K_pq <- g_prqs D_rs (the HF exchange build) and the Coulomb order; extents chosen so the
# K range is cut into 4 pieces (one partial) or 8 and more (two or more partials), with
# pieces of unequal length.
REDUCE = [
    ("pq", "prqs", "rs", dict(p=32, q=32, r=32, s=32)),
    ("pq", "prqs", "rs", dict(p=48, q=48, r=48, s=48)),
    ("pq", "prqs", "rs", dict(p=20, q=30, r=37, s=41)),
    ("pq", "prqs", "rs", dict(p=10, q=10, r=50, s=61)),
    ("pq", "pqrs", "rs", dict(p=24, q=40, r=45, s=47)),
    ("q", "rqs", "sr", dict(q=7, r=300, s=90)),
    ("pq", "rspq", "sr", dict(p=12, q=9, r=40, s=33)),
    ("pq", "prqs", "rs", dict(p=40, q=36, r=6, s=20)),   # one wave per target (K < 1024)
]


# Even (FP64) or multiple-of-4 (FP32) innermost link extents take 16-byte vector loads.
@pytest.mark.parametrize("pat", REDUCE, ids=lambda p: f"{p[0]}<-{p[1]},{p[2]}-" + "x".join(map(str, p[3].values())))
@pytest.mark.parametrize("cp", [0.0, 0.5])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 1e-5)])
def test_reduction_cut_k_vs_oracle(core, oracle, pat, cp, dtype, tol):
    Ci, Ai, Bi, ext = pat
    got, ref, _ = run_both(core, oracle, cp, Ci, Ai, Bi, ext, dtype=dtype, generic_only=True)
    assert rel_fro(got, ref) < tol


def _lib():
    from einsums_amd import _lib as L
    return L


# (m, n, lda): long contiguous columns (16-byte loads) with one to four waves per column,
# odd m (scalar tail), few columns (cut across block rows into partials), and a strided lda.
GEMV_T = [(4608, 1200, 4608), (4097, 64, 4098), (20000, 7, 20000), (100000, 3, 100000), (600, 5000, 600),
          (9000, 40, 9001), (300, 2, 300)]


@pytest.mark.parametrize("m,n,lda", GEMV_T)
@pytest.mark.parametrize("beta", [0.0, -2.0])
def test_gemv_t_cut_columns(m, n, lda, beta):
    L = _lib()
    rng = np.random.default_rng(m + 13 * n)
    A = rng.uniform(-1, 1, (n, lda))  # column j of the column-major matrix is A[j]
    x = rng.uniform(-1, 1, m)
    y0 = rng.uniform(-1, 1, n)
    ref = 0.5 * (A[:, :m].astype(np.longdouble) @ x.astype(np.longdouble)) + beta * y0
    dA, dx, dy = dev(A), dev(x), dev(y0)
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    L.check(L.lib.ein_gemv_f64(b"T", m, n, 0.5, vp(dA), lda, vp(dx), 1, beta, vp(dy), 1, None))
    assert rel_fro(host(dy), ref.astype(np.float64)) < 1e-13
