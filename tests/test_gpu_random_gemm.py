"""Randomised GPU parity sweep of the dense GEMM (linear_algebra::gemm) and the CRT split
GEMM against an FP64 host product of the same values.

Each case draws m, n, k (0..700, with a bias towards tile edges: 63..65, 95..97,
127..129, 255..257), both transposes, padded leading dimensions, alpha and beta (beta = 0
with a NaN-filled C, which must not be read) and the element type, from a seed fixed by
the case number. Tolerance: relative Frobenius 1e-13 (FP64) and 1e-5 (FP32) against the
host product; the CRT split with the automatic or a drawn moduli count L, max(8 * 2^-kb,
2e-14) of |alpha * op(A) op(B)| with kb the bits it keeps per element at that k.
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import host

pytestmark = pytest.mark.gpu

EDGES = (1, 2, 63, 64, 65, 95, 96, 97, 127, 128, 129, 255, 256, 257)


def _extent(rng):
    r = rng.random()
    if r < 0.35:
        return int(rng.choice(EDGES))
    if r < 0.38:
        return 0
    if r < 0.7:
        return int(rng.integers(1, 64))
    return int(rng.integers(64, 701))


def make_case(idx, split=False):
    rng = np.random.default_rng(zlib.crc32(f"random-gemm|{split}|{idx}".encode()))
    m, n, k = _extent(rng), _extent(rng), _extent(rng)
    if split and k == 0:
        k = 1
    ta, tb = str(rng.choice(("N", "T"))), str(rng.choice(("N", "T")))
    pads = [int(rng.choice((0, 0, 1, 3, 16))) for _ in range(3)]
    alpha = float(rng.choice((1.0, -0.5, 2.0)))
    beta = float(rng.choice((0.0, 1.0, 0.25)))
    dtype = np.float32 if (not split and idx % 3 == 2) else np.float64
    nmod = "auto" if idx % 2 == 0 else int(rng.integers(8, 17))
    return dict(m=m, n=n, k=k, ta=ta, tb=tb, pads=pads, alpha=alpha, beta=beta, dtype=dtype, nmod=nmod, seed=idx)


CASES = [make_case(i) for i in range(240)]
SPLIT_CASES = [make_case(i, split=True) for i in range(60)]


def _ids(c):
    return f"{c['seed']}:{c['m']}x{c['n']}x{c['k']}:{c['ta']}{c['tb']}:{np.dtype(c['dtype']).name}"


def _padded(rng, rows, cols, pad, dtype):
    """A rows x cols row-major device matrix whose row stride is cols + pad, and its values."""
    pad = max(pad, 1) if cols == 0 else pad  # a leading dimension of at least 1
    full = rng.uniform(-1, 1, (max(rows, 1), cols + pad)).astype(dtype)
    # as_strided keeps the row stride of an empty matrix (a plain slice would drop it)
    t = torch.from_numpy(full).to("cuda").as_strided((rows, cols), (cols + pad, 1))
    return t, full[:rows, :cols].astype(np.float64)


def _operands(case):
    rng = np.random.default_rng(zlib.crc32(f"random-gemm-data|{case['seed']}|{case['nmod']}".encode()))
    m, n, k, dtype = case["m"], case["n"], case["k"], case["dtype"]
    A, a = _padded(rng, *((k, m) if case["ta"] == "T" else (m, k)), case["pads"][0], dtype)
    B, b = _padded(rng, *((n, k) if case["tb"] == "T" else (k, n)), case["pads"][1], dtype)
    C, c = _padded(rng, m, n, case["pads"][2], dtype)
    if case["beta"] == 0.0:
        C.fill_(float("nan"))
        c = np.zeros((m, n))
    opa = a.T if case["ta"] == "T" else a
    opb = b.T if case["tb"] == "T" else b
    prod = case["alpha"] * (opa @ opb) if k else np.zeros((m, n))
    return A, B, C, prod, prod + case["beta"] * c


def _rel(got, ref, scale):
    den = np.linalg.norm(scale.ravel())
    num = np.linalg.norm((np.asarray(got, dtype=np.float64) - ref).ravel())
    return float(num / den) if den > 0 else float(num)


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_random_gemm(core, case):
    A, B, C, prod, ref = _operands(case)
    core.gemm(case["ta"], case["tb"], case["alpha"], A, B, case["beta"], C)
    got = host(C)
    assert not np.isnan(got).any()
    assert _rel(got, ref, ref) < (1e-13 if case["dtype"] == np.float64 else 1e-5)


@pytest.mark.parametrize("case", SPLIT_CASES, ids=lambda c: _ids(c) + f":crt{c['nmod']}")
def test_random_crt_gemm(core, case):
    A, B, C, prod, ref = _operands(case)
    nmod = case["nmod"]
    core.gemm_split(case["ta"], case["tb"], case["alpha"], A, B, case["beta"], C, nmod, "crt")
    got = host(C)
    assert not np.isnan(got).any()
    L = core.crt_moduli(case["k"]) if nmod == "auto" else nmod
    tol = max(8.0 * 2.0 ** -core.crt_bits(L, case["k"]), 2e-14)
    assert _rel(got, ref, prod) <= tol
