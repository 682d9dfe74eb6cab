"""Randomised GPU parity sweep of the GEMM with C in host memory (gemm_host_c).

The shapes, transposes, padded leading dimensions, alpha and beta of the randomised GEMM
sweep, with C a pinned or pageable host matrix, panels of a drawn size (one column, a few
columns, 64 KiB or the default), and on a quarter of the FP64 cases the CRT split with the
automatic moduli count. Expected: an FP64 host product of the same values. Tolerance: as
in the GEMM sweep (1e-13 FP64, 1e-5 FP32; the CRT bound of |alpha * op(A) op(B)|).
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.test_gpu_random_gemm import CASES as GEMM_CASES, _padded

pytestmark = pytest.mark.gpu

CASES = GEMM_CASES[:80]


def _ids(c):
    return f"{c['seed']}:{c['m']}x{c['n']}x{c['k']}:{c['ta']}{c['tb']}:{np.dtype(c['dtype']).name}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_random_gemm_host_c(core, case):
    rng = np.random.default_rng(zlib.crc32(f"random-hostc|{case['seed']}".encode()))
    m, n, k, dtype = case["m"], case["n"], case["k"], case["dtype"]
    A, a = _padded(rng, *((k, m) if case["ta"] == "T" else (m, k)), case["pads"][0], dtype)
    B, b = _padded(rng, *((n, k) if case["tb"] == "T" else (k, n)), case["pads"][1], dtype)
    ldc = n + case["pads"][2] + (1 if n == 0 else 0)
    pinned = case["seed"] % 3 != 1
    full = torch.from_numpy(rng.uniform(-1, 1, (max(m, 1), ldc)).astype(dtype))
    if pinned:
        full = full.pin_memory()
    C = full.as_strided((m, n), (ldc, 1))
    c0 = C.numpy().astype(np.float64).copy()
    if case["beta"] == 0.0:
        C.fill_(float("nan"))
        c0 = np.zeros((m, n))
    opa = a.T if case["ta"] == "T" else a
    opb = b.T if case["tb"] == "T" else b
    prod = case["alpha"] * (opa @ opb) if k else np.zeros((m, n))
    ref = prod + case["beta"] * c0
    outside = full.numpy()[:, n:].copy()

    col = max(n, 1) * np.dtype(dtype).itemsize  # one panel column: a row of the row-major C
    panel = [0, col, col * int(rng.integers(2, 9)), 64 << 10][case["seed"] % 4]
    split = "auto" if dtype == np.float64 and case["seed"] % 4 == 1 and k > 0 else 0
    core.gemm_host_c(case["ta"], case["tb"], case["alpha"], A, B, case["beta"], C, panel_bytes=panel, split=split,
                     slice_type="crt")
    got = C.numpy().astype(np.float64)
    assert not np.isnan(got).any()
    scale = np.linalg.norm((prod if split else ref).ravel())
    err = np.linalg.norm((got - ref).ravel())
    if split:
        tol = max(8.0 * 2.0 ** -core.crt_bits(core.crt_moduli(k), k), 2e-14)
    else:
        tol = 1e-13 if dtype == np.float64 else 1e-5
    assert (err / scale if scale > 0 else err) <= tol
    assert np.array_equal(full.numpy()[:, n:], outside)  # the padding columns are not written
