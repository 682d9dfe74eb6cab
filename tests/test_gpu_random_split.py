"""Randomised GPU parity sweep of split-precision einsums (the CRT scheme on int8 MFMA).

The FP64 cases of the randomised einsum sweep (label roles, extents, diagonals, strided
views, prefactors) run through compile_plan/execute with split="auto" (the automatic
moduli count) or a drawn count of 8..16 moduli, plus one refinement step on every fifth
case. Products that map to a GEMM run in split precision; the others take their FP64
kernels. Expected: the oracle's long-double loop on the same values. Tolerance: relative
Frobenius of the error to |AB_prefactor * A B|, at most max(8 * 2^-kb, 2e-14) with kb the
bits the scheme keeps per element at the contraction length (taken as the product of the
extents of every summed label, which can only loosen the bound), and 4e-14 with the
refinement step.
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.test_gpu_random_einsum import CASES as EINSUM_CASES, _operand

pytestmark = pytest.mark.gpu

CASES = [c for c in EINSUM_CASES if c["dtype"] == np.float64][:200]


def _ids(c):
    return f"{c['seed']}:{c['C']}<-{c['A']},{c['B']}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_random_split_einsum(core, oracle, case):
    rng = np.random.default_rng(zlib.crc32(f"random-split|{case['seed']}".encode()))
    Ci, Ai, Bi, ext = case["C"], case["A"], case["B"], case["ext"]
    a_base, a_t, a_idx = _operand(rng, Ai, ext, case["views"][0], np.float64)
    b_base, b_t, b_idx = _operand(rng, Bi, ext, case["views"][1], np.float64)
    c_base, c_t, c_idx = _operand(rng, Ci, ext, case["c_view"], np.float64)
    a = np.array(a_base[a_idx], dtype=np.float64, copy=True, order="C")
    b = np.array(b_base[b_idx], dtype=np.float64, copy=True, order="C")
    c0 = np.array(c_base[c_idx], dtype=np.float64, copy=True, order="C")

    prod = np.zeros_like(c0)
    oracle.einsum_generic(0.0, Ci, prod, case["ab"], Ai, a, Bi, b, extended=True)
    ref = c0.copy()
    oracle.einsum_generic(case["cp"], Ci, ref, case["ab"], Ai, a, Bi, b, extended=True)

    nmod = "auto" if case["seed"] % 2 == 0 else 8 + case["seed"] % 9
    refine = case["seed"] % 5 == 0
    core.compile_plan(Ci, Ai, Bi).execute(case["cp"], c_t[c_idx], case["ab"], a_t[a_idx], b_t[b_idx], split=nmod,
                                          slice_type="crt+refine" if refine else "crt")
    torch.cuda.synchronize()
    got = c_t.cpu().numpy()

    k = 1
    for l in (set(Ai) | set(Bi)) - set(Ci):  # every summed label, links and one-operand labels
        k *= ext[l]
    k = max(k, 1)
    L = core.crt_moduli(k) if nmod == "auto" else nmod
    tol = 4e-14 if refine else max(8.0 * 2.0 ** -core.crt_bits(L, k), 2e-14)
    den = np.linalg.norm(prod.ravel())
    err = np.linalg.norm((got[c_idx] - ref).ravel())
    assert (err / den if den > 0 else err) <= tol
    if case["c_view"] != "dense":
        outside = np.ones(c_base.shape, dtype=bool)
        outside[c_idx] = False
        assert np.array_equal(got[outside], c_base[outside])
