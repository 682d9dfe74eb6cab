"""Randomised GPU parity sweep of complex einsums.

The label structures of the randomised einsum sweep whose output labels all occur in an
input (roles, extents 0..40, diagonals, padded and stepped views), with complex128 or
complex64 operands and real or complex prefactors, through compile_plan/execute and
through the generic contraction. Expected: numpy.einsum in complex128 on the same values.
Tolerance: relative complex Frobenius 1e-13 (complex128), 1e-5 (complex64).
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.test_gpu_random_einsum import CASES as EINSUM_CASES

pytestmark = pytest.mark.gpu

CASES = [c for c in EINSUM_CASES if set(c["C"]) <= set(c["A"]) | set(c["B"])][:240]
PREFACTORS = [(0.0, 1.0), (1.0, -1.25), (0.5 - 0.25j, 1.0), (0.0, 0.5 + 1.0j), (-1.0j, 2.0 - 0.5j)]


def _ids(c):
    return f"{c['seed']}:{c['C']}<-{c['A']},{c['B']}:{c['route']}:{'c64' if c['dtype'] == np.float32 else 'c128'}"


def _operand(rng, labels, ext, view, dtype):
    shape = [ext[l] for l in labels]
    if view == "pad" and shape:
        lo = [int(rng.integers(0, 3)) for _ in shape]
        full = [s + o + int(rng.integers(0, 3)) for s, o in zip(shape, lo)]
        idx = tuple(slice(o, o + s) for o, s in zip(lo, shape))
    elif view == "step" and shape:
        d = int(rng.integers(len(shape)))
        full = [s * 2 if i == d else s for i, s in enumerate(shape)]
        idx = tuple(slice(None, None, 2) if i == d else slice(None) for i in range(len(shape)))
    else:
        full = list(shape)
        idx = tuple(slice(None) for _ in shape)
    base = np.asarray(rng.uniform(-1, 1, full) + 1j * rng.uniform(-1, 1, full)).astype(dtype)
    return base, torch.from_numpy(base.copy()).to("cuda"), idx


def _crel(x, ref):
    den = np.linalg.norm(np.asarray(ref, dtype=np.complex128).ravel())
    num = np.linalg.norm((np.asarray(x, dtype=np.complex128) - ref).ravel())
    return float(num / den) if den > 0 else float(num)


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_random_complex_einsum(core, case):
    rng = np.random.default_rng(zlib.crc32(f"random-complex|{case['seed']}".encode()))
    Ci, Ai, Bi, ext = case["C"], case["A"], case["B"], case["ext"]
    dtype = np.complex64 if case["dtype"] == np.float32 else np.complex128
    a_base, a_t, a_idx = _operand(rng, Ai, ext, case["views"][0], dtype)
    b_base, b_t, b_idx = _operand(rng, Bi, ext, case["views"][1], dtype)
    c_base, c_t, c_idx = _operand(rng, Ci, ext, case["c_view"], dtype)
    cp, ab = PREFACTORS[case["seed"] % len(PREFACTORS)]

    a = np.array(a_base[a_idx], dtype=np.complex128)
    b = np.array(b_base[b_idx], dtype=np.complex128)
    ref = cp * np.array(c_base[c_idx], dtype=np.complex128) + ab * np.einsum(f"{Ai},{Bi}->{Ci}", a, b)

    A, B, C = a_t[a_idx], b_t[b_idx], c_t[c_idx]
    if case["route"] == "plan":
        core.compile_plan(Ci, Ai, Bi).execute(cp, C, ab, A, B)
    else:
        core.contract(Ci, C, Ai, A, Bi, B, C_prefactor=cp, AB_prefactor=ab)
    torch.cuda.synchronize()
    out = c_t.cpu().numpy()
    assert _crel(out[c_idx], ref) < (1e-13 if dtype == np.complex128 else 1e-5)
    if case["c_view"] != "dense":
        outside = np.ones(c_base.shape, dtype=bool)
        outside[c_idx] = False
        assert np.array_equal(out[outside], c_base[outside])
