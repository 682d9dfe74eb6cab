"""Randomised GPU parity sweep of the einsum autograd layer (the adjoint kernels).

The cases are those of the randomised einsum sweep (label roles, extents 0..40,
diagonals, strided input views, FP64/FP32) whose output labels all occur in an input.
Forward C = alpha * contraction(A, B) and the gradients of sum(W * C) for a random W are
checked against torch.einsum's autograd in FP64 on the host: relative Frobenius 1e-13
(FP64) and 1e-5 (FP32).
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import rel_fro
from tests.test_gpu_random_einsum import CASES as EINSUM_CASES, _operand

pytestmark = pytest.mark.gpu

CASES = [c for c in EINSUM_CASES if set(c["C"]) <= set(c["A"]) | set(c["B"])][:300]


def _ids(c):
    return f"{c['seed']}:{c['C']}<-{c['A']},{c['B']}:{np.dtype(c['dtype']).name}"


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_random_einsum_gradients(case):
    from einsums_amd import autograd

    rng = np.random.default_rng(zlib.crc32(f"random-autograd|{case['seed']}".encode()))
    Ci, Ai, Bi, ext, dtype = case["C"], case["A"], case["B"], case["ext"], case["dtype"]
    a_base, a_t, a_idx = _operand(rng, Ai, ext, case["views"][0], dtype)
    b_base, b_t, b_idx = _operand(rng, Bi, ext, case["views"][1], dtype)
    W = rng.uniform(-1, 1, [ext[l] for l in Ci])
    alpha = case["ab"]

    A = a_t[a_idx].detach().requires_grad_(True)
    B = b_t[b_idx].detach().requires_grad_(True)
    C = autograd.einsum(Ci, Ai, A, Bi, B, alpha=alpha)
    (C * torch.from_numpy(np.asarray(W)).to(C)).sum().backward()

    hA = torch.from_numpy(np.array(a_base[a_idx], dtype=np.float64)).requires_grad_(True)
    hB = torch.from_numpy(np.array(b_base[b_idx], dtype=np.float64)).requires_grad_(True)
    hC = alpha * torch.einsum(f"{Ai},{Bi}->{Ci}", hA, hB)
    (hC * torch.from_numpy(np.asarray(W, dtype=np.float64))).sum().backward()

    tol = 1e-13 if dtype == np.float64 else 1e-5
    assert rel_fro(C.detach().cpu().numpy(), hC.detach().numpy()) < tol
    assert rel_fro(A.grad.cpu().numpy(), hA.grad.numpy()) < tol
    assert rel_fro(B.grad.cpu().numpy(), hB.grad.numpy()) < tol
