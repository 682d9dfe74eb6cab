"""GEMM with C in host memory, streamed through the GPU in column panels (the out-of-core
output path): against torch's GEMM of the same operands. Panels are made small so that
many of them, and a short last one, pass through both device buffers. Tolerances (relative
Frobenius): FP64 1e-13, split precision 1e-11, FP32 1e-5."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import rel_fro

pytestmark = pytest.mark.gpu


def _case(m, n, k, ta, tb, dtype, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    A = torch.rand((k, m) if ta else (m, k), generator=g, dtype=dtype) * 2 - 1
    B = torch.rand((n, k) if tb else (k, n), generator=g, dtype=dtype) * 2 - 1
    C0 = torch.rand((m, n), generator=g, dtype=dtype) * 2 - 1
    return A.cuda(), B.cuda(), C0


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("beta", [0.0, -0.5])
def test_host_c_panels_fp64(core, ta, tb, beta):
    m, n, k = 300, 1000, 200
    A, B, C0 = _case(m, n, k, ta == "T", tb == "T", torch.float64, 17)
    C = torch.empty((m, n), dtype=torch.float64, pin_memory=True)
    C.copy_(C0)
    # row-major C is streamed as column-major C^T (n = 1000 rows): 88 000 B = 11 columns
    # of C^T per panel -> 28 panels, the last of 3
    core.gemm_host_c(ta, tb, 1.5, A, B, beta, C, panel_bytes=88_000)
    opA = A.t() if ta == "T" else A
    opB = B.t() if tb == "T" else B
    ref = (1.5 * (opA.contiguous() @ opB.contiguous()) + beta * C0.cuda()).cpu()
    assert rel_fro(C.numpy(), ref.numpy()) < 1e-13


def test_host_c_split_and_fp32_and_pageable(core):
    m, n, k = 256, 384, 320
    A, B, C0 = _case(m, n, k, False, False, torch.float64, 5)
    C = torch.empty((m, n), dtype=torch.float64, pin_memory=True)
    C.copy_(C0)
    core.gemm_host_c("N", "N", 1.0, A, B, 1.0, C, panel_bytes=100_000, split=6, slice_type="int8")
    ref = (A @ B + C0.cuda()).cpu().numpy()
    err = rel_fro(C.numpy(), ref)
    assert 1e-16 < err < 1e-11  # the split kernel ran (not bit-identical to FP64)
    A32, B32, C32 = _case(m, n, k, False, True, torch.float32, 6)
    C = C32.clone().pin_memory()
    core.gemm_host_c("N", "T", 2.0, A32, B32, 0.25, C, panel_bytes=50_000)
    ref = (2.0 * (A32.double() @ B32.double().t()) + 0.25 * C32.double().cuda()).cpu().numpy()
    assert rel_fro(C.numpy(), ref) < 1e-5
    # pageable host memory, one panel
    Cp = C0.clone()
    core.gemm_host_c("N", "N", 1.0, A, B, 0.0, Cp)
    assert rel_fro(Cp.numpy(), (A @ B).cpu().numpy()) < 1e-13


def test_host_c_argument_errors(core):
    A = torch.rand(4, 5, dtype=torch.float64, device="cuda")
    B = torch.rand(5, 6, dtype=torch.float64, device="cuda")
    with pytest.raises(core.DimensionError):
        core.gemm_host_c("N", "N", 1.0, A, B, 0.0, torch.empty(4, 7, dtype=torch.float64))
    with pytest.raises(ValueError):  # C on the device
        core.gemm_host_c("N", "N", 1.0, A, B, 0.0, torch.empty(4, 6, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):  # bad trans character: argument 2 of the column-major call
        core.gemm_host_c("X", "N", 1.0, A, torch.rand(4, 6, dtype=torch.float64, device="cuda"), 0.0,
                         torch.empty(5, 6, dtype=torch.float64))
