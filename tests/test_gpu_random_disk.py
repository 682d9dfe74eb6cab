"""Randomised GPU sweep of the GEMM with C in an HDF5 disk tensor (gemm_disk_c).

Each case draws C's dims (rank 1..4; the leading dims multiply to C's rows m, the rest to
its columns n), k, the transposes, alpha and beta, the panel size (from one leading slice
up) and FP64/FP32, from a seed fixed by the case number. Expected: the FP64 host product
of the same values plus beta times the old C. Tolerance: relative Frobenius 1e-13 (FP64),
1e-5 (FP32).
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import rel_fro

pytestmark = pytest.mark.gpu

NCASES = 24


def make_case(idx):
    rng = np.random.default_rng(zlib.crc32(f"random-disk|{idx}".encode()))
    rank = int(rng.integers(1, 5))
    dims = [int(rng.integers(1, 13)) for _ in range(rank)]
    p = int(rng.integers(1, rank + 1))  # dims[:p] -> m, dims[p:] -> n
    return dict(dims=dims, p=p, k=int(rng.integers(1, 40)), ta=str(rng.choice(("N", "T"))),
                tb=str(rng.choice(("N", "T"))), alpha=float(rng.choice((1.0, -0.5))),
                beta=float(rng.choice((0.0, 1.0, 0.5))), slices=int(rng.integers(1, 4)),
                dtype=np.float32 if idx % 4 == 3 else np.float64, seed=idx)


CASES = [make_case(i) for i in range(NCASES)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['seed']}:{'x'.join(map(str, c['dims']))}|{c['p']}:k{c['k']}")
def test_random_gemm_disk(tmp_path, case):
    from einsums_amd import disk

    rng = np.random.default_rng(zlib.crc32(f"random-disk-data|{case['seed']}".encode()))
    dims, p, k, dtype = case["dims"], case["p"], case["k"], case["dtype"]
    m, n = int(np.prod(dims[:p])), int(np.prod(dims[p:]))
    a = rng.uniform(-1, 1, (k, m) if case["ta"] == "T" else (m, k)).astype(dtype)
    b = rng.uniform(-1, 1, (n, k) if case["tb"] == "T" else (k, n)).astype(dtype)
    c0 = rng.uniform(-1, 1, dims).astype(dtype)
    C = disk.DiskTensor(tmp_path / "c.h5", "C", dims, dtype=dtype)
    C.put(c0)
    slice_bytes = (m * n // dims[0]) * np.dtype(dtype).itemsize  # one leading slice
    disk.gemm_disk_c(case["ta"], case["tb"], case["alpha"], torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(),
                     case["beta"], C, panel_bytes=slice_bytes * case["slices"])
    opa = a.astype(np.float64).T if case["ta"] == "T" else a.astype(np.float64)
    opb = b.astype(np.float64).T if case["tb"] == "T" else b.astype(np.float64)
    ref = case["alpha"] * (opa @ opb) + case["beta"] * c0.astype(np.float64).reshape(m, n)
    assert rel_fro(C.get().reshape(m, n), ref) < (1e-13 if dtype == np.float64 else 1e-5)
