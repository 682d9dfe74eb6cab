"""Randomised GPU sweep of the device eigensolver (syev) and LU solve (gesv).

syev: 40 symmetric matrices (n 1..300) built as Q diag(lam) Q^T from drawn spectra --
uniform, clustered (groups of equal or nearly equal eigenvalues), graded over eight
decades, and low rank -- checked through the invariants that fix the result: eigenvalues
sorted and equal to numpy.linalg.eigvalsh, A V = V diag(W) and V^T V = I (tolerances
relative to |A|: FP64 1e-12, FP32 1e-4, eigenvalues 1e-5 in FP32). gesv: 40 systems (n
1..300, 1..8 right-hand sides) with condition numbers 1..1e6, checked through the backward
error |A X - B| / (|A| |X|) <= 1e-13 sqrt(n) (FP64) and 1e-5 sqrt(n) (FP32).
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host

pytestmark = pytest.mark.gpu


def _spectrum(rng, n, kind):
    if kind == "uniform":
        return rng.uniform(-5, 5, n)
    if kind == "clustered":
        centres = rng.uniform(-3, 3, max(1, n // 6))
        lam = rng.choice(centres, n)
        return lam + (rng.uniform(-1e-10, 1e-10, n) if rng.random() < 0.5 else 0.0)
    if kind == "graded":
        return rng.choice((-1.0, 1.0), n) * 10.0 ** rng.uniform(-8, 0, n)
    lam = np.zeros(n)  # low rank
    r = int(rng.integers(0, max(1, n // 3) + 1))
    lam[:r] = rng.uniform(-4, 4, r)
    return rng.permutation(lam)


SYEV = []
for i in range(40):
    r = np.random.default_rng(zlib.crc32(f"random-syev|{i}".encode()))
    SYEV.append(dict(n=int(r.choice((1, 2, 3, int(r.integers(4, 64)), int(r.integers(64, 301))))),
                     kind=("uniform", "clustered", "graded", "lowrank")[i % 4],
                     dtype=np.float32 if i % 5 == 4 else np.float64, seed=i))
GESV = []
for i in range(40):
    r = np.random.default_rng(zlib.crc32(f"random-gesv|{i}".encode()))
    GESV.append(dict(n=int(r.choice((1, 2, int(r.integers(3, 64)), int(r.integers(64, 301))))),
                     nrhs=int(r.integers(1, 9)), logcond=float(r.uniform(0, 6)),
                     dtype=np.float32 if i % 5 == 4 else np.float64, seed=i))


@pytest.mark.parametrize("case", SYEV, ids=lambda c: f"{c['seed']}:n{c['n']}:{c['kind']}:{np.dtype(c['dtype']).name}")
def test_random_syev(core, case):
    rng = np.random.default_rng(zlib.crc32(f"random-syev-data|{case['seed']}".encode()))
    n, dtype = case["n"], case["dtype"]
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Q * _spectrum(rng, n, case["kind"])) @ Q.T
    A = ((A + A.T) / 2).astype(dtype)
    dA = dev(A)
    W = torch.empty(n, dtype=dA.dtype, device="cuda")
    core.syev(dA, W)
    vals, vecs, A64 = host(W).astype(np.float64), host(dA).astype(np.float64), A.astype(np.float64)
    tol = 1e-12 if dtype == np.float64 else 1e-4
    scale = max(1.0, np.linalg.norm(A64))
    assert np.all(np.diff(vals) >= 0)
    assert np.abs(vals - np.linalg.eigvalsh(A64)).max() <= (tol if dtype == np.float64 else 1e-5) * scale
    assert np.linalg.norm(A64 @ vecs - vecs * vals) <= tol * scale * max(1, n) ** 0.5
    assert np.abs(vecs.T @ vecs - np.eye(n)).max() <= tol * max(1, n) ** 0.5


@pytest.mark.parametrize("case", GESV, ids=lambda c: f"{c['seed']}:n{c['n']}x{c['nrhs']}:{np.dtype(c['dtype']).name}")
def test_random_gesv(core, case):
    rng = np.random.default_rng(zlib.crc32(f"random-gesv-data|{case['seed']}".encode()))
    n, nrhs, dtype = case["n"], case["nrhs"], case["dtype"]
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = ((U * np.logspace(0, -case["logcond"], n)) @ V.T).astype(dtype)
    B = rng.standard_normal((n, nrhs)).astype(dtype)
    dA, dB = dev(A), dev(B)
    core.gesv(dA, dB)
    X = host(dB).astype(np.float64)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    berr = np.linalg.norm(A64 @ X - B64) / (np.linalg.norm(A64) * max(np.linalg.norm(X), 1e-300))
    assert berr <= (1e-13 if dtype == np.float64 else 1e-5) * max(1, n) ** 0.5
