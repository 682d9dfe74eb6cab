"""The fused Fock build (ein_fock_jk_f64: J and K from one pass over g) against the CPU
oracle's two einsum calls of the reference's Fock build (QChemTests.cpp:299-304):
  F = beta*F + alpha_j * einsum(ij <- ijkl g, kl D) + alpha_k * einsum(ij <- ikjl g, kl D).
Cases: the HF workload at nbf = 48 (8-fold symmetric g, idempotent D), unsymmetric random g
and D at extents that exercise every lane/row split of the kernel (n = 1, 7, 17, 63, 64, 65,
96, 128), a padded leading dimension, beta != 0, bit-for-bit repeatability, and the
two-element-load kernel (even n <= 64, even leading dimension, aligned D) beside the
one-element ones it falls back to (odd n, n > 64, odd leading dimension, unaligned D).
Tolerance: FP64 relative Frobenius 1e-13."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from einsums_amd import workloads
from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu


def oracle_fock(oracle, g, D, alpha_j, alpha_k, beta, F0):
    F = F0.copy()
    oracle.einsum(beta, "ij", F, alpha_j, "kl", D, "ijkl", g)
    oracle.einsum(1.0, "ij", F, alpha_k, "kl", D, "ikjl", g)
    return F


def test_fock_hf_nbf48(core, oracle):
    g, D = workloads.hf_inputs(48, 12, seed=2024)
    dg, dD = dev(g), dev(D)
    G = torch.full((48, 48), float("nan"), dtype=torch.float64, device="cuda")  # beta = 0 clears NaN
    core.fock_jk(dg, dD, G)  # G = 2J - K
    ref = oracle_fock(oracle, g, D, 2.0, -1.0, 0.0, np.zeros((48, 48)))
    hG = host(G)
    assert rel_fro(hG, ref) < 1e-13
    assert rel_fro(hG, hG.T) < 1e-13  # symmetric g and D give a symmetric G
    # J alone and K alone
    J = torch.zeros((48, 48), dtype=torch.float64, device="cuda")
    K = torch.zeros((48, 48), dtype=torch.float64, device="cuda")
    core.fock_jk(dg, dD, J, alpha_j=1.0, alpha_k=0.0)
    core.fock_jk(dg, dD, K, alpha_j=0.0, alpha_k=1.0)
    rJ, rK = np.zeros((48, 48)), np.zeros((48, 48))
    oracle.einsum(0.0, "pq", rJ, 1.0, "pqrs", g, "rs", D)
    oracle.einsum(0.0, "pq", rK, 1.0, "prqs", g, "rs", D)
    assert rel_fro(host(J), rJ) < 1e-13 and rel_fro(host(K), rK) < 1e-13


@pytest.mark.parametrize("n", [1, 7, 17, 63, 64, 65, 96, 128])
def test_fock_random(core, oracle, n):
    rng = np.random.default_rng(900 + n)
    g = rng.uniform(-1, 1, (n, n, n, n))
    D = rng.uniform(-1, 1, (n, n))
    F0 = rng.uniform(-1, 1, (n, n))
    F = dev(F0)
    core.fock_jk(dev(g), dev(D), F, alpha_j=0.75, alpha_k=-1.25, beta=0.5)
    ref = oracle_fock(oracle, g, D, 0.75, -1.25, 0.5, F0)
    assert rel_fro(host(F), ref) < 1e-13


def test_fock_padded_and_repeatable(core, oracle):
    n = 40
    rng = np.random.default_rng(5)
    g = rng.standard_normal((n, n, n, n))
    Dbig = dev(rng.standard_normal((n, n + 9)))
    Fbig = torch.zeros((n, n + 3), dtype=torch.float64, device="cuda")
    D, F = Dbig[:, :n], Fbig[:, :n]
    dg = dev(g)
    core.fock_jk(dg, D, F)
    first = Fbig.clone()
    core.fock_jk(dg, D, F)
    torch.cuda.synchronize()
    assert torch.equal(first, Fbig)  # same bits
    assert torch.all(Fbig[:, n:] == 0)  # padding untouched
    ref = oracle_fock(oracle, g, host(D).copy(), 2.0, -1.0, 0.0, np.zeros((n, n)))
    assert rel_fro(host(F), ref) < 1e-13


@pytest.mark.parametrize("n,pad,shift", [(2, 0, 0), (32, 8, 0), (40, 8, 0), (48, 2, 1), (64, 0, 0), (66, 2, 0),
                                         (96, 4, 0), (96, 0, 1), (128, 2, 0)])
def test_fock_two_element_loads(core, oracle, n, pad, shift):
    # even n <= 64 with an even leading dimension and 16-byte aligned D reads two elements
    # per lane; shift = 1 moves D by one element and n > 64 runs the one-element kernel
    rng = np.random.default_rng(70 + n)
    g = rng.uniform(-1, 1, (n, n, n, n))
    Dbuf = dev(rng.uniform(-1, 1, (n * (n + pad) + shift,)))
    D = Dbuf[shift:].view(n, n + pad)[:, :n]
    F0 = rng.uniform(-1, 1, (n, n))
    F = dev(F0)
    core.fock_jk(dev(g), D, F, alpha_j=1.5, alpha_k=-0.5, beta=-1.0)
    ref = oracle_fock(oracle, g, host(D).copy(), 1.5, -0.5, -1.0, F0)
    assert rel_fro(host(F), ref) < 1e-13


def test_fock_rejects_large_n(core):
    from einsums_amd import _lib

    g = torch.zeros(1, dtype=torch.float64, device="cuda")
    # rejected on the host before any launch
    assert _lib.lib.ein_fock_jk_f64(129, g.data_ptr(), g.data_ptr(), 129, 1.0, 1.0, 0.0, g.data_ptr(), 129, None) == 1005
