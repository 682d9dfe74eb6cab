"""Gradients through the adjoint einsum kernels (torch.autograd.Function), checked against
gradients assembled from CPU oracle contractions, central finite differences
(h = 1e-6 on 64 random entries) and torch.autograd.gradcheck on small cases.
Tolerances: adjoint vs oracle gradient, relative Frobenius 1e-12 (FP64) / 1e-5 (FP32);
finite differences, 1e-6 relative per entry."""
import math
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu


def seeded(*parts):
    return np.random.default_rng(zlib.crc32("|".join(map(str, parts)).encode()))


def antisym(x):
    x = (x - x.transpose(1, 0, 2, 3)) / 2
    return (x - x.transpose(0, 1, 3, 2)) / 2


def diff_inputs(o=32, v=128, seed=11):
    rng = seeded("diff", o, v, seed)
    t = antisym(0.01 * rng.standard_normal((o, o, v, v)))
    g = antisym(rng.standard_normal((o, o, v, v)) / math.sqrt(v))
    W = antisym(rng.standard_normal((o, o, o, o)) / math.sqrt(v))
    return t, g, W


def L_of(ag, t, g, W):
    """L(t) = 0.25 sum t_ijab g_ijab + 0.125 sum t_ijab t_klab W_klij."""
    e1 = ag.einsum("", "ijab", t, "ijab", g, alpha=0.25)
    X = ag.einsum("ijkl", "ijab", t, "klab", t)
    e2 = ag.einsum("", "ijkl", X, "klij", W, alpha=0.125)
    return e1 + e2


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_ccd_functional_gradient_vs_oracle(oracle, dtype):
    from einsums_amd import autograd as ag

    t, g, W = diff_inputs()
    dt = dev(t).to(dtype).requires_grad_(True)
    L = L_of(ag, dt, dev(g).to(dtype), dev(W).to(dtype))
    L.backward()
    got = host(dt.grad).astype(np.float64)
    # dL/dt_pqcd = 0.25 g_pqcd + 0.125 (sum_kl t_klcd W_klpq + sum_ij t_ijcd W_pqij)
    ref = 0.25 * g.copy()
    oracle.einsum(1.0, "pqcd", ref, 0.125, "klcd", t, "klpq", W)
    oracle.einsum(1.0, "pqcd", ref, 0.125, "ijcd", t, "pqij", W)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert rel_fro(got, ref) < tol


def test_ccd_functional_finite_differences():
    from einsums_amd import autograd as ag

    t, g, W = diff_inputs()
    dg, dW = dev(g), dev(W)
    dt = dev(t).requires_grad_(True)
    L_of(ag, dt, dg, dW).backward()
    grad = host(dt.grad)
    rng = seeded("fd")
    h = 1e-6
    with torch.no_grad():
        base = dev(t)
        for _ in range(64):
            idx = tuple(int(rng.integers(0, n)) for n in t.shape)
            tp, tm = base.clone(), base.clone()
            tp[idx] += h
            tm[idx] -= h
            fd = (float(L_of(ag, tp, dg, dW)) - float(L_of(ag, tm, dg, dW))) / (2 * h)
            assert abs(fd - grad[idx]) <= 1e-6 * max(abs(grad[idx]), 1e-3), (idx, fd, grad[idx])


def test_wmbej_backward_both_inputs(oracle):
    """Backward of the generic-layout Wmbej contraction W_mbej = -0.5 t_jnfb g_mnef at the
    differentiation size o=32, v=128 with a seeded N(0,1) cotangent: both input gradients
    against the oracle's gradient contractions on 16 sampled slabs each (the full contractions
    are 1.4e11 flops through the oracle's generic loop), then central differences on 64
    random entries of each input."""
    from einsums_amd import autograd as ag

    o, v = 32, 128
    rng = seeded("wmbej-bwd")
    t = rng.standard_normal((o, o, v, v)) * 0.01
    g = rng.standard_normal((o, o, v, v)) / math.sqrt(v)
    up = rng.standard_normal((o, v, v, o))
    dt = dev(t).requires_grad_(True)
    dg = dev(g).requires_grad_(True)
    W = ag.einsum("mbej", "jnfb", dt, "mnef", dg, alpha=-0.5)
    W.backward(dev(up))
    gt, gg = host(dt.grad), host(dg.grad)
    for _ in range(16):
        j, n, m = (int(x) for x in rng.integers(0, o, 3))
        # dL/dt[j, n, f, b] = -0.5 sum_me up[m, b, e, j] g[m, n, e, f]
        rt = np.zeros((v, v))
        oracle.einsum(0.0, "fb", rt, -0.5, "mbe", np.ascontiguousarray(up[:, :, :, j]), "mef",
                      np.ascontiguousarray(g[:, n]))
        assert rel_fro(gt[j, n], rt) < 1e-12
        # dL/dg[m, n, e, f] = -0.5 sum_bj up[m, b, e, j] t[j, n, f, b]
        rg = np.zeros((v, v))
        oracle.einsum(0.0, "ef", rg, -0.5, "bej", np.ascontiguousarray(up[m]), "jfb", np.ascontiguousarray(t[:, n]))
        assert rel_fro(gg[m, n], rg) < 1e-12
    # central differences of <up, W(t, g)> on 64 entries of each input, h = 1e-6. f is
    # bilinear, so the difference has no truncation error; what remains is the FP64 rounding
    # of the two evaluations of f (a sum of 16.8M products), about eps*|f|/h, which the
    # tolerance adds to the 1e-6 relative bound.
    h = 1e-6
    dup, t0, g0 = dev(up), dev(t), dev(g)
    with torch.no_grad():
        f = lambda a, b: float((ag.einsum("mbej", "jnfb", a, "mnef", b, alpha=-0.5) * dup).sum())
        fd_noise = 16 * np.finfo(np.float64).eps * abs(f(t0, g0)) / h
        for which, base, grad in (("t", t0, gt), ("g", g0, gg)):
            for _ in range(64):
                idx = tuple(int(rng.integers(0, n)) for n in base.shape)
                p, m = base.clone(), base.clone()
                p[idx] += h
                m[idx] -= h
                fd = (f(p, g0) - f(m, g0)) / (2 * h) if which == "t" else (f(t0, p) - f(t0, m)) / (2 * h)
                assert abs(fd - grad[idx]) <= 1e-6 * max(abs(grad[idx]), 1e-3) + fd_noise, (which, idx, fd, fd_noise)


@pytest.mark.parametrize("pat", [("ij", "ik", "kj"), ("i", "ii", "i"), ("", "ij", "ji"), ("ijk", "ij", "k"),
                                 ("ii", "i", "i"), ("ab", "ac", "cbd"), ("", "i", "i"), ("mbej", "jnfb", "mnef")],
                         ids=lambda p: f"{p[0]}<-{p[1]},{p[2]}")
def test_gradcheck_small_patterns(pat):
    from einsums_amd import autograd as ag

    Ci, Ai, Bi = pat
    rng = seeded("gc", *pat)
    ext = {ch: int(rng.integers(2, 5)) for ch in set(Ci + Ai + Bi)}
    A = dev(rng.standard_normal([ext[c] for c in Ai])).requires_grad_(True)
    B = dev(rng.standard_normal([ext[c] for c in Bi])).requires_grad_(True)
    assert torch.autograd.gradcheck(lambda a, b: ag.einsum(Ci, Ai, a, Bi, b, alpha=0.7), (A, B), eps=1e-6, atol=1e-8,
                                    rtol=1e-6)
