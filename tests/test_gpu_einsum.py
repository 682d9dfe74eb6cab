"""GPU parity of compile_plan/execute (and the generic contraction kernels) against the
CPU oracle on the same seeded inputs. FP64 tolerance: relative Frobenius 1e-13;
FP32: 1e-5; integer-valued known answers exact."""
import json
import zlib
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
KA = json.loads((GOLD / "reference_known_answers.json").read_text())
DISPATCH = json.loads((GOLD / "dispatch_table.json").read_text())


def seeded(*parts):
    return np.random.default_rng(zlib.crc32("|".join(map(str, parts)).encode()))


def run_both(core, oracle, cp, Ci, Ai, Bi, ext, dtype=np.float64, seed=0, generic_only=False):
    rng = seeded(Ci, Ai, Bi, seed, dtype.__name__)
    A = rng.uniform(-1, 1, [ext[c] for c in Ai]).astype(dtype)
    B = rng.uniform(-1, 1, [ext[c] for c in Bi]).astype(dtype)
    C0 = (rng.uniform(-1, 1, [ext[c] for c in Ci]) if Ci else np.array(rng.uniform(-1, 1))).astype(dtype)
    ref = C0.astype(np.float64).copy()
    oracle.einsum_generic(cp, Ci, ref, -1.25, Ai, A.astype(np.float64), Bi, B.astype(np.float64))
    dC = dev(C0)
    if generic_only:
        core.contract(Ci, dC, Ai, dev(A), Bi, dev(B), C_prefactor=cp, AB_prefactor=-1.25)
        alg = None
    else:
        plan = core.compile_plan(Ci, Ai, Bi)
        plan.execute(cp, dC, -1.25, dev(A), dev(B))
        alg = plan.last_algorithm
    return host(dC), ref, alg


@pytest.mark.parametrize("case", ["einsum1_ik_ij_jk", "einsum1_il_ijk_jkl"])
def test_einsum1_integer_known_answers(core, case):
    c = KA[case]
    fill = lambda s: np.arange(1, int(np.prod(s)) + 1, dtype=np.float64).reshape(s)
    dC = dev(np.full((3, 3), np.nan))
    plan = core.compile_plan(c["C"], c["A"], c["B"])
    plan.execute(0.0, dC, 1.0, dev(fill(c["shape_A"])), dev(fill(c["shape_B"])))
    assert core.algorithm_name(plan.last_algorithm) == c["algorithm"]
    assert np.array_equal(host(dC), np.array(c["expected_C"], dtype=np.float64))


@pytest.mark.parametrize("case", DISPATCH["cases"], ids=lambda c: f"{c['C']}<-{c['A']},{c['B']}")
@pytest.mark.parametrize("cp", [0.0, 0.5])
def test_dispatch_cases_vs_oracle(core, oracle, case, cp):
    rng = seeded(case["C"], case["A"], case["B"])
    ext = {ch: int(rng.integers(3, 40)) for ch in set(case["C"] + case["A"] + case["B"])}
    got, ref, alg = run_both(core, oracle, cp, case["C"], case["A"], case["B"], ext)
    assert core.algorithm_name(alg) == case["alg"]
    assert rel_fro(got, ref) < 1e-13


GENERIC = [
    ("i", "ii", "i"), ("", "ii", "jj"), ("ij", "ij", "ji"), ("ii", "i", "i"), ("iaW", "iW", "aW"), ("aX", "QaX", "QX"),
    ("i", "ijk", "jik"), ("ijk", "ij", "k"), ("ijk", "i", "jk"), ("i", "ij", "k"), ("ijk", "ik", "kj"),
    ("", "ijkl", "klij"), ("bij", "bik", "bkj"), ("mbej", "jnfb", "mnef"), ("pq", "prqs", "rs"), ("ij", "ji", "ij"),
    ("ijkl", "ijpl", "pk"), ("pqks", "pqrs", "rk"), ("QaX", "QYX", "aY"), ("ab", "ab", "b"), ("abc", "ac", "cb"),
    ("iab", "iaj", "jb"), ("", "ab", "ba"), ("kl", "ij", "kl"),
]


@pytest.mark.parametrize("pat", GENERIC, ids=lambda p: f"{p[0]}<-{p[1]},{p[2]}")
@pytest.mark.parametrize("size", ["small", "mfma"])
@pytest.mark.parametrize("cp", [0.0, -0.5])
def test_generic_patterns_vs_oracle(core, oracle, pat, size, cp):
    Ci, Ai, Bi = pat
    rng = seeded(*pat, size)
    labels = sorted(set(Ci + Ai + Bi))
    hi = 9 if size == "small" else (48 if len(labels) <= 4 else 20)
    ext = {ch: int(rng.integers(2, hi)) for ch in labels}
    got, ref, _ = run_both(core, oracle, cp, Ci, Ai, Bi, ext, generic_only=True)
    assert rel_fro(got, ref) < 1e-13


@pytest.mark.parametrize("pat", [("ij", "ik", "kj"), ("mbej", "jnfb", "mnef"), ("", "ijab", "ijab"), ("pq", "prqs", "rs")])
def test_fp32_vs_oracle(core, oracle, pat):
    rng = seeded(*pat, "f32")
    ext = {ch: int(rng.integers(8, 40)) for ch in set("".join(pat))}
    got, ref, _ = run_both(core, oracle, 0.25, *pat, ext, dtype=np.float32)
    assert rel_fro(got, ref) < 1e-5


def test_strided_views_fall_back_and_match(core, oracle):
    rng = seeded("views")
    naux, nocc, nvirt, ur = 30, 9, 20, 25
    Bq = rng.uniform(-1, 1, (naux, nocc + nvirt, nocc + nvirt))
    iaX = rng.uniform(-1, 1, (nocc, nvirt, ur))
    dBq = dev(Bq)
    Qov = dBq[:, 0:nocc, nocc:]
    dC = torch.zeros((naux, ur), dtype=torch.float64, device="cuda")
    plan = core.compile_plan("QX", "Qia", "iaX")
    assert type(plan) is core.EinsumGemmPlan
    plan.execute(0.0, dC, 1.0, Qov, dev(iaX))
    assert core.algorithm_name(plan.last_algorithm) == "GENERIC"
    ref = np.zeros((naux, ur))
    oracle.einsum(0.0, "QX", ref, 1.0, "Qia", Bq[:, 0:nocc, nocc:], "iaX", iaX)
    assert rel_fro(host(dC), ref) < 1e-13
    # a strided output view: every other row of a bigger C
    big = torch.full((2 * naux, ur), 7.0, dtype=torch.float64, device="cuda")
    plan.execute(0.0, big[::2], 1.0, Qov, dev(iaX))
    out = host(big)
    assert rel_fro(out[::2], ref) < 1e-13
    assert np.all(out[1::2] == 7.0)


def test_prefactor_zero_clears_nan_in_c(core):
    for pat in (("ij", "ik", "kj"), ("mbej", "jnfb", "mnef"), ("i", "ij", "j"), ("ij", "ij", "ij")):
        ext = {ch: 12 for ch in set("".join(pat))}
        dC = torch.full([ext[c] for c in pat[0]], float("nan"), dtype=torch.float64, device="cuda")
        A = torch.ones([ext[c] for c in pat[1]], dtype=torch.float64, device="cuda")
        B = torch.ones([ext[c] for c in pat[2]], dtype=torch.float64, device="cuda")
        core.einsum(pat[0], dC, pat[1], A, pat[2], B)
        assert torch.isfinite(dC).all(), pat


def test_dimension_mismatch_raises(core):
    A = torch.zeros((3, 4), dtype=torch.float64, device="cuda")
    B = torch.zeros((5, 6), dtype=torch.float64, device="cuda")
    C = torch.zeros((3, 6), dtype=torch.float64, device="cuda")
    with pytest.raises(core.DimensionError):
        core.einsum("ij", C, "ik", A, "kj", B)


def test_lists_and_numpy_operands(core, oracle):
    rng = seeded("lists")
    As = [rng.uniform(-1, 1, (13, 10)) for _ in range(4)]
    Bs = [rng.uniform(-1, 1, (10, 13)) for _ in range(4)]
    Cs = [np.zeros((10, 13)) for _ in range(4)]
    plan = core.compile_plan("ij", "ji", "ij")
    assert type(plan) is core.EinsumGenericPlan
    plan.execute(0.0, Cs, 1.0, As, Bs)  # host numpy operands are staged through device memory
    for a, b, c in zip(As, Bs, Cs):
        assert rel_fro(c, a.T * b) < 1e-15
    with pytest.raises(ValueError):
        plan.execute(0.0, Cs[:2], 1.0, As, Bs)


def test_gpuview_copy_and_map(core):
    rng = seeded("gpuview")
    A, B = rng.uniform(-1, 1, (20, 30)), rng.uniform(-1, 1, (30, 20))
    C = np.zeros((20, 20))
    vA, vB, vC = core.GPUView(A, core.COPY), core.GPUView(B, core.COPY), core.GPUView(C, core.COPY)
    core.compile_plan("ij", "ik", "kj").execute(0.0, vC, 1.0, vA, vB)
    assert np.all(C == 0)
    vC.update_D2H()
    assert rel_fro(C, A @ B) < 1e-14
    C2 = np.zeros((20, 20))
    core.compile_plan("ij", "ik", "kj").execute(0.0, core.GPUView(C2, core.MAP), 1.0, vA, vB)
    assert rel_fro(C2, A @ B) < 1e-14


# Patterns whose innermost row digit (ordered by A) is far from unit stride in C, so the
# generic kernel writes a dense buffer and adds it into C with the tiled permute.
STAGED = [
    ("zija", "zaik", "zkj", dict(z=3, i=40, j=36, a=24, k=48)),     # batch group, C unit stride on a
    ("mbej", "jnfb", "mnef", dict(m=20, b=40, e=40, j=20, n=20, f=40)),  # the Wmbej layout
    ("jiij", "ika", "akj", dict(i=32, j=48, k=24, a=20)),            # repeated C labels (diagonal)
]


@pytest.mark.parametrize("pat", STAGED, ids=lambda p: f"{p[0]}<-{p[1]},{p[2]}")
@pytest.mark.parametrize("cp", [0.0, 0.75])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_staged_epilogue_vs_oracle(core, oracle, pat, cp, dtype):
    Ci, Ai, Bi, ext = pat
    got, ref, _ = run_both(core, oracle, cp, Ci, Ai, Bi, ext, dtype=dtype, generic_only=True)
    assert rel_fro(got, ref) < (1e-13 if dtype == np.float64 else 1e-5)
