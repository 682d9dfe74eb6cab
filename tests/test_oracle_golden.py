"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the parity checker.

Checks, in order: the integer known answers of the reference's own tests (exact), the
algorithm choices the reference tests assert, the water STO-3G RHF/MP2 energies of
QChemTests.cpp, an independent second oracle (numpy.einsum) on random patterns, closed
forms (permute round trips, dot = vec.vec), and planted mistakes that must be caught.
"""
import json
import zlib
from pathlib import Path

import numpy as np
import pytest

from tests.conftest import rel_fro

GOLD = Path(__file__).resolve().parent / "golden"
KA = json.loads((GOLD / "reference_known_answers.json").read_text())
DISPATCH = json.loads((GOLD / "dispatch_table.json").read_text())


def _filled(shape):
    return np.arange(1, int(np.prod(shape)) + 1, dtype=np.float64).reshape(shape)


@pytest.mark.parametrize("case", ["einsum1_ik_ij_jk", "einsum1_il_ijk_jkl"])
def test_einsum1_integer_known_answers(oracle, case):
    c = KA[case]
    A, B = _filled(c["shape_A"]), _filled(c["shape_B"])
    C = np.zeros((3, 3))
    alg = oracle.einsum(0.0, c["C"], C, 1.0, c["A"], A, c["B"], B)
    assert alg == c["algorithm"]
    assert np.array_equal(C, np.array(c["expected_C"], dtype=np.float64))
    # the generic loop nest must give the same exact integers
    C2 = np.full((3, 3), np.nan)
    oracle.einsum_generic(0.0, c["C"], C2, 1.0, c["A"], A, c["B"], B)
    assert np.array_equal(C2, C)


@pytest.mark.parametrize("trans", ["NN", "TN", "NT", "TT"])
def test_gemm_3x3_known_answers(oracle, trans):
    g = KA["gemm_3x3"]
    A = np.array(g["A"], dtype=np.float64)
    B = np.array(g["B"], dtype=np.float64)
    C = np.zeros(9)
    # row-major C = op(A) op(B)  <=>  column-major C^T = op(B)^T op(A)^T on the same buffers
    oracle.gemm_colmajor(trans[1], trans[0], 3, 3, 3, 1.0, B, 3, A, 3, 0.0, C, 3)
    assert C.tolist() == g["expected"][trans]


@pytest.mark.parametrize("case", DISPATCH["cases"], ids=lambda c: f"{c['C']}<-{c['A']},{c['B']}")
def test_dispatch_table(oracle, case):
    assert oracle.classify(case["C"], case["A"], case["B"]) == case["alg"]


def test_dispatch_view_falls_back_to_generic(oracle):
    """Einsum7.cpp:197-209: a sub-view that does not fold runs the generic algorithm."""
    rng = np.random.default_rng(0)
    naux, nocc, nvirt, urank = 6, 3, 4, 5
    Bq = rng.uniform(-1, 1, (naux, nocc + nvirt, nocc + nvirt))
    Qov = Bq[:, 0:nocc, nocc:]
    iaX = rng.uniform(-1, 1, (nocc, nvirt, urank))
    C = np.zeros((naux, urank))
    assert oracle.einsum(0.0, "QX", C, 1.0, "Qia", Qov, "iaX", iaX) == "GENERIC"
    assert rel_fro(C, np.einsum("Qia,iaX->QX", Qov, iaX)) < 1e-14


def test_water_rhf_mp2_energies(oracle):
    from tests.helpers import read_water
    from tests.qchem import NumpyOracleBackend, rhf_mp2

    w = KA["water_sto3g"]
    S, T, V, G = read_water(GOLD / "water_sto3g")
    res = rhf_mp2(NumpyOracleBackend(oracle), S, T, V, G, nocc=w["nocc"], enuc=w["E_nuc"])
    assert res["cycles"] < 50
    assert abs(res["E_HF"] - w["E_HF"]) < w["tolerance_abs"]
    assert abs(res["E_MP2"] - w["E_MP2"]) < w["tolerance_abs"]
    assert abs(res["E_total"] - w["E_total"]) < w["tolerance_abs"]


PATTERNS = [
    # (C, A, B, shapes by label)
    ("ij", "ik", "kj", None), ("ij", "ki", "jk", None), ("mnij", "ijef", "mnef", None), ("abef", "mnef", "mnab", None),
    ("mbej", "jnfb", "mnef", None), ("", "ijab", "ijab", None), ("pq", "pqrs", "rs", None), ("pq", "prqs", "rs", None),
    ("i", "ii", "i", None), ("", "ii", "jj", None), ("ij", "ij", "ji", None), ("ii", "i", "i", None),
    ("ij", "i", "j", None), ("iaW", "iW", "aW", None), ("aX", "QaX", "QX", None), ("i", "ijk", "jik", None),
    ("ijk", "ij", "k", None), ("ijk", "i", "jk", None), ("i", "ij", "k", None), ("ijk", "ik", "kj", None),
    ("", "ijkl", "klij", None), ("bij", "bik", "bkj", None),
]


def _shapes(C, A, B, rng):
    ext = {ch: int(rng.integers(2, 6)) for ch in set(C + A + B)}
    return ext


@pytest.mark.parametrize("pat", PATTERNS, ids=lambda p: f"{p[0]}<-{p[1]},{p[2]}")
def test_generic_vs_numpy_einsum(oracle, pat):
    """Independent second oracle: numpy.einsum (same semantics: labels absent from C are
    summed, repeated labels address diagonals)."""
    Ci, Ai, Bi, _ = pat
    rng = np.random.default_rng(zlib.crc32("|".join(pat[:3]).encode()))
    ext = _shapes(Ci, Ai, Bi, rng)
    A = rng.uniform(-1, 1, [ext[c] for c in Ai])
    B = rng.uniform(-1, 1, [ext[c] for c in Bi])
    C0 = rng.uniform(-1, 1, [ext[c] for c in Ci]) if Ci else np.array(rng.uniform(-1, 1))
    uniq = "".join(dict.fromkeys(Ci))
    for cp, abp in ((0.0, 1.0), (0.5, -2.0)):
        ref = np.einsum(f"{Ai},{Bi}->{uniq}", A, B) * abp
        if uniq != Ci:
            # repeated output label: the contraction lands on the diagonal of a full C
            expect = (C0 * cp).copy()
            for idx in np.ndindex(*ref.shape):
                expect[tuple(idx[uniq.index(ch)] for ch in Ci)] += ref[idx]
        else:
            expect = ref + cp * C0
        C = C0.copy()
        oracle.einsum_generic(cp, Ci, C, abp, Ai, A, Bi, B)
        assert rel_fro(C, expect) < 1e-13
        C = C0.copy()
        oracle.einsum(cp, Ci, C, abp, Ai, A, Bi, B)
        assert rel_fro(C, expect) < 1e-13


def test_permute_roundtrip_and_numpy(oracle):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3, 4, 5, 6))
    C = np.empty((3, 5, 4, 6))
    oracle.permute(0.0, "iajb", C, 1.0, "ijab", A)
    assert np.array_equal(C, A.transpose(0, 2, 1, 3))
    back = np.empty_like(A)
    oracle.permute(0.0, "ijab", back, 1.0, "iajb", C)
    assert np.array_equal(back, A)
    C2 = np.ones((6, 5, 4, 3))
    oracle.permute(2.0, "bajI".replace("I", "i"), C2, 0.5, "ijab", A)
    assert rel_fro(C2, 2.0 + 0.5 * A.transpose(3, 2, 1, 0)) < 1e-15


def test_fastdiv_matches_integer_division(oracle):
    rng = np.random.default_rng(1)
    divisors = list(range(1, 300)) + [2**k for k in range(1, 31)] + [2**k - 1 for k in range(2, 31)] + \
        rng.integers(1, 2**31 - 1, 200).tolist()
    for d in divisors:
        ns = [0, 1, d - 1, d, d + 1, 2**31 - 1, 2**31 - 2] + rng.integers(0, 2**31 - 1, 50).tolist()
        for n in ns:
            if n < 0 or n >= 2**31:
                continue
            assert oracle.fastdiv(int(n), int(d)) == n // d


def test_planted_mistakes_are_caught(oracle):
    """A transposed operand or a dropped prefactor must fail the known-answer checks."""
    c = KA["einsum1_ik_ij_jk"]
    A = _filled(c["shape_A"])
    expected = np.array(c["expected_C"], dtype=np.float64)
    C = np.zeros((3, 3))
    oracle.einsum(0.0, "ij", C, 1.0, "ki", A, "kj", A)  # transposed A
    assert not np.array_equal(C, expected)
    C = np.zeros((3, 3))
    oracle.einsum(0.0, "ij", C, -1.0, "ik", A, "kj", A)  # wrong sign
    assert not np.array_equal(C, expected)
    from tests.helpers import read_water
    from tests.qchem import NumpyOracleBackend, rhf_mp2

    S, T, V, G = read_water(GOLD / "water_sto3g")
    res = rhf_mp2(NumpyOracleBackend(oracle), S, T, V, G.transpose(0, 2, 1, 3).copy())  # wrong integral layout
    assert abs(res["E_HF"] - KA["water_sto3g"]["E_HF"]) > 1e-3
