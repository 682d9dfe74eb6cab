"""CPU checks of the C ABI: the library loads, exports every declared entry point, and
its host-side analysis (pattern classification, argument and dimension checks) agrees
with the oracle -- none of these calls touch a GPU."""
import ctypes
import json
import re
import zlib
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "einsums_amd.h"
DISPATCH = json.loads((ROOT / "tests" / "golden" / "dispatch_table.json").read_text())


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(ein_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from einsums_amd import _lib

    syms = declared_symbols()
    assert len(syms) > 30
    missing = [s for s in syms if not hasattr(_lib.lib, s)]
    assert missing == []
    assert set(_lib.SYMBOLS) == set(syms)


def test_version_string():
    from einsums_amd import _lib

    assert "gfx950" in _lib.version()


@pytest.mark.parametrize("case", DISPATCH["cases"], ids=lambda c: f"{c['C']}<-{c['A']},{c['B']}")
def test_compile_plan_types(core, case):
    plan = core.compile_plan(case["C"], case["A"], case["B"])
    expected = {"GENERIC": core.EinsumGenericPlan, "DOT": core.EinsumDotPlan, "DIRECT": core.EinsumDirectProductPlan,
                "GER": core.EinsumGerPlan, "GEMV": core.EinsumGemvPlan, "GEMM": core.EinsumGemmPlan}[case["alg"]]
    assert type(plan) is expected


def _random_pattern(rng):
    letters = "abcdefgh"
    n = int(rng.integers(1, 6))
    labs = letters[:n]

    def pick(k):
        repl = bool(rng.integers(0, 4) == 0) or k > n
        return "".join(rng.choice(list(labs), size=k, replace=repl))

    return pick(int(rng.integers(0, 4))), pick(int(rng.integers(1, 5))), pick(int(rng.integers(1, 5)))


def test_classification_agrees_with_oracle(core, oracle):
    rng = np.random.default_rng(zlib.crc32(b"classify"))
    seen = set()
    for _ in range(3000):
        C, A, B = _random_pattern(rng)
        if (C, A, B) in seen:
            continue
        seen.add((C, A, B))
        got = core.algorithm_name(core.classify(C, A, B))
        assert got == oracle.classify(C, A, B), (C, A, B)


def test_dimension_mismatch_is_reported_without_gpu():
    from einsums_amd import _lib

    C = _lib.make_desc("ij", [3, 4], [4, 1])
    A = _lib.make_desc("ik", [3, 5], [5, 1])
    B = _lib.make_desc("kj", [6, 4], [4, 1])  # k: 5 vs 6
    st = _lib.lib.ein_einsum_f64(0.0, ctypes.byref(C), None, 1.0, ctypes.byref(A), None, ctypes.byref(B), None, None,
                                 None)
    assert st == _lib.EIN_ERROR_DIMENSION
    with pytest.raises(_lib.DimensionError):
        _lib.check(st)


def test_gemm_argument_errors_without_gpu():
    from einsums_amd import _lib

    f = _lib.lib.ein_gemm_f64
    assert f(b"X", b"N", 2, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2, None) == -1
    assert f(b"N", b"X", 2, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2, None) == -2
    assert f(b"N", b"N", -1, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2, None) == -3
    assert f(b"N", b"N", 4, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 4, None) == -8
    assert f(b"N", b"N", 2, 2, 4, 1.0, None, 2, None, 2, 0.0, None, 2, None) == -10
    assert f(b"N", b"N", 4, 2, 2, 1.0, None, 4, None, 2, 0.0, None, 2, None) == -13
    # quick return: nothing to do for an empty problem
    assert f(b"N", b"N", 0, 0, 0, 1.0, None, 1, None, 1, 0.0, None, 1, None) == 0


def test_permute_rejects_non_permutation_without_gpu():
    from einsums_amd import _lib

    I64, I32 = ctypes.c_int64 * 2, ctypes.c_int32 * 2
    st = _lib.lib.ein_permute_f64(2, I64(2, 2), I32(0, 0), 1.0, None, I64(2, 1), 0.0, None, I64(2, 1), None)
    assert st == -3


def test_rank_mismatch_in_python_layer(core):
    import torch

    with pytest.raises(core.RankError):
        core._desc(torch.zeros(2, 2), "ijk")


def test_contraction_route_setter_without_gpu(core):
    from einsums_amd import _lib

    lib = _lib.lib
    assert lib.ein_set_contraction_route(_lib.ROUTE_TTGT) == _lib.ROUTE_AUTO
    assert lib.ein_set_contraction_route(7) == -1            # unknown: rejected, route kept
    assert lib.ein_set_contraction_route(_lib.ROUTE_AUTO) == _lib.ROUTE_TTGT
    with core.contraction_route("gather"):
        assert lib.ein_set_contraction_route(_lib.ROUTE_GATHER) == _lib.ROUTE_GATHER
    assert lib.ein_set_contraction_route(_lib.ROUTE_AUTO) == _lib.ROUTE_AUTO
    with pytest.raises(ValueError):
        with core.contraction_route("fast"):
            pass


def test_noalloc_dot_limit_without_gpu():
    """ein_dot_f64_noalloc rejects a unit-stride length whose partial sums would not fit its
    static scratch on the host, before any device call (so this runs without a GPU)."""
    from einsums_amd import _lib

    f = _lib.lib.ein_dot_f64_noalloc
    max_n = 65536 * 8192  # EIN_DOT_NOALLOC_MAX_N
    fake = 4096  # never dereferenced: the call returns before any launch
    assert f(max_n + 1, 1.0, fake, 1, fake, 1, 0.0, fake, None) == 1005  # EIN_ERROR_UNSUPPORTED
    assert b"static scratch" in _lib.lib.ein_last_error_string()
    assert f(-1, 1.0, fake, 1, fake, 1, 0.0, fake, None) == -1
    assert f(10, 1.0, fake, 1, fake, 1, 0.0, None, None) == -8


def test_fock_argument_errors_without_gpu():
    """ein_fock_jk_f64 checks its arguments on the host before any device call."""
    from einsums_amd import _lib

    f = _lib.lib.ein_fock_jk_f64
    fake = 4096  # never dereferenced
    assert f(-1, fake, fake, 1, 2.0, -1.0, 0.0, fake, 1, None) == -1
    assert f(8, fake, fake, 7, 2.0, -1.0, 0.0, fake, 8, None) == -4
    assert f(8, fake, fake, 8, 2.0, -1.0, 0.0, fake, 7, None) == -9
    assert f(129, fake, fake, 129, 2.0, -1.0, 0.0, fake, 129, None) == 1005
    assert b"exceeds 128" in _lib.lib.ein_last_error_string()
    assert f(0, None, None, 1, 2.0, -1.0, 0.0, None, 1, None) == 0  # empty: nothing to do


def test_split_crt_bits_host_only():
    """ein_split_crt_bits: the largest kb with 4 k 2^(2 kb) <= product of the first L moduli
    (256, 255, 253, ...), at most 53; -1 outside 8..16 moduli or for k < 0. Checked against
    exact integer arithmetic here; no device call."""
    from math import gcd, prod

    from einsums_amd import _lib

    mods = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191]
    assert all(gcd(a, b) == 1 for i, a in enumerate(mods) for b in mods[i + 1:])  # pairwise coprime
    f = _lib.lib.ein_split_crt_bits
    for L in range(8, 17):
        M = prod(mods[:L])
        for k in (1, 7, 64, 1000, 2304, 12288, 65536, 70000, 1 << 20):
            kb = max(b for b in range(0, 54) if 4 * k * 2 ** (2 * b) <= M)
            assert f(L, k) == kb, (L, k)
    assert f(13, 2304) == 44 and f(13, 65536) == 42
    assert f(7, 100) == -1 and f(17, 100) == -1 and f(13, -1) == -1


def test_slice_type_codes():
    """Split-precision slice names map to the ABI codes; "+refine" sets EIN_SPLIT_REFINE."""
    from einsums_amd import _lib, core

    assert core._slice_code("bf16") == _lib.SLICE_BF16 and core._slice_code("crt") == _lib.SLICE_CRT
    assert core._slice_code("int8+refine") == _lib.SLICE_INT8 | _lib.SPLIT_REFINE == 0x102
    for bad in ("bf8", "crt+", "crt+refine+x", "int8+fast"):
        with pytest.raises(ValueError):
            core._slice_code(bad)


def test_split_crt_auto_moduli_host_only():
    """nslices = 0 with EIN_SLICE_CRT: the fewest moduli (8..16) keeping >= 40 bits; checked
    against ein_split_crt_bits (itself checked against exact integers above)."""
    from einsums_amd import _lib

    for k in (1, 64, 1000, 2304, 12288, 65536, 70000, 1 << 22):
        L = _lib.lib.ein_split_crt_moduli(k)
        bits = [_lib.lib.ein_split_crt_bits(n, k) for n in range(8, 17)]
        want = next((n for n, b in zip(range(8, 17), bits) if b >= 40), 16)
        assert L == want, (k, L, want)
        assert _lib.lib.ein_split_crt_bits(0, k) == _lib.lib.ein_split_crt_bits(L, k)
    assert _lib.lib.ein_split_crt_moduli(2304) == 12 and _lib.lib.ein_split_crt_moduli(65536) == 13
    assert _lib.lib.ein_split_crt_moduli(-1) == -1
