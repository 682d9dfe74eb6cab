"""Randomised GPU parity sweep of einsum against the oracle.

Each case draws its labels' roles (batch, M, N, K, summed in one operand only, or in C
only), extents, the label order inside every operand, an occasional repeated label in A or
B (a diagonal), strided operand and output views, prefactors and the element type, all
from a seed fixed by the case number, so a failure names a reproducible case. One case in
four uses MFMA-sized extents (8..40), the rest 1..7 (extents of 1 included, and in about
one small case in twelve an extent of 0). Half of the cases run through
compile_plan/execute; those whose every label is a target or a link are checked against
the oracle's dispatch on views with the same strides (the result, and on dense operands
with no empty extent the algorithm it picks), the others against the oracle's
long-double generic loop (labels found in one operand are summed, labels found in C only
broadcast). The rest run through the generic contraction alone, against that loop. C
outside a strided output view must stay untouched. Tolerance: relative Frobenius 1e-13
(FP64), 1e-5 (FP32).
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import host, rel_fro

pytestmark = pytest.mark.gpu

ROLES = ("batch", "m", "n", "k", "a_only", "b_only", "c_only")
WEIGHTS = (0.15, 0.22, 0.22, 0.25, 0.06, 0.06, 0.04)
NCASES = 600


def make_case(idx):
    rng = np.random.default_rng(zlib.crc32(f"random-einsum|{idx}".encode()))
    big = idx % 4 == 3
    while True:
        nl = int(rng.integers(2, 5 if big else 7))
        labels = list(rng.permutation(list("abcdefghpq"))[:nl])
        roles = rng.choice(ROLES, size=nl, p=WEIGHTS)
        A = [l for l, r in zip(labels, roles) if r in ("batch", "m", "k", "a_only")]
        B = [l for l, r in zip(labels, roles) if r in ("batch", "n", "k", "b_only")]
        C = [l for l, r in zip(labels, roles) if r in ("batch", "m", "n", "c_only")]
        if A and B:
            break
    ext = {l: int(rng.integers(8, 41)) if big else int(rng.integers(1, 8)) for l in labels}
    if not big and rng.random() < 0.08:  # an empty extent
        ext[labels[int(rng.integers(nl))]] = 0
    if not big and rng.random() < 0.15:  # a diagonal: one label twice in one operand
        side = A if rng.random() < 0.5 else B
        side.append(side[int(rng.integers(len(side)))])
    A = "".join(rng.permutation(A))
    B = "".join(rng.permutation(B))
    C = "".join(rng.permutation(C)) if C else ""
    views = [str(rng.choice(("dense", "pad", "step"), p=(0.6, 0.25, 0.15))) for _ in range(2)]
    c_view = str(rng.choice(("dense", "pad"), p=(0.75, 0.25))) if C else "dense"
    dtype = np.float32 if idx % 5 == 4 else np.float64
    cp = (0.0, 1.0, -0.5)[idx % 3]
    ab = (1.0, -1.25, 0.5)[(idx // 3) % 3]
    route = "plan" if (idx // 2) % 2 == 0 else "generic"
    return dict(C=C, A=A, B=B, ext=ext, views=views, c_view=c_view, dtype=dtype, cp=cp, ab=ab, route=route, seed=idx)


CASES = [make_case(i) for i in range(NCASES)]


def _case_id(c):
    return f"{c['seed']}:{c['C']}<-{c['A']},{c['B']}:{c['route']}:{np.dtype(c['dtype']).name}"


def _operand(rng, labels, ext, view, dtype):
    """A device tensor (possibly a strided view of a bigger one) and its values on the host."""
    shape = [ext[l] for l in labels]
    if view == "pad" and shape:
        lo = [int(rng.integers(0, 3)) for _ in shape]
        full = [s + o + int(rng.integers(0, 3)) for s, o in zip(shape, lo)]
        base = rng.uniform(-1, 1, full).astype(dtype)
        idx = tuple(slice(o, o + s) for o, s in zip(lo, shape))
    elif view == "step" and shape:
        d = int(rng.integers(len(shape)))
        full = list(shape)
        full[d] *= 2
        base = rng.uniform(-1, 1, full).astype(dtype)
        idx = tuple(slice(None, None, 2) if i == d else slice(None) for i in range(len(shape)))
    else:
        base = np.asarray(rng.uniform(-1, 1, shape)).astype(dtype)
        idx = tuple(slice(None) for _ in shape)
    t = torch.from_numpy(base.copy()).to("cuda")
    return base, t, idx


@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_random_einsum_vs_oracle(core, oracle, case):
    rng = np.random.default_rng(zlib.crc32(f"random-einsum-data|{case['seed']}".encode()))
    Ci, Ai, Bi, ext, dtype = case["C"], case["A"], case["B"], case["ext"], case["dtype"]
    a_base, a_t, a_idx = _operand(rng, Ai, ext, case["views"][0], dtype)
    b_base, b_t, b_idx = _operand(rng, Bi, ext, case["views"][1], dtype)
    c_base, c_t, c_idx = _operand(rng, Ci, ext, case["c_view"], dtype)
    A, B, C = a_t[a_idx], b_t[b_idx], c_t[c_idx]

    # labels in one operand only, or in C only: the BLAS-shaped routes do not take such a
    # product (the reference's folded views, restated by the oracle's dispatch, would use
    # one slice along them), so it is checked against the generic loop
    plain = all(sum(l in s for s in (Ci, Ai, Bi)) >= 2 for l in set(Ci + Ai + Bi))
    if case["route"] == "plan" and plain:
        # the oracle's dispatch on views with the same strides: like the reference's, its
        # choice (and a fallback to the generic loop) depends on them
        ref_full = c_base.copy()
        ref_view = ref_full[c_idx + (Ellipsis,)]
        expect_alg = oracle.einsum(case["cp"], Ci, ref_view, case["ab"], Ai, a_base[a_idx + (Ellipsis,)], Bi,
                                   b_base[b_idx + (Ellipsis,)])
        ref = np.array(ref_view, dtype=np.float64)
        plan = core.compile_plan(Ci, Ai, Bi)
        plan.execute(case["cp"], C, case["ab"], A, B)
        # the algorithm is compared on dense operands with no empty extent: the device's
        # BLAS-shaped routes need one stride per folded group (other views run as the
        # generic contraction, with the same result), and empty products run as the generic
        # contraction
        if case["views"] == ["dense", "dense"] and case["c_view"] == "dense" and 0 not in ext.values():
            assert core.algorithm_name(plan.last_algorithm) == expect_alg
    else:
        ref = np.array(c_base[c_idx], dtype=np.float64, copy=True, order="C")
        oracle.einsum_generic(case["cp"], Ci, ref, case["ab"], Ai,
                              np.array(a_base[a_idx], dtype=np.float64, copy=True, order="C"), Bi,
                              np.array(b_base[b_idx], dtype=np.float64, copy=True, order="C"), extended=True)
        if case["route"] == "plan":
            core.compile_plan(Ci, Ai, Bi).execute(case["cp"], C, case["ab"], A, B)
        else:
            core.contract(Ci, C, Ai, A, Bi, B, C_prefactor=case["cp"], AB_prefactor=case["ab"])
    out = host(c_t)
    tol = 1e-13 if dtype == np.float64 else 1e-5
    assert rel_fro(out[c_idx], ref) < tol
    if case["c_view"] != "dense":
        outside = np.ones(c_base.shape, dtype=bool)
        outside[c_idx] = False
        assert np.array_equal(out[outside], c_base[outside])


BIG = [c for c in CASES if c["seed"] % 4 == 3]


@pytest.mark.parametrize("route", ["gather", "ttgt"])
@pytest.mark.parametrize("case", BIG, ids=_case_id)
def test_random_einsum_forced_routes(core, oracle, case, route):
    """The MFMA-sized cases through each generic-contraction route forced (the strided-gather
    MFMA kernel, and packing by the permute + dense GEMM), against the oracle's loop."""
    rng = np.random.default_rng(zlib.crc32(f"random-einsum-route|{case['seed']}".encode()))
    Ci, Ai, Bi, ext, dtype = case["C"], case["A"], case["B"], case["ext"], case["dtype"]
    a_base, a_t, a_idx = _operand(rng, Ai, ext, case["views"][0], dtype)
    b_base, b_t, b_idx = _operand(rng, Bi, ext, case["views"][1], dtype)
    c_base, c_t, c_idx = _operand(rng, Ci, ext, case["c_view"], dtype)
    ref = np.array(c_base[c_idx], dtype=np.float64, copy=True, order="C")
    oracle.einsum_generic(case["cp"], Ci, ref, case["ab"], Ai,
                          np.array(a_base[a_idx], dtype=np.float64, copy=True, order="C"), Bi,
                          np.array(b_base[b_idx], dtype=np.float64, copy=True, order="C"), extended=True)
    with core.contraction_route(route):
        core.contract(Ci, c_t[c_idx], Ai, a_t[a_idx], Bi, b_t[b_idx], C_prefactor=case["cp"], AB_prefactor=case["ab"])
    out = host(c_t)
    assert rel_fro(out[c_idx], ref) < (1e-13 if dtype == np.float64 else 1e-5)
    if case["c_view"] != "dense":
        outside = np.ones(c_base.shape, dtype=bool)
        outside[c_idx] = False
        assert np.array_equal(out[outside], c_base[outside])
