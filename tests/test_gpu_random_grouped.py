"""Randomised GPU parity sweep of grouped plans (GroupedEinsum) against the oracle.

Each plan is a list of 1..24 contractions drawn from a seed fixed by the plan number. A
term either uses fresh tensors, writes again into the output of an earlier term (with
C_prefactor 1, 0.5 or 0), or reads an earlier term's output as its A; label roles are
batch, M, N and K, plus labels summed in one operand only and labels present only in C,
with extents 1..48. The plan's result must equal running the terms one after another in
list order with the oracle's long-double loop, for every tensor the plan touches
(relative Frobenius 1e-13). A third of the plans are autotuned first, and every plan is
executed twice with its outputs reset between, so a replay of the same plan is checked as
well.
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import host, rel_fro

pytestmark = pytest.mark.gpu

NPLANS = 120
LETTERS = "abcdefgh"


def _term(rng, pool):
    """One term (labels, operand indices into pool, prefactors); new tensors join pool."""
    mode = rng.choice(("fresh", "reuse_c", "reuse_a"), p=(0.5, 0.25, 0.25)) if pool else "fresh"
    ext = {}
    if mode == "reuse_c":
        ci = int(rng.integers(len(pool)))
        shape = pool[ci]["shape"]
        c_lab = list(LETTERS[: len(shape)])
        ext.update(zip(c_lab, shape))
        roles = {l: str(rng.choice(("batch", "m", "n", "c_only"), p=(0.25, 0.35, 0.35, 0.05))) for l in c_lab}
        start = len(shape)
    elif mode == "reuse_a":
        ai = int(rng.integers(len(pool)))
        shape = pool[ai]["shape"]
        a_lab = list(LETTERS[: len(shape)])
        ext.update(zip(a_lab, shape))
        roles = {l: str(rng.choice(("batch", "m", "k", "a_only"), p=(0.2, 0.35, 0.35, 0.1))) for l in a_lab}
        start = len(shape)
    else:
        roles, start = {}, 0
    # new labels must stay off the reused operand: it keeps its own axes
    choices = {"fresh": ("batch", "m", "n", "k", "k", "a_only", "b_only", "c_only"),
               "reuse_c": ("k", "k", "a_only", "b_only"), "reuse_a": ("n", "n", "b_only")}[mode]
    for l in LETTERS[start: start + int(rng.integers(1 if start == 0 else 0, 4 - min(start, 2)))]:
        roles[l] = str(rng.choice(choices))
        ext[l] = int(rng.integers(1, 49))
    if float(np.prod(list(ext.values()))) > 2e6:  # keep the oracle's loop short
        return None
    A = [l for l, r in roles.items() if r in ("batch", "m", "k", "a_only")]
    B = [l for l, r in roles.items() if r in ("batch", "n", "k", "b_only")]
    C = [l for l, r in roles.items() if r in ("batch", "m", "n", "c_only")]
    nxt = LETTERS[len(roles)]
    if not B:  # every operand needs a label: a K label shared by A and B
        roles[nxt], ext[nxt] = "k", int(rng.integers(1, 9))
        A.append(nxt)
        B.append(nxt)
        if mode == "reuse_a":
            return None  # A's axes are fixed: redraw the term
    if not A:
        return None
    if mode != "reuse_a":
        A = list(rng.permutation(A))
    B = list(rng.permutation(B))
    if mode != "reuse_c":
        C = list(rng.permutation(C))
    else:
        C = list(LETTERS[: len(pool[ci]["shape"])])
    if mode == "reuse_a":
        A = list(LETTERS[: len(pool[ai]["shape"])])

    def new(labels):
        pool.append({"shape": [ext[l] for l in labels]})
        return len(pool) - 1

    c_idx = ci if mode == "reuse_c" else new(C)
    a_idx = ai if mode == "reuse_a" else new(A)
    b_idx = new(B)
    cp = float(rng.choice((1.0, 0.5, 0.0))) if mode == "reuse_c" else float(rng.choice((0.0, 0.5)))
    ab = float(rng.choice((1.0, -1.25, 0.5)))
    return ("".join(C), c_idx, "".join(A), a_idx, "".join(B), b_idx, cp, ab)


def make_plan(idx):
    rng = np.random.default_rng(zlib.crc32(f"random-grouped|{idx}".encode()))
    pool, terms = [], []
    n = int(rng.integers(1, 25))
    while len(terms) < n:
        t = _term(rng, pool)
        if t is not None:
            terms.append(t)
    return dict(pool=[p["shape"] for p in pool], terms=terms, autotune=idx % 3 == 0, seed=idx)


PLANS = [make_plan(i) for i in range(NPLANS)]


@pytest.mark.parametrize("plan", PLANS, ids=lambda p: f"{p['seed']}:{len(p['terms'])}terms")
def test_random_grouped_plan(core, oracle, plan):
    from einsums_amd.blocks import GroupedEinsum

    rng = np.random.default_rng(zlib.crc32(f"random-grouped-data|{plan['seed']}".encode()))
    init = [rng.uniform(-1, 1, shape) for shape in plan["pool"]]
    dev = [torch.from_numpy(x.copy()).to("cuda") for x in init]
    expect = [x.copy() for x in init]
    for ci, c, ai, a, bi, b, cp, ab in plan["terms"]:
        oracle.einsum_generic(cp, ci, expect[c], ab, ai, expect[a].copy(), bi, expect[b].copy(), extended=True)
    g = GroupedEinsum([(ci, dev[c], ai, dev[a], bi, dev[b], cp, ab) for ci, c, ai, a, bi, b, cp, ab in plan["terms"]],
                      autotune=plan["autotune"])
    for rep in range(2):
        if rep:
            for d, x in zip(dev, init):
                d.copy_(torch.from_numpy(x))
        g.execute()
        torch.cuda.synchronize()
        for k, (d, e) in enumerate(zip(dev, expect)):
            assert rel_fro(host(d), e) < 1e-13, f"tensor {k}, execution {rep}"


@pytest.mark.parametrize("plan", PLANS[:40], ids=lambda p: f"{p['seed']}:{len(p['terms'])}terms:f32")
def test_random_grouped_plan_fp32(core, oracle, plan):
    """The first 40 plans in FP32 (one execution), against the oracle's long-double loop on
    the same FP32 values: relative Frobenius 1e-4. A plan's terms feed one another (up to 24
    in a row, each a sum of up to ~10^3 products rounded to FP32), so the rounding compounds
    past the 1e-5 of a single FP32 contraction (measured up to 1.8e-5)."""
    from einsums_amd.blocks import GroupedEinsum

    rng = np.random.default_rng(zlib.crc32(f"random-grouped-data-f32|{plan['seed']}".encode()))
    init = [rng.uniform(-1, 1, shape).astype(np.float32) for shape in plan["pool"]]
    dev = [torch.from_numpy(x.copy()).to("cuda") for x in init]
    expect = [x.astype(np.float64) for x in init]
    for ci, c, ai, a, bi, b, cp, ab in plan["terms"]:
        oracle.einsum_generic(cp, ci, expect[c], ab, ai, expect[a].copy(), bi, expect[b].copy(), extended=True)
    g = GroupedEinsum([(ci, dev[c], ai, dev[a], bi, dev[b], cp, ab) for ci, c, ai, a, bi, b, cp, ab in plan["terms"]])
    g.execute()
    torch.cuda.synchronize()
    for k, (d, e) in enumerate(zip(dev, expect)):
        assert rel_fro(host(d), e) < 1e-4, f"tensor {k}"
