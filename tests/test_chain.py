"""Order search and index bookkeeping of the multi-operand contraction planner (CPU only).

The dynamic programme is checked against an exhaustive search over every sequence of
pairwise contractions, and each planned chain is executed pairwise with numpy.einsum and
compared with numpy.einsum of the whole expression."""
import itertools
import math

import numpy as np
import pytest

from einsums_amd.chain import parse_subscripts, plan_order, run_steps


def test_parse_subscripts():
    assert parse_subscripts("ij, jk ->ik") == (["ij", "jk"], "ik")
    assert parse_subscripts("ijab,klab,klij->") == (["ijab", "klab", "klij"], "")
    for bad in ("ij,jk", "ij,jk->ii", "ij,jk->iz", "i1,jk->ik"):
        with pytest.raises(ValueError):
            parse_subscripts(bad)


def test_matrix_chain_known_answer():
    # A 10x100, B 100x5, C 5x50: (AB)C costs 2*(10*100*5 + 10*5*50) = 15000 flops,
    # A(BC) costs 2*(100*5*50 + 10*100*50) = 150000.
    sizes = dict(i=10, k=100, l=5, j=50)
    steps = plan_order(["ik", "kl", "lj"], "ij", sizes)
    assert sum(s.flops for s in steps) == 15000
    assert (steps[0].left, steps[0].right) == (0, 1) and steps[0].result == "il"
    assert steps[-1].result == "ij"


def _exhaustive(inputs, output, sizes):
    """Cheapest cost over every sequence of pairwise contractions."""
    best = math.inf

    def rec(work, cost):
        nonlocal best
        if cost >= best:
            return
        if len(work) == 1:
            best = cost
            return
        for x, y in itertools.combinations(range(len(work)), 2):
            rest = [w for k, w in enumerate(work) if k not in (x, y)]
            outside = set(output).union(*[set(w) for w in rest]) if rest else set(output)
            res = "".join(sorted((set(work[x]) | set(work[y])) & outside))
            rec(rest + [res], cost + 2 * math.prod(sizes[c] for c in set(work[x]) | set(work[y])))

    rec(list(inputs), 0)
    return best


def _random_chain(rng, n):
    labels = "abcdefghij"
    nl = int(rng.integers(n, min(len(labels), 2 * n + 1)))
    use = labels[:nl]
    inputs = []
    for _ in range(n):
        r = int(rng.integers(1, 4))
        inputs.append("".join(rng.choice(list(use), size=r, replace=False)))
    present = sorted(set("".join(inputs)))
    out = "".join(c for c in present if rng.random() < 0.3)
    sizes = {c: int(rng.integers(2, 6)) for c in present}
    return inputs, out, sizes


@pytest.mark.parametrize("seed", range(40))
def test_dynamic_programme_is_optimal(seed):
    rng = np.random.default_rng(seed)
    n = 3 + seed % 3
    inputs, out, sizes = _random_chain(rng, n)
    steps = plan_order(inputs, out, sizes)
    assert len(steps) == n - 1
    assert sum(s.flops for s in steps) == _exhaustive(inputs, out, sizes)


@pytest.mark.parametrize("seed", range(40))
@pytest.mark.parametrize("limit", [10, 0])
def test_chain_matches_numpy_einsum(seed, limit):
    rng = np.random.default_rng(1000 + seed)
    n = 3 + seed % 4
    inputs, out, sizes = _random_chain(rng, n)
    ops = [rng.standard_normal([sizes[c] for c in t]) for t in inputs]
    steps = plan_order(inputs, out, sizes, optimal_limit=limit)
    pair = lambda st, x, y: np.einsum(f"{st.left_indices},{st.right_indices}->{st.result}", x, y)
    got = run_steps(steps, ops, pair, pair)
    want = np.einsum(",".join(inputs) + "->" + out, *ops)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10 * max(1.0, float(np.abs(want).max(initial=0))))


def test_repeated_labels_and_single_operand_sums():
    # a diagonal (ii), a label only on one operand (x) and a label on three operands (b)
    inputs, out = ["bii", "bix", "bj", "j"], "b"
    sizes = dict(b=3, i=4, x=5, j=6)
    rng = np.random.default_rng(7)
    ops = [rng.standard_normal([sizes[c] for c in t]) for t in inputs]
    pair = lambda st, x, y: np.einsum(f"{st.left_indices},{st.right_indices}->{st.result}", x, y)
    for limit in (10, 0):
        steps = plan_order(inputs, out, sizes, optimal_limit=limit)
        got = run_steps(steps, ops, pair, pair)
        np.testing.assert_allclose(got, np.einsum("bii,bix,bj,j->b", *ops), rtol=1e-12)
