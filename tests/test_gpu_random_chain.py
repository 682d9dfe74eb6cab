"""Randomised GPU parity sweep of multi-operand contractions (einsums_amd.chain).

Each case draws 3..6 operands over 3..8 labels (1..3 labels per operand, extents 1..10,
every label in at least one operand), an output made of a random subset of the labels,
and the prefactors, from a seed fixed by the case number. The planned chain runs every
pairwise step through the einsum dispatch on the GPU; the result must equal numpy.einsum
of the whole expression in FP64 (relative Frobenius 1e-12).
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import host, rel_fro

pytestmark = pytest.mark.gpu

NCASES = 60


def make_case(idx):
    rng = np.random.default_rng(zlib.crc32(f"random-chain|{idx}".encode()))
    nl = int(rng.integers(3, 9))
    labels = list("abcdefghij"[:nl])
    ext = {l: int(rng.integers(1, 11)) for l in labels}
    nops = int(rng.integers(3, 7))
    ops = [list(rng.choice(labels, size=int(rng.integers(1, 4)), replace=False)) for _ in range(nops)]
    for l in labels:  # every label in some operand
        if not any(l in o for o in ops):
            ops[int(rng.integers(nops))].append(l)
    out = [l for l in labels if rng.random() < 0.3]
    out = list(rng.permutation(out)) if out else []
    subs = ",".join("".join(o) for o in ops) + "->" + "".join(out)
    cp = (0.0, 0.5)[idx % 2]
    ab = (1.0, -1.5)[(idx // 2) % 2]
    return dict(subs=subs, ext=ext, cp=cp, ab=ab, seed=idx)


CASES = [make_case(i) for i in range(NCASES)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['seed']}:{c['subs']}")
def test_random_chain(core, case):
    from einsums_amd.chain import compile_chain

    rng = np.random.default_rng(zlib.crc32(f"random-chain-data|{case['seed']}".encode()))
    ins, out = case["subs"].split("->")
    ins = ins.split(",")
    ext = case["ext"]
    ops = [rng.uniform(-1, 1, [ext[c] for c in t]) for t in ins]
    C0 = np.asarray(rng.uniform(-1, 1, [ext[c] for c in out]))
    plan = compile_chain(case["subs"], *[o.shape for o in ops])
    dC = torch.from_numpy(C0.copy()).to("cuda")
    plan.execute(case["cp"], dC, case["ab"], *[torch.from_numpy(o).to("cuda") for o in ops])
    got = host(dC)
    whole = case["cp"] * C0 + case["ab"] * np.einsum(case["subs"], *ops, optimize=True)
    assert rel_fro(got, whole) < 1e-12
