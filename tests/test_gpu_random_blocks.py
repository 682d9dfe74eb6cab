"""Randomised GPU parity sweep of block einsums (blocks.einsum over BlockTensors).

Each case draws the block extents (1..6 blocks of 0..20), the labels of A, B and C (ranks
1..3 over up to three labels; C may be a scalar), which of A and B is a BlockTensor and
which a dense tensor cut into the diagonal sub-block of every block's range, whether C is
a BlockTensor or a dense tensor (only its diagonal sub-blocks change), and the prefactors,
from a seed fixed by the case number. Expected: the oracle's long-double loop applied
block by block (a scalar C sums the blocks' contributions). Relative Frobenius 1e-13; a
dense C must keep every element outside its diagonal sub-blocks bit for bit.
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import rel_fro

pytestmark = pytest.mark.gpu

NCASES = 80


def make_case(idx):
    rng = np.random.default_rng(zlib.crc32(f"random-blocks|{idx}".encode()))
    labels = list("ijk"[: int(rng.integers(1, 4))])
    A = "".join(rng.permutation(labels)[: int(rng.integers(1, len(labels) + 1))])
    B = "".join(rng.permutation(labels)[: int(rng.integers(1, len(labels) + 1))])
    used = sorted(set(A) | set(B))
    C = "".join(l for l in rng.permutation(used) if rng.random() < 0.6)
    dims = [int(x) for x in rng.integers(0, 21, int(rng.integers(1, 7)))]
    kinds = str(rng.choice(("bb", "db", "bd")))
    c_kind = "scalar" if not C else str(rng.choice(("block", "dense"), p=(0.6, 0.4)))
    cp = (0.0, 1.0, 0.5)[idx % 3]
    ab = (1.0, -0.75)[(idx // 3) % 2]
    return dict(A=A, B=B, C=C, dims=dims, kinds=kinds, c_kind=c_kind, cp=cp, ab=ab, seed=idx)


CASES = [make_case(i) for i in range(NCASES)]


def _operand(blocks, rng, kind, labels, dims):
    """(operand, per-block host arrays) for a BlockTensor or a dense tensor."""
    if kind == "b":
        T = blocks.BlockTensor("X", dims, len(labels))
        hs = []
        for i, n in enumerate(dims):
            h = rng.uniform(-1, 1, (n,) * len(labels))
            T.blocks[i].copy_(torch.from_numpy(h))
            hs.append(h)
        return T, hs
    N = sum(dims)
    h = rng.uniform(-1, 1, (N,) * len(labels))
    offs = np.cumsum([0] + dims)
    hs = [h[tuple(slice(offs[i], offs[i] + n) for _ in labels)] for i, n in enumerate(dims)]
    return torch.from_numpy(h).to("cuda"), hs


@pytest.mark.parametrize("case", CASES,
                         ids=lambda c: f"{c['seed']}:{c['C'] or 'scalar'}<-{c['A']},{c['B']}:{c['kinds']}:{c['c_kind']}")
def test_random_block_einsum(oracle, case):
    from einsums_amd import blocks

    rng = np.random.default_rng(zlib.crc32(f"random-blocks-data|{case['seed']}".encode()))
    Ci, Ai, Bi, dims = case["C"], case["A"], case["B"], case["dims"]
    A, hA = _operand(blocks, rng, case["kinds"][0], Ai, dims)
    B, hB = _operand(blocks, rng, case["kinds"][1], Bi, dims)
    cp, ab = case["cp"], case["ab"]
    if case["c_kind"] == "scalar":
        h0 = float(rng.uniform(-1, 1))
        C = torch.tensor(h0, dtype=torch.float64, device="cuda")
        expect = np.array(cp * h0)
        for i, n in enumerate(dims):
            if n:
                part = np.zeros(())
                oracle.einsum_generic(0.0, "", part, ab, Ai, np.ascontiguousarray(hA[i]), Bi,
                                      np.ascontiguousarray(hB[i]), extended=True)
                expect = expect + part
        blocks.einsum(cp, "", C, ab, Ai, A, Bi, B)
        got = C.cpu().numpy()
        assert rel_fro(got, expect) < 1e-13
        return
    C, hC = _operand(blocks, rng, "b" if case["c_kind"] == "block" else "d", Ci, dims)
    before = [np.array(h, copy=True) for h in hC]
    dense_before = C.cpu().numpy().copy() if case["c_kind"] == "dense" else None
    blocks.einsum(cp, Ci, C, ab, Ai, A, Bi, B)
    torch.cuda.synchronize()
    offs = np.cumsum([0] + dims)
    got_dense = C.cpu().numpy() if case["c_kind"] == "dense" else None
    for i, n in enumerate(dims):
        if n == 0:
            continue
        ref = before[i].copy()
        oracle.einsum_generic(cp, Ci, ref, ab, Ai, np.ascontiguousarray(hA[i]), Bi, np.ascontiguousarray(hB[i]),
                              extended=True)
        got = C.blocks[i].cpu().numpy() if case["c_kind"] == "block" else \
            got_dense[tuple(slice(offs[i], offs[i] + n) for _ in Ci)]
        assert rel_fro(got, ref) < 1e-13, f"block {i}"
    if case["c_kind"] == "dense":
        inside = np.zeros(got_dense.shape, dtype=bool)
        for i, n in enumerate(dims):
            inside[tuple(slice(offs[i], offs[i] + n) for _ in Ci)] = True
        assert np.array_equal(got_dense[~inside], dense_before[~inside])
