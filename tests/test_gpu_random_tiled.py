"""Randomised GPU parity sweep of tiled einsums (blocks.einsum over TiledTensors).

Each case draws label roles (batch, M, N, K; at most three labels per tensor), a tile
partition per label (1..4 tiles of extent 0..20, shared by every tensor using the label),
which tiles of A, B and C are present, a tiled or scalar C and the prefactors, from a seed
fixed by the case number. Absent tiles are zero, so the dense view of the result must equal
C_prefactor * C + AB_prefactor * A B on the dense views, computed by the oracle's
long-double loop (relative Frobenius 1e-13). A C tile that a contribution lands on is
created; every call is made twice (the second reuses the cached plan) with C reset between.
"""
import itertools
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import rel_fro

pytestmark = pytest.mark.gpu

NCASES = 80


def make_case(idx):
    rng = np.random.default_rng(zlib.crc32(f"random-tiled|{idx}".encode()))
    while True:
        nl = int(rng.integers(2, 5))
        labels = list("ijklmn"[:nl])
        roles = {l: str(rng.choice(("batch", "m", "n", "k"), p=(0.15, 0.3, 0.3, 0.25))) for l in labels}
        A = [l for l in labels if roles[l] in ("batch", "m", "k")]
        B = [l for l in labels if roles[l] in ("batch", "n", "k")]
        C = [l for l in labels if roles[l] in ("batch", "m", "n")]
        if A and B and max(len(A), len(B), len(C)) <= 3:
            break
    parts = {l: [int(x) for x in rng.integers(0, 21, int(rng.integers(1, 5)))] for l in labels}
    for l in labels:
        if sum(parts[l]) == 0:
            parts[l][0] = int(rng.integers(1, 9))
    A = "".join(rng.permutation(A))
    B = "".join(rng.permutation(B))
    C = "".join(rng.permutation(C)) if C else ""
    fill = (float(rng.uniform(0.5, 1.0)), float(rng.uniform(0.5, 1.0)), float(rng.uniform(0.2, 0.8)))
    cp = (0.0, 1.0, 0.5)[idx % 3]
    ab = (1.0, -0.75)[(idx // 3) % 2]
    return dict(C=C, A=A, B=B, parts=parts, fill=fill, cp=cp, ab=ab, seed=idx)


CASES = [make_case(i) for i in range(NCASES)]


def _tiled(blocks, rng, name, labels, parts, p_present):
    T = blocks.TiledTensor(name, [parts[l] for l in labels])
    for coord in itertools.product(*[range(len(parts[l])) for l in labels]):
        if rng.random() < p_present:
            shape = T.tile_shape(coord)
            T.add_tile(coord, torch.from_numpy(rng.uniform(-1, 1, shape)).to("cuda"))
    return T


@pytest.mark.parametrize("case", CASES,
                         ids=lambda c: f"{c['seed']}:{c['C'] or 'scalar'}<-{c['A']},{c['B']}")
def test_random_tiled_einsum(oracle, case):
    from einsums_amd import blocks

    rng = np.random.default_rng(zlib.crc32(f"random-tiled-data|{case['seed']}".encode()))
    Ci, Ai, Bi, parts = case["C"], case["A"], case["B"], case["parts"]
    A = _tiled(blocks, rng, "A", Ai, parts, case["fill"][0])
    B = _tiled(blocks, rng, "B", Bi, parts, case["fill"][1])
    hA, hB = A.to_dense().cpu().numpy(), B.to_dense().cpu().numpy()
    if Ci:
        C = _tiled(blocks, rng, "C", Ci, parts, case["fill"][2])
        init = {c: t.clone() for c, t in C.tiles.items()}
        hC = C.to_dense().cpu().numpy()
    else:
        hC = np.array(rng.uniform(-1, 1))
        C = torch.tensor(float(hC), dtype=torch.float64, device="cuda")
    expect = hC.copy()
    oracle.einsum_generic(case["cp"], Ci, expect, case["ab"], Ai, hA, Bi, hB, extended=True)
    for rep in range(2):
        if rep:  # reset C and run again: the second call reuses the cached plan
            if Ci:
                for c, t in C.tiles.items():
                    t.copy_(init[c]) if c in init else t.zero_()
            else:
                C.fill_(float(hC))
        blocks.einsum(case["cp"], Ci, C, case["ab"], Ai, A, Bi, B)
        torch.cuda.synchronize()
        got = (C.to_dense() if Ci else C).cpu().numpy()
        assert rel_fro(got, expect) < 1e-13, f"call {rep}"
