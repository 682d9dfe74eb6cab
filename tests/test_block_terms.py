"""Host logic of the block/tiled executor (no GPU): which per-block / per-tile contractions
an einsum over BlockTensors, TiledTensors and dense tensors becomes, checked by rebuilding
each einsum from its terms with numpy.einsum and comparing with numpy.einsum of the dense
tensors (absent tiles zero, a dense C changed only on its diagonal blocks)."""
import itertools
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from einsums_amd import blocks


def run_terms(terms, scalar=False):
    """Apply terms in order with numpy on the (CPU) torch tensors; returns the scalar sum."""
    total = 0.0
    for ci, C, ai, A, bi, B, cp, abp in terms:
        r = abp * np.einsum(f"{ai},{bi}->{ci}", A.numpy(), B.numpy())
        if scalar:
            total += float(r)
        else:
            C.copy_(torch.from_numpy(cp * C.numpy() + r))
    return total


def cpu_block(name, dims, rank, rng):
    T = blocks.BlockTensor(name, dims, rank, device="cpu")
    for i, n in enumerate(dims):
        T.blocks[i].copy_(torch.from_numpy(rng.uniform(-1, 1, (n,) * rank)))
    return T


def test_block_terms_dense_c_and_vector():
    rng = np.random.default_rng(3)
    dims = [3, 0, 5, 2]
    A = cpu_block("A", dims, 2, rng)
    B = cpu_block("B", dims, 2, rng)
    N = sum(dims)
    C = torch.from_numpy(rng.uniform(-1, 1, (N, N)))
    C0 = C.clone()
    terms = blocks.block_terms(0.5, "ij", C, 2.0, "ik", A, "kj", B)
    assert len(terms) == 3  # the empty block is skipped
    run_terms(terms)
    Ad, Bd = A.to_dense().numpy(), B.to_dense().numpy()
    want = C0.numpy().copy()
    off = 0
    for n in dims:
        s = slice(off, off + n)
        want[s, s] = 0.5 * want[s, s] + 2.0 * Ad[s, s] @ Bd[s, s]
        off += n
    assert np.allclose(C.numpy(), want, rtol=0, atol=1e-14)
    x = torch.from_numpy(rng.uniform(-1, 1, N))
    y = torch.zeros(N, dtype=torch.float64)
    run_terms(blocks.block_terms(0.0, "i", y, 1.0, "ij", A, "j", x))
    assert np.allclose(y.numpy(), Ad @ x.numpy(), rtol=0, atol=1e-14)
    s = run_terms(blocks.block_terms(0.0, "", None, 1.0, "ij", A, "ij", B), scalar=True)
    assert abs(s - float(np.sum(Ad * Bd))) < 1e-13
    with pytest.raises(ValueError):
        blocks.block_terms(0.0, "ij", C, 1.0, "ik", A, "kj", cpu_block("B", dims[:3], 2, rng))


def test_tiled_terms_create_missing_c_tiles_and_mix_blocks():
    rng = np.random.default_rng(4)
    td = [[2, 0, 3], [4, 1]]
    A = blocks.TiledTensor("A", [td[0], td[1]], device="cpu")
    B = blocks.TiledTensor("B", [td[1], td[0]], device="cpu")
    C = blocks.TiledTensor("C", [td[0], td[0]], device="cpu")
    for T, present in ((A, [(0, 0), (2, 1), (1, 0), (2, 0)]), (B, [(0, 2), (1, 0), (1, 2)])):
        for c in present:
            T.add_tile(c, torch.from_numpy(rng.uniform(-1, 1, T.tile_shape(c))))
    C.add_tile((0, 0), torch.from_numpy(rng.uniform(-1, 1, (2, 2))))
    B.add_tile((0, 0), torch.from_numpy(rng.uniform(-1, 1, B.tile_shape((0, 0)))))  # C(2,0) gets two triples
    Ad, Bd, Cd = A.to_dense().numpy(), B.to_dense().numpy(), C.to_dense().numpy()
    terms, scale = blocks.tiled_terms(0.5, "ij", C, 1.0, "ik", A, "kj", B)
    assert len(scale) == 1 and set(C.tiles) == {(0, 0), (0, 2), (2, 0), (2, 2)}
    for t in scale:
        t.mul_(0.5)
    run_terms(terms)
    assert np.allclose(C.to_dense().numpy(), 0.5 * Cd + Ad @ Bd, rtol=0, atol=1e-14)
    # a BlockTensor in the tile grid contributes its diagonal tiles only
    dims = [2, 0, 3]
    D = cpu_block("D", dims, 2, rng)
    E = blocks.TiledTensor("E", [dims, dims], device="cpu")
    terms, scale = blocks.tiled_terms(0.0, "ij", E, 1.0, "ik", D, "kj", B if False else D)
    assert len(terms) == 2 and all(t[3].shape == t[5].shape for t in terms)


def brute_force_triples(C_idx, C, A_idx, A, B_idx, B):
    """The reference's walk over the whole tile grid (TileAlgebra.hpp:59-92)."""
    labels = list(dict.fromkeys(C_idx + A_idx + B_idx))
    grid = {}
    for T, idx in ((C, C_idx), (A, A_idx), (B, B_idx)):
        for d, ch in enumerate(idx):
            grid[ch] = T.grid_size(d)
    out = []
    for combo in itertools.product(*[range(grid[ch]) for ch in labels]):
        pos = dict(zip(labels, combo))
        ac = tuple(pos[ch] for ch in A_idx)
        bc = tuple(pos[ch] for ch in B_idx)
        if A.has_tile(ac) and B.has_tile(bc):
            cc = tuple(pos[ch] for ch in C_idx)
            if all(e > 0 for e in C.tile_shape(cc)):
                out.append((cc, ac, bc))
    return out


@pytest.mark.parametrize("pattern", [("ij", "ik", "kj"), ("abij", "abef", "efij"), ("ijc", "iik", "kjj"),
                                     ("ijc", "ik", "kj"), ("ab", "ab", "ab")])
def test_tile_join_matches_grid_walk(pattern):
    """The join of present tiles gives the grid walk's triples in the grid walk's order, for
    sparse (irrep-like) tilings, repeated labels and labels found only in C."""
    rng = np.random.default_rng(zlib.crc32("|".join(pattern).encode()))
    ci, ai, bi = pattern
    ext = {ch: [int(x) for x in rng.integers(0, 4, 3)] for ch in set(ci + ai + bi)}

    def make(name, idx, density):
        T = blocks.TiledTensor(name, [ext[ch] for ch in idx], device="cpu")
        for c in itertools.product(*[range(3)] * len(idx)):
            if rng.uniform() < density:
                T.add_tile(c, torch.zeros(T.tile_shape(c), dtype=torch.float64))
        return T

    A, B = make("A", ai, 0.4), make("B", bi, 0.4)
    C = make("C", ci, 1.0)  # every C tile present: the walk and the join meet the same tiles
    want = brute_force_triples(ci, C, ai, A, bi, B)
    terms, _ = blocks.tiled_terms(1.0, ci, C, 1.0, ai, A, bi, B)
    ident = {id(t): c for T in (A, B, C) for c, t in T.tiles.items()}
    got = [(ident[id(t[1])], ident[id(t[3])], ident[id(t[5])]) for t in terms]
    assert got == want
