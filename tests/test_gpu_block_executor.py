"""Block and tiled einsums beyond block x block -> block, against the CPU oracle: the operand
kinds of the reference's special dispatch (BlockAlgebra.hpp:23-428, TileAlgebra.hpp:18-158,
BlockTileAlgebra.hpp:18-173):
  * block x block -> dense C (diagonal sub-blocks only; the rest of C untouched, bit for bit)
    and -> scalar C (the CCD energy summed over irreps);
  * dense vector x block and block x dense vector -> block, dense and scalar C;
  * tiled x tiled -> tiled with an absent C tile created where a contribution lands, and
    -> scalar C;
  * block x tiled and tiled x block -> tiled and block C.
Oracle: the reference CPU einsum restated (oracle/), applied per block, or on the dense
tensors (absent tiles zero). FP64 relative Frobenius 1e-13."""
import itertools
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu


def seeded(*parts):
    return np.random.default_rng(zlib.crc32("|".join(map(str, parts)).encode()))


def make_block(blocks, name, dims, rank, rng):
    T = blocks.BlockTensor(name, dims, rank)
    hs = []
    for i, n in enumerate(dims):
        h = rng.uniform(-1, 1, (n,) * rank)
        T.blocks[i].copy_(dev(h))
        hs.append(h)
    return T, hs


DIMS = [5, 0, 17, 33, 1, 64, 12, 40]  # includes an empty and a 1 x 1 block


def test_block_block_to_dense(oracle):
    from einsums_amd import blocks

    rng = seeded("bb-dense")
    A, hA = make_block(blocks, "A", DIMS, 2, rng)
    B, hB = make_block(blocks, "B", DIMS, 2, rng)
    N = sum(DIMS)
    hC = rng.uniform(-1, 1, (N, N))
    C = dev(hC)
    blocks.einsum(0.5, "ij", C, -1.5, "ik", A, "kj", B)
    got = host(C)
    expect = hC.copy()
    off = 0
    for i, n in enumerate(DIMS):
        if n:
            ref = hC[off:off + n, off:off + n].copy()
            oracle.einsum(0.5, "ij", ref, -1.5, "ik", hA[i], "kj", hB[i])
            expect[off:off + n, off:off + n] = ref
        off += n
    diag = np.zeros((N, N), bool)
    off = 0
    for n in DIMS:
        diag[off:off + n, off:off + n] = True
        off += n
    assert rel_fro(got[diag], expect[diag]) < 1e-13
    assert np.array_equal(got[~diag], hC[~diag])  # outside the diagonal blocks: untouched


@pytest.mark.parametrize("cp", [0.0, 1.0])
def test_block_block_to_scalar_energy(oracle, cp):
    """E = cp*E + 0.25 * sum over irreps of t_ijab g_ijab (o_h, v_h per irrep)."""
    from einsums_amd import blocks

    rng = seeded("energy", cp)
    o = [int(x) for x in rng.integers(2, 25, 8)]
    v = [int(x) for x in rng.integers(8, 97, 8)]
    # rank-4 blocks of a BlockTensor are hypercubic: use per-irrep 4-index tiles as blocks of
    # a block tensor of extent n_h = o_h*v_h flattened to rank 2 (same dot product)
    dims = [oo * vv for oo, vv in zip(o, v)]
    T, hT = make_block(blocks, "t", dims, 2, rng)
    G, hG = make_block(blocks, "g", dims, 2, rng)
    E0 = float(rng.uniform(-1, 1))
    E = torch.tensor(float("nan") if cp == 0.0 else E0, dtype=torch.float64, device="cuda")
    blocks.einsum(cp, "", E, 0.25, "ia", T, "ia", G)
    ref = np.zeros(())
    ref[()] = 0.0 if cp == 0.0 else E0
    total = np.zeros(())
    for i in range(8):
        part = np.zeros(())
        oracle.einsum(0.0, "", part, 0.25, "ia", hT[i], "ia", hG[i])
        total[()] += part[()]
    expect = cp * ref[()] + total[()] if cp else total[()]
    assert abs(float(E.item()) - expect) <= 1e-13 * abs(expect)


def test_vector_block_cases(oracle):
    from einsums_amd import blocks

    rng = seeded("vec-block")
    N = sum(DIMS)
    offs = [0] + list(itertools.accumulate(DIMS))
    A, hA = make_block(blocks, "A", DIMS, 2, rng)
    hx = rng.uniform(-1, 1, N)
    x = dev(hx)
    # block x vector -> dense vector: y_i = 2 y_i + sum_j A_ij x_j per block (block-diagonal matvec)
    hy = rng.uniform(-1, 1, N)
    y = dev(hy)
    blocks.einsum(2.0, "i", y, 1.0, "ij", A, "j", x)
    ref = hy.copy()
    for i, n in enumerate(DIMS):
        if n:
            r = ref[offs[i]:offs[i + 1]].copy()
            oracle.einsum(2.0, "i", r, 1.0, "ij", hA[i], "j", hx[offs[i]:offs[i + 1]].copy())
            ref[offs[i]:offs[i + 1]] = r
    assert rel_fro(host(y), ref) < 1e-13
    # vector x block -> block: C_ij = x_i A_ij (a direct-product broadcast per block)
    C, hC = make_block(blocks, "C", DIMS, 2, rng)
    blocks.einsum(0.0, "ij", C, 3.0, "i", x, "ij", A)
    for i, n in enumerate(DIMS):
        if n:
            r = np.zeros((n, n))
            oracle.einsum(0.0, "ij", r, 3.0, "i", hx[offs[i]:offs[i + 1]].copy(), "ij", hA[i])
            assert rel_fro(host(C.blocks[i]), r) < 1e-13, i
    # block x vector -> scalar: s = sum_i x_i A_ii over blocks (a trace weighted by x)
    s = torch.tensor(0.5, dtype=torch.float64, device="cuda")
    blocks.einsum(1.0, "", s, 1.0, "ii", A, "i", x)
    tot = 0.5
    for i, n in enumerate(DIMS):
        if n:
            part = np.zeros(())
            oracle.einsum(0.0, "", part, 1.0, "ii", hA[i], "i", hx[offs[i]:offs[i + 1]].copy())
            tot += part[()]
    assert abs(s.item() - tot) <= 1e-13 * abs(tot)
    # vector x block -> dense matrix: D_ij = x_i A_ij on the diagonal blocks
    hD = rng.uniform(-1, 1, (N, N))
    Dm = dev(hD)
    blocks.einsum(1.0, "ij", Dm, 1.0, "i", x, "ij", A)
    got = host(Dm)
    for i, n in enumerate(DIMS):
        if n:
            r = hD[offs[i]:offs[i + 1], offs[i]:offs[i + 1]].copy()
            oracle.einsum(1.0, "ij", r, 1.0, "i", hx[offs[i]:offs[i + 1]].copy(), "ij", hA[i])
            assert rel_fro(got[offs[i]:offs[i + 1], offs[i]:offs[i + 1]], r) < 1e-13, i


def tiled(blocks, name, tile_dims, present, rng):
    T = blocks.TiledTensor(name, tile_dims)
    for c in present:
        T.add_tile(c, dev(rng.uniform(-1, 1, T.tile_shape(c))))
    return T


def test_tiled_missing_c_tile_and_scalar(oracle):
    from einsums_amd import blocks

    rng = seeded("tiled-more")
    td = [[7, 0, 19, 12], [5, 16, 9]]
    A = tiled(blocks, "A", [td[0], td[1]], [(0, 0), (0, 2), (2, 1), (3, 0), (3, 1), (1, 1)], rng)
    B = tiled(blocks, "B", [td[1], td[0]], [(0, 0), (1, 2), (2, 3), (0, 3), (1, 0)], rng)
    C = tiled(blocks, "C", [td[0], td[0]], [(0, 0), (2, 2)], rng)  # most C tiles absent
    hA, hB, hC = (host(X.to_dense()) for X in (A, B, C))
    blocks.einsum(0.5, "ij", C, 2.0, "ik", A, "kj", B)
    ref = hC.copy()
    oracle.einsum(0.5, "ij", ref, 2.0, "ik", hA, "kj", hB)
    assert rel_fro(host(C.to_dense()), ref) < 1e-13
    assert (3, 3) in C.tiles and (1, 1) not in C.tiles  # created where needed; empty ones not
    s = torch.tensor(float("nan"), dtype=torch.float64, device="cuda")
    blocks.einsum(0.0, "", s, 1.0, "ik", A, "ki", B)
    part = np.zeros(())
    oracle.einsum(0.0, "", part, 1.0, "ik", hA, "ki", hB)
    assert abs(s.item() - part[()]) <= 1e-13 * abs(part[()])


def test_mixed_block_and_tiled(oracle):
    from einsums_amd import blocks

    rng = seeded("mixed")
    dims = [6, 11, 0, 23]
    A, hA = make_block(blocks, "A", dims, 2, rng)
    B = tiled(blocks, "B", [dims, dims], [(0, 0), (0, 1), (1, 3), (3, 3), (3, 0)], rng)
    C = tiled(blocks, "C", [dims, dims], [(0, 0), (3, 3)], rng)
    hAd = host(A.to_dense())
    hB, hC = host(B.to_dense()), host(C.to_dense())
    blocks.einsum(0.25, "ij", C, 1.0, "ik", A, "kj", B)  # block x tiled -> tiled
    ref = hC.copy()
    oracle.einsum(0.25, "ij", ref, 1.0, "ik", hAd, "kj", hB)
    assert rel_fro(host(C.to_dense()), ref) < 1e-13
    # tiled x block -> block: only the diagonal positions of C exist
    D, hD = make_block(blocks, "D", dims, 2, rng)
    blocks.einsum(1.0, "ij", D, -1.0, "ik", B, "kj", A)
    offs = [0] + list(itertools.accumulate(dims))
    full = np.zeros((sum(dims), sum(dims)))
    oracle.einsum(0.0, "ij", full, -1.0, "ik", hB, "kj", hAd)
    for i, n in enumerate(dims):
        if n:
            r = hD[i] + full[offs[i]:offs[i + 1], offs[i]:offs[i + 1]]
            assert rel_fro(host(D.blocks[i]), r) < 1e-13, i


def test_plan_cache_reuse_and_invalidation(oracle):
    """A repeated tiled einsum reuses its triple list and device plan; adding a tile
    invalidates it. Every call's result is checked against the oracle."""
    from einsums_amd import blocks

    rng = seeded("cache")
    dims = [[9, 14, 6], [11, 7]]
    A = tiled(blocks, "A", [dims[0], dims[1]], [(0, 0), (1, 1), (2, 0)], rng)
    B = tiled(blocks, "B", [dims[1], dims[0]], [(0, 0), (1, 2), (0, 1)], rng)
    C = tiled(blocks, "C", [dims[0], dims[0]], [(0, 0)], rng)
    p1 = blocks.einsum(0.5, "ij", C, 1.0, "ik", A, "kj", B)
    hC = host(C.to_dense())
    p2 = blocks.einsum(0.5, "ij", C, 1.0, "ik", A, "kj", B)
    assert p2 is p1  # the cached plan
    ref = hC.copy()
    oracle.einsum(0.5, "ij", ref, 1.0, "ik", host(A.to_dense()), "kj", host(B.to_dense()))
    assert rel_fro(host(C.to_dense()), ref) < 1e-13
    A.add_tile((2, 1), dev(rng.uniform(-1, 1, A.tile_shape((2, 1)))))
    hC = host(C.to_dense())
    p3 = blocks.einsum(0.5, "ij", C, 1.0, "ik", A, "kj", B)
    assert p3 is not p1
    ref = hC.copy()
    oracle.einsum(0.5, "ij", ref, 1.0, "ik", host(A.to_dense()), "kj", host(B.to_dense()))
    assert rel_fro(host(C.to_dense()), ref) < 1e-13
