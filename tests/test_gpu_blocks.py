"""Block-sparse einsums as grouped persistent launches, against the CPU oracle per block:
BlockTensor (BlockAlgebra.hpp:23-52), TiledTensor (TileAlgebra.hpp:18-93), per-irrep CCD
W intermediates in one launch, 4096 uneven GEMMs, the autotuner, and hipGraph replay.
FP64 tolerance: relative Frobenius 1e-13 (FP32: 1e-5)."""
import math
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu


def seeded(*parts):
    return np.random.default_rng(zlib.crc32("|".join(map(str, parts)).encode()))


def test_blocktensor_gemm_and_generic(core, oracle):
    from einsums_amd import blocks

    rng = seeded("block")
    dims = [int(x) for x in rng.integers(8, 321, 8)]
    for pat in (("ij", "ik", "kj"), ("ij", "ji", "ij"), ("ij", "ki", "jk")):
        A = blocks.BlockTensor("A", dims, 2)
        B = blocks.BlockTensor("B", dims, 2)
        C = blocks.BlockTensor("C", dims, 2)
        hA = [rng.uniform(-1, 1, (n, n)) for n in dims]
        hB = [rng.uniform(-1, 1, (n, n)) for n in dims]
        hC = [rng.uniform(-1, 1, (n, n)) for n in dims]
        for i in range(8):
            A.blocks[i].copy_(dev(hA[i]))
            B.blocks[i].copy_(dev(hB[i]))
            C.blocks[i].copy_(dev(hC[i]))
        plan = blocks.einsum(0.5, pat[0], C, -1.0, pat[1], A, pat[2], B)
        assert plan.info()["waves"] == 1
        for i in range(8):
            ref = hC[i].copy()
            oracle.einsum(0.5, pat[0], ref, -1.0, pat[1], hA[i], pat[2], hB[i])
            assert rel_fro(host(C.blocks[i]), ref) < 1e-13, (pat, i)


def ccd_irrep_terms(rng, nirrep=8):
    """Per-irrep CCD blocks: o_h in [2,24], v_h in [8,96]."""
    o = [int(x) for x in rng.integers(2, 25, nirrep)]
    v = [int(x) for x in rng.integers(8, 97, nirrep)]
    host_t, terms, refs = [], [], []
    for h in range(nirrep):
        oo, vv = o[h], v[h]
        t = rng.uniform(-0.01, 0.01, (oo, oo, vv, vv))
        g = rng.uniform(-1, 1, (oo, oo, vv, vv)) / math.sqrt(vv)
        Wm = rng.uniform(-1, 1, (oo, oo, oo, oo))
        Wa = rng.uniform(-1, 1, (vv, vv, vv, vv))
        Wb = rng.uniform(-1, 1, (oo, vv, vv, oo))
        dt, dg, dWm, dWa, dWb = dev(t), dev(g), dev(Wm), dev(Wa), dev(Wb)
        terms += [("mnij", dWm, "ijef", dt, "mnef", dg, 1.0, 0.25), ("abef", dWa, "mnef", dg, "mnab", dt, 1.0, 0.25),
                  ("mbej", dWb, "jnfb", dt, "mnef", dg, 1.0, -0.5)]
        refs += [(Wm, "mnij", "ijef", t, "mnef", g, 0.25, dWm), (Wa, "abef", "mnef", g, "mnab", t, 0.25, dWa),
                 (Wb, "mbej", "jnfb", t, "mnef", g, -0.5, dWb)]
    return terms, refs


def test_ccd_irreps_one_grouped_launch(oracle):
    from einsums_amd import blocks

    rng = seeded("ccd-irreps")
    terms, refs = ccd_irrep_terms(rng)
    plan = blocks.GroupedEinsum(terms)
    assert plan.info()["waves"] == 1  # every output differs: one persistent launch
    plan.execute()
    for W0, ci, ai, A, bi, B, abp, dW in refs:
        ref = W0.copy()
        oracle.einsum(1.0, ci, ref, abp, ai, A, bi, B)
        assert rel_fro(host(dW), ref) < 1e-13, ci


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_4096_uneven_gemms_autotuned(oracle, dtype):
    from einsums_amd import blocks

    rng = seeded("uneven", dtype.__name__)
    shapes = np.exp(rng.uniform(np.log(4), np.log(512), (4096, 3))).astype(int)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    terms, checks = [], []
    for k, (m, n, kk) in enumerate(shapes):
        A = rng.uniform(-1, 1, (m, kk)).astype(dtype)
        B = rng.uniform(-1, 1, (kk, n)).astype(dtype)
        dA, dB = dev(A), dev(B)
        dC = torch.full((m, n), float("nan"), dtype=tdt, device="cuda")
        terms.append(("ij", dC, "ik", dA, "kj", dB, 0.0, 1.0))
        checks.append((A, B, dC))  # every output is checked
    plan = blocks.GroupedEinsum(terms, autotune=True)
    info = plan.info()
    assert info["waves"] == 1 and info["tile_rows"] in (64, 128)
    plan.execute()
    for A, B, dC in checks:
        ref = (A.astype(np.float64) @ B.astype(np.float64))
        assert rel_fro(host(dC), ref) < (1e-13 if dtype == np.float64 else 1e-5)


def test_tiledtensor_vs_dense(oracle):
    from einsums_amd import blocks

    rng = seeded("tiled")
    td = [[3, 5, 4], [2, 6, 3]]  # i and k tiles; j tiles below
    tj = [4, 3, 5]
    A = blocks.TiledTensor("A", [td[0], td[1]])      # ik
    B = blocks.TiledTensor("B", [td[1], tj])         # kj
    C = blocks.TiledTensor("C", [td[0], tj])         # ij
    for c in [(0, 0), (1, 1), (2, 2), (0, 2), (2, 1)]:
        A.add_tile(c, dev(rng.uniform(-1, 1, A.tile_shape(c))))
    for c in [(0, 0), (1, 1), (2, 2), (1, 0), (2, 1)]:
        B.add_tile(c, dev(rng.uniform(-1, 1, B.tile_shape(c))))
    for c in [(0, 0), (1, 1), (2, 2), (0, 1), (2, 1)]:
        C.add_tile(c, dev(rng.uniform(-1, 1, C.tile_shape(c))))
    dA, dB, dC0 = host(A.to_dense()), host(B.to_dense()), host(C.to_dense())
    plan = blocks.einsum(0.5, "ij", C, 2.0, "ik", A, "kj", B)
    assert plan.info()["waves"] >= 2  # several k tiles accumulate into the same C tile
    # reference: C scaled, then only the present tiles of C receive contributions
    full = 0.5 * dC0 + 2.0 * dA @ dB
    got = host(C.to_dense())
    offs_i = np.cumsum([0] + td[0])
    offs_j = np.cumsum([0] + tj)
    for (ti, tj_), t in C.tiles.items():
        blk = full[offs_i[ti]:offs_i[ti + 1], offs_j[tj_]:offs_j[tj_ + 1]]
        assert rel_fro(host(t), blk) < 1e-13


def test_graph_capture_replays_match_eager(core):
    from einsums_amd import blocks

    rng = seeded("graph")
    terms, refs = ccd_irrep_terms(rng, 4)
    plan = blocks.GroupedEinsum(terms)
    t0 = refs[0][3]
    dt, dg = terms[0][3], terms[0][5]
    E = torch.zeros((), dtype=torch.float64, device="cuda")
    P = torch.empty((t0.shape[0], t0.shape[2], t0.shape[1], t0.shape[3]), dtype=torch.float64, device="cuda")
    outs = [r[-1] for r in refs]
    init = [o.clone() for o in outs]
    s = torch.cuda.Stream()

    def iteration():
        plan.execute(s)
        core.einsum("", E, "ijab", dt, "ijab", dg, AB_prefactor=0.25, stream=s)
        core.permute(0.0, "iajb", P, 1.0, "ijab", dt, stream=s)

    # eager reference trajectory: three iterations from the initial state
    with torch.cuda.stream(s):
        iteration()
    s.synchronize()
    for o, i0 in zip(outs, init):
        o.copy_(i0)
    torch.cuda.synchronize()
    eager = []
    with torch.cuda.stream(s):
        for _ in range(3):
            iteration()
            s.synchronize()
            eager.append(([o.clone() for o in outs], E.clone(), P.clone()))
    for o, i0 in zip(outs, init):
        o.copy_(i0)
    torch.cuda.synchronize()
    g = blocks.Graph(iteration, s)
    for r in range(3):
        g.replay()
        s.synchronize()
        x = torch.ones(1000, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        ws, e, p = eager[r]
        for o, w in zip(outs, ws):
            assert torch.equal(o, w)
        assert torch.equal(E, e) and torch.equal(P, p)


def test_grouped_terms_keep_list_order(oracle):
    """Terms that read or rescale an earlier term's output run after it (sequential result)."""
    from einsums_amd import blocks

    rng = seeded("order")
    n = 96
    hA, hB, hX0, hY0 = (rng.uniform(-1, 1, (n, n)) for _ in range(4))
    A, B, X, Y = dev(hA), dev(hB), dev(hX0), dev(hY0)
    Z = dev(rng.uniform(-1, 1, (n, n)))
    hZ0 = host(Z)
    terms = [
        ("ij", X, "ik", A, "kj", B, 0.5, 1.0),   # X = 0.5 X + A B
        ("ij", Y, "ik", X, "kj", B, 0.0, 1.0),   # Y = X B           (reads X)
        ("ij", X, "ik", A, "kj", A, -2.0, 1.0),  # X = -2 X + A A    (rescales X after Y read it)
        ("ii", Z, "i", A[0], "i", B[0], 3.0, 1.0),  # Z = 3 Z, then Z_ii += A0_i B0_i (repeated C labels)
    ]
    plan = blocks.GroupedEinsum(terms)
    assert plan.info()["waves"] == 3
    plan.execute()
    rX = 0.5 * hX0 + hA @ hB
    rY = rX @ hB
    rX = -2.0 * rX + hA @ hA
    rZ = 3.0 * hZ0
    np.fill_diagonal(rZ, np.diag(rZ) + hA[0] * hB[0])
    assert rel_fro(host(X), rX) < 1e-13
    assert rel_fro(host(Y), rY) < 1e-13
    assert rel_fro(host(Z), rZ) < 1e-13
