"""Permute on the GPU: bit-exact against numpy's transpose (alpha=1, beta=0) and the
CPU oracle, on every kernel route (row copy with 16-byte vectors, LDS tile transpose with
64- and 32-wide tiles, element-wise strided), plus alpha/beta blending and inverse round
trips."""
import itertools
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu


def seeded(*parts):
    return np.random.default_rng(zlib.crc32("|".join(map(str, parts)).encode()))


def labels(n):
    return "ijklmnop"[:n]


CASES = [
    ((24, 24, 96, 96), "ikjl"),   # CCD ijab -> iajb: innermost axis kept (row copy, 2-wide vectors)
    ((24, 24, 96, 96), "klij"),   # ijab -> abij: innermost axis changes (LDS tile transpose)
    ((3, 130, 66), "ikj"),        # tile transpose with partial edge tiles on both axes
    ((2, 128, 192), "ikj"),       # full tiles only, 16-byte vectors for both widths
    ((2, 136, 200), "ikj"),       # full (vector) and edge (element) tiles in one launch
    ((24, 24, 96, 96), "iljk"),   # Wmbej pack jnfb -> jbnf: a 96-long axis takes 32-wide tiles
    ((2, 96, 160), "ikj"),        # 32-wide tiles on both axes (neither a multiple of 64)
    ((5, 7, 9), "kji"),
    ((64, 65), "ji"),
    ((1, 130, 1, 67), "lkji"),
    ((3, 4, 5, 6, 7), "mlkji"),
    ((333,), "i"),
    ((17, 19, 3), "kij"),
]


@pytest.mark.parametrize("shape,out", CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_permute_bit_exact(core, oracle, shape, out, dtype):
    rng = seeded(shape, out, dtype.__name__)
    inp = labels(len(shape))
    A = rng.standard_normal(shape).astype(dtype)
    A.flat[:: max(1, A.size // 7)] = -0.0  # signed zeros must survive
    perm = [inp.index(c) for c in out]
    expect = np.ascontiguousarray(A.transpose(perm))
    dC = torch.full(expect.shape, float("nan"), dtype=torch.float64 if dtype == np.float64 else torch.float32,
                    device="cuda")
    core.permute(0.0, out, dC, 1.0, inp, dev(A))
    got = host(dC)
    assert got.tobytes() == expect.tobytes()
    ref = np.full(expect.shape, np.nan, dtype=dtype)
    oracle.permute(0.0, out, ref, 1.0, inp, A)
    assert ref.tobytes() == got.tobytes()


def test_permute_all_rank4_permutations_roundtrip(core):
    rng = seeded("rank4")
    A = rng.standard_normal((6, 5, 4, 3))
    dA = dev(A)
    for p in itertools.permutations(range(4)):
        out = "".join("ijkl"[q] for q in p)
        dC = torch.empty([A.shape[q] for q in p], dtype=torch.float64, device="cuda")
        core.permute(0.0, out, dC, 1.0, "ijkl", dA)
        back = torch.empty_like(dA)
        core.permute(0.0, "ijkl", back, 1.0, out, dC)
        assert torch.equal(back, dA)
        assert host(dC).tobytes() == np.ascontiguousarray(A.transpose(p)).tobytes()


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-15), (np.float32, 1e-6)])
@pytest.mark.parametrize("shape", [(2, 136, 200), (3, 96, 160)])
def test_permute_tile_alpha_beta(core, oracle, dtype, tol, shape):
    """C = beta*C + alpha*perm(A) through the tile transpose: 64-wide vector and edge tiles,
    and 32-wide tiles."""
    rng = seeded("tile-ab", dtype.__name__, shape)
    A = rng.standard_normal(shape).astype(dtype)
    C0 = rng.standard_normal((shape[0], shape[2], shape[1])).astype(dtype)
    dC = dev(C0)
    core.permute(0.5, "ikj", dC, -2.0, "ijk", dev(A))
    ref = C0.astype(np.float64)
    oracle.permute(0.5, "ikj", ref, -2.0, "ijk", A.astype(np.float64))
    assert rel_fro(host(dC), ref) < tol


def test_permute_alpha_beta_and_strided_views(core, oracle):
    rng = seeded("ab")
    A = rng.standard_normal((40, 70, 30))
    C0 = rng.standard_normal((30, 70, 40))
    dC = dev(C0)
    core.permute(0.5, "kji", dC, -2.0, "ijk", dev(A))
    ref = C0.copy()
    oracle.permute(0.5, "kji", ref, -2.0, "ijk", A)
    assert rel_fro(host(dC), ref) < 1e-15
    # strided source and destination views
    big = dev(rng.standard_normal((80, 70, 60)))
    src = big[::2, :, 1::2]
    dst_full = torch.zeros((30, 140, 40), dtype=torch.float64, device="cuda")
    dst = dst_full[:, ::2, :]
    core.permute(0.0, "kji", dst, 1.0, "ijk", src)
    assert torch.equal(dst, src.permute(2, 1, 0))
    assert torch.all(dst_full[:, 1::2, :] == 0)
