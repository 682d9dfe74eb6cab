"""Randomised GPU parity sweep of the permute against numpy's transpose and the oracle.

Each case draws a rank (1..6), extents (0..150 on at most two axes so that the tiled routes
see full, partial and 32/64-wide tiles; 1..9 elsewhere), the output order, padded or
stepped views of the source and the destination, the element type and a prefactor pair,
from a seed fixed by the case number. With C_prefactor 0 and A_prefactor 1 the result
must equal the transpose bit for bit (signed zeros included); otherwise it is checked
against the oracle to 1e-15 (FP64) / 1e-6 (FP32) relative Frobenius. C outside a
destination view must stay untouched.
"""
import zlib

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import host, rel_fro

pytestmark = pytest.mark.gpu

NCASES = 300
LABELS = "ijklmn"


def make_case(idx):
    rng = np.random.default_rng(zlib.crc32(f"random-permute|{idx}".encode()))
    rank = int(rng.integers(1, 7))
    shape = [int(rng.integers(1, 10)) for _ in range(rank)]
    for ax in rng.permutation(rank)[: min(rank, 2)]:  # up to two long axes
        shape[int(ax)] = int(rng.choice((int(rng.integers(10, 151)), 32, 64, 96, 128)))
    while int(np.prod(shape)) > 2_000_000:  # at most 16 MB per operand
        big = int(np.argmax(shape))
        shape[big] = (shape[big] + 1) // 2
    if rng.random() < 0.05:
        shape[int(rng.integers(rank))] = 0
    perm = [int(p) for p in rng.permutation(rank)]
    views = [str(rng.choice(("dense", "pad", "step"), p=(0.6, 0.25, 0.15))) for _ in range(2)]
    dtype = np.float32 if idx % 3 == 2 else np.float64
    pref = [(0.0, 1.0), (0.5, -2.0), (1.0, 1.0), (0.0, 0.75)][idx % 4]
    return dict(shape=shape, perm=perm, views=views, dtype=dtype, pref=pref, seed=idx)


CASES = [make_case(i) for i in range(NCASES)]


def _case_id(c):
    return f"{c['seed']}:{'x'.join(map(str, c['shape']))}:{''.join(map(str, c['perm']))}:{np.dtype(c['dtype']).name}"


def _tensor(rng, shape, view, dtype):
    if view == "pad" and shape:
        lo = [int(rng.integers(0, 3)) for _ in shape]
        full = [s + o + int(rng.integers(0, 3)) for s, o in zip(shape, lo)]
        idx = tuple(slice(o, o + s) for o, s in zip(lo, shape))
    elif view == "step" and shape:
        d = int(rng.integers(len(shape)))
        full = [s * 2 if i == d else s for i, s in enumerate(shape)]
        idx = tuple(slice(None, None, 2) if i == d else slice(None) for i in range(len(shape)))
    else:
        full = list(shape)
        idx = tuple(slice(None) for _ in shape)
    base = rng.standard_normal(full).astype(dtype)
    if base.size:
        base.flat[:: max(1, base.size // 7)] = -0.0
    return base, torch.from_numpy(base.copy()).to("cuda"), idx


@pytest.mark.parametrize("case", CASES, ids=_case_id)
def test_random_permute(core, oracle, case):
    rng = np.random.default_rng(zlib.crc32(f"random-permute-data|{case['seed']}".encode()))
    shape, perm, dtype = case["shape"], case["perm"], case["dtype"]
    inp = LABELS[: len(shape)]
    out = "".join(inp[p] for p in perm)
    a_base, a_t, a_idx = _tensor(rng, shape, case["views"][0], dtype)
    c_base, c_t, c_idx = _tensor(rng, [shape[p] for p in perm], case["views"][1], dtype)
    cp, ap = case["pref"]
    core.permute(cp, out, c_t[c_idx], ap, inp, a_t[a_idx])
    got = host(c_t)
    A = np.array(a_base[a_idx], copy=True, order="C")
    if (cp, ap) == (0.0, 1.0):
        assert np.ascontiguousarray(got[c_idx]).tobytes() == np.ascontiguousarray(A.transpose(perm)).tobytes()
    else:
        ref = np.array(c_base[c_idx], dtype=np.float64, copy=True, order="C")
        oracle.permute(cp, out, ref, ap, inp, A.astype(np.float64))
        assert rel_fro(got[c_idx], ref) < (1e-15 if dtype == np.float64 else 1e-6)
    outside = np.ones(c_base.shape, dtype=bool)
    outside[c_idx] = False
    assert np.array_equal(got[outside], c_base[outside])
