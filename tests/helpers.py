"""Shared helpers for the GPU parity tests."""
from __future__ import annotations

import numpy as np
import torch


def dev(a: np.ndarray) -> torch.Tensor:
    a = np.asarray(a)
    return torch.from_numpy(a.copy(order="C")).to("cuda")


def host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def rel_fro(x, ref) -> float:
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((x - ref).ravel())
    return float(num / den) if den > 0 else float(num)


def random_tensor(rng, shape, dtype=np.float64):
    return rng.uniform(-1.0, 1.0, size=shape).astype(dtype)


def read_water(path):
    """Read the water STO-3G integral files (1-based indices, unique elements only),
    filling the symmetric partners as QChemTests.cpp:33-89 does."""
    import pathlib

    p = pathlib.Path(path)
    S = np.zeros((7, 7)); T = np.zeros((7, 7)); V = np.zeros((7, 7)); G = np.zeros((7, 7, 7, 7))
    for name, M in (("S", S), ("T", T), ("V", V)):
        for line in (p / f"{name}.dat").read_text().split("\n"):
            f = line.split()
            if len(f) < 3:
                continue
            i, j, val = int(f[0]) - 1, int(f[1]) - 1, float(f[2])
            M[i, j] = val
            M[j, i] = val
    for line in (p / "TEI.dat").read_text().split("\n"):
        f = line.split()
        if len(f) < 5:
            continue
        i, j, k, l = (int(x) - 1 for x in f[:4])
        val = float(f[4])
        for a, b, c, d in ((i, j, k, l), (i, j, l, k), (j, i, k, l), (j, i, l, k), (k, l, i, j), (l, k, i, j),
                           (k, l, j, i), (l, k, j, i)):
            G[a, b, c, d] = val
    return S, T, V, G
