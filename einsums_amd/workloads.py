"""Seeded synthetic inputs for the benchmark workloads (host numpy arrays, row-major).

* Hartree-Fock G build, nbf basis functions: g[p,q,r,s] ~ N(0,1) averaged over the
  8 ERI permutations (pq<->qp, rs<->sr, pq<->rs) and scaled by 1/nbf; C (nbf x nocc)
  = Q of a seeded N(0,1) matrix; D = C C^T.
  J = einsum(pq <- pqrs, rs) [GEMV], K = einsum(pq <- prqs, rs) [GENERIC], G = 2J - K.
* CCD W intermediates at (o, v): t_oovv ~ N(0, 0.01^2); g_oooo, g_oovv, g_vvvv,
  g_ovvo ~ N(0,1)/sqrt(v), antisymmetrised in bra and ket pairs where both indices of a
  pair are in the same space. Wmnij += 0.25 t_ijef g_mnef, Wabef += 0.25 g_mnef t_mnab,
  Wmbej += -0.5 t_jnfb g_mnef, E = 0.25 t.g, and the permute ijab -> iajb.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def hf_inputs(nbf: int = 48, nocc: int = 12, seed: int = 2024):
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((nbf, nbf, nbf, nbf))
    g = (g + g.transpose(1, 0, 2, 3)) / 2
    g = (g + g.transpose(0, 1, 3, 2)) / 2
    g = (g + g.transpose(2, 3, 0, 1)) / 2
    g = np.ascontiguousarray(g / nbf)
    q, _ = np.linalg.qr(rng.standard_normal((nbf, nocc)))
    D = np.ascontiguousarray(q @ q.T)
    return g, D


def _antisym_pair(x: np.ndarray, a: int, b: int) -> np.ndarray:
    return (x - np.swapaxes(x, a, b)) / 2


@dataclass
class CCDInputs:
    o: int
    v: int
    t_oovv: np.ndarray
    g_oooo: np.ndarray
    g_oovv: np.ndarray
    g_vvvv: np.ndarray | None
    g_ovvo: np.ndarray


def ccd_inputs(o: int = 24, v: int = 96, seed: int = 7, with_vvvv: bool = True) -> CCDInputs:
    rng = np.random.default_rng(seed)
    s = 1.0 / np.sqrt(v)
    t = 0.01 * rng.standard_normal((o, o, v, v))
    t = _antisym_pair(_antisym_pair(t, 0, 1), 2, 3)
    goooo = s * rng.standard_normal((o, o, o, o))
    goooo = _antisym_pair(_antisym_pair(goooo, 0, 1), 2, 3)
    goovv = s * rng.standard_normal((o, o, v, v))
    goovv = _antisym_pair(_antisym_pair(goovv, 0, 1), 2, 3)
    gvvvv = None
    if with_vvvv:
        gvvvv = s * rng.standard_normal((v, v, v, v))
        gvvvv = _antisym_pair(_antisym_pair(gvvvv, 0, 1), 2, 3)
    # ovvo: bra (o,v) and ket (v,o) mix spaces, so no antisymmetrisation
    govvo = s * rng.standard_normal((o, v, v, o))
    c = np.ascontiguousarray
    return CCDInputs(o, v, c(t), c(goooo), c(goovv), None if gvvvv is None else c(gvvvv), c(govvo))


# (name, C labels, A name, A labels, B name, B labels, C_prefactor, AB_prefactor, W init)
CCD_CONTRACTIONS = [
    ("Wmnij", "mnij", "t_oovv", "ijef", "g_oovv", "mnef", 1.0, 0.25, "g_oooo"),
    ("Wabef", "abef", "g_oovv", "mnef", "t_oovv", "mnab", 1.0, 0.25, "g_vvvv"),
    ("Wmbej", "mbej", "t_oovv", "jnfb", "g_oovv", "mnef", 1.0, -0.5, "g_ovvo"),
]


def ccd_flops(o: int, v: int) -> dict:
    return {
        "Wmnij": 2.0 * o**4 * v**2,
        "Wabef": 2.0 * v**4 * o**2,
        "Wmbej": 2.0 * o**3 * v**3,
        "E": 2.0 * o**2 * v**2,
        "permute": 0.0,
    }


def ccd_bytes(o: int, v: int, elem: int = 8) -> dict:
    """Algorithmic HBM bytes: each operand read once, C read and written (beta != 0)."""
    t = o * o * v * v
    return {
        "Wmnij": elem * (t + t + 2 * o**4),
        "Wabef": elem * (t + t + 2 * v**4),
        "Wmbej": elem * (t + t + 2 * o * v * v * o),
        "E": elem * (2 * t),
        "permute": elem * (2 * t),
    }


def sample_rows(nrows: int, count: int = 64, seed: int = 97) -> np.ndarray:
    """`count` distinct row indices out of `nrows`, sorted, from a seeded generator, with
    the first and last row always included: the rows on which a large result is compared
    with a host computation."""
    rng = np.random.default_rng(seed)
    count = min(count, nrows)
    inner = rng.choice(np.arange(1, nrows - 1), size=max(0, count - 2), replace=False) if nrows > 2 else []
    return np.unique(np.concatenate([[0, nrows - 1], np.asarray(inner, dtype=np.int64)]))[:count].astype(np.int64)


def ccd_inputs_device(o: int, v: int, seed: int = 11, with_vvvv: bool = True, device: str = "cuda") -> CCDInputs:
    """The CCD inputs of `ccd_inputs` (same distributions, antisymmetrised the same way),
    drawn on the device with torch's counter-based generator so that o=48, v=256 (g_vvvv
    of 34.4 GB) is made in milliseconds. Fields are contiguous float64 device tensors."""
    import torch

    gen = torch.Generator(device=device).manual_seed(seed)

    def normal(shape, scale):
        return torch.randn(shape, dtype=torch.float64, device=device, generator=gen).mul_(scale)

    def anti(x):
        x = (x - x.transpose(0, 1)).mul_(0.5)
        return (x - x.transpose(2, 3)).mul_(0.5)

    s = 1.0 / float(np.sqrt(v))
    t = anti(normal((o, o, v, v), 0.01))
    goooo = anti(normal((o, o, o, o), s))
    goovv = anti(normal((o, o, v, v), s))
    gvvvv = anti(normal((v, v, v, v), s)) if with_vvvv else None
    govvo = normal((o, v, v, o), s)
    return CCDInputs(o, v, t, goooo, goovv, gvvvv, govvo)
