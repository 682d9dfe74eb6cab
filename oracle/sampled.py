"""Oracle results on sampled output rows of the large CCD contractions -- TEST INFRASTRUCTURE.

At o=48, v=256 the full CPU oracle of Wabef (19.8 TFLOP, 34 GB output) and of the generic
Wmbej (3.7 TFLOP through the loop nest) does not finish in a test. For those the oracle
contracts the same operands restricted to sampled output rows: the sampled rows become a
label `s` of a smaller einsum, which the oracle dispatches like any other (BLAS GEMM for
Wabef and Wmnij, the generic loop nest for Wmbej, as the reference would). Used by tests/
and by bench.py's CPU leg, never by the product path.

Contractions (README.md:92-118 of the reference):
  Wmnij += 0.25 t_ijef g_mnef    rows = (m, n) pairs
  Wabef += 0.25 g_mnef t_mnab    rows = (a, b) pairs
  Wmbej += -0.5 t_jnfb g_mnef    rows = (m, b) pairs
Each function returns the contraction part alone (without the initial W), shape
(len(rows), <row shape>).
"""
from __future__ import annotations

import numpy as np

from oracle import pyoracle


def wmnij_rows(t: np.ndarray, g: np.ndarray, rows: np.ndarray) -> np.ndarray:
    """0.25 sum_ef t_ijef g_mnef for the sampled (m, n) = divmod(row, o)."""
    o, v = t.shape[0], t.shape[2]
    gs = np.ascontiguousarray(g.reshape(o * o, v, v)[rows])  # (s, e, f)
    out = np.zeros((len(rows), o, o))
    pyoracle.einsum(0.0, "sij", out, 0.25, "ijef", t, "sef", gs)
    return out


def wabef_rows(t: np.ndarray, g: np.ndarray, rows: np.ndarray) -> np.ndarray:
    """0.25 sum_mn g_mnef t_mnab for the sampled (a, b) = divmod(row, v)."""
    o, v = t.shape[0], t.shape[2]
    ts = np.ascontiguousarray(t.reshape(o, o, v * v)[:, :, rows].transpose(2, 0, 1))  # (s, m, n)
    out = np.zeros((len(rows), v, v))
    pyoracle.einsum(0.0, "sef", out, 0.25, "smn", ts, "mnef", g)
    return out


def wmbej_rows(t: np.ndarray, g: np.ndarray, rows: np.ndarray) -> np.ndarray:
    """-0.5 sum_nf t_jnfb g_mnef for the sampled (m, b) = divmod(row, v)."""
    o, v = t.shape[0], t.shape[2]
    m, b = rows // v, rows % v
    ts = np.ascontiguousarray(np.moveaxis(t[:, :, :, b], 3, 0))  # (s, j, n, f)
    gs = np.ascontiguousarray(g[m])  # (s, n, e, f)
    out = np.zeros((len(rows), v, o))
    pyoracle.einsum(0.0, "sej", out, -0.5, "sjnf", ts, "snef", gs)
    return out


ROWS_OF = {"Wmnij": wmnij_rows, "Wabef": wabef_rows, "Wmbej": wmbej_rows}


def row_count(name: str, o: int, v: int) -> int:
    return {"Wmnij": o * o, "Wabef": v * v, "Wmbej": o * v}[name]


def rel_err(got: np.ndarray, ref: np.ndarray) -> float:
    den = float(np.linalg.norm(ref.ravel()))
    num = float(np.linalg.norm((np.asarray(got, dtype=np.float64) - ref).ravel()))
    return num / den if den > 0 else num
