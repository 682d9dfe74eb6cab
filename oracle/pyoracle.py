"""Python access to the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/ as the parity checker and by bench.py for the CPU-baseline timing.
Product code never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

MAXR = 16
ALG = {0: "GENERIC", 1: "DOT", 2: "DIRECT", 3: "GER", 4: "GEMV", 5: "GEMM"}
ALG_ID = {v: k for k, v in ALG.items()}


class Desc(ctypes.Structure):
    _fields_ = [("rank", ctypes.c_int32), ("labels", ctypes.c_int32 * MAXR), ("dims", ctypes.c_int64 * MAXR),
                ("strides", ctypes.c_int64 * MAXR)]


def build() -> Path:
    """Compile the oracle with its Makefile (g++ -O3 -fopenmp)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _load():
    if not LIB.exists():
        build()
    lib = ctypes.CDLL(str(LIB))
    dp, fp = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
    descp = ctypes.POINTER(Desc)
    lib.oracle_blas_name.restype = ctypes.c_char_p
    lib.oracle_set_threads.argtypes = [ctypes.c_int]
    lib.oracle_classify.argtypes = [descp, descp, descp]
    for name, T, P in (("f64", ctypes.c_double, dp), ("f32", ctypes.c_float, fp), ("ld", ctypes.c_double, dp)):
        fn = getattr(lib, f"oracle_einsum_generic_{name}")
        fn.argtypes = [T, descp, P, T, descp, P, descp, P]
    for name, T, P in (("f64", ctypes.c_double, dp), ("f32", ctypes.c_float, fp)):
        fn = getattr(lib, f"oracle_einsum_{name}")
        fn.argtypes = [T, descp, P, T, descp, P, descp, P, ctypes.POINTER(ctypes.c_int)]
        fn = getattr(lib, f"oracle_permute_{name}")
        i64p, i32p = ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int)
        fn.argtypes = [ctypes.c_int, i64p, i32p, T, P, i64p, T, P, i64p]
    lib.oracle_gemm_f64.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_double, dp, ctypes.c_int64, dp, ctypes.c_int64, ctypes.c_double, dp,
                                    ctypes.c_int64]
    lib.oracle_fastdiv.restype = ctypes.c_uint32
    lib.oracle_fastdiv.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        if "ORACLE_OPENBLAS64" not in os.environ:
            libs = Path(np.__file__).resolve().parent.parent / "numpy.libs"
            cands = sorted(libs.glob("libscipy_openblas64_*.so")) if libs.exists() else []
            if cands:
                os.environ["ORACLE_OPENBLAS64"] = str(cands[0])
        _lib = _load()
    return _lib


def desc(labels: str, arr: np.ndarray) -> Desc:
    d = Desc()
    d.rank = arr.ndim
    if len(labels) != arr.ndim:
        raise ValueError(f"labels {labels!r} do not match rank {arr.ndim}")
    for i in range(arr.ndim):
        d.labels[i] = ord(labels[i])
        d.dims[i] = arr.shape[i]
        d.strides[i] = arr.strides[i] // arr.itemsize
    return d


def _ptr(arr: np.ndarray):
    if arr.dtype == np.float64:
        return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def classify(C: str, A: str, B: str) -> str:
    z = lambda s: np.zeros([1] * len(s))
    return ALG[lib().oracle_classify(desc(C, z(C)), desc(A, z(A)), desc(B, z(B)))]


def einsum(C_prefactor, C_labels, C, AB_prefactor, A_labels, A, B_labels, B) -> str:
    """Reference CPU einsum (dispatch + folded BLAS or generic loop), in place on C."""
    fn = lib().oracle_einsum_f64 if C.dtype == np.float64 else lib().oracle_einsum_f32
    alg = ctypes.c_int(-1)
    st = fn(C_prefactor, desc(C_labels, C), _ptr(C), AB_prefactor, desc(A_labels, A), _ptr(A), desc(B_labels, B),
            _ptr(B), ctypes.byref(alg))
    if st:
        raise ValueError(f"oracle einsum failed with status {st}")
    return ALG[alg.value]


def einsum_generic(C_prefactor, C_labels, C, AB_prefactor, A_labels, A, B_labels, B, extended=False) -> None:
    """Generic loop nest only (the reference's OnlyUseGenericAlgorithm), in place on C."""
    if extended:
        fn = lib().oracle_einsum_generic_ld
    else:
        fn = lib().oracle_einsum_generic_f64 if C.dtype == np.float64 else lib().oracle_einsum_generic_f32
    st = fn(C_prefactor, desc(C_labels, C), _ptr(C), AB_prefactor, desc(A_labels, A), _ptr(A), desc(B_labels, B),
            _ptr(B))
    if st:
        raise ValueError(f"oracle generic failed with status {st}")


def permute(C_prefactor, C_labels, C, A_prefactor, A_labels, A) -> None:
    r = A.ndim
    perm = [A_labels.index(ch) for ch in C_labels]
    I64, I32 = ctypes.c_int64 * max(r, 1), ctypes.c_int * max(r, 1)
    fn = lib().oracle_permute_f64 if A.dtype == np.float64 else lib().oracle_permute_f32
    fn(r, I64(*A.shape), I32(*perm), A_prefactor, _ptr(A), I64(*[s // A.itemsize for s in A.strides]), C_prefactor,
       _ptr(C), I64(*[s // C.itemsize for s in C.strides]))


def gemm_colmajor(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc) -> None:
    lib().oracle_gemm_f64(ta.encode(), tb.encode(), m, n, k, alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(C), ldc)


def blas_name() -> str:
    return lib().oracle_blas_name().decode()


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(n)


def fastdiv(n: int, d: int) -> int:
    return lib().oracle_fastdiv(n, d)
