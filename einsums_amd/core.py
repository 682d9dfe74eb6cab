"""Python surface of the einsum path, shaped like ``einsums.core`` of EinsumsPy.

``compile_plan(C, A, B)`` analyses three index strings and returns a plan whose class
names the algorithm, exactly like the reference
(libs/EinsumsPy/TensorAlgebra/src/PyTensorAlgebra.cpp:169-300, classes exported at
libs/EinsumsPy/TensorAlgebra/src/Export.cpp:14-146). ``plan.execute(C_prefactor, C,
AB_prefactor, A, B)`` runs it on the GPU through the C ABI. Operands are torch tensors
on the GPU (any strides), ``GPUView`` wrappers of host arrays, host numpy arrays (staged
through device memory), or equally long lists of those (one call per entry).

There is no CPU path: every execute runs HIP kernels from libeinsums_amd.so.
"""
from __future__ import annotations

import contextlib
import ctypes
from typing import Any, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import (DOT, DIRECT, GEMM, GEMV, GENERIC, GER, INDETERMINATE, ALGORITHM_NAMES, DimensionError,
                   EinsumsError, RankError, check, lib, make_desc)

__all__ = [
    "compile_plan", "EinsumGenericPlan", "EinsumDotPlan", "EinsumDirectProductPlan", "EinsumGerPlan",
    "EinsumGemvPlan", "EinsumGemmPlan", "GPUView", "COPY", "MAP", "DEVICE_TENSOR", "gpu_enabled", "einsum",
    "permute", "stream_handle", "DimensionError", "RankError", "EinsumsError", "contraction_route", "syev", "gesv",
    "gemm_host_c", "fock_jk", "gemm_tuning", "save_gemm_tuning", "load_gemm_tuning", "clear_gemm_tuning",
    "gemm_tuning_lookup",
]

# GPUView modes, numbered as the reference's PyViewMode
# (libs/EinsumsPy/GPU/include/EinsumsPy/GPU/PyGPUView.hpp:27-46)
MAP, COPY, DEVICE_TENSOR = 0, 1, 2

_SUPPORTED = (torch.float64, torch.float32)
_COMPLEX = (torch.complex128, torch.complex64)


def gpu_enabled() -> bool:
    """True when a HIP device is usable (the reference's ein.core.gpu_enabled)."""
    return torch.cuda.is_available()


def stream_handle(stream: torch.cuda.Stream | None = None) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class GPUView:
    """Device view of a buffer (reference PyGPUView, PyGPUView.hpp:27-120).

    COPY: a device copy of a host array, refreshed only by ``update_H2D``/``update_D2H``.
    MAP: a host array kept in step automatically after every execute that writes it.
    DEVICE_TENSOR: a tensor already in device memory (a CUDA torch tensor), used in place;
    the update calls do nothing. Passing a CUDA tensor selects this mode.
    The reference maps a MAP buffer into the device address space by registering the host
    memory (hipHostRegister); registering existing host memory is refused on the systems
    this backend targets, so MAP keeps a device copy and writes it back after each execute.
    """

    def __init__(self, array, mode: int = COPY):
        if isinstance(array, torch.Tensor) and array.is_cuda:
            if array.dtype not in _SUPPORTED + _COMPLEX:
                raise TypeError(f"GPUView: unsupported dtype {array.dtype}")
            self._host, self._array, self.mode, self.tensor = None, array, DEVICE_TENSOR, array
            return
        if mode == DEVICE_TENSOR:
            raise ValueError("GPUView: DEVICE_TENSOR mode views a device tensor; host buffers take COPY or MAP")
        if mode not in (COPY, MAP):
            raise ValueError(f"GPUView: unknown mode {mode}")
        if isinstance(array, torch.Tensor):
            host = array
        else:
            host = torch.from_numpy(np.asarray(array))
        if host.dtype not in _SUPPORTED + _COMPLEX:
            raise TypeError(f"GPUView: unsupported dtype {host.dtype}")
        self._host = host
        self._array = array
        self.mode = mode
        self.tensor = host.to("cuda")

    def update_H2D(self) -> None:
        if self.mode != DEVICE_TENSOR:
            self.tensor.copy_(self._host)

    def update_D2H(self) -> None:
        if self.mode != DEVICE_TENSOR:
            self._host.copy_(self.tensor.cpu())

    @property
    def shape(self):
        return tuple(self.tensor.shape)

    def dim(self, i: int) -> int:
        return self.tensor.shape[i]

    def rank(self) -> int:
        return self.tensor.dim()


def _as_device(x, write: bool):
    """Returns (device tensor, finaliser)."""
    if isinstance(x, GPUView):
        fin = (lambda: x.update_D2H()) if (write and x.mode == MAP) else None
        return x.tensor, fin
    if isinstance(x, torch.Tensor):
        if x.is_cuda:
            return x, None
        dev = x.to("cuda")
        return dev, ((lambda: x.copy_(dev.cpu())) if write else None)
    if isinstance(x, np.ndarray):
        host = torch.from_numpy(x)
        dev = host.to("cuda")
        return dev, ((lambda: host.copy_(dev.cpu())) if write else None)
    raise TypeError(f"unsupported operand type {type(x).__name__}")


def _desc(t: torch.Tensor, idx: str):
    if t.dim() != len(idx):
        raise RankError(f"rank of tensor ({t.dim()}) does not match index string {idx!r}")
    return make_desc(idx, t.shape, t.stride())


def _fn(name: str, dtype: torch.dtype):
    return getattr(lib, f"{name}_{'f64' if dtype == torch.float64 else 'f32'}")


def _labels_only_desc(idx: str):
    return make_desc(idx, [1] * len(idx), [1] * len(idx))


def classify(C_indices: str, A_indices: str, B_indices: str) -> int:
    out = ctypes.c_int32(-1)
    check(lib.ein_einsum_classify(ctypes.byref(_labels_only_desc(C_indices)), ctypes.byref(_labels_only_desc(A_indices)),
                                  ctypes.byref(_labels_only_desc(B_indices)), ctypes.byref(out)), "classify")
    return out.value


class EinsumGenericPlan:
    """Plan for an arbitrary contraction (reference PyEinsumGenericPlan)."""

    algorithm = GENERIC

    def __init__(self, C_indices: str, A_indices: str, B_indices: str):
        self.C_indices = C_indices
        self.A_indices = A_indices
        self.B_indices = B_indices
        self.last_algorithm = INDETERMINATE

    def __repr__(self) -> str:
        return f"{type(self).__name__}({self.C_indices!r} <- {self.A_indices!r} * {self.B_indices!r})"

    def execute(self, C_prefactor, C, AB_prefactor, A, B, stream: torch.cuda.Stream | None = None,
                split: int = 0, slice_type: str = "bf16") -> None:
        """Run the plan on ``stream`` (torch's current stream when None). ``split=s`` (3..6)
        runs an FP64 GEMM-mapped contraction with the split-precision (Ozaki) GEMM, cutting
        each operand into s slices of ``slice_type`` "bf16", "fp16" or "int8", or, with
        ``slice_type="crt"``, uses s moduli (8..16) of the integer modular scheme."""
        if isinstance(C, (list, tuple)):
            if not (isinstance(A, (list, tuple)) and isinstance(B, (list, tuple))) or not (len(C) == len(A) == len(B)):
                raise ValueError("The lengths of the collections need to be the same!")
            for c, a, b in zip(C, A, B):
                self.execute(C_prefactor, c, AB_prefactor, a, b, stream, split, slice_type)
            return
        # every torch op of the call (staging copies, temporaries, the copy back) runs on the
        # kernel's stream, so they are ordered with it and the caching allocator reuses the
        # temporaries only after the kernel
        with torch.cuda.stream(stream):
            self._execute_one(C_prefactor, C, AB_prefactor, A, B, split, slice_type)

    def _execute_one(self, C_prefactor, C, AB_prefactor, A, B, split, slice_type) -> None:
        stream = None  # the current stream, set by execute()
        dC, finC = _as_device(C, write=True)
        dA, _ = _as_device(A, write=False)
        dB, _ = _as_device(B, write=False)
        dtype = dC.dtype
        if dtype in _COMPLEX and dA.dtype == dtype and dB.dtype == dtype:
            self._execute_complex(C_prefactor, dC, AB_prefactor, dA, dB, split, slice_type)
            if finC is not None:
                torch.cuda.current_stream().synchronize()
                finC()
            return
        if dtype not in _SUPPORTED or dA.dtype != dtype or dB.dtype != dtype:
            raise TypeError(f"execute: operands must share float64, float32, complex128 or complex64, got {dC.dtype}, "
                            f"{dA.dtype}, {dB.dtype}")
        C_prefactor, AB_prefactor = _real_prefactor(C_prefactor), _real_prefactor(AB_prefactor)
        cd, ad, bd = _desc(dC, self.C_indices), _desc(dA, self.A_indices), _desc(dB, self.B_indices)
        used = ctypes.c_int32(-1)
        if split:
            if dtype != torch.float64:
                raise TypeError("split precision applies to float64 operands")
            st = _slice_code(slice_type)
            status = lib.ein_einsum_split_f64(_split_count(split), st, float(C_prefactor), ctypes.byref(cd), _ptr(dC),
                                              float(AB_prefactor), ctypes.byref(ad), _ptr(dA), ctypes.byref(bd), _ptr(dB),
                                              ctypes.byref(used), stream_handle(stream))
        else:
            fn = _fn("ein_einsum_execute", dtype)
            status = fn(self.algorithm, float(C_prefactor), ctypes.byref(cd), _ptr(dC), float(AB_prefactor),
                        ctypes.byref(ad), _ptr(dA), ctypes.byref(bd), _ptr(dB), ctypes.byref(used), stream_handle(stream))
        check(status, f"einsum {self.C_indices} <- {self.A_indices},{self.B_indices}")
        self.last_algorithm = used.value
        if finC is not None:
            torch.cuda.current_stream().synchronize()
            finC()


    def _execute_complex(self, C_prefactor, dC, AB_prefactor, dA, dB, split, slice_type) -> None:
        """Complex operands as real einsums on the real and imaginary parts (strided real
        views, the same kernels): with P = sum A B, Re P = Re A Re B - Im A Im B and
        Im P = Re A Im B + Im A Re B, and C = c C + a P for complex c and a. A complex c
        first forms c C from a copy of C (identity permutes), then the products accumulate.
        For a GEMM plan the parts of A and B are first packed into contiguous planes."""
        c, a = complex(C_prefactor), complex(AB_prefactor)
        Cr, Ci = _parts(dC)
        Ar, Ai = _parts(dA)
        Br, Bi = _parts(dB)
        if self.algorithm == GEMM:
            # the GEMM stages stride-2 operands through its slow element path: each part is
            # first copied into a contiguous plane by the permute
            Ar, Ai = _packed(Ar, self.A_indices), _packed(Ai, self.A_indices)
            Br, Bi = _packed(Br, self.B_indices), _packed(Bi, self.B_indices)
        if c.imag != 0.0:
            _complex_scale(c, dC, Cr, Ci)
            beta = 1.0
        else:
            beta = c.real
        # (target part, [(prefactor, A part, B part)]): Re C += a_r Re P - a_i Im P, Im C += a_r Im P + a_i Re P
        terms = [(Cr, [(a.real, Ar, Br), (-a.real, Ai, Bi), (-a.imag, Ar, Bi), (-a.imag, Ai, Br)]),
                 (Ci, [(a.real, Ar, Bi), (a.real, Ai, Br), (a.imag, Ar, Br), (-a.imag, Ai, Bi)])]
        for target, parts in terms:
            first = True
            for pre, x, y in parts:
                if pre == 0.0 and not first:
                    continue
                self._execute_one(beta if first else 1.0, target, pre, x, y, split, slice_type)
                first = False


def _real_prefactor(x) -> float:
    if isinstance(x, complex):
        if x.imag != 0.0:
            raise TypeError("a complex prefactor needs complex operands")
        return x.real
    return float(x)


def _parts(t: torch.Tensor):
    """Real and imaginary parts of a complex tensor as strided real views of its storage."""
    r = torch.view_as_real(t)
    return r[..., 0], r[..., 1]


def _packed(x: torch.Tensor, idx: str) -> torch.Tensor:
    """A contiguous copy of the real view x, made by the identity permute."""
    out = torch.empty(x.shape, dtype=x.dtype, device=x.device)
    _permute_one(0.0, idx, out, 1.0, idx, x)
    return out


def _complex_scale(c: complex, dC: torch.Tensor, Cr: torch.Tensor, Ci: torch.Tensor) -> None:
    """C = c C for complex c, through identity permutes of the parts (and a copy of C)."""
    idx = "abcdefghijklmnop"[:dC.dim()]
    T = torch.empty_like(dC)
    Tr, Ti = _parts(T)
    _permute_one(0.0, idx, Tr, 1.0, idx, Cr)
    _permute_one(0.0, idx, Ti, 1.0, idx, Ci)
    _permute_one(0.0, idx, Cr, c.real, idx, Tr)
    _permute_one(1.0, idx, Cr, -c.imag, idx, Ti)
    _permute_one(0.0, idx, Ci, c.real, idx, Ti)
    _permute_one(1.0, idx, Ci, c.imag, idx, Tr)


class EinsumDotPlan(EinsumGenericPlan):
    algorithm = DOT


class EinsumDirectProductPlan(EinsumGenericPlan):
    algorithm = DIRECT


class EinsumGerPlan(EinsumGenericPlan):
    algorithm = GER


class EinsumGemvPlan(EinsumGenericPlan):
    algorithm = GEMV


class EinsumGemmPlan(EinsumGenericPlan):
    algorithm = GEMM


_PLAN_CLASSES = {GENERIC: EinsumGenericPlan, DOT: EinsumDotPlan, DIRECT: EinsumDirectProductPlan, GER: EinsumGerPlan,
                 GEMV: EinsumGemvPlan, GEMM: EinsumGemmPlan}


def compile_plan(C_indices: str, A_indices: str, B_indices: str) -> EinsumGenericPlan:
    """Choose the algorithm for ``C[C_indices] <- A[A_indices] * B[B_indices]``."""
    alg = classify(C_indices, A_indices, B_indices)
    return _PLAN_CLASSES[alg](C_indices, A_indices, B_indices)


def einsum(C_indices: str, C, A_indices: str, A, B_indices: str, B, C_prefactor: float = 0.0,
           AB_prefactor: float = 1.0, stream=None, split: int = 0, slice_type: str = "bf16") -> int:
    """C = C_prefactor*C + AB_prefactor * contraction; returns the algorithm that ran."""
    plan = compile_plan(C_indices, A_indices, B_indices)
    plan.execute(C_prefactor, C, AB_prefactor, A, B, stream, split, slice_type)
    return plan.last_algorithm


def contract(C_indices: str, C, A_indices: str, A, B_indices: str, B, C_prefactor: float = 0.0,
             AB_prefactor: float = 1.0, stream=None) -> None:
    """Always the generic contraction kernel (the reference's OnlyUseGenericAlgorithm)."""
    plan = EinsumGenericPlan(C_indices, A_indices, B_indices)
    plan.execute(C_prefactor, C, AB_prefactor, A, B, stream)


def permute(C_prefactor: float, C_indices: str, C: torch.Tensor, A_prefactor: float, A_indices: str, A: torch.Tensor,
            stream=None) -> None:
    """C = C_prefactor*C + A_prefactor*A with axes relabelled (reference
    ``permute(beta, Indices{...}, &C, alpha, Indices{...}, A)``, Permute.hpp:330-407)."""
    if sorted(C_indices) != sorted(A_indices) or len(set(A_indices)) != len(A_indices):
        raise ValueError(f"permute: {C_indices!r} is not a permutation of {A_indices!r}")
    with torch.cuda.stream(stream):
        _permute_one(C_prefactor, C_indices, C, A_prefactor, A_indices, A)


def _permute_one(C_prefactor, C_indices, C, A_prefactor, A_indices, A) -> None:
    stream = None  # the current stream, set by permute()
    dC, finC = _as_device(C, write=True)
    dA, _ = _as_device(A, write=False)
    if dC.dtype == dA.dtype and dC.dtype in _COMPLEX:
        # the parts move independently; a complex C_prefactor or A_prefactor mixes them
        c, a = complex(C_prefactor), complex(A_prefactor)
        Cr, Ci = _parts(dC)
        Ar, Ai = _parts(dA)
        if c.imag != 0.0:
            _complex_scale(c, dC, Cr, Ci)
            beta = 1.0
        else:
            beta = c.real
        for target, parts in ((Cr, [(a.real, Ar), (-a.imag, Ai)]), (Ci, [(a.real, Ai), (a.imag, Ar)])):
            _permute_one(beta, C_indices, target, parts[0][0], A_indices, parts[0][1])
            if parts[1][0] != 0.0:
                _permute_one(1.0, C_indices, target, parts[1][0], A_indices, parts[1][1])
        if finC is not None:
            torch.cuda.current_stream().synchronize()
            finC()
        return
    if dC.dtype != dA.dtype or dC.dtype not in _SUPPORTED:
        raise TypeError("permute: operands must share float64, float32, complex128 or complex64")
    C_prefactor, A_prefactor = _real_prefactor(C_prefactor), _real_prefactor(A_prefactor)
    r = dA.dim()
    if dC.dim() != r or len(A_indices) != r:
        raise RankError("permute: rank mismatch")
    perm = [A_indices.index(ch) for ch in C_indices]
    for i in range(r):
        if dC.shape[i] != dA.shape[perm[i]]:
            raise DimensionError("permute: extents do not match")
    I64 = ctypes.c_int64 * max(r, 1)
    I32 = ctypes.c_int32 * max(r, 1)
    status = _fn("ein_permute", dC.dtype)(r, I64(*dA.shape), I32(*perm), float(A_prefactor), _ptr(dA), I64(*dA.stride()),
                                          float(C_prefactor), _ptr(dC), I64(*dC.stride()), stream_handle(stream))
    check(status, "permute")
    if finC is not None:
        torch.cuda.current_stream().synchronize()
        finC()


def _square_check(A: torch.Tensor, name: str) -> int:
    if A.dim() != 2:
        raise RankError(f"{name}: the matrix must be rank 2")
    if A.shape[0] != A.shape[1]:
        raise DimensionError(f"{name}: the matrix must be square")
    if A.dtype not in _SUPPORTED or A.device.type != "cuda":
        raise TypeError(f"{name}: float64 or float32 CUDA tensors only")
    return int(A.shape[0])


def syev(A: torch.Tensor, W: torch.Tensor, compute_vectors: bool = True, stream=None, check_info: bool = True):
    """Symmetric eigendecomposition on the GPU (reference ``linear_algebra::syev(&A, &W)``,
    LinearAlgebra.hpp:362): the upper triangle of A (n x n) is read; A is overwritten with
    orthonormal eigenvectors as columns (``A[:, j]`` belongs to ``W[j]``) and W with the
    eigenvalues in ascending order. Returns the device ``info`` tensor; with ``check_info``
    it is read (a synchronisation) and a failure to converge raises, as in the reference."""
    with torch.cuda.stream(stream):
        return _syev_body(A, W, compute_vectors, check_info)


def _syev_body(A, W, compute_vectors, check_info):
    stream = None  # the current stream, set by the caller
    n = _square_check(A, "syev")
    if W.dim() != 1:
        raise RankError("syev: W must be rank 1")
    if W.shape[0] != n:
        raise DimensionError("syev: W must have n entries")
    if W.dtype != A.dtype or not W.is_contiguous():
        raise TypeError("syev: W must be a contiguous vector of A's dtype")
    info = torch.zeros((), dtype=torch.int32, device=A.device)
    job = b"V" if compute_vectors else b"N"
    fn = _fn("ein_syev", A.dtype)
    if A.stride(0) == 1 and A.stride(1) >= max(1, n):
        # column-major storage: the upper triangle as the solver sees it
        check(fn(job, b"U", n, _ptr(A), A.stride(1), _ptr(W), _ptr(info), stream_handle(stream)), "syev")
    elif A.stride(1) == 1 and A.stride(0) >= max(1, n):
        # row-major storage is the column-major transpose: its upper triangle is the solver's
        # lower one, and the eigenvector columns come out as rows
        check(fn(job, b"L", n, _ptr(A), A.stride(0), _ptr(W), _ptr(info), stream_handle(stream)), "syev")
        if compute_vectors:
            A.copy_(A.t().contiguous())
    else:
        tmp = A.contiguous()
        syev(tmp, W, compute_vectors, stream, check_info=False)
        A.copy_(tmp)
    if check_info and int(info.item()) != 0:
        raise EinsumsError("syev: the Jacobi iteration did not converge")
    return info


def gesv(A: torch.Tensor, B: torch.Tensor, stream=None, check_info: bool = True):
    """Solve A X = B on the GPU (reference ``linear_algebra::gesv(&A, &B)``,
    LinearAlgebra.hpp:495): A (n x n) is overwritten with its LU factors (unit L below the
    diagonal, U on and above), B (n or n x nrhs) with X. Returns ``(info, ipiv)``: ipiv is
    1-based on the device; info is an int read from the device when ``check_info``
    (0, or i > 0 when U(i,i) is exactly zero, as from LAPACK), else the device tensor."""
    with torch.cuda.stream(stream):
        return _gesv_body(A, B, check_info)


def _gesv_body(A, B, check_info):
    stream = None  # the current stream, set by the caller
    n = _square_check(A, "gesv")
    if B.dim() not in (1, 2) or B.shape[0] != n:
        raise DimensionError("gesv: B must have n rows")
    if B.dtype != A.dtype:
        raise TypeError("gesv: A and B must share a dtype")
    info = torch.zeros((), dtype=torch.int32, device=A.device)
    ipiv = torch.zeros(max(n, 1), dtype=torch.int64, device=A.device)
    Ac = A.t().contiguous().t()  # column-major copy
    B2 = B.reshape(n, 1) if B.dim() == 1 else B
    Bc = B2.t().contiguous().t()
    nrhs = int(B2.shape[1])
    fn = _fn("ein_gesv", A.dtype)
    check(fn(n, nrhs, _ptr(Ac), max(1, n), _ptr(ipiv), _ptr(Bc), max(1, n), _ptr(info), stream_handle(stream)), "gesv")
    A.copy_(Ac)
    B2.copy_(Bc)
    if check_info:
        return int(info.item()), ipiv[:n]
    return info, ipiv[:n]


def gemm(transA: str, transB: str, alpha: float, A: torch.Tensor, B: torch.Tensor, beta: float, C: torch.Tensor,
         stream=None) -> None:
    """Row-major C = alpha*op(A) op(B) + beta*C on 2-D tensors (linear_algebra::gemm)."""
    with torch.cuda.stream(stream):
        return _gemm_body(transA, transB, alpha, A, B, beta, C)


def _gemm_body(transA, transB, alpha, A, B, beta, C):
    stream = None  # the current stream, set by the caller
    for t in (A, B, C):
        if t.dim() != 2 or t.stride(1) != 1:
            raise ValueError("gemm: 2-D row-major tensors with unit column stride required")
    ta, tb = transA.upper() != "N", transB.upper() != "N"
    m = A.shape[1] if ta else A.shape[0]
    k = A.shape[0] if ta else A.shape[1]
    n = B.shape[0] if tb else B.shape[1]
    # row-major C = op(A) op(B)  <=>  column-major C^T = op(B)^T op(A)^T
    status = _fn("ein_gemm", C.dtype)(transB.encode(), transA.encode(), n, m, k, float(alpha), _ptr(B), B.stride(0),
                                      _ptr(A), A.stride(0), float(beta), _ptr(C), C.stride(0), stream_handle(stream))
    check(status, "gemm")


def gemm_split(transA: str, transB: str, alpha: float, A: torch.Tensor, B: torch.Tensor, beta: float, C: torch.Tensor,
               nslices: int = 5, slice_type: str = "bf16", stream=None) -> None:
    """Row-major C = alpha*op(A) op(B) + beta*C in FP64 through the split-precision GEMM."""
    with torch.cuda.stream(stream):
        return _gemm_split_body(transA, transB, alpha, A, B, beta, C, nslices, slice_type)


SLICE_TYPES = {"bf16": _lib.SLICE_BF16, "fp16": _lib.SLICE_FP16, "int8": _lib.SLICE_INT8, "crt": _lib.SLICE_CRT}


def _slice_code(slice_type: str) -> int:
    """ABI code of a split-precision slice type: "bf16", "fp16" (nslices 3..6 slices),
    "int8" (3..6 slices) or "crt" (nslices = 8..16 moduli of the integer modular scheme);
    a "+refine" suffix (e.g. "crt+refine") adds one iterative-refinement correction."""
    base, plus, flag = slice_type.partition("+")
    if base not in SLICE_TYPES or (plus and flag != "refine"):
        raise ValueError(f"slice_type must be one of {sorted(SLICE_TYPES)}, optionally with '+refine', "
                         f"got {slice_type!r}")
    return SLICE_TYPES[base] | (_lib.SPLIT_REFINE if plus else 0)


def _split_count(split) -> int:
    """nslices for the ABI: an int, or "auto" (0: the CRT scheme's automatic moduli count)."""
    return 0 if split == "auto" else int(split)


def crt_moduli(k: int) -> int:
    """Moduli the "crt" split uses with split="auto" at contraction length k: the fewest that
    keep at least 40 bits per element."""
    return lib.ein_split_crt_moduli(int(k))


def crt_bits(nmoduli: int, k: int) -> int:
    """Bits kept per operand element by the "crt" split GEMM with `nmoduli` moduli at
    contraction length k (the relative rounding of every A and B element is 2^-bits)."""
    b = lib.ein_split_crt_bits(int(nmoduli), int(k))
    if b < 0:
        raise ValueError(f"crt_bits: {nmoduli} moduli at k = {k} is not a valid CRT split")
    return b


def _gemm_split_body(transA, transB, alpha, A, B, beta, C, nslices, slice_type):
    stream = None  # the current stream, set by the caller
    for t in (A, B, C):
        if t.dim() != 2 or t.stride(1) != 1 or t.dtype != torch.float64:
            raise ValueError("gemm_split: 2-D row-major float64 tensors with unit column stride required")
    ta, tb = transA.upper() != "N", transB.upper() != "N"
    m = A.shape[1] if ta else A.shape[0]
    k = A.shape[0] if ta else A.shape[1]
    n = B.shape[0] if tb else B.shape[1]
    st = _slice_code(slice_type)
    status = lib.ein_gemm_split_f64(transB.encode(), transA.encode(), n, m, k, float(alpha), _ptr(B), B.stride(0), _ptr(A),
                                    A.stride(0), float(beta), _ptr(C), C.stride(0), _split_count(nslices), st,
                                    stream_handle(stream))
    check(status, "gemm_split")


def fock_jk(g: torch.Tensor, D: torch.Tensor, F: torch.Tensor, alpha_j: float = 2.0, alpha_k: float = -1.0,
            beta: float = 0.0, stream=None) -> None:
    """F = beta*F + alpha_j*J + alpha_k*K with J(a,b) = sum_cd g(a,b,c,d) D(c,d) and
    K(a,c) = sum_bd g(a,b,c,d) D(b,d), both from one pass over g (n^4, contiguous, FP64).
    Equals the reference's pair of calls
    ``einsum(1, "ij", F, alpha_j, "kl", D, "ijkl", g)`` and ``einsum(1, "ij", F, alpha_k, "kl", D, "ikjl", g)``
    (QChemTests.cpp:299-304) after ``F *= beta``; the defaults give G = 2J - K. n <= 128."""
    with torch.cuda.stream(stream):
        n = int(g.shape[0]) if g.dim() == 4 else -1
        if g.dim() != 4 or tuple(g.shape) != (n,) * 4 or not g.is_contiguous():
            raise ValueError("fock_jk: g must be a contiguous n x n x n x n tensor")
        for name, t in (("D", D), ("F", F)):
            if t.dim() != 2 or tuple(t.shape) != (n, n) or t.stride(1) != 1:
                raise ValueError(f"fock_jk: {name} must be n x n with unit column stride")
        for t in (g, D, F):
            if t.dtype != torch.float64 or t.device.type != "cuda":
                raise TypeError("fock_jk: float64 CUDA tensors only")
        status = lib.ein_fock_jk_f64(n, _ptr(g), _ptr(D), D.stride(0), float(alpha_j), float(alpha_k), float(beta),
                                     _ptr(F), F.stride(0), stream_handle(None))
        check(status, "fock_jk")


def gemm_host_c(transA: str, transB: str, alpha: float, A: torch.Tensor, B: torch.Tensor, beta: float, C: torch.Tensor,
                panel_bytes: int = 0, split: int = 0, slice_type: str = "int8", stream=None) -> None:
    """Row-major C = alpha*op(A) op(B) + beta*C with A, B on the GPU and C a host tensor
    (pin it: ``torch.empty(..., pin_memory=True)``), streamed through the GPU in panels of
    about ``panel_bytes``; ``split=s`` runs each panel in split precision (FP64 only).
    Returns when C holds the result."""
    with torch.cuda.stream(stream):
        return _gemm_host_c_body(transA, transB, alpha, A, B, beta, C, panel_bytes, split, slice_type)


def _gemm_host_c_body(transA, transB, alpha, A, B, beta, C, panel_bytes, split, slice_type):
    stream = None  # the current stream, set by the caller
    for t in (A, B):
        if t.dim() != 2 or t.stride(1) != 1 or t.device.type != "cuda":
            raise ValueError("gemm_host_c: A and B must be 2-D row-major CUDA tensors")
    if C.dim() != 2 or C.stride(1) != 1 or C.device.type != "cpu":
        raise ValueError("gemm_host_c: C must be a 2-D row-major host tensor")
    if not (A.dtype == B.dtype == C.dtype) or A.dtype not in _SUPPORTED:
        raise TypeError("gemm_host_c: operands must share float64 or float32")
    ta, tb = transA.upper() != "N", transB.upper() != "N"
    m = A.shape[1] if ta else A.shape[0]
    k = A.shape[0] if ta else A.shape[1]
    n = B.shape[0] if tb else B.shape[1]
    if tuple(C.shape) != (m, n) or (B.shape[1] if tb else B.shape[0]) != k:
        raise DimensionError("gemm_host_c: shapes do not match")
    cptr = ctypes.c_void_p(C.data_ptr())
    # row-major C = op(A) op(B) is column-major C^T = op(B)^T op(A)^T: swap the operands
    if A.dtype == torch.float64:
        st = _slice_code(slice_type)
        status = lib.ein_gemm_host_c_f64(transB.encode(), transA.encode(), n, m, k, float(alpha), _ptr(B), B.stride(0),
                                         _ptr(A), A.stride(0), float(beta), cptr, C.stride(0), int(panel_bytes),
                                         crt_moduli(k) if split == "auto" else int(split), st, stream_handle(stream))
    else:
        if split:
            raise TypeError("gemm_host_c: split precision applies to float64")
        status = lib.ein_gemm_host_c_f32(transB.encode(), transA.encode(), n, m, k, float(alpha), _ptr(B), B.stride(0),
                                         _ptr(A), A.stride(0), float(beta), cptr, C.stride(0), int(panel_bytes),
                                         stream_handle(stream))
    check(status, "gemm_host_c")


_ROUTES = {"auto": _lib.ROUTE_AUTO, "gather": _lib.ROUTE_GATHER, "ttgt": _lib.ROUTE_TTGT}


@contextlib.contextmanager
def contraction_route(route: str):
    """Within the block, generic contractions of GEMM size on this thread take ``route``:
    "auto" (cost model), "gather" (MFMA strided-gather kernel) or "ttgt" (operands packed
    by the permute, then the dense MFMA GEMM). The previous route is restored after."""
    if route not in _ROUTES:
        raise ValueError(f"contraction_route: unknown route {route!r}")
    prev = lib.ein_set_contraction_route(_ROUTES[route])
    try:
        yield
    finally:
        lib.ein_set_contraction_route(prev)


@contextlib.contextmanager
def gemm_tuning():
    """Tuning mode for the MFMA GEMM (every GEMM-mapped einsum and ``gemm`` call inside the
    block): a problem not yet in the tuning table is timed over all candidate
    configurations on its own operands and the fastest is stored (``save_gemm_tuning``)."""
    prev = lib.ein_gemm_tuning_mode(1)
    try:
        yield
    finally:
        lib.ein_gemm_tuning_mode(prev)


def save_gemm_tuning(path) -> None:
    check(lib.ein_gemm_tuning_save(str(path).encode()), "save gemm tuning")


def load_gemm_tuning(path) -> None:
    check(lib.ein_gemm_tuning_load(str(path).encode()), "load gemm tuning")


def clear_gemm_tuning() -> None:
    check(lib.ein_gemm_tuning_clear(), "clear gemm tuning")


def gemm_tuning_lookup(elem_bytes: int, m: int, n: int, k: int, batch: int, a_kcontig: bool, b_kcontig: bool):
    """The stored configuration {bm, bn, data_parallel, blocks_per_cu} of a normalised problem
    (C contiguous along m), or None."""
    cfg = (ctypes.c_int32 * 4)()
    if lib.ein_gemm_tuning_lookup(elem_bytes, m, n, k, batch, int(a_kcontig), int(b_kcontig), cfg) != 0:
        return None
    return {"bm": cfg[0], "bn": cfg[1], "data_parallel": cfg[2], "blocks_per_cu": cfg[3]}


def algorithm_name(alg: int) -> str:
    return ALGORITHM_NAMES.get(alg, "?")
