"""Tensors kept on disk as HDF5 datasets, and the GEMM whose output C lives in one.

``DiskTensor(path, name, dims)`` follows the reference's DiskTensor
(libs/Einsums/Tensor/include/Einsums/Tensor/DiskTensor.hpp:200-306): one chunked dataset per
tensor, its dims in row-major order, created when absent and opened (extents and element
type checked) when present; ``get()`` reads it whole and ``put()`` writes it whole. The
reference takes an open HDF5 file handle; this one takes the file's path. HDF5 comes from
the libhdf5 the backend loads at first use (see include/einsums_amd.h); no Python HDF5
package is needed.

``gemm_disk_c`` is ``core.gemm_host_c`` with C a DiskTensor: the out-of-core case of the
v^4 CCD intermediates (Wabef += g_mnef t_mnab with Wabef a (v, v, v, v) DiskTensor), streamed
through the GPU in panels of whole leading slices.
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Sequence

import numpy as np
import torch

from ._lib import DimensionError, check, lib
from .core import _SUPPORTED, _ptr, _slice_code, crt_moduli, stream_handle

__all__ = ["DiskTensor", "gemm_disk_c"]

_DTYPES = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}


class DiskTensor:
    """A rank >= 1 FP64 or FP32 tensor stored as dataset ``name`` of the HDF5 file ``path``."""

    def __init__(self, path, name: str, dims: Sequence[int], dtype=np.float64):
        self.path = str(Path(path))
        self.name = str(name)
        self.shape = tuple(int(d) for d in dims)
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DTYPES:
            raise TypeError("DiskTensor: float64 or float32 only")
        if not self.shape:
            raise ValueError("DiskTensor: rank must be at least 1")
        fn = getattr(lib, f"ein_disk_create_{_DTYPES[self.dtype]}")
        check(fn(self.path.encode(), self.name.encode(), len(self.shape), self._dims()), "DiskTensor")

    def _dims(self):
        return (ctypes.c_int64 * len(self.shape))(*self.shape)

    @property
    def rank(self) -> int:
        return len(self.shape)

    def get(self) -> np.ndarray:
        """The whole tensor as a new array."""
        out = np.empty(self.shape, dtype=self.dtype)
        fn = getattr(lib, f"ein_disk_read_{_DTYPES[self.dtype]}")
        check(fn(self.path.encode(), self.name.encode(), self.rank, self._dims(), out.ctypes.data_as(ctypes.c_void_p)),
              "DiskTensor.get")
        return out

    def put(self, data) -> None:
        """Writes ``data`` (this tensor's shape) over the whole tensor."""
        a = np.ascontiguousarray(data, dtype=self.dtype)
        if a.shape != self.shape:
            raise DimensionError(f"DiskTensor.put: shape {a.shape} is not {self.shape}")
        fn = getattr(lib, f"ein_disk_write_{_DTYPES[self.dtype]}")
        check(fn(self.path.encode(), self.name.encode(), self.rank, self._dims(), a.ctypes.data_as(ctypes.c_void_p)),
              "DiskTensor.put")

    def __repr__(self) -> str:
        return f"DiskTensor({self.path!r}, {self.name!r}, {self.shape}, {self.dtype})"


def gemm_disk_c(transA: str, transB: str, alpha: float, A: torch.Tensor, B: torch.Tensor, beta: float, C: DiskTensor,
                panel_bytes: int = 0, split=0, slice_type: str = "crt", stream=None) -> None:
    """Row-major C = alpha*op(A) op(B) + beta*C with A, B 2-D row-major CUDA tensors and C a
    DiskTensor whose leading dims multiply to C's rows (m) and the rest to its columns (n).
    ``split`` > 0 (or "auto" with CRT) runs each panel in split precision (FP64 only).
    Returns when C is on disk."""
    with torch.cuda.stream(stream):
        for t in (A, B):
            if t.dim() != 2 or t.stride(1) != 1 or t.device.type != "cuda":
                raise ValueError("gemm_disk_c: A and B must be 2-D row-major CUDA tensors")
        if not isinstance(C, DiskTensor):
            raise TypeError("gemm_disk_c: C must be a DiskTensor")
        if A.dtype != B.dtype or A.dtype not in _SUPPORTED or np.dtype(str(A.dtype).split(".")[1]) != C.dtype:
            raise TypeError("gemm_disk_c: operands must share float64 or float32 with C")
        ta, tb = transA.upper() != "N", transB.upper() != "N"
        m = A.shape[1] if ta else A.shape[0]
        k = A.shape[0] if ta else A.shape[1]
        n = B.shape[0] if tb else B.shape[1]
        if (B.shape[1] if tb else B.shape[0]) != k or int(np.prod(C.shape)) != m * n:
            raise DimensionError("gemm_disk_c: shapes do not match")
        # row-major C = op(A) op(B) is column-major C^T = op(B)^T op(A)^T: swap the operands
        path, name = C.path.encode(), C.name.encode()
        if A.dtype == torch.float64:
            status = lib.ein_gemm_disk_c_f64(transB.encode(), transA.encode(), n, m, k, float(alpha), _ptr(B),
                                             B.stride(0), _ptr(A), A.stride(0), float(beta), path, name,
                                             int(panel_bytes), crt_moduli(k) if split == "auto" else int(split),
                                             _slice_code(slice_type), stream_handle(None))
        else:
            if split:
                raise TypeError("gemm_disk_c: split precision applies to float64")
            status = lib.ein_gemm_disk_c_f32(transB.encode(), transA.encode(), n, m, k, float(alpha), _ptr(B),
                                             B.stride(0), _ptr(A), A.stride(0), float(beta), path, name,
                                             int(panel_bytes), stream_handle(None))
        check(status, "gemm_disk_c")
