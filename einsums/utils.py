"""Tensor factories of the reference's ``einsums.utils`` (random_tensor_factory /
tensor_factory), as its Python tests use them: ``array="numpy"`` gives a numpy array,
``array="einsums"`` a host torch tensor. Either kind is staged to the GPU by ``execute``
and written back when it is the output. Complex element types are not provided."""
from __future__ import annotations

import numpy as np
import torch


def _check(dtype):
    dtype = np.dtype(dtype)
    if dtype not in (np.float64, np.float32):
        raise TypeError(f"einsums.utils: element type {dtype} is not provided (float64 and float32 only)")
    return dtype


def tensor_factory(name: str, dims, dtype=np.float64, array: str = "einsums"):
    """A zero tensor of extents `dims`."""
    dtype = _check(dtype)
    a = np.zeros(list(dims), dtype=dtype)
    return torch.from_numpy(a) if array == "einsums" else a


def random_tensor_factory(name: str, dims, dtype=np.float64, array: str = "einsums", seed: int | None = None):
    """A tensor of extents `dims` with entries uniform in [-1, 1)."""
    dtype = _check(dtype)
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, list(dims)).astype(dtype)
    return torch.from_numpy(a) if array == "einsums" else a
