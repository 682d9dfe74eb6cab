"""ctypes binding of ``libeinsums_amd.so`` (the C ABI declared in include/einsums_amd.h).

Loading the library never touches the GPU; every compute entry point runs HIP kernels
on the caller's stream. There is no CPU fallback: a missing library raises ImportError.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("EINSUMS_AMD_LIB", _PKG / "libeinsums_amd.so"))

EIN_MAX_RANK = 16

EIN_SUCCESS = 0
EIN_ERROR_HIP = 1001
EIN_ERROR_ALLOC = 1002
EIN_ERROR_DIMENSION = 1003
EIN_ERROR_RANK = 1004
EIN_ERROR_UNSUPPORTED = 1005
EIN_ERROR_WORKSPACE = 1006
EIN_ERROR_VALUE = 1007

# detail::AlgorithmChoice numbering (TensorAlgebra/Detail/Utilities.hpp:26-34)
GENERIC, DOT, DIRECT, GER, GEMV, GEMM, INDETERMINATE = range(7)
SLICE_BF16, SLICE_FP16, SLICE_INT8, SLICE_CRT = 0, 1, 2, 3
SPLIT_REFINE = 0x100
ROUTE_AUTO, ROUTE_GATHER, ROUTE_TTGT = 0, 1, 2
ALGORITHM_NAMES = {GENERIC: "GENERIC", DOT: "DOT", DIRECT: "DIRECT", GER: "GER", GEMV: "GEMV", GEMM: "GEMM",
                   INDETERMINATE: "INDETERMINATE"}


class EinsumsError(RuntimeError):
    """A failure reported by the backend."""


class DimensionError(EinsumsError, ValueError):
    """Extents of a shared index label disagree (reference: einsums::dimension_error)."""


class RankError(EinsumsError, ValueError):
    """Rank / label-count mismatch (reference: einsums::rank_error)."""


class WorkspaceError(EinsumsError):
    """Scratch space too small while a stream is being captured into a graph."""


class TensorDesc(ctypes.Structure):
    _fields_ = [
        ("rank", ctypes.c_int32),
        ("labels", ctypes.c_int32 * EIN_MAX_RANK),
        ("dims", ctypes.c_int64 * EIN_MAX_RANK),
        ("strides", ctypes.c_int64 * EIN_MAX_RANK),
    ]


class ContractionItem(ctypes.Structure):
    _fields_ = [
        ("C", TensorDesc), ("A", TensorDesc), ("B", TensorDesc),
        ("c", ctypes.c_void_p), ("a", ctypes.c_void_p), ("b", ctypes.c_void_p),
        ("c_prefactor", ctypes.c_double), ("ab_prefactor", ctypes.c_double),
    ]


def make_desc(labels, dims, strides) -> TensorDesc:
    d = TensorDesc()
    r = len(dims)
    if len(labels) != r:
        raise RankError(f"{len(labels)} index labels given for a rank-{r} tensor")
    if r > EIN_MAX_RANK:
        raise RankError(f"rank {r} above {EIN_MAX_RANK}")
    d.rank = r
    for i in range(r):
        lab = labels[i]
        d.labels[i] = ord(lab) if isinstance(lab, str) else int(lab)
        d.dims[i] = int(dims[i])
        d.strides[i] = int(strides[i])
    return d


_c_int = ctypes.c_int
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_dp = ctypes.c_void_p  # device pointers are passed as void*
_desc_p = ctypes.POINTER(TensorDesc)


def _proto(lib, name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"einsums_amd: {LIB_PATH} is missing; build it with `python -m einsums_amd.build` "
            "(the package has no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    sig = {}
    sig["ein_version"] = (ctypes.c_char_p, [])
    sig["ein_last_error_string"] = (ctypes.c_char_p, [])
    sig["ein_init"] = (_c_int, [])
    sig["ein_malloc"] = (_c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t])
    sig["ein_free"] = (_c_int, [_vp])
    sig["ein_free_async"] = (_c_int, [_vp, _vp])
    sig["ein_workspace_reserve"] = (_c_int, [_vp, ctypes.c_size_t])
    sig["ein_workspace_size"] = (ctypes.c_size_t, [_vp])
    sig["ein_stream_synchronize"] = (_c_int, [_vp])
    sig["ein_graph_begin"] = (_c_int, [_vp])
    sig["ein_graph_end"] = (_c_int, [_vp, ctypes.POINTER(ctypes.c_void_p)])
    sig["ein_graph_launch"] = (_c_int, [_vp, _vp])
    sig["ein_graph_destroy"] = (_c_int, [_vp])
    sig["ein_fock_jk_f64"] = (_c_int, [_i64, _dp, _dp, _i64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                       _i64, _vp])
    sig["ein_gemm_tuning_mode"] = (_c_int, [ctypes.c_int32])
    sig["ein_gemm_tuning_clear"] = (_c_int, [])
    sig["ein_gemm_tuning_count"] = (_i64, [])
    sig["ein_gemm_tuning_lookup"] = (_c_int, [ctypes.c_int32, _i64, _i64, _i64, _i64, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.POINTER(ctypes.c_int32)])
    sig["ein_gemm_tuning_save"] = (_c_int, [ctypes.c_char_p])
    sig["ein_gemm_tuning_load"] = (_c_int, [ctypes.c_char_p])
    sig["ein_dot_f64_noalloc"] = (_c_int, [_i64, ctypes.c_double, _dp, _i64, _dp, _i64, ctypes.c_double, _dp, _vp])
    for sfx, T in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
        sig[f"ein_gemm_{sfx}"] = (_c_int, [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, T, _dp, _i64, _dp, _i64, T,
                                           _dp, _i64, _vp])
        sig[f"ein_gemm_strided_batched_{sfx}"] = (
            _c_int, [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, T, _dp, _i64, _i64, _dp, _i64, _i64, T, _dp, _i64,
                     _i64, _i64, _vp])
        sig[f"ein_gemv_{sfx}"] = (_c_int, [ctypes.c_char, _i64, _i64, T, _dp, _i64, _dp, _i64, T, _dp, _i64, _vp])
        sig[f"ein_dot_{sfx}"] = (_c_int, [_i64, T, _dp, _i64, _dp, _i64, T, _dp, _vp])
        sig[f"ein_ger_{sfx}"] = (_c_int, [_i64, _i64, T, _dp, _i64, _dp, _i64, _dp, _i64, _vp])
        sig[f"ein_dirprod_{sfx}"] = (_c_int, [_i64, T, _dp, _i64, _dp, _i64, T, _dp, _i64, _vp])
        sig[f"ein_scale_{sfx}"] = (_c_int, [_i64, T, _dp, _i64, _vp])
        sig[f"ein_axpby_{sfx}"] = (_c_int, [_i64, T, _dp, _i64, T, _dp, _i64, _vp])
        sig[f"ein_tensor_scale_{sfx}"] = (_c_int, [_desc_p, _dp, T, _vp])
        sig[f"ein_permute_{sfx}"] = (_c_int, [ctypes.c_int32, ctypes.POINTER(_i64), ctypes.POINTER(ctypes.c_int32), T,
                                              _dp, ctypes.POINTER(_i64), T, _dp, ctypes.POINTER(_i64), _vp])
        sig[f"ein_contract_{sfx}"] = (_c_int, [T, _desc_p, _dp, T, _desc_p, _dp, _desc_p, _dp, _vp])
        sig[f"ein_einsum_{sfx}"] = (_c_int, [T, _desc_p, _dp, T, _desc_p, _dp, _desc_p, _dp,
                                             ctypes.POINTER(ctypes.c_int32), _vp])
        sig[f"ein_einsum_adjoint_{sfx}"] = (_c_int, [ctypes.c_int32, T, _desc_p, _dp, _desc_p, _dp, _desc_p, _dp, T,
                                                     _desc_p, _dp, ctypes.POINTER(ctypes.c_int32), _vp])
        sig[f"ein_einsum_execute_{sfx}"] = (_c_int, [ctypes.c_int32, T, _desc_p, _dp, T, _desc_p, _dp, _desc_p, _dp,
                                                     ctypes.POINTER(ctypes.c_int32), _vp])
        sig[f"ein_syev_{sfx}"] = (_c_int, [ctypes.c_char, ctypes.c_char, _i64, _dp, _i64, _dp, _vp, _vp])
        sig[f"ein_gesv_{sfx}"] = (_c_int, [_i64, _i64, _dp, _i64, _vp, _dp, _i64, _vp, _vp])
    sig["ein_gemm_split_f64"] = (_c_int, [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_double, _dp, _i64, _dp,
                                          _i64, ctypes.c_double, _dp, _i64, ctypes.c_int32, ctypes.c_int32, _vp])
    sig["ein_split_crt_bits"] = (_c_int, [ctypes.c_int32, _i64])
    sig["ein_split_crt_moduli"] = (_c_int, [_i64])
    sig["ein_gemm_host_c_f64"] = (_c_int, [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_double, _dp, _i64,
                                           _dp, _i64, ctypes.c_double, _vp, _i64, ctypes.c_size_t, ctypes.c_int32,
                                           ctypes.c_int32, _vp])
    sig["ein_gemm_host_c_f32"] = (_c_int, [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_float, _dp, _i64,
                                           _dp, _i64, ctypes.c_float, _vp, _i64, ctypes.c_size_t, _vp])
    _i64p = ctypes.POINTER(ctypes.c_int64)
    for sfx in ("f64", "f32"):
        sig[f"ein_disk_create_{sfx}"] = (_c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, _i64p])
        sig[f"ein_disk_write_{sfx}"] = (_c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, _i64p, _vp])
        sig[f"ein_disk_read_{sfx}"] = (_c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32, _i64p, _vp])
    sig["ein_gemm_disk_c_f64"] = (_c_int, [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_double, _dp, _i64,
                                           _dp, _i64, ctypes.c_double, ctypes.c_char_p, ctypes.c_char_p,
                                           ctypes.c_size_t, ctypes.c_int32, ctypes.c_int32, _vp])
    sig["ein_gemm_disk_c_f32"] = (_c_int, [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, ctypes.c_float, _dp, _i64,
                                           _dp, _i64, ctypes.c_float, ctypes.c_char_p, ctypes.c_char_p,
                                           ctypes.c_size_t, _vp])
    sig["ein_einsum_split_f64"] = (_c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _desc_p, _dp, ctypes.c_double,
                                            _desc_p, _dp, _desc_p, _dp, ctypes.POINTER(ctypes.c_int32), _vp])
    sig["ein_grouped_plan_create"] = (_c_int, [ctypes.c_int32, ctypes.POINTER(ContractionItem), ctypes.c_int32,
                                               ctypes.POINTER(ctypes.c_void_p)])
    sig["ein_grouped_plan_execute"] = (_c_int, [_vp, _vp])
    sig["ein_grouped_plan_autotune"] = (_c_int, [_vp, _vp])
    sig["ein_grouped_plan_info"] = (_c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                             ctypes.POINTER(ctypes.c_int64)])
    sig["ein_grouped_plan_destroy"] = (_c_int, [_vp])
    sig["ein_set_contraction_route"] = (ctypes.c_int32, [ctypes.c_int32])
    sig["ein_einsum_classify"] = (_c_int, [_desc_p, _desc_p, _desc_p, ctypes.POINTER(ctypes.c_int32)])
    for name, (res, args) in sig.items():
        _proto(lib, name, res, args)
    return lib, sorted(sig)


lib, SYMBOLS = _load()


def check(status: int, what: str = "") -> None:
    """Raise the Python exception matching a backend status code."""
    if status == EIN_SUCCESS:
        return
    msg = lib.ein_last_error_string().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if status < 0:
        raise ValueError(f"{text} (argument {-status})")
    if status == EIN_ERROR_DIMENSION:
        raise DimensionError(text)
    if status == EIN_ERROR_RANK:
        raise RankError(text)
    if status == EIN_ERROR_WORKSPACE:
        raise WorkspaceError(text)
    if status in (EIN_ERROR_VALUE,):
        raise ValueError(text)
    raise EinsumsError(f"{text} (status {status})")


def version() -> str:
    return lib.ein_version().decode()
