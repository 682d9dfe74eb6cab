"""Block-sparse tensors and grouped execution on the GPU.

``BlockTensor`` mirrors the reference's BlockTensor (libs/Einsums/Tensor/include/Einsums/
Tensor/BlockTensor.hpp:46,119-131): a list of hypercubic blocks, one extent per block.
``TiledTensor`` mirrors TiledTensor (TiledTensor.hpp:56,93-133): per-axis tile extents and a
map from tile coordinates to tiles; absent tiles are zero.

Their einsums follow BlockAlgebra.hpp:23-52 (block i of C from block i of A and B) and
TileAlgebra.hpp:18-93 (C scaled once, then every tile triple of the tile grid accumulates
with prefactor 1). Instead of one library call per block, all the per-block contractions
become one ``GroupedEinsum`` -- a plan executed as persistent grouped launches -- that can be
replayed and captured into a hipGraph.
"""
from __future__ import annotations

import ctypes
import itertools
import weakref
from typing import Iterable, Sequence

import torch

from . import _lib
from ._lib import ContractionItem, check, lib, make_desc
from .core import _ptr, stream_handle


class GroupedEinsum:
    """Many independent contractions ``C_k = cp_k C_k + abp_k * A_k B_k`` as one plan.

    ``terms`` is a sequence of ``(C_labels, C, A_labels, A, B_labels, B, C_prefactor,
    AB_prefactor)`` with device tensors. The result is that of running the terms one after
    another in list order: a term whose output overlaps an earlier term's output or inputs,
    or whose inputs overlap an earlier term's output, runs in a later wave of the plan. The
    plan keeps pointers to the tensors: keep them alive and do not reallocate them while
    the plan is used."""

    def __init__(self, terms: Sequence[tuple], autotune: bool = False, stream=None):
        if not terms:
            raise ValueError("GroupedEinsum needs at least one term")
        dtype = terms[0][1].dtype
        if dtype not in (torch.float64, torch.float32):
            raise TypeError("GroupedEinsum: float64 or float32 tensors")
        items = (ContractionItem * len(terms))()
        self._keep = []
        for k, (ci, C, ai, A, bi, B, cp, abp) in enumerate(terms):
            for t in (C, A, B):
                if t.dtype != dtype or not t.is_cuda:
                    raise TypeError("GroupedEinsum: all operands must be device tensors of one dtype")
            items[k].C = make_desc(ci, C.shape, C.stride())
            items[k].A = make_desc(ai, A.shape, A.stride())
            items[k].B = make_desc(bi, B.shape, B.stride())
            items[k].c, items[k].a, items[k].b = C.data_ptr(), A.data_ptr(), B.data_ptr()
            items[k].c_prefactor, items[k].ab_prefactor = float(cp), float(abp)
            self._keep.extend((C, A, B))
        self._plan = ctypes.c_void_p()
        check(lib.ein_grouped_plan_create(8 if dtype == torch.float64 else 4, items, len(terms),
                                          ctypes.byref(self._plan)), "grouped plan")
        self.dtype = dtype
        self.count = len(terms)
        if autotune:
            self.autotune(stream)

    def execute(self, stream=None) -> None:
        check(lib.ein_grouped_plan_execute(self._plan, stream_handle(stream)), "grouped execute")

    def autotune(self, stream=None) -> None:
        """Time the candidate tile heights on scratch operands and keep the faster."""
        check(lib.ein_grouped_plan_autotune(self._plan, stream_handle(stream)), "grouped autotune")

    def info(self) -> dict:
        rows, waves, tiles = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        check(lib.ein_grouped_plan_info(self._plan, ctypes.byref(rows), ctypes.byref(waves), ctypes.byref(tiles)))
        return {"tile_rows": rows.value, "waves": waves.value, "tiles": tiles.value}

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            lib.ein_grouped_plan_destroy(plan)
            self._plan = ctypes.c_void_p()


class BlockTensor:
    """Block-diagonal tensor: block i is an (n_i,)*rank device tensor."""

    def __init__(self, name: str, block_dims: Sequence[int], rank: int, dtype=torch.float64, device="cuda"):
        self.name = name
        self.rank = rank
        self.block_dims = list(int(n) for n in block_dims)
        self.blocks = [torch.zeros((n,) * rank, dtype=dtype, device=device) for n in self.block_dims]

    def num_blocks(self) -> int:
        return len(self.blocks)

    def __getitem__(self, i: int) -> torch.Tensor:
        return self.blocks[i]

    def dim(self) -> int:
        return sum(self.block_dims)

    def to_dense(self) -> torch.Tensor:
        n = self.dim()
        out = torch.zeros((n,) * self.rank, dtype=self.blocks[0].dtype, device=self.blocks[0].device)
        off = 0
        for b, nb in zip(self.blocks, self.block_dims):
            out[tuple(slice(off, off + nb) for _ in range(self.rank))] = b
            off += nb
        return out


class TiledTensor:
    """Tiled tensor: axis d is cut into tiles of extents tile_dims[d]; ``tiles`` maps a tile
    coordinate tuple to a device tensor of those extents (absent = zero)."""

    def __init__(self, name: str, tile_dims: Sequence[Sequence[int]], dtype=torch.float64, device="cuda"):
        self.name = name
        self.tile_dims = [list(int(x) for x in d) for d in tile_dims]
        self.rank = len(self.tile_dims)
        self.tiles: dict[tuple, torch.Tensor] = {}
        self.dtype, self.device = dtype, device
        self.version = 0  # bumped whenever the set of tiles changes (plan cache key)

    def tile_shape(self, coord: tuple) -> tuple:
        return tuple(self.tile_dims[d][coord[d]] for d in range(self.rank))

    def add_tile(self, coord: tuple, value: torch.Tensor | None = None) -> torch.Tensor:
        shape = self.tile_shape(coord)
        t = torch.zeros(shape, dtype=self.dtype, device=self.device) if value is None else value
        if tuple(t.shape) != shape:
            raise ValueError(f"tile {coord} must have shape {shape}")
        self.tiles[tuple(coord)] = t
        self.version += 1
        return t

    def has_tile(self, coord: tuple) -> bool:
        t = self.tiles.get(tuple(coord))
        return t is not None and t.numel() > 0

    def grid_size(self, d: int) -> int:
        return len(self.tile_dims[d])

    def to_dense(self) -> torch.Tensor:
        offs = [[0] + list(itertools.accumulate(d))[:-1] for d in self.tile_dims]
        out = torch.zeros([sum(d) for d in self.tile_dims], dtype=self.dtype, device=self.device)
        for c, t in self.tiles.items():
            out[tuple(slice(offs[d][c[d]], offs[d][c[d]] + t.shape[d]) for d in range(self.rank))] = t
        return out


def _diag_view(t: torch.Tensor, off: int, n: int) -> torch.Tensor:
    """The sub-block [off, off+n) on every axis of a dense tensor (the reference's
    ``std::apply(T, view_index)`` with every Range equal to one block's range)."""
    return t[tuple(slice(off, off + n) for _ in range(t.dim()))]


def _kind(x) -> str:
    if x is None:  # a scalar C whose partials the caller places
        return "scalar"
    if isinstance(x, BlockTensor):
        return "block"
    if isinstance(x, TiledTensor):
        return "tiled"
    if isinstance(x, torch.Tensor):
        return "scalar" if x.dim() == 0 else "dense"
    raise TypeError(f"blocks.einsum: unsupported operand type {type(x).__name__}")


def block_terms(C_prefactor, C_indices: str, C, AB_prefactor, A_indices: str, A, B_indices: str, B) -> list:
    """Per-block contraction terms of an einsum with BlockTensor operands
    (BlockAlgebra.hpp:23-428). A and B are BlockTensors, or one of them is a dense tensor
    (the reference's vector case) that is cut into the diagonal sub-block of every block's
    range; C is a BlockTensor (block i of C), a dense tensor (its diagonal sub-block i, the
    rest untouched) or -- handled by ``einsum`` -- a scalar. Blocks of extent 0 are skipped."""
    kinds = (_kind(C), _kind(A), _kind(B))
    blk = A if kinds[1] == "block" else B
    if kinds[1] != "block" and kinds[2] != "block":
        raise TypeError("block_terms: A or B must be a BlockTensor")
    nb = blk.num_blocks()
    for X, k in zip((C, A, B), kinds):
        if k == "block" and X.num_blocks() != nb:
            raise ValueError("block counts differ")  # tensor_compat_error
    offs = [0] + list(itertools.accumulate(blk.block_dims))
    terms = []
    for i in range(nb):
        n = blk.block_dims[i]
        for X, k in zip((C, A, B), kinds):
            if k == "block" and X.block_dims[i] != n:
                raise _lib.DimensionError(f"block {i} extents differ")
        if n == 0:
            continue

        def part(X, k):
            return X[i] if k == "block" else _diag_view(X, offs[i], n)

        terms.append((C_indices, part(C, kinds[0]) if kinds[0] != "scalar" else None, A_indices, part(A, kinds[1]),
                      B_indices, part(B, kinds[2]), C_prefactor, AB_prefactor))
    return terms


def _tile_views(X, kind: str, rank: int):
    """(grid extents per axis, {coord: tile}) of a TiledTensor, or of a BlockTensor seen as a
    tiled tensor whose only tiles are the diagonal ones (BlockTileAlgebra.hpp:18-173)."""
    if kind == "tiled":
        return [X.grid_size(d) for d in range(X.rank)], {c: t for c, t in X.tiles.items() if t.numel() > 0}
    nb = X.num_blocks()
    return [nb] * rank, {(i,) * rank: X[i] for i in range(nb) if X.block_dims[i] > 0}


def tiled_terms(C_prefactor, C_indices: str, C, AB_prefactor, A_indices: str, A, B_indices: str, B) -> tuple[list, list]:
    """Tile contraction terms of an einsum over TiledTensors, or over a mix of TiledTensors
    and BlockTensors (TileAlgebra.hpp:18-158, BlockTileAlgebra.hpp:18-173). The tile grid of
    every distinct label is walked; a triple whose A or B tile is absent or empty is skipped,
    and a TiledTensor C gets a zero tile where a triple lands on an absent one (the
    reference's ``C->tile(index)``). Returns (terms accumulating with prefactor 1, C tiles to
    scale by C_prefactor first); a scalar C is handled by ``einsum``."""
    kinds = (_kind(C), _kind(A), _kind(B))
    views = {}
    for name, X, k, idx in (("C", C, kinds[0], C_indices), ("A", A, kinds[1], A_indices), ("B", B, kinds[2], B_indices)):
        if k in ("tiled", "block"):
            views[name] = _tile_views(X, k, len(idx))
    labels = list(dict.fromkeys(C_indices + A_indices + B_indices))
    grid = {}
    for name, idx in (("C", C_indices), ("A", A_indices), ("B", B_indices)):
        if name not in views:
            continue
        for d, ch in enumerate(idx):
            n = views[name][0][d]
            if grid.setdefault(ch, n) != n:
                raise _lib.DimensionError(f"tile grids disagree on index {ch!r}")
    scale = list(C.tiles.values()) if kinds[0] == "tiled" else (C.blocks if kinds[0] == "block" else [])
    # Join of the present A and B tiles on their shared labels (instead of a walk over the
    # whole tile grid: with irrep-blocked tensors most grid positions have no tile), then the
    # labels found only in C; the triples are ordered as the grid walk of TileAlgebra.hpp:59-92
    # would meet them (lexicographic over the distinct labels), so the accumulation order into
    # every C tile is the reference's.
    shared = [ch for ch in dict.fromkeys(A_indices) if ch in B_indices]
    c_only = [ch for ch in dict.fromkeys(C_indices) if ch not in A_indices and ch not in B_indices]
    by_key = {}
    for bc, tb in views["B"][1].items():
        by_key.setdefault(tuple(bc[B_indices.index(ch)] for ch in shared), []).append((bc, tb))
    found = []
    for ac, ta in views["A"][1].items():
        if any(ac[d] != ac[A_indices.index(ch)] for d, ch in enumerate(A_indices)):
            continue  # a repeated label takes one tile index
        for bc, tb in by_key.get(tuple(ac[A_indices.index(ch)] for ch in shared), ()):
            if any(bc[d] != bc[B_indices.index(ch)] for d, ch in enumerate(B_indices)):
                continue
            pos = {ch: ac[A_indices.index(ch)] for ch in A_indices}
            pos.update({ch: bc[B_indices.index(ch)] for ch in B_indices})
            for extra in itertools.product(*[range(grid[ch]) for ch in c_only]):
                pos.update(zip(c_only, extra))
                found.append((tuple(pos[ch] for ch in labels), ta, tb))
    found.sort(key=lambda f: f[0])
    terms = []
    for combo, ta, tb in found:
        pos = dict(zip(labels, combo))
        tc = None
        if kinds[0] != "scalar":
            cc = tuple(pos[ch] for ch in C_indices)
            tc = views["C"][1].get(cc)
            if tc is None:
                if kinds[0] == "block":
                    continue  # off-diagonal position of a block C: no such block
                if any(e == 0 for e in C.tile_shape(cc)):
                    continue
                tc = C.add_tile(cc)  # zero tile, created where a contribution lands
                views["C"][1][cc] = tc  # later triples on this position use the same tile
        terms.append((C_indices, tc, A_indices, ta, B_indices, tb, 1.0, AB_prefactor))
    return terms, scale


def _scalar_einsum(C_prefactor, C: torch.Tensor, terms: list, autotune, stream) -> GroupedEinsum | None:
    """C = C_prefactor*C + sum over the terms' full contractions: each term writes its own
    partial, then the partials are added in term order by the deterministic dot (the
    reference's scalar-C block/tile dispatch, BlockAlgebra.hpp:105-180, TileAlgebra.hpp:95-158:
    C zeroed when C_prefactor == 0, scaled otherwise). The returned plan recomputes the
    partials only; capture the whole call in a ``Graph`` to replay the sum as well."""
    nt = len(terms)
    partial = torch.zeros(max(nt, 1), dtype=C.dtype, device=C.device)
    ones = torch.ones_like(partial)
    if stream is not None:  # allocated on the current stream, used on `stream`
        partial.record_stream(stream)
        ones.record_stream(stream)
    plan = None
    if nt:
        plan = GroupedEinsum([(t[0], partial[k], t[2], t[3], t[4], t[5], 0.0, t[7]) for k, t in enumerate(terms)],
                             autotune, stream)
        plan.execute(stream)
        plan._keep.extend((partial, ones))
    from .core import _fn
    check(_fn("ein_dot", C.dtype)(nt, 1.0, _ptr(partial), 1, _ptr(ones), 1, float(C_prefactor), _ptr(C),
                                  stream_handle(stream)), "block scalar sum")
    return plan


# Plans of earlier block/tiled einsums, reused while the same operand objects with the same
# tile sets come back (the reference rebuilds its tile loop on every call).
_PLAN_CACHE: dict = {}
_PLAN_CACHE_MAX = 64


def _cache_key(C_prefactor, C_indices, C, AB_prefactor, A_indices, A, B_indices, B, autotune):
    def ident(X):
        if isinstance(X, TiledTensor):
            return ("tiled", id(X), X.version)
        if isinstance(X, BlockTensor):
            return ("block", id(X), tuple(b.data_ptr() for b in X.blocks))
        return ("tensor", id(X), X.data_ptr(), tuple(X.shape), tuple(X.stride()), X.dtype)

    return (float(C_prefactor), C_indices, ident(C), float(AB_prefactor), A_indices, ident(A), B_indices, ident(B),
            bool(autotune))


def einsum(C_prefactor, C_indices: str, C, AB_prefactor, A_indices: str, A, B_indices: str, B,
           stream=None, autotune: bool = False, cache: bool = True) -> GroupedEinsum | None:
    """Block/tiled einsum in one grouped plan; returns the plan so it can be replayed (None
    when nothing is left to contract). Operand kinds, as the reference's special dispatch:
      * BlockTensor x BlockTensor -> BlockTensor, dense tensor or scalar (BlockAlgebra.hpp:23-180);
      * dense tensor x BlockTensor, BlockTensor x dense tensor -> BlockTensor, dense or scalar
        (the dense operand cut by every block's range; BlockAlgebra.hpp:182-428);
      * TiledTensor x TiledTensor -> TiledTensor or scalar (TileAlgebra.hpp:18-158);
      * any mix of BlockTensor and TiledTensor (BlockTileAlgebra.hpp:18-173).
    A scalar C is a 0-dimensional device tensor. With ``cache`` (the default) the triple list
    and its device-resident plan are built once per operand set and reused by later calls
    with the same operand objects, labels, prefactors and tile sets."""
    kinds = (_kind(C), _kind(A), _kind(B))
    if "tiled" not in kinds and kinds[1] != "block" and kinds[2] != "block":
        raise TypeError("blocks.einsum expects BlockTensor or TiledTensor operands")
    if "tiled" in kinds and "dense" in kinds:
        raise TypeError("blocks.einsum: dense operands combine with BlockTensors only")
    key = None
    entry = None
    if cache and kinds[0] != "scalar":
        key = _cache_key(C_prefactor, C_indices, C, AB_prefactor, A_indices, A, B_indices, B, autotune)
        entry = _PLAN_CACHE.get(key)
        if entry is not None and not all(ref() is X for ref, X in zip(entry[0], (C, A, B))):
            entry = None  # an id was reused by another object
    if entry is None:
        if "tiled" in kinds:
            terms, scale = tiled_terms(C_prefactor, C_indices, C, AB_prefactor, A_indices, A, B_indices, B)
        else:
            terms, scale = block_terms(C_prefactor, C_indices, C, AB_prefactor, A_indices, A, B_indices, B), []
        if kinds[0] == "scalar":
            return _scalar_einsum(C_prefactor, C, terms, autotune, stream)
        plan = GroupedEinsum(terms, autotune, stream) if terms else None
        if key is not None:
            if isinstance(C, TiledTensor):  # tiles created by the triple list are part of the key
                key = _cache_key(C_prefactor, C_indices, C, AB_prefactor, A_indices, A, B_indices, B, autotune)
            if len(_PLAN_CACHE) >= _PLAN_CACHE_MAX:
                _PLAN_CACHE.pop(next(iter(_PLAN_CACHE)))
            _PLAN_CACHE[key] = ((weakref.ref(C), weakref.ref(A), weakref.ref(B)), plan, bool(scale))
    else:
        plan = entry[1]
        # every C tile or block present now, the ones the first call created included
        scale = (list(C.tiles.values()) if isinstance(C, TiledTensor) else list(C.blocks)) if entry[2] else []
    if scale:
        from .core import _fn
        for t in scale:  # C scaled once by C_prefactor, then every tile triple accumulates
            d = make_desc("".join(chr(97 + i) for i in range(t.dim())), t.shape, t.stride())
            check(_fn("ein_tensor_scale", t.dtype)(ctypes.byref(d), _ptr(t), float(C_prefactor), stream_handle(stream)))
    if plan is not None:
        plan.execute(stream)
    return plan


class Graph:
    """hipGraph of a sequence of backend calls on one stream: ``g = Graph(fn, stream)``
    captures ``fn()`` once; ``g.replay()`` relaunches it. The backend's per-stream
    workspace must already be large enough (run ``fn()`` once before capturing)."""

    def __init__(self, fn, stream: torch.cuda.Stream):
        self.stream = stream
        h = stream_handle(stream)
        self._exec = ctypes.c_void_p()
        with torch.cuda.stream(stream):
            check(lib.ein_graph_begin(h), "graph begin")
            try:
                fn()
            finally:
                st = lib.ein_graph_end(h, ctypes.byref(self._exec))
        check(st, "graph end")

    def replay(self) -> None:
        check(lib.ein_graph_launch(self._exec, stream_handle(self.stream)), "graph launch")

    def __del__(self):
        if getattr(self, "_exec", None) is not None and self._exec.value:
            lib.ein_graph_destroy(self._exec)
            self._exec = ctypes.c_void_p()
