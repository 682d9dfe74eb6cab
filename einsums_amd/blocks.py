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

    def tile_shape(self, coord: tuple) -> tuple:
        return tuple(self.tile_dims[d][coord[d]] for d in range(self.rank))

    def add_tile(self, coord: tuple, value: torch.Tensor | None = None) -> torch.Tensor:
        shape = self.tile_shape(coord)
        t = torch.zeros(shape, dtype=self.dtype, device=self.device) if value is None else value
        if tuple(t.shape) != shape:
            raise ValueError(f"tile {coord} must have shape {shape}")
        self.tiles[tuple(coord)] = t
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


def block_terms(C_prefactor, C_indices: str, C: BlockTensor, AB_prefactor, A_indices: str, A: BlockTensor,
                B_indices: str, B: BlockTensor) -> list:
    """Per-block contraction terms of a BlockTensor einsum (BlockAlgebra.hpp:23-52)."""
    if not (C.num_blocks() == A.num_blocks() == B.num_blocks()):
        raise ValueError("block counts differ")
    terms = []
    for i in range(C.num_blocks()):
        if C.block_dims[i] != A.block_dims[i] or C.block_dims[i] != B.block_dims[i]:
            raise _lib.DimensionError(f"block {i} extents differ")
        if C.block_dims[i] == 0:
            continue
        terms.append((C_indices, C[i], A_indices, A[i], B_indices, B[i], C_prefactor, AB_prefactor))
    return terms


def tiled_terms(C_prefactor, C_indices: str, C: TiledTensor, AB_prefactor, A_indices: str, A: TiledTensor,
                B_indices: str, B: TiledTensor) -> tuple[list, list]:
    """Tile contraction terms of a TiledTensor einsum (TileAlgebra.hpp:18-93): the tile grid of
    every distinct label is walked; triples with a missing or empty tile are skipped. Returns
    (terms accumulating with prefactor 1, C tiles to scale by C_prefactor first)."""
    labels = list(dict.fromkeys(C_indices + A_indices + B_indices))
    grid = {}
    for T, idx in ((C, C_indices), (A, A_indices), (B, B_indices)):
        for d, ch in enumerate(idx):
            n = T.grid_size(d)
            if grid.setdefault(ch, n) != n:
                raise _lib.DimensionError(f"tile grids disagree on index {ch!r}")
    scale = list(C.tiles.values())
    terms = []
    for combo in itertools.product(*[range(grid[ch]) for ch in labels]):
        pos = dict(zip(labels, combo))
        cc = tuple(pos[ch] for ch in C_indices)
        ac = tuple(pos[ch] for ch in A_indices)
        bc = tuple(pos[ch] for ch in B_indices)
        if not (C.has_tile(cc) and A.has_tile(ac) and B.has_tile(bc)):
            continue
        terms.append((C_indices, C.tiles[cc], A_indices, A.tiles[ac], B_indices, B.tiles[bc], 1.0, AB_prefactor))
    return terms, scale


def einsum(C_prefactor, C_indices: str, C, AB_prefactor, A_indices: str, A, B_indices: str, B,
           stream=None, autotune: bool = False) -> GroupedEinsum:
    """Block/tiled einsum in one grouped plan; returns the plan so it can be replayed."""
    if isinstance(C, BlockTensor):
        plan = GroupedEinsum(block_terms(C_prefactor, C_indices, C, AB_prefactor, A_indices, A, B_indices, B),
                             autotune, stream)
        plan.execute(stream)
        return plan
    if isinstance(C, TiledTensor):
        terms, scale = tiled_terms(C_prefactor, C_indices, C, AB_prefactor, A_indices, A, B_indices, B)
        from .core import _fn
        for t in scale:  # C scaled once by C_prefactor, then every tile triple accumulates
            d = make_desc("".join(chr(97 + i) for i in range(t.dim())), t.shape, t.stride())
            check(_fn("ein_tensor_scale", t.dtype)(ctypes.byref(d), _ptr(t), float(C_prefactor), stream_handle(stream)))
        if not terms:
            return None
        plan = GroupedEinsum(terms, autotune, stream)
        plan.execute(stream)
        return plan
    raise TypeError("blocks.einsum expects BlockTensor or TiledTensor operands")


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
