"""Contractions over more than two operands as a planned sequence of pairwise einsums.

The reference contracts two operands per call (``detail::einsum``,
libs/Einsums/TensorAlgebra/include/Einsums/TensorAlgebra/Detail/Dispatch.hpp:859-957); a
product of three or more tensors is written there as a chain of calls with user-chosen
intermediates. This module chooses the chain. ``compile_chain("ik,kl,lj->ij", shapes)``
fixes the pairwise order and the index order of every intermediate once; ``execute`` then
runs each pair through ``compile_plan`` (so each step takes the GEMM / GEMV / DOT / generic
route the two-operand dispatch picks) with intermediates in device memory.

Order search: every way of contracting the operands is a binary tree; its cost is the sum
over internal nodes of 2 * (product of the extents of every index of the two children),
the multiply-add count of that pairwise contraction. Up to ``optimal_limit`` operands the
cheapest tree is found exactly by dynamic programming over operand subsets (3^n splits);
beyond it, greedily, by contracting the cheapest pair first.

Index bookkeeping: a subset's result keeps the indices that occur outside the subset or in
the output (summed indices disappear as soon as no other operand carries them; an index on
a single operand that is absent from the output is summed in the step that consumes that
operand). An intermediate lists the kept indices of its left child in that child's order,
then the new ones of its right child, so the digits of each side stay adjacent and fold
into one GEMM dimension where the children allow it.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Sequence

__all__ = ["ChainStep", "ContractionChain", "parse_subscripts", "plan_order", "compile_chain", "einsum_chain"]


def parse_subscripts(subscripts: str) -> tuple[list[str], str]:
    """``"ij,jk->ik"`` -> (["ij", "jk"], "ik"). The output part is required: with more than
    two operands an implicit output would be ambiguous about which repeated labels to keep."""
    s = subscripts.replace(" ", "")
    if "->" not in s:
        raise ValueError("einsum_chain: subscripts need an explicit output ('...->out')")
    lhs, out = s.split("->")
    ins = lhs.split(",")
    for t in ins + [out]:
        if not all(c.isalpha() for c in t):
            raise ValueError(f"einsum_chain: index labels must be letters, got {t!r}")
    if len(set(out)) != len(out):
        raise ValueError(f"einsum_chain: the output repeats a label: {out!r}")
    missing = set(out) - set("".join(ins))
    if missing:
        raise ValueError(f"einsum_chain: output labels {sorted(missing)} appear in no operand")
    return ins, out


def _kept(labels: str, outside: set[str]) -> str:
    seen, kept = set(), []
    for c in labels:
        if c in outside and c not in seen:
            seen.add(c)
            kept.append(c)
    return "".join(kept)


def _pair_cost(a: str, b: str, sizes: dict[str, int]) -> int:
    return 2 * math.prod(sizes[c] for c in set(a) | set(b))


@dataclass(frozen=True)
class ChainStep:
    """Contract working operands ``left`` and ``right`` (positions in the working list,
    which holds the inputs first, then each step's result in order) into ``result``."""
    left: int
    right: int
    left_indices: str
    right_indices: str
    result: str
    flops: int


def plan_order(inputs: Sequence[str], output: str, sizes: dict[str, int],
               optimal_limit: int = 10) -> list[ChainStep]:
    """Pairwise steps contracting ``inputs`` into ``output`` (the last step's result is the
    output itself, in the output's order)."""
    n = len(inputs)
    if n < 2:
        raise ValueError("einsum_chain: needs at least two operands")
    for t in inputs:
        for c in t:
            if c not in sizes:
                raise ValueError(f"einsum_chain: no extent for label {c!r}")
    full = (1 << n) - 1
    labels_of = [set(t) for t in inputs]

    def outside(mask: int) -> set[str]:
        s = set(output)
        for i in range(n):
            if not mask >> i & 1:
                s |= labels_of[i]
        return s

    def result_of(left: str, right: str, mask: int) -> str:
        out = outside(mask)
        if mask == full:
            return output
        return _kept(left + right, out)

    if n <= optimal_limit:
        # best[mask] = (cost, indices of the contracted subset, tree)
        best: dict[int, tuple[int, str, object]] = {}
        for i in range(n):
            best[1 << i] = (0, inputs[i], i)
        for mask in range(1, full + 1):
            if mask & (mask - 1) == 0:
                continue
            cand = None
            low = mask & -mask
            sub = (mask - 1) & mask
            while sub:
                if sub & low:  # each split once: the part that holds the lowest operand is on the left
                    rest = mask ^ sub
                    cl, il, tl = best[sub]
                    cr, ir, tr = best[rest]
                    c = cl + cr + _pair_cost(il, ir, sizes)
                    if cand is None or c < cand[0]:
                        cand = (c, result_of(il, ir, mask), (tl, tr, il, ir))
                sub = (sub - 1) & mask
            best[mask] = cand
        steps: list[ChainStep] = []

        def emit(tree) -> int:
            if isinstance(tree, int):
                return tree
            tl, tr, il, ir = tree
            pl, pr = emit(tl), emit(tr)
            mask = _mask_of(tree)
            res = result_of(il, ir, mask)
            steps.append(ChainStep(pl, pr, il, ir, res, _pair_cost(il, ir, sizes)))
            return n + len(steps) - 1

        def _mask_of(tree) -> int:
            if isinstance(tree, int):
                return 1 << tree
            return _mask_of(tree[0]) | _mask_of(tree[1])

        emit(best[full][2])
        return steps

    # greedy: repeatedly contract the pair with the smallest cost (ties: smallest result)
    work = [(1 << i, inputs[i], i) for i in range(n)]
    steps = []
    pos = n
    while len(work) > 1:
        pick = None
        for x in range(len(work)):
            for y in range(x + 1, len(work)):
                mx, ix, _ = work[x]
                my, iy, _ = work[y]
                mask = mx | my
                res = result_of(ix, iy, mask)
                key = (_pair_cost(ix, iy, sizes), math.prod(sizes[c] for c in res))
                if pick is None or key < pick[0]:
                    pick = (key, x, y, res)
        _, x, y, res = pick
        mx, ix, px = work[x]
        my, iy, py = work[y]
        steps.append(ChainStep(px, py, ix, iy, res, _pair_cost(ix, iy, sizes)))
        work = [w for k, w in enumerate(work) if k not in (x, y)] + [(mx | my, res, pos)]
        pos += 1
    return steps


def run_steps(steps: Sequence[ChainStep], operands: Sequence, pair: Callable, final: Callable):
    """Drive a chain: ``pair(step, X, Y)`` returns a new intermediate, ``final(step, X, Y)``
    contracts the last pair into the caller's output."""
    work = list(operands)
    for k, st in enumerate(steps):
        x, y = work[st.left], work[st.right]
        if k == len(steps) - 1:
            return final(st, x, y)
        work.append(pair(st, x, y))
        # drop references the remaining steps no longer need, so intermediates are freed early
        work[st.left] = work[st.right] = None
    return None


class ContractionChain:
    """A planned chain: ``steps`` (ChainStep list), ``flops`` (total multiply-add count x2),
    and ``execute(C_prefactor, C, AB_prefactor, *operands)``:
    C = C_prefactor * C + AB_prefactor * contraction."""

    def __init__(self, inputs: Sequence[str], output: str, sizes: dict[str, int], optimal_limit: int = 10):
        self.inputs = list(inputs)
        self.output = output
        self.sizes = dict(sizes)
        self.steps = plan_order(self.inputs, output, self.sizes, optimal_limit)
        self.flops = sum(s.flops for s in self.steps)
        from . import core

        self._plans = [core.compile_plan(s.result, s.left_indices, s.right_indices) for s in self.steps]
        self.algorithms = [p.algorithm for p in self._plans]

    def __repr__(self) -> str:
        return f"ContractionChain({','.join(self.inputs)}->{self.output}, {len(self.steps)} steps, {self.flops:.3g} flop)"

    def execute(self, C_prefactor, C, AB_prefactor, *operands, stream=None) -> None:
        """Run the chain on ``stream`` (None = torch's current stream). Intermediates are torch
        allocations, so with another stream run the call under ``torch.cuda.stream(stream)``
        to keep the caching allocator's reuse ordered with the kernels."""
        import torch

        if len(operands) != len(self.inputs):
            raise ValueError(f"execute: {len(self.inputs)} operands planned, {len(operands)} given")
        for t, idx in zip(operands, self.inputs):
            if tuple(t.shape) != tuple(self.sizes[c] for c in idx):
                raise ValueError(f"execute: operand {idx!r} has shape {tuple(t.shape)}, planned "
                                 f"{tuple(self.sizes[c] for c in idx)}")
        ref = operands[0]
        plans = iter(self._plans)

        def pair(st: ChainStep, x, y):
            z = torch.empty([self.sizes[c] for c in st.result], dtype=ref.dtype, device=ref.device)
            next(plans).execute(0.0, z, 1.0, x, y, stream)
            return z

        def final(st: ChainStep, x, y):
            next(plans).execute(C_prefactor, C, AB_prefactor, x, y, stream)

        run_steps(self.steps, operands, pair, final)


def compile_chain(subscripts: str, *shapes: Sequence[int], optimal_limit: int = 10) -> ContractionChain:
    """Plan ``subscripts`` (e.g. ``"ijab,klab,klij->"``) for operands of the given shapes."""
    ins, out = parse_subscripts(subscripts)
    if len(shapes) != len(ins):
        raise ValueError(f"compile_chain: {len(ins)} operands in {subscripts!r}, {len(shapes)} shapes given")
    sizes: dict[str, int] = {}
    for idx, shp in zip(ins, shapes):
        if len(idx) != len(shp):
            raise ValueError(f"compile_chain: operand {idx!r} has rank {len(shp)}")
        for c, e in zip(idx, shp):
            if sizes.setdefault(c, int(e)) != int(e):
                raise ValueError(f"compile_chain: label {c!r} has extents {sizes[c]} and {e}")
    return ContractionChain(ins, out, sizes, optimal_limit)


def einsum_chain(subscripts: str, C, *operands, C_prefactor: float = 0.0, AB_prefactor: float = 1.0,
                 stream=None) -> ContractionChain:
    """Plan and run in one call; returns the plan (reuse it to skip the search)."""
    plan = compile_chain(subscripts, *[tuple(t.shape) for t in operands])
    plan.execute(C_prefactor, C, AB_prefactor, *operands, stream=stream)
    return plan
