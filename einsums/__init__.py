"""``import einsums`` for the MI355X backend: the reference's Python package name
(einsums/__init__.py of the reference, which loads the compiled ``core`` module and its
``utils``). ``einsums.core`` is ``einsums_amd.core``: ``compile_plan`` returns the same
plan classes (EinsumDotPlan, EinsumGemmPlan, ...), ``execute(C_prefactor, C, AB_prefactor,
A, B)`` runs the HIP kernels, and ``GPUView`` with ``MAP``/``COPY``/``DEVICE_TENSOR`` wraps
operands. There is no CPU execution: a missing backend library raises ImportError."""
from einsums_amd import core  # noqa: F401

from . import utils  # noqa: E402,F401

__all__ = ["core", "utils"]
