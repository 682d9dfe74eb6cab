"""einsums_amd -- the Einsums ``tensor_algebra::einsum`` path on MI355X (gfx950).

The compute path is libeinsums_amd.so (hand-written HIP kernels behind a C ABI, see
include/einsums_amd.h). This package is the Python surface, shaped like EinsumsPy's
``einsums.core`` (compile_plan / Einsum*Plan.execute / GPUView), plus a
torch.autograd layer in ``einsums_amd.autograd`` and a planner for contractions over more
than two operands in ``einsums_amd.chain``.
"""
from . import _lib
from ._lib import version

__all__ = ["core", "version"]


def __getattr__(name):
    # core imports torch; keep `import einsums_amd` light for symbol checks
    if name in ("core", "autograd", "blocks", "chain"):
        import importlib

        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
