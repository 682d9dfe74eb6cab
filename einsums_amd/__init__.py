"""einsums_amd -- the Einsums ``tensor_algebra::einsum`` path on MI355X (gfx950).

The compute path is libeinsums_amd.so (hand-written HIP kernels behind a C ABI, see
include/einsums_amd.h). This package is the Python surface, shaped like EinsumsPy's
``einsums.core`` (compile_plan / Einsum*Plan.execute / GPUView), plus a
torch.autograd layer in ``einsums_amd.autograd``, HDF5 disk tensors in ``einsums_amd.disk``
and a planner for contractions over more than two operands in ``einsums_amd.chain``.
"""
__all__ = ["core", "version"]


def __getattr__(name):
    # submodules load on first use: core imports torch, and _lib binds the built library,
    # which `einsums_amd.build` must be able to replace without loading the old one
    import importlib

    if name in ("core", "autograd", "blocks", "chain", "disk", "_lib"):
        return importlib.import_module(f".{name}", __name__)
    if name == "version":
        return importlib.import_module("._lib", __name__).version
    raise AttributeError(name)
