"""The ``einsums`` package name on the CPU: import, plan classes chosen by the pattern
analysis (no device work), the utils factories and GPUView's mode checks."""
import numpy as np
import pytest


def test_import_and_plan_classes():
    import einsums as ein

    expect = {("", "i", "i"): "EinsumDotPlan", ("ij", "ij", "ij"): "EinsumDirectProductPlan",
              ("ij", "ik", "kj"): "EinsumGemmPlan", ("i", "ij", "j"): "EinsumGemvPlan",
              ("i", "j", "ij"): "EinsumGemvPlan", ("ij", "i", "j"): "EinsumGerPlan",
              ("ij", "ji", "ij"): "EinsumGenericPlan"}
    for (c, a, b), cls in expect.items():
        assert type(ein.core.compile_plan(c, a, b)) is getattr(ein.core, cls)


def test_utils_factories():
    import torch

    import einsums as ein

    a = ein.utils.random_tensor_factory("A", [3, 4], np.float64, "numpy", seed=1)
    t = ein.utils.random_tensor_factory("A", [3, 4], np.float64, "einsums", seed=1)
    assert isinstance(a, np.ndarray) and isinstance(t, torch.Tensor) and np.array_equal(a, t.numpy())
    assert np.all(ein.utils.tensor_factory("C", [2, 2], np.float32, "numpy") == 0)
    with pytest.raises(TypeError):
        ein.utils.tensor_factory("C", [2], np.complex128, "numpy")


def test_gpuview_rejects_device_tensor_mode_for_host_buffers():
    import einsums as ein

    with pytest.raises(ValueError):
        ein.core.GPUView(np.zeros(3), ein.core.DEVICE_TENSOR)
    with pytest.raises(ValueError):
        ein.core.GPUView(np.zeros(3), 7)
