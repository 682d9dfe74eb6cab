"""``import einsums as ein`` used the way the reference's Python tests use it
(tests/test_dot_product.py, test_direct_product.py, test_mat_product.py,
test_mat_vec_product.py, test_outer_product.py, test_generic.py of the reference): plan
classes from ``ein.core.compile_plan``, ``execute`` on numpy arrays and on the package's
own ("einsums") tensors, lists of operands, and ``GPUView`` in COPY, MAP and DEVICE_TENSOR
modes. Every result is checked against the CPU oracle (float64 and float32; the
reference's complex128 cases have no counterpart here)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import einsums as ein
from tests.helpers import rel_fro

pytestmark = pytest.mark.gpu

# (C, A, B, plan class, A dims, B dims, C dims) for extents a, b
CASES = [
    ("", "i", "i", "EinsumDotPlan", ("a",), ("a",), ()),
    ("ij", "ij", "ij", "EinsumDirectProductPlan", ("a", "b"), ("a", "b"), ("a", "b")),
    ("ij", "ik", "kj", "EinsumGemmPlan", ("a", "b"), ("b", "a"), ("a", "a")),
    ("i", "ij", "j", "EinsumGemvPlan", ("a", "b"), ("b",), ("a",)),
    ("i", "j", "ij", "EinsumGemvPlan", ("b",), ("a", "b"), ("a",)),
    ("ij", "i", "j", "EinsumGerPlan", ("a",), ("b",), ("a", "b")),
    ("ij", "j", "i", "EinsumGerPlan", ("b",), ("a",), ("a", "b")),
    ("ij", "ji", "ij", "EinsumGenericPlan", ("b", "a"), ("a", "b"), ("a", "b")),
]


def _dims(spec, a, b):
    return [{"a": a, "b": b}[x] for x in spec]


def _oracle(oracle, cp, Ci, C, abp, Ai, A, Bi, B):
    ref = np.array(C, dtype=np.float64, copy=True)
    oracle.einsum(cp, Ci, ref, abp, Ai, np.asarray(A, dtype=np.float64), Bi, np.asarray(B, dtype=np.float64))
    return ref


@pytest.mark.parametrize("array", ["numpy", "einsums"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("a,b", [(10, 10), (11, 13), (300, 257)])
@pytest.mark.parametrize("case", CASES, ids=[f"{c[0] or 'scalar'}={c[1]},{c[2]}" for c in CASES])
def test_plan_execute(oracle, case, a, b, dtype, array):
    Ci, Ai, Bi, cls, ad, bd, cd = case
    A = ein.utils.random_tensor_factory("A", _dims(ad, a, b), dtype, array, seed=a * 7 + b)
    B = ein.utils.random_tensor_factory("B", _dims(bd, a, b), dtype, array, seed=a * 11 + b)
    C = ein.utils.tensor_factory("C", _dims(cd, a, b) or [], dtype, array)
    plan = ein.core.compile_plan(Ci, Ai, Bi)
    assert type(plan) is getattr(ein.core, cls)
    ref = _oracle(oracle, 0.0, Ci, np.asarray(C), 0.75, Ai, A, Bi, B)
    plan.execute(0.0, C, 0.75, A, B)
    tol = 1e-13 if dtype == np.float64 else 2e-5
    assert rel_fro(np.asarray(C, dtype=np.float64), ref) <= tol


@pytest.mark.parametrize("case", CASES[:4], ids=[c[3] for c in CASES[:4]])
def test_plan_execute_lists(oracle, case):
    Ci, Ai, Bi, cls, ad, bd, cd = case
    a, b = 11, 13
    A = [ein.utils.random_tensor_factory(f"A {i}", _dims(ad, a, b), np.float64, "numpy", seed=i) for i in range(10)]
    B = [ein.utils.random_tensor_factory(f"B {i}", _dims(bd, a, b), np.float64, "einsums", seed=50 + i) for i in range(10)]
    C = [ein.utils.random_tensor_factory(f"C {i}", _dims(cd, a, b) or [], np.float64, "numpy", seed=90 + i)
         for i in range(10)]
    refs = [_oracle(oracle, 0.5, Ci, c, 1.0, Ai, x, Bi, y) for c, x, y in zip(C, A, B)]
    ein.core.compile_plan(Ci, Ai, Bi).execute(0.5, C, 1.0, A, B)
    for c, r in zip(C, refs):
        assert rel_fro(c, r) <= 1e-13
    with pytest.raises(ValueError):
        ein.core.compile_plan(Ci, Ai, Bi).execute(0.5, C, 1.0, A[:-1], B)


@pytest.mark.parametrize("mode", ["COPY", "MAP", "DEVICE_TENSOR"])
def test_gpuview_modes(oracle, mode):
    rng = np.random.default_rng(5)
    A, B = rng.uniform(-1, 1, (64, 48)), rng.uniform(-1, 1, (48, 40))
    C = np.zeros((64, 40))
    ref = _oracle(oracle, 0.0, "ij", C, 1.0, "ik", A, "kj", B)
    if mode == "DEVICE_TENSOR":
        dA, dB, dC = (torch.from_numpy(x).cuda() for x in (A, B, C))
        vA, vB, vC = ein.core.GPUView(dA), ein.core.GPUView(dB), ein.core.GPUView(dC)
        assert vC.mode == ein.core.DEVICE_TENSOR
        ein.core.compile_plan("ij", "ik", "kj").execute(0.0, vC, 1.0, vA, vB)
        vC.update_D2H()  # nothing to do: the view is the device tensor
        assert rel_fro(dC.cpu().numpy(), ref) <= 1e-13
        with pytest.raises(ValueError):
            ein.core.GPUView(C, ein.core.DEVICE_TENSOR)
        return
    m = getattr(ein.core, mode)
    vA, vB, vC = ein.core.GPUView(A, m), ein.core.GPUView(B, m), ein.core.GPUView(C, m)
    ein.core.compile_plan("ij", "ik", "kj").execute(0.0, vC, 1.0, vA, vB)
    if mode == "COPY":
        assert np.all(C == 0)  # host copy untouched until update_D2H
        vC.update_D2H()
    assert rel_fro(C, ref) <= 1e-13


def test_mode_numbering_matches_reference():
    # PyViewMode: MAP, COPY, DEVICE_TENSOR (PyGPUView.hpp:27-46)
    assert (ein.core.MAP, ein.core.COPY, ein.core.DEVICE_TENSOR) == (0, 1, 2)
    assert ein.core.gpu_enabled()
