"""Complex einsums and permutes (complex128 and complex64 operands, real or complex
prefactors), run as real einsums on the real and imaginary parts, against numpy.einsum
in complex128 on the same inputs (relative complex Frobenius error): every dispatch pattern (GEMM, generic, DOT, DIRECT,
GEMV, GER), strided views, C_prefactor = 0 clearing NaN, host arrays through GPUView, and
permutes (bit-exact for alpha = 1, beta = 0). Tolerances (relative Frobenius): complex128
1e-13, complex64 1e-5."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


pytestmark = pytest.mark.gpu

CASES = [  # (C, A, B, extents)
    ("ij", "ik", "kj", dict(i=37, j=29, k=41)),           # GEMM
    ("mbej", "jnfb", "mnef", dict(m=5, b=7, e=6, j=4, n=3, f=8)),  # generic
    ("", "ij", "ij", dict(i=33, j=21)),                   # DOT
    ("ij", "ij", "ij", dict(i=19, j=23)),                 # DIRECT
    ("i", "ij", "j", dict(i=31, j=17)),                   # GEMV
    ("ij", "i", "j", dict(i=13, j=11)),                   # GER
]
PREFACTORS = [(0.0, 1.0), (0.5, -2.0), (0.5 + 0.25j, 1.0 - 0.5j), (1j, 2j)]


def crel(x, ref) -> float:
    """Relative Frobenius error in the complex norm."""
    x, ref = np.ravel(np.asarray(x, dtype=np.complex128)), np.ravel(np.asarray(ref, dtype=np.complex128))
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def crand(rng, shape, dtype):
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(dtype)


def dev(a):
    return torch.from_numpy(np.array(a, copy=True, order="C")).to("cuda")  # keeps 0-d arrays 0-d


@pytest.mark.parametrize("dtype,tol", [(np.complex128, 1e-13), (np.complex64, 1e-5)])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0] or 'scalar'}<-{c[1]},{c[2]}")
@pytest.mark.parametrize("c,a", PREFACTORS)
def test_complex_einsum(core, dtype, tol, case, c, a):
    Ci, Ai, Bi, ext = case
    rng = np.random.default_rng(abs(hash((Ci, Ai, Bi, str(c), str(a), dtype.__name__))) % 2**32)
    shp = lambda idx: tuple(ext[x] for x in idx)
    A, B, C0 = crand(rng, shp(Ai), dtype), crand(rng, shp(Bi), dtype), crand(rng, shp(Ci), dtype)
    plan = core.compile_plan(Ci, Ai, Bi)
    C = dev(C0)
    assert C.dim() == len(Ci)
    if c == 0.0:
        C.fill_(complex("nan+nanj"))  # C_prefactor = 0 clears C
    plan.execute(c, C, a, dev(A), dev(B))
    ref = c * C0.astype(np.complex128) + a * np.einsum(f"{Ai},{Bi}->{Ci}", A.astype(np.complex128),
                                                           B.astype(np.complex128))
    got = C.cpu().numpy()
    assert np.isfinite(got).all()
    # relative to the complex norm: a part can cancel to far below the other (a scalar of
    # -0.008 + 20.5i), so each part's own relative error says nothing about the kernels
    assert crel(got, ref) < tol
    # the parts are stride-2 views: GEMM, DOT and DIRECT keep their kernels; GEMV and GER,
    # whose kernels want a unit-stride axis, run as the generic contraction
    if Ai != "i" and Bi != "j":
        assert plan.last_algorithm == core.classify(Ci, Ai, Bi)


def test_complex_strided_views_and_gpuview(core):
    rng = np.random.default_rng(7)
    A = crand(rng, (40, 30), np.complex128)
    B = crand(rng, (40, 50), np.complex128)
    dA = dev(A).t()  # (30, 40) transposed view
    C = torch.zeros((30, 50), dtype=torch.complex128, device="cuda")
    core.einsum("ij", C, "ik", dA, "kj", dev(B), 0.0, 1.0 + 1.0j)
    ref = (1.0 + 1.0j) * (A.T @ B)
    assert crel(C.cpu().numpy(), ref) < 1e-13
    # host arrays through GPUView (COPY) and a host C written back
    Ch = np.zeros((30, 50), dtype=np.complex128)
    core.einsum("ij", Ch, "ik", core.GPUView(np.ascontiguousarray(A.T)), "kj", B, 0.0, 2.0)
    assert crel(Ch, 2.0 * (A.T @ B)) < 1e-13


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_complex_permute(core, dtype):
    rng = np.random.default_rng(11)
    A = crand(rng, (6, 7, 8), dtype)
    C = torch.empty((8, 6, 7), dtype=torch.complex128 if dtype == np.complex128 else torch.complex64, device="cuda")
    core.permute(0.0, "kij", C, 1.0, "ijk", dev(A))
    assert C.cpu().numpy().tobytes() == np.ascontiguousarray(A.transpose(2, 0, 1)).tobytes()  # bit-exact
    C0 = crand(rng, (8, 6, 7), dtype)
    C = dev(C0)
    core.permute(0.5 - 1j, "kij", C, 2.0 + 0.5j, "ijk", dev(A))
    ref = (0.5 - 1j) * C0.astype(np.complex128) + (2.0 + 0.5j) * A.transpose(2, 0, 1).astype(np.complex128)
    tol = 1e-15 if dtype == np.complex128 else 1e-6
    assert crel(C.cpu().numpy(), ref) < tol


def test_complex_prefactor_with_real_operands_is_refused(core):
    A = torch.rand(4, 4, dtype=torch.float64, device="cuda")
    C = torch.zeros(4, 4, dtype=torch.float64, device="cuda")
    with pytest.raises(TypeError):
        core.einsum("ij", C, "ik", A, "kj", A, 0.0, 1j)
