"""GPU parity of the two GEMM-size routes of the generic contraction -- the MFMA
strided-gather kernel and transpose-transpose-GEMM (operands that do not fold into a matrix
packed by the permute, then the dense MFMA GEMM) -- against the CPU oracle, each route
forced through ein_set_contraction_route. FP64 tolerance: relative Frobenius 1e-13;
FP32: 1e-5."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro
from tests.test_gpu_einsum import GENERIC, STAGED, run_both, seeded

pytestmark = pytest.mark.gpu

ROUTES = ["gather", "ttgt"]


@pytest.mark.parametrize("pat", GENERIC, ids=lambda p: f"{p[0]}<-{p[1]},{p[2]}")
@pytest.mark.parametrize("route", ROUTES)
@pytest.mark.parametrize("cp", [0.0, -0.5])
def test_routes_vs_oracle(core, oracle, pat, route, cp):
    Ci, Ai, Bi = pat
    rng = seeded(*pat, "route")
    labels = sorted(set(Ci + Ai + Bi))
    hi = 48 if len(labels) <= 4 else 20
    ext = {ch: int(rng.integers(16, hi + 1)) for ch in labels}
    with core.contraction_route(route):
        got, ref, _ = run_both(core, oracle, cp, Ci, Ai, Bi, ext, generic_only=True)
    assert rel_fro(got, ref) < 1e-13


@pytest.mark.parametrize("pat", STAGED, ids=lambda p: f"{p[0]}<-{p[1]},{p[2]}")
@pytest.mark.parametrize("cp", [0.0, 0.75])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ttgt_staged_output_vs_oracle(core, oracle, pat, cp, dtype):
    """Outputs whose M/N digits do not fold are formed in a dense buffer and added into C."""
    Ci, Ai, Bi, ext = pat
    with core.contraction_route("ttgt"):
        got, ref, _ = run_both(core, oracle, cp, Ci, Ai, Bi, ext, dtype=dtype, generic_only=True)
    assert rel_fro(got, ref) < (1e-13 if dtype == np.float64 else 1e-5)


# Only one operand (or only C) needs packing / staging, and a batch (Hadamard) group.
PARTIAL = [
    ("jmi", "ijk", "km", dict(i=48, j=40, k=64, m=72)),      # A and B fold, C does not
    ("ijm", "jik", "km", dict(i=40, j=48, k=64, m=72)),      # A does not fold (i, j swapped)
    ("ijm", "ijk", "mk", dict(i=40, j=48, k=64, m=72)),      # all fold: a plain GEMM through the route
    ("zim", "izk", "zkm", dict(z=5, i=64, k=48, m=56)),      # batch digit, A strided over it
    ("iim", "ik", "km", dict(i=64, k=48, m=40)),              # C diagonal
]


@pytest.mark.parametrize("pat", PARTIAL, ids=lambda p: f"{p[0]}<-{p[1]},{p[2]}")
@pytest.mark.parametrize("cp", [0.0, 0.5])
def test_ttgt_partial_packing_vs_oracle(core, oracle, pat, cp):
    Ci, Ai, Bi, ext = pat
    with core.contraction_route("ttgt"):
        got, ref, _ = run_both(core, oracle, cp, Ci, Ai, Bi, ext, generic_only=True)
    assert rel_fro(got, ref) < 1e-13


def test_auto_route_at_ccd_size_matches_gather(core):
    """Wmbej at o=24, v=96 (the cost model picks the packed route there): the automatic
    route agrees with the gather route to 1e-13, and 64 sampled entries agree with a
    direct sum over the links in numpy."""
    o, v = 24, 96
    rng = np.random.default_rng(11)
    t = rng.standard_normal((o, o, v, v)) * 0.01   # t_jnfb
    g = rng.standard_normal((o, o, v, v)) / np.sqrt(v)  # g_mnef
    W0 = rng.standard_normal((o, v, v, o))         # W_mbej
    out = {}
    for route in ("auto", "gather"):
        dW = dev(W0)
        with core.contraction_route(route):
            core.contract("mbej", dW, "jnfb", dev(t), "mnef", dev(g), C_prefactor=1.0, AB_prefactor=-0.5)
        out[route] = host(dW)
    assert rel_fro(out["auto"], out["gather"]) < 1e-13
    for _ in range(64):
        m, b, e, j = (int(rng.integers(0, x)) for x in (o, v, v, o))
        want = W0[m, b, e, j] - 0.5 * np.sum(t[j, :, :, b] * g[m, :, e, :])
        assert abs(out["auto"][m, b, e, j] - want) <= 1e-13 * max(1.0, abs(want))


@pytest.mark.parametrize("slice_type,ns,tol", [("int8", 6, 1e-11), ("bf16", 5, 1e-11)])
def test_ttgt_split_precision_vs_oracle(core, oracle, slice_type, ns, tol):
    """In split mode the packed route runs its GEMM on the split-precision kernel."""
    Ci, Ai, Bi = "mbej", "jnfb", "mnef"
    ext = dict(m=20, b=40, e=40, j=20, n=20, f=40)
    rng = seeded(Ci, Ai, Bi, "split")
    A = rng.uniform(-1, 1, [ext[c] for c in Ai])
    B = rng.uniform(-1, 1, [ext[c] for c in Bi])
    C0 = rng.uniform(-1, 1, [ext[c] for c in Ci])
    ref = C0.copy()
    oracle.einsum_generic(0.5, Ci, ref, -0.5, Ai, A, Bi, B)
    dC = dev(C0)
    with core.contraction_route("ttgt"):
        core.compile_plan(Ci, Ai, Bi).execute(0.5, dC, -0.5, dev(A), dev(B), split=ns, slice_type=slice_type)
    err = rel_fro(host(dC), ref)
    assert err < tol
    # the split kernel ran: slicing leaves ~1e-13, a plain FP64 GEMM would match to ~1e-16
    assert err > 1e-15
