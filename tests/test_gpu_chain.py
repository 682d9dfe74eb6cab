"""Multi-operand contractions on the GPU: every pairwise step runs through the einsum
dispatch (GEMM / GEMV / DOT / generic kernels) and the chain's result is compared with the
CPU oracle applied pair by pair along the same plan, and with numpy.einsum of the whole
expression. Tolerances (relative Frobenius): FP64 1e-12, FP32 1e-5."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu

CHAINS = [
    ("ik,kl,lj->ij", dict(i=70, k=300, l=20, j=90)),             # matrix chain, GEMMs
    ("ijab,klab,klij->", dict(i=12, j=12, k=12, l=12, a=20, b=20)),   # rank-0 result
    ("bij,bjk,bkl,l->bi", dict(b=6, i=30, j=40, k=50, l=60)),   # a label on four operands
    ("pqrs,rs,pq->", dict(p=20, q=20, r=20, s=20)),             # HF energy: J build + dot
    ("ijx,jk,kl->il", dict(i=40, j=30, x=7, k=50, l=60)),       # x summed where its operand is consumed
    ("mnef,ijef,mnab->ijab", dict(m=8, n=8, i=8, j=8, e=24, f=24, a=24, b=24)),  # CCD-like ladder
]


@pytest.mark.parametrize("subs,sizes", CHAINS, ids=[c[0] for c in CHAINS])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_chain_vs_oracle_and_numpy(core, oracle, subs, sizes, dtype, tol):
    from einsums_amd.chain import compile_chain, run_steps

    ins, out = subs.split("->")
    ins = ins.split(",")
    rng = np.random.default_rng(len(subs) + 31 * len(ins))
    ops = [rng.uniform(-1, 1, [sizes[c] for c in t]).astype(dtype) for t in ins]
    C0 = np.asarray(rng.uniform(-1, 1, [sizes[c] for c in out])).astype(dtype)
    plan = compile_chain(subs, *[o.shape for o in ops])
    dC = dev(C0)
    plan.execute(0.5, dC, -1.5, *[dev(o) for o in ops])
    got = host(dC)

    # the oracle along the same steps, in float64
    def opair(st, x, y):
        z = np.zeros([sizes[c] for c in st.result])
        oracle.einsum_generic(0.0, st.result, z, 1.0, st.left_indices, x, st.right_indices, y)
        return z

    def ofinal(st, x, y):
        z = C0.astype(np.float64).copy()
        oracle.einsum_generic(0.5, st.result, z, -1.5, st.left_indices, x, st.right_indices, y)
        return z

    ref = run_steps(plan.steps, [o.astype(np.float64) for o in ops], opair, ofinal)
    assert rel_fro(got, ref) < tol
    whole = 0.5 * C0.astype(np.float64) - 1.5 * np.einsum(subs, *[o.astype(np.float64) for o in ops])
    assert rel_fro(got, whole) < tol


def test_chain_plan_reuse_and_shape_check(core):
    from einsums_amd.chain import compile_chain

    plan = compile_chain("ik,kl,lj->ij", (8, 9), (9, 10), (10, 11))
    a, b, c = (torch.rand(s, dtype=torch.float64, device="cuda") for s in ((8, 9), (9, 10), (10, 11)))
    C = torch.empty(8, 11, dtype=torch.float64, device="cuda")
    for _ in range(2):
        plan.execute(0.0, C, 1.0, a, b, c)
        assert torch.allclose(C, a @ b @ c, rtol=1e-12, atol=1e-12)
    with pytest.raises(ValueError):
        plan.execute(0.0, C, 1.0, a, b, torch.rand(10, 12, dtype=torch.float64, device="cuda"))
