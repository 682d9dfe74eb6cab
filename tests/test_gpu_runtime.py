"""Runtime hazards on the GPU, each a regression test of the fixed path:

* a graph captured on a stream, then a larger eager call on the same stream that grows the
  stream's scratch buffer, then replays: the replays must equal the eager results bit for
  bit (the buffer the graph points at is retired, not freed);
* grouped plans created, executed and dropped back to back on a side stream: the second
  plan's descriptors must not land in blocks the first plan's queued kernel still reads;
* syev, gesv and an einsum with a host (numpy) output on a side stream: every torch op of
  the call runs on that stream, so the results match the references.
References: the CPU oracle (einsums) and numpy/LAPACK (solvers)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu


def test_graph_survives_workspace_growth(core):
    from einsums_amd import blocks

    rng = np.random.default_rng(41)
    s = torch.cuda.Stream()
    A, B = dev(rng.uniform(-1, 1, (1024, 4096))), dev(rng.uniform(-1, 1, (4096, 1024)))
    x, y = dev(rng.uniform(-1, 1, 3_000_000)), dev(rng.uniform(-1, 1, 3_000_000))
    C = torch.zeros((1024, 1024), dtype=torch.float64, device="cuda")
    E = torch.zeros((), dtype=torch.float64, device="cuda")

    def step():
        core.einsum("ij", C, "ik", A, "kj", B, stream=s)  # GEMM (stream-K partials in the workspace)
        core.einsum("", E, "i", x, "i", y, stream=s)      # DOT (partials in the workspace)

    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        step()
    s.synchronize()
    C_eager, E_eager = C.clone(), E.clone()
    ws_before = int(core.lib.ein_workspace_size(core.stream_handle(s)))
    g = blocks.Graph(step, s)
    # eager calls on the same stream that grow its workspace: an explicit reservation past
    # the current size (torch hands out pooled streams, so this one may already hold a large
    # buffer from earlier work), then a split GEMM that uses the grown buffer
    h = core.stream_handle(s)
    assert core.lib.ein_workspace_reserve(h, 2 * ws_before + (1 << 20)) == 0
    assert int(core.lib.ein_workspace_size(h)) > ws_before
    big = dev(rng.uniform(-1, 1, (2048, 2048)))
    out = torch.empty_like(big)
    core.gemm_split("N", "N", 1.0, big, big, 0.0, out, nslices=5, slice_type="bf16", stream=s)
    s.synchronize()
    # fresh device allocations written with garbage, then replays
    junk = [torch.full((1 << 20,), float("nan"), dtype=torch.float64, device="cuda") for _ in range(8)]
    for _ in range(3):
        C.zero_()
        E.zero_()
        torch.cuda.synchronize()
        g.replay()
        s.synchronize()
        assert torch.equal(C, C_eager) and torch.equal(E, E_eager)
    del junk


def test_dropped_plans_back_to_back_on_side_stream(oracle):
    from einsums_amd import blocks

    rng = np.random.default_rng(42)
    s = torch.cuda.Stream()
    results = []
    for round_ in range(2):
        hA = [rng.uniform(-1, 1, (n, 512)) for n in (768, 640, 512)]
        hB = [rng.uniform(-1, 1, (512, n)) for n in (768, 640, 512)]
        A = [dev(a) for a in hA]
        B = [dev(b) for b in hB]
        C = [torch.zeros((a.shape[0], b.shape[1]), dtype=torch.float64, device="cuda") for a, b in zip(hA, hB)]
        torch.cuda.synchronize()
        plan = blocks.GroupedEinsum([("ij", c, "ik", a, "kj", b, 0.0, 1.0) for c, a, b in zip(C, A, B)])
        plan.execute(s)
        del plan  # destroyed while its kernel may still be queued on s
        results.append((C, hA, hB))
    s.synchronize()
    for C, hA, hB in results:
        for c, a, b in zip(C, hA, hB):
            ref = np.zeros((a.shape[0], b.shape[1]))
            oracle.einsum(0.0, "ij", ref, 1.0, "ik", a, "kj", b)
            assert rel_fro(host(c), ref) < 1e-13


def test_side_stream_solvers_and_host_output(core, oracle):
    rng = np.random.default_rng(43)
    s = torch.cuda.Stream()
    n = 48
    M = rng.standard_normal((n, n))
    S = (M + M.T) / 2
    A = dev(S)
    W = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    core.syev(A, W, stream=s)
    s.synchronize()
    assert np.allclose(host(W), np.linalg.eigvalsh(S), atol=1e-12)
    V = host(A)
    assert rel_fro(S @ V, V * host(W)) < 1e-12

    G = rng.standard_normal((n, n)) + n * np.eye(n)
    rhs = rng.standard_normal((n, 3))
    dG, dR = dev(G), dev(rhs)
    torch.cuda.synchronize()
    info, _ = core.gesv(dG, dR, stream=s)
    s.synchronize()
    assert info == 0 and rel_fro(host(dR), np.linalg.solve(G, rhs)) < 1e-12

    a, b = rng.uniform(-1, 1, (64, 96)), rng.uniform(-1, 1, (96, 80))
    c = np.ones((64, 80))
    ref = c.copy()
    oracle.einsum(0.5, "ij", ref, 2.0, "ik", a, "kj", b)
    core.compile_plan("ij", "ik", "kj").execute(0.5, c, 2.0, a, b, stream=s)
    assert rel_fro(c, ref) < 1e-13
