"""The GEMM autotuner on the GPU: tuning mode times the candidate configurations of a new
problem on its own operands and stores the fastest; the tuned call and later calls match
the oracle, repeat bit for bit, and survive a save / clear / load of the table."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu


def test_tuning_mode_on_gemm_and_einsum(core, oracle, tmp_path):
    rng = np.random.default_rng(55)
    core.clear_gemm_tuning()
    # Wmnij-shaped einsum (576 x 576 x 9216 at o=24, v=96 is too slow for the oracle here:
    # o=12, v=48 gives 144 x 144 x 2304) and a plain GEMM with uneven extents
    t = rng.uniform(-1, 1, (12, 12, 48, 48))
    g = rng.uniform(-1, 1, (12, 12, 48, 48))
    W0 = rng.uniform(-1, 1, (12, 12, 12, 12))
    A = rng.uniform(-1, 1, (700, 333))
    B = rng.uniform(-1, 1, (333, 1100))
    dt, dg, dA, dB = dev(t), dev(g), dev(A), dev(B)

    def run():
        W = dev(W0)
        core.einsum("mnij", W, "ijef", dt, "mnef", dg, C_prefactor=1.0, AB_prefactor=0.25)
        C = torch.zeros((700, 1100), dtype=torch.float64, device="cuda")
        core.gemm("N", "N", 1.0, dA, dB, 0.0, C)
        return host(W), host(C)

    with core.gemm_tuning():
        W1, C1 = run()
    assert core.lib.ein_gemm_tuning_count() == 2
    Wr = W0.copy()
    oracle.einsum(1.0, "mnij", Wr, 0.25, "ijef", t, "mnef", g)
    assert rel_fro(W1, Wr) < 1e-13 and rel_fro(C1, A @ B) < 1e-13
    W2, C2 = run()  # the stored configurations, outside tuning mode
    assert np.array_equal(W1, W2) and np.array_equal(C1, C2)
    f = tmp_path / "tuned.txt"
    core.save_gemm_tuning(f)
    core.clear_gemm_tuning()
    core.load_gemm_tuning(f)
    assert core.lib.ein_gemm_tuning_count() == 2
    W3, C3 = run()
    assert np.array_equal(W1, W3) and np.array_equal(C1, C3)
    core.clear_gemm_tuning()


@pytest.mark.parametrize("m,n,k", [(576, 576, 9216), (100, 300, 5000), (144, 144, 2304)])
@pytest.mark.parametrize("tile", [(64, 64), (64, 128), (128, 128)])
@pytest.mark.parametrize("bpc", [0, 1])
def test_split_k_schedule_from_table(core, oracle, tmp_path, m, n, k, tile, bpc):
    """Schedule 2 (split-K over XCD-local contraction chunks) forced through a loaded table
    entry: matches the oracle and repeats bit for bit; the fix-up adds the chunks in order."""
    rng = np.random.default_rng(m + n + k)
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    C0 = rng.uniform(-1, 1, (m, n))
    dA, dB = dev(A), dev(B)
    core.clear_gemm_tuning()
    with core.gemm_tuning():  # creates the problem's key
        C = dev(C0)
        core.gemm("N", "N", 0.5, dA, dB, -1.0, C)
    f = tmp_path / "t.txt"
    core.save_gemm_tuning(f)
    lines = [ln for ln in f.read_text().splitlines() if ln and not ln.startswith("#")]
    assert len(lines) == 1
    key = lines[0].split(":")[0]
    f.write_text(f"{key}: {tile[0]} {tile[1]} 2 {bpc}\n")
    core.clear_gemm_tuning()
    core.load_gemm_tuning(f)
    ref = C0.copy().ravel()
    # row-major C = A B is column-major C^T = B^T A^T
    oracle.gemm_colmajor("N", "N", n, m, k, 0.5, B.ravel(), n, A.ravel(), k, -1.0, ref, n)
    outs = []
    for _ in range(2):
        C = dev(C0)
        core.gemm("N", "N", 0.5, dA, dB, -1.0, C)
        outs.append(host(C))
    core.clear_gemm_tuning()
    assert rel_fro(outs[0].ravel(), ref) < 1e-13
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("sched", [0, 1, 2])
@pytest.mark.parametrize("o,v", [(12, 48), (24, 96), (7, 20)])
def test_96_tile_wmnij_from_table(core, oracle, tmp_path, sched, o, v):
    """The 96 x 96 tile (both operands contiguous along k) forced through a loaded table entry
    on the Wmnij pattern mnij <- ijef, mnef: matches the oracle, repeats bit for bit."""
    rng = np.random.default_rng(o * 100 + v)
    t = rng.uniform(-1, 1, (o, o, v, v))
    g = rng.uniform(-1, 1, (o, o, v, v))
    W0 = rng.uniform(-1, 1, (o, o, o, o))
    dt, dg = dev(t), dev(g)
    core.clear_gemm_tuning()
    with core.gemm_tuning():
        W = dev(W0)
        core.einsum("mnij", W, "ijef", dt, "mnef", dg, C_prefactor=1.0, AB_prefactor=0.25)
    f = tmp_path / "t.txt"
    core.save_gemm_tuning(f)
    lines = [ln for ln in f.read_text().splitlines() if ln and not ln.startswith("#")]
    assert len(lines) == 1
    f.write_text(lines[0].split(":")[0] + f": 96 96 {sched} 0\n")
    core.clear_gemm_tuning()
    core.load_gemm_tuning(f)
    outs = []
    for _ in range(2):
        W = dev(W0)
        core.einsum("mnij", W, "ijef", dt, "mnef", dg, C_prefactor=1.0, AB_prefactor=0.25)
        outs.append(host(W))
    core.clear_gemm_tuning()
    ref = W0.copy()
    oracle.einsum(1.0, "mnij", ref, 0.25, "ijef", t, "mnef", g)
    assert rel_fro(outs[0], ref) < 1e-13
    assert np.array_equal(outs[0], outs[1])
