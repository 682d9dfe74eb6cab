"""GEMM with C in a disk tensor (an HDF5 dataset), streamed through the GPU in panels of
whole leading slices: the Wabef update of a (v, v, v, v) DiskTensor against the oracle's
einsum, plain 2-D GEMMs against torch's GEMM of the same operands, split precision and
FP32. Panels are made small so that several pass through both buffers, with a short last
one. Tolerances (relative Frobenius): FP64 1e-13, split precision 1e-11, FP32 1e-5."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tests.helpers import dev, rel_fro

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def disk():
    from einsums_amd import disk as d

    return d


def test_wabef_update_on_disk_vs_oracle(disk, oracle, tmp_path):
    # Wabef += 0.25 t_mnab g_mnef with Wabef a (v, v, v, v) disk tensor: row-major
    # C[ab, ef] = op(A) op(B), A = t as (o^2, v^2) transposed, B = g as (o^2, v^2)
    o, v = 4, 12
    rng = np.random.default_rng(11)
    t = rng.uniform(-1, 1, (o, o, v, v))
    g = rng.uniform(-1, 1, (o, o, v, v))
    W0 = rng.uniform(-1, 1, (v, v, v, v))
    W = disk.DiskTensor(tmp_path / "ccd.h5", "Wabef", (v, v, v, v))
    W.put(W0)
    # one leading slice (a) is v^3 doubles = 13 824 B: 30 000 B -> 2 slices a panel, 6 panels
    disk.gemm_disk_c("T", "N", 0.25, dev(t.reshape(o * o, v * v)), dev(g.reshape(o * o, v * v)), 1.0, W,
                     panel_bytes=30_000)
    ref = W0.copy()
    oracle.einsum(1.0, "abef", ref, 0.25, "mnab", t, "mnef", g)
    assert rel_fro(W.get(), ref) < 1e-13


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_gemm_disk_2d_fp64(disk, tmp_path, ta, tb):
    m, n, k = 70, 90, 33
    gen = torch.Generator(device="cpu").manual_seed(3)
    A = (torch.rand((k, m) if ta == "T" else (m, k), generator=gen, dtype=torch.float64) * 2 - 1).cuda()
    B = (torch.rand((n, k) if tb == "T" else (k, n), generator=gen, dtype=torch.float64) * 2 - 1).cuda()
    C = disk.DiskTensor(tmp_path / "g.h5", "C", (m, n))
    C.put(np.full((m, n), 7.0))  # beta = 0: the old values are not read
    disk.gemm_disk_c(ta, tb, -1.5, A, B, 0.0, C, panel_bytes=9 * n * 8 + 1)  # 9 rows a panel, last of 7
    opA = A.t() if ta == "T" else A
    opB = B.t() if tb == "T" else B
    ref = (-1.5 * (opA.contiguous() @ opB.contiguous())).cpu().numpy()
    assert rel_fro(C.get(), ref) < 1e-13


def test_gemm_disk_split_and_fp32(disk, tmp_path):
    m, n, k = 96, 128, 200
    rng = np.random.default_rng(5)
    A, B, C0 = rng.uniform(-1, 1, (m, k)), rng.uniform(-1, 1, (k, n)), rng.uniform(-1, 1, (m, n))
    C = disk.DiskTensor(tmp_path / "s.h5", "C", (4, 24, 128))  # leading dims 4 x 24 = m
    C.put(C0.reshape(4, 24, 128))
    disk.gemm_disk_c("N", "N", 1.0, dev(A), dev(B), 1.0, C, panel_bytes=24 * 128 * 8, split="auto", slice_type="crt")
    ref = A @ B + C0
    err = rel_fro(C.get().reshape(m, n), ref)
    assert 1e-17 < err < 1e-11  # the split kernel ran (not bit-identical to FP64)
    A32, B32 = A.astype(np.float32), B.astype(np.float32)
    F = disk.DiskTensor(tmp_path / "s.h5", "F", (m, n), dtype=np.float32)
    F.put(C0)
    disk.gemm_disk_c("N", "N", 2.0, dev(A32), dev(B32), 0.5, F, panel_bytes=10_000)
    ref32 = 2.0 * (A32.astype(np.float64) @ B32.astype(np.float64)) + 0.5 * C0.astype(np.float32)
    assert rel_fro(F.get(), ref32) < 1e-5


def test_gemm_disk_k0_and_empty(disk, tmp_path):
    # k = 0: C = beta * C through the panels; m = 0: nothing to do
    C0 = np.random.default_rng(8).uniform(-1, 1, (30, 20))
    C = disk.DiskTensor(tmp_path / "k0.h5", "C", (30, 20))
    C.put(C0)
    A = torch.zeros((30, 0), dtype=torch.float64, device="cuda")
    B = torch.zeros((0, 20), dtype=torch.float64, device="cuda")
    disk.gemm_disk_c("N", "N", 1.0, A, B, -2.0, C, panel_bytes=7 * 20 * 8)
    assert np.array_equal(C.get(), -2.0 * C0)
    E = disk.DiskTensor(tmp_path / "k0.h5", "E", (0, 20))
    disk.gemm_disk_c("N", "N", 1.0, torch.zeros((0, 4), dtype=torch.float64, device="cuda"),
                     torch.zeros((4, 20), dtype=torch.float64, device="cuda"), 0.0, E)
    assert E.get().shape == (0, 20)


def test_gemm_disk_refusals(disk, core, tmp_path):
    A = torch.rand(4, 5, dtype=torch.float64, device="cuda")
    B = torch.rand(5, 6, dtype=torch.float64, device="cuda")
    C = disk.DiskTensor(tmp_path / "e.h5", "C", (4, 7))
    with pytest.raises(core.DimensionError):
        disk.gemm_disk_c("N", "N", 1.0, A, B, 0.0, C)
    C = disk.DiskTensor(tmp_path / "e.h5", "D", (6, 4))  # 24 elements, but leading 6 != m = 4
    with pytest.raises(core.DimensionError):
        disk.gemm_disk_c("N", "N", 1.0, A, B, 0.0, C)
    with pytest.raises(TypeError):
        disk.gemm_disk_c("N", "N", 1.0, A.float(), B.float(), 0.0, disk.DiskTensor(tmp_path / "e.h5", "E", (4, 6)))
