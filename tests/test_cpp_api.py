"""The C++ header API (include/einsums): compile-time dispatch checks (CPU) and the
DeviceTensor einsum/permute program (GPU). The programs are built by
__graft_entry__.build() from tests/cpp/Makefile."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "bin"


def _run(name, timeout, *args):
    exe = BIN / name
    if not exe.exists():
        subprocess.run(["make", "-s", "-C", str(exe.parent.parent)], check=True)
    res = subprocess.run([str(exe), *args], capture_output=True, text=True, timeout=timeout)
    print(res.stdout)
    print(res.stderr)
    return res


def test_compile_time_dispatch_program():
    res = _run("test_compile_time", 60)
    assert res.returncode == 0 and "passed" in res.stdout


def test_streamk_schedule_program():
    """Host check of the stream-K unit ownership, piece slots and fix-up coverage."""
    res = _run("test_streamk_schedule", 120)
    assert res.returncode == 0 and "0 failures" in res.stdout


def test_disk_tensor_program_cpu():
    """DiskTensor create / put / get round trip and refusals (HDF5, no GPU)."""
    res = _run("test_disk_tensor", 60, "cpu")
    assert res.returncode == 0 and "PASS test_disk_tensor cpu" in res.stdout, res.stdout + res.stderr


@pytest.mark.gpu
def test_disk_tensor_program_gpu():
    """linear_algebra::gemm into a (v, v, v, v) DiskTensor in 5 panels, against a
    long-double triple loop."""
    res = _run("test_disk_tensor", 120, "gpu")
    assert res.returncode == 0 and "PASS test_disk_tensor gpu" in res.stdout, res.stdout + res.stderr


@pytest.mark.gpu
def test_complex_program():
    """Complex einsum (GEMM with conjugated operands, generic, dot), permute and scaling on
    DeviceTensor<std::complex<T>> and a host Tensor einsum, against long-double loops."""
    res = _run("test_complex", 120)
    assert res.returncode == 0 and "PASS test_complex" in res.stdout, res.stdout + res.stderr


@pytest.mark.gpu
def test_device_einsum_program():
    res = _run("test_device_einsum", 120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PASS test_device_einsum" in res.stdout


@pytest.mark.gpu
def test_device_blocks_program():
    """BlockDeviceTensor / TiledDeviceTensor einsums through grouped plans."""
    res = _run("test_device_blocks", 120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PASS test_device_blocks" in res.stdout


@pytest.mark.gpu
def test_device_linalg_program():
    """linear_algebra::syev / gesv on row-major DeviceTensors."""
    res = _run("test_device_linalg", 120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PASS test_device_linalg" in res.stdout


def test_lds_layout_program():
    """Host check of the direct-to-LDS slab layouts: exact coverage, LDS positions the MFMA
    fragment reads expect, and conflict-free fragment reads."""
    res = _run("test_lds_layout", 60)
    assert res.returncode == 0 and "0 failures" in res.stdout


def test_pool_book_program():
    """Host check of the pool's reuse conditions (stream events, device synchronisation) and
    of the workspace growth rule (buffers used inside a graph capture are never freed)."""
    res = _run("test_pool_book", 60)
    assert res.returncode == 0 and "0 failures" in res.stdout


@pytest.mark.gpu
def test_api_overloads_program():
    """The reference's einsum spellings (MAKE_INDEX, host Tensor/TensorView, scalar C,
    shared_ptr operands, batched collections, host permute) against the CPU oracle."""
    res = _run("test_api_overloads", 120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PASS test_api_overloads" in res.stdout


@pytest.mark.gpu
def test_block_executor_program():
    """Block / tiled einsums with dense, vector and scalar operands and mixed block/tiled
    operands in C++, against host loop nests."""
    res = _run("test_block_executor", 120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "PASS test_block_executor" in res.stdout
