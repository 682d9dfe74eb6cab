"""CPU checks of the disk tensors (HDF5 datasets read and written through the backend's
C ABI, no GPU needed): create / open / put / get round trips, the file is HDF5, and
mismatched extents, element types, ranks and missing datasets are refused."""
import ctypes

import numpy as np
import pytest

disk = pytest.importorskip("einsums_amd.disk")
from einsums_amd import _lib


def test_round_trip_and_reopen(tmp_path):
    p = tmp_path / "t.h5"
    t = disk.DiskTensor(p, "Wabef", (4, 3, 5, 2))
    assert not t.get().any()  # a new dataset reads as zeros
    a = np.random.default_rng(1).standard_normal((4, 3, 5, 2))
    t.put(a)
    assert np.array_equal(disk.DiskTensor(p, "Wabef", (4, 3, 5, 2)).get(), a)  # bit-exact
    f = disk.DiskTensor(p, "v32", (7,), dtype=np.float32)
    f.put(np.arange(7) / 3)
    assert np.array_equal(f.get(), (np.arange(7) / 3).astype(np.float32))
    assert np.array_equal(t.get(), a)  # a second dataset leaves the first alone
    assert p.read_bytes()[:8] == b"\x89HDF\r\n\x1a\n"  # HDF5 file signature


def test_large_leading_slices_are_chunked(tmp_path):
    # one leading slice of 48 x 48 x 1000 doubles is 18 MB: its chunk is cut below it
    t = disk.DiskTensor(tmp_path / "c.h5", "x", (3, 48, 48, 1000))
    a = np.random.default_rng(2).standard_normal(t.shape)
    t.put(a)
    assert np.array_equal(t.get(), a)


def test_small_slices_and_rank_one(tmp_path):
    # slices of a few bytes are grouped into larger chunks; the data round-trips either way
    p = tmp_path / "s.h5"
    a = np.random.default_rng(3).standard_normal((200_000, 2))
    t = disk.DiskTensor(p, "pairs", a.shape)
    t.put(a)
    assert np.array_equal(t.get(), a)
    v = np.arange(300_001, dtype=np.float64)
    r = disk.DiskTensor(p, "vec", v.shape)
    r.put(v)
    assert np.array_equal(r.get(), v)
    assert p.stat().st_size < 2 * (a.nbytes + v.nbytes)  # no per-element chunk overhead


def test_empty_and_single_element(tmp_path):
    p = tmp_path / "e.h5"
    t = disk.DiskTensor(p, "e", (0, 5))
    t.put(np.zeros((0, 5)))
    assert disk.DiskTensor(p, "e", (0, 5)).get().shape == (0, 5)
    one = disk.DiskTensor(p, "one", (1,))
    one.put(np.array([3.5]))
    assert one.get().tolist() == [3.5]


def test_refusals(tmp_path):
    p = tmp_path / "r.h5"
    disk.DiskTensor(p, "x", (4, 6))
    with pytest.raises(_lib.DimensionError):
        disk.DiskTensor(p, "x", (4, 7))
    with pytest.raises(_lib.RankError):
        disk.DiskTensor(p, "x", (4, 6, 1))
    with pytest.raises(ValueError):
        disk.DiskTensor(p, "x", (4, 6), dtype=np.float32)
    with pytest.raises(_lib.DimensionError):
        disk.DiskTensor(p, "x", (4, 6)).put(np.zeros((6, 4)))
    with pytest.raises(ValueError):
        disk.DiskTensor(p, "x", ())
    dims = (ctypes.c_int64 * 2)(4, 6)
    out = np.empty((4, 6))
    st = _lib.lib.ein_disk_read_f64(str(p).encode(), b"absent", 2, dims, out.ctypes.data_as(ctypes.c_void_p))
    assert st == _lib.EIN_ERROR_VALUE and "absent" in _lib.lib.ein_last_error_string().decode()
    st = _lib.lib.ein_disk_read_f64(str(tmp_path / "none.h5").encode(), b"x", 2, dims, out.ctypes.data_as(ctypes.c_void_p))
    assert st == _lib.EIN_ERROR_VALUE
