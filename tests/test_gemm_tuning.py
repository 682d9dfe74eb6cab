"""The GEMM tuning table on the host (no GPU): save, load, lookup, clear, and rejection of
malformed files."""
import pytest


def test_tuning_table_round_trip(tmp_path, core):
    core.clear_gemm_tuning()
    assert core.lib.ein_gemm_tuning_count() == 0
    f = tmp_path / "tuning.txt"
    f.write_text("# header\n8 576 576 9216 1 1 1 : 64 64 0 0\n4 9216 9216 576 1 0 0 : 128 128 1 0\n")
    core.load_gemm_tuning(f)
    assert core.lib.ein_gemm_tuning_count() == 2
    assert core.gemm_tuning_lookup(8, 576, 576, 9216, 1, True, True) == {"bm": 64, "bn": 64, "data_parallel": 0,
                                                                         "blocks_per_cu": 0}
    assert core.gemm_tuning_lookup(4, 9216, 9216, 576, 1, False, False)["data_parallel"] == 1
    assert core.gemm_tuning_lookup(8, 576, 576, 9216, 1, False, True) is None  # other staging class
    g = tmp_path / "saved.txt"
    core.save_gemm_tuning(g)
    core.clear_gemm_tuning()
    assert core.lib.ein_gemm_tuning_count() == 0
    core.load_gemm_tuning(g)
    assert core.lib.ein_gemm_tuning_count() == 2
    assert core.gemm_tuning_lookup(8, 576, 576, 9216, 1, True, True)["bm"] == 64
    core.clear_gemm_tuning()


@pytest.mark.parametrize("bad", ["8 576 576 9216 1 1 1 : 96 64 0 0\n", "8 576 576 9216 1 1 1 64 64 0 0\n",
                                 "3 576 576 9216 1 1 1 : 64 64 0 0\n", "8 -1 576 9216 1 1 1 : 64 64 0 0\n"])
def test_tuning_table_rejects_malformed(tmp_path, core, bad):
    core.clear_gemm_tuning()
    f = tmp_path / "bad.txt"
    f.write_text(bad)
    with pytest.raises(ValueError, match="malformed line 1"):
        core.load_gemm_tuning(f)
    assert core.lib.ein_gemm_tuning_count() == 0  # nothing of a bad file is kept
