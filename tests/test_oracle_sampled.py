"""The sampled-row oracle (oracle/sampled.py) against the full oracle and numpy.einsum at
small sizes: every sampled row must equal the matching row of the full contraction, so a
wrong row decode (a swapped pair, a transposed operand) fails here on the CPU."""
import numpy as np
import pytest

from einsums_amd import workloads
from oracle import sampled


@pytest.fixture(scope="module")
def small():
    return workloads.ccd_inputs(5, 7, seed=3, with_vvvv=False)


@pytest.mark.parametrize("name", ["Wmnij", "Wabef", "Wmbej"])
def test_sampled_rows_match_full_oracle(oracle, small, name):
    I = small
    o, v = I.o, I.v
    spec = {c[0]: c for c in workloads.CCD_CONTRACTIONS}[name]
    _, Ci, An, Ai, Bn, Bi, _, abp, _ = spec
    arrs = {"t_oovv": I.t_oovv, "g_oovv": I.g_oovv}
    shape = {"Wmnij": (o, o, o, o), "Wabef": (v, v, v, v), "Wmbej": (o, v, v, o)}[name]
    full = np.zeros(shape)
    oracle.einsum(0.0, Ci, full, abp, Ai, arrs[An], Bi, arrs[Bn])
    ref_np = abp * np.einsum(f"{Ai},{Bi}->{Ci}", arrs[An], arrs[Bn])
    assert np.allclose(full, ref_np, rtol=1e-13, atol=1e-15)
    nrows = sampled.row_count(name, o, v)
    rows = workloads.sample_rows(nrows, count=9, seed=5)
    assert len(rows) == 9 and rows[0] == 0 and rows[-1] == nrows - 1 and len(set(rows.tolist())) == 9
    got = sampled.ROWS_OF[name](I.t_oovv, I.g_oovv, rows)
    want = full.reshape(nrows, -1)[rows].reshape(got.shape)
    assert sampled.rel_err(got, want) < 1e-14


def test_sample_rows_caps_and_is_deterministic():
    a = workloads.sample_rows(100, 64, seed=1)
    assert np.array_equal(a, workloads.sample_rows(100, 64, seed=1))
    assert len(workloads.sample_rows(10, 64)) == 10
    assert np.all(np.diff(a) > 0)
