"""Split-precision (Ozaki) FP64 GEMM on BF16/FP16 and int8 MFMA against the FP64 CPU oracle.

Tolerances (relative Frobenius error against the FP64 oracle):
  BF16/FP16 nslices=5 (40 bits) and int8 nslices=6 (42 bits):
             <= 1e-11 for U(-1,1) data and the CCD contractions,
             <= 1e-10 for the stressed set U(-1,1)*2^e, e in [-20, 20];
  BF16 nslices=3 (24 bits): <= 1e-6; int8 nslices=3 (21 bits): <= 1e-5;
  CRT with L moduli (kb = ein_split_crt_bits(L, k) bits per element): <= 8 * 2^-kb for U(-1,1)
  data (about 1e-12 with 13 moduli);
  integer-valued known answers: exact.
"""
import ctypes
import json
import zlib
from pathlib import Path

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from einsums_amd import workloads
from tests.helpers import dev, host, rel_fro

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


def vp(t):
    return ctypes.c_void_p(t.data_ptr())


def seeded(*parts):
    return np.random.default_rng(zlib.crc32("|".join(map(str, parts)).encode()))


@pytest.mark.parametrize("slice_type", [0, 1, 2, 3])
@pytest.mark.parametrize("trans", ["NN", "TN", "NT", "TT"])
def test_split_gemm_3x3_known_answers(slice_type, trans):
    from einsums_amd import _lib

    g = json.loads((GOLD / "reference_known_answers.json").read_text())["gemm_3x3"]
    A, B = dev(np.array(g["A"], dtype=np.float64)), dev(np.array(g["B"], dtype=np.float64))
    C = dev(np.full(9, np.nan))
    _lib.check(_lib.lib.ein_gemm_split_f64(trans[1].encode(), trans[0].encode(), 3, 3, 3, 1.0, vp(B), 3, vp(A), 3, 0.0,
                                           vp(C), 3, {2: 6, 3: 13}.get(slice_type, 5), slice_type, None))
    assert host(C).tolist() == g["expected"][trans]


SHAPES = [(1, 1, 1), (5, 7, 3), (64, 64, 32), (100, 130, 257), (257, 129, 1000), (64, 64, 65536), (1024, 1024, 2304),
          (33, 300, 70000)]


@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
@pytest.mark.parametrize("slice_type,nslices,tol", [(0, 5, 1e-11), (0, 3, 1e-6), (0, 4, 2e-9), (0, 6, 1e-13),
                                                     (1, 4, 2e-9), (1, 6, 1e-13), (2, 6, 1e-11), (2, 3, 1e-5),
                                                     (2, 4, 2e-7), (2, 5, 2e-10)],
                         ids=["bf16x5", "bf16x3", "bf16x4", "bf16x6", "fp16x4", "fp16x6", "int8x6", "int8x3", "int8x4",
                              "int8x5"])
def test_split_gemm_vs_oracle(oracle, m, n, k, ta, tb, slice_type, nslices, tol):
    split_vs_oracle(oracle, m, n, k, ta, tb, slice_type, nslices, tol)


@pytest.mark.parametrize("m,n,k", SHAPES + [(300, 257, 513), (513, 1000, 64)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("nmod", [8, 10, 13, 16])
def test_crt_gemm_vs_oracle(oracle, m, n, k, ta, tb, nmod):
    """The CRT split: ragged edges of the 256 x 256 residue tiles, k = 1 .. 70000 (k > 65536
    runs as two residue launches that add), every transpose, 8..16 moduli."""
    from einsums_amd import core

    tol = max(8.0 * 2.0 ** -core.crt_bits(nmod, k), 2e-14)  # FP64 oracle rounding below that
    split_vs_oracle(oracle, m, n, k, ta, tb, 3, nmod, tol)


def split_vs_oracle(oracle, m, n, k, ta, tb, slice_type, nslices, tol):
    from einsums_amd import _lib

    rng = seeded(m, n, k, ta, tb)
    lda = (m if ta == "N" else k) + 1
    ldb = (k if tb == "N" else n) + 2
    ldc = m + 3
    A = rng.uniform(-1, 1, (k if ta == "N" else m) * lda)
    B = rng.uniform(-1, 1, (n if tb == "N" else k) * ldb)
    C0 = rng.uniform(-1, 1, n * ldc)
    ref = C0.copy()
    oracle.gemm_colmajor(ta, tb, m, n, k, 0.75, A, lda, B, ldb, -0.5, ref, ldc)
    dA, dB, dC = dev(A), dev(B), dev(C0)
    _lib.check(_lib.lib.ein_gemm_split_f64(ta.encode(), tb.encode(), m, n, k, 0.75, vp(dA), lda, vp(dB), ldb, -0.5, vp(dC),
                                           ldc, nslices, slice_type, None))
    got = host(dC).reshape(n, ldc)
    ref = ref.reshape(n, ldc)
    assert np.array_equal(got[:, m:], C0.reshape(n, ldc)[:, m:])
    err = rel_fro(got[:, :m], ref[:, :m])
    print(f"split type {slice_type} x{nslices} {m}x{n}x{k} {ta}{tb}: rel err {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("slice_type", ["bf16", "fp16", "int8", "crt"])
@pytest.mark.parametrize("n", [1024, 2048])
def test_split_dgemm_uniform_and_stressed(core, oracle, slice_type, n):
    rng = seeded("stress", n)
    A = rng.uniform(-1, 1, (n, n))
    B = rng.uniform(-1, 1, (n, n))
    for stressed in (False, True):
        if stressed:
            A = A * np.exp2(rng.integers(-20, 21, (n, n)))
            B = B * np.exp2(rng.integers(-20, 21, (n, n)))
        ref = np.zeros(n * n)
        # row-major C = A B  <=>  column-major C^T = B^T A^T
        oracle.gemm_colmajor("N", "N", n, n, n, 1.0, np.ascontiguousarray(B).ravel(), n, np.ascontiguousarray(A).ravel(),
                             n, 0.0, ref, n)
        dC = torch.zeros((n, n), dtype=torch.float64, device="cuda")
        ns = {"int8": 6, "crt": 13}.get(slice_type, 5)
        core.gemm_split("N", "N", 1.0, dev(A), dev(B), 0.0, dC, nslices=ns, slice_type=slice_type)
        err = rel_fro(host(dC).ravel(), ref)
        print(f"{slice_type} n={n} stressed={stressed}: rel err {err:.3e}")
        assert err <= (1e-10 if stressed else 1e-11)


@pytest.mark.parametrize("slice_type,nslices", [("bf16", 5), ("int8", 6), ("crt", 13)])
@pytest.mark.parametrize("name", ["Wmnij", "Wabef"])
def test_split_ccd_contractions(core, oracle, name, slice_type, nslices):
    I = workloads.ccd_inputs(24, 96, seed=7)
    spec = {c[0]: c for c in workloads.CCD_CONTRACTIONS}[name]
    _, Ci, An, Ai, Bn, Bi, cp, abp, init = spec
    arrs = {"t_oovv": I.t_oovv, "g_oovv": I.g_oovv, "g_oooo": I.g_oooo, "g_vvvv": I.g_vvvv}
    A, B, W0 = arrs[An], arrs[Bn], arrs[init]
    ref = W0.copy()
    oracle.einsum(cp, Ci, ref, abp, Ai, A, Bi, B)
    dW = dev(W0)
    alg = core.einsum(Ci, dW, Ai, dev(A), Bi, dev(B), C_prefactor=cp, AB_prefactor=abp, split=nslices,
                      slice_type=slice_type)
    assert core.algorithm_name(alg) == "GEMM"
    # error of the contraction part alone (W0 is added exactly in FP64)
    err = rel_fro(host(dW) - W0, ref - W0)
    print(f"{name} {slice_type} x{nslices} rel err of the contraction {err:.3e}")
    assert err <= 1e-11


REFINE_CASES = [("crt", 13), ("crt", 10), ("crt", 8)]


@pytest.mark.parametrize("m,n,k", [(257, 129, 1000), (64, 64, 65536), (1024, 1024, 2304), (33, 300, 70000)])
@pytest.mark.parametrize("st,ns", REFINE_CASES, ids=[f"{s}x{n}" for s, n in REFINE_CASES])
@pytest.mark.parametrize("stressed", [False, True])
def test_split_refine_vs_oracle(core, oracle, m, n, k, st, ns, stressed):
    """One iterative-refinement correction ("crt+refine"): the FP64 residuals of the kb-bit
    representation are multiplied by CRT GEMMs with 8 moduli (kb8 bits) and added. Error
    against the oracle after refinement <= max(8 * 2^-(kb + kb8), 4e-14), FP64 level unless
    both are short (8 moduli at k = 65536: 22 + 22 bits); the unrefined error is about 2^-kb.
    Uniform and stressed (2^+-20) data, 8..13 moduli."""
    rng = seeded("refine", m, n, k, st, ns, stressed)
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    if stressed:
        A = A * np.exp2(rng.integers(-20, 21, (m, k)))
        B = B * np.exp2(rng.integers(-20, 21, (k, n)))
    C0 = rng.uniform(-1, 1, (m, n))
    ref = C0.copy().ravel()
    # row-major C = A B is column-major C^T = B^T A^T
    oracle.gemm_colmajor("N", "N", n, m, k, 0.75, np.ascontiguousarray(B).ravel(), n, np.ascontiguousarray(A).ravel(), k,
                         -0.5, ref, n)
    dA, dB = dev(A), dev(B)
    errs = {}
    for mode in (st, st + "+refine"):
        dC = dev(C0)
        core.gemm_split("N", "N", 0.75, dA, dB, -0.5, dC, nslices=ns, slice_type=mode)
        errs[mode] = rel_fro(host(dC).ravel(), ref)
    tol = max(8.0 * 2.0 ** -(core.crt_bits(ns, k) + core.crt_bits(8, k)), 4e-14)
    print(f"{st}x{ns} {m}x{n}x{k} stressed={stressed}: {errs[st]:.3e} -> refined {errs[st + '+refine']:.3e} (tol {tol:.1e})")
    assert errs[st + "+refine"] <= tol
    assert errs[st + "+refine"] <= max(errs[st], 4e-14)


@pytest.mark.parametrize("st", ["bf16", "fp16", "int8"])
def test_split_refine_rejected_for_slices(core, st):
    C = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    A = torch.ones((8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="EIN_SLICE_CRT only"):
        core.gemm_split("N", "N", 1.0, A, A, 0.0, C, nslices=5, slice_type=st + "+refine")


def test_split_refine_einsum(core, oracle):
    """The refinement through the einsum entry point (GEMM-mapped Wabef at o=12, v=48)."""
    I = workloads.ccd_inputs(12, 48, seed=3)
    spec = {c[0]: c for c in workloads.CCD_CONTRACTIONS}["Wabef"]
    _, Ci, An, Ai, Bn, Bi, cp, abp, init = spec
    arrs = {"t_oovv": I.t_oovv, "g_oovv": I.g_oovv, "g_vvvv": I.g_vvvv}
    ref = arrs[init].copy()
    oracle.einsum(cp, Ci, ref, abp, Ai, arrs[An], Bi, arrs[Bn])
    dW = dev(arrs[init])
    alg = core.einsum(Ci, dW, Ai, dev(arrs[An]), Bi, dev(arrs[Bn]), C_prefactor=cp, AB_prefactor=abp, split=13,
                      slice_type="crt+refine")
    assert core.algorithm_name(alg) == "GEMM"
    err = rel_fro(host(dW) - arrs[init], ref - arrs[init])
    print(f"Wabef o=12 v=48 crt x13 + refine: {err:.3e}")
    assert err <= 1e-14


@pytest.mark.parametrize("m,n,k", [(300, 257, 2304), (64, 80, 12288), (100, 64, 65536)])
def test_crt_auto_moduli(core, oracle, m, n, k):
    """split "auto" (nslices 0) runs the CRT scheme with ein_split_crt_moduli(k) moduli: same
    bits as that explicit count, error within its tolerance."""
    rng = seeded("auto", m, n, k)
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    L = core.crt_moduli(k)
    outs = []
    for ns in (0, L):
        C = torch.zeros((m, n), dtype=torch.float64, device="cuda")
        core.gemm_split("N", "N", 1.0, dev(A), dev(B), 0.0, C, nslices=ns, slice_type="crt")
        outs.append(host(C))
    assert np.array_equal(outs[0], outs[1])
    ref = np.zeros(m * n)
    oracle.gemm_colmajor("N", "N", n, m, k, 1.0, B.ravel(), n, A.ravel(), k, 0.0, ref, n)
    assert rel_fro(outs[0].ravel(), ref) <= max(8.0 * 2.0 ** -core.crt_bits(L, k), 2e-14)


def test_crt_panels_overlapped(core, oracle):
    """A CRT GEMM whose residue planes exceed one panel (65536 x 16384 x 64, 13 moduli: 14 GB
    of planes in 4 GiB panels): the reconstruction of each panel runs on the side stream while
    the next panel's residue GEMM runs. Checked on 64 sampled rows (first and last included)
    against the oracle, on a non-default stream, and repeated bit for bit."""
    m, n, k = 65536, 16384, 64
    g = torch.Generator(device="cuda").manual_seed(77)
    A = torch.rand((m, k), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B = torch.rand((k, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    C = torch.empty((m, n), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    core.gemm_split("N", "N", 1.0, A, B, 0.0, C, nslices=13, slice_type="crt", stream=s)
    s.synchronize()
    rows = workloads.sample_rows(m, 64)
    got = C[torch.from_numpy(rows).cuda()].cpu().numpy()
    Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
    ref = np.zeros(len(rows) * n)
    oracle.gemm_colmajor("N", "N", n, len(rows), k, 1.0, Bh.ravel(), n, np.ascontiguousarray(Ah[rows]).ravel(), k, 0.0,
                         ref, n)
    assert rel_fro(got.ravel(), ref) <= max(8.0 * 2.0 ** -core.crt_bits(13, k), 2e-14)
    first = C[torch.from_numpy(rows).cuda()].clone()
    C.fill_(float("nan"))
    core.gemm_split("N", "N", 1.0, A, B, 0.0, C, nslices=13, slice_type="crt", stream=s)
    s.synchronize()
    assert torch.equal(C[torch.from_numpy(rows).cuda()], first)
    assert not torch.isnan(C).any()
    del A, B, C
    torch.cuda.empty_cache()
