"""Precision sweep at full size on the GPU against the FP64 CPU oracle.

Cases:
  * the CCD W intermediates at o=48, v=256 (t_oovv, g_oovv 1.2 GB; g_vvvv 34.4 GB),
    inputs drawn on the device from a fixed seed (workloads.ccd_inputs_device), copied to
    the host for the oracle;
  * square DGEMM n in {4096, 8192, 16384}, entries U(-1,1) and the stressed set
    U(-1,1)*2^e with e uniform in [-20, 20].
Each runs as plain FP64, pure FP32 (operands rounded to FP32) and split precision with
3..6 BF16, FP16 and int8 slices and 12..13 CRT moduli (DGEMM: the headline modes BF16x5,
FP16x5, int8x6, CRTx13).

Oracle coverage (stated per case):
  * Wmnij (2304 x 2304 output) and DGEMM n <= 8192: the full oracle result;
  * Wabef (65536 x 65536 output) and DGEMM n = 16384: 64 sampled full output rows
    (workloads.sample_rows, seed 97; first and last row included);
  * Wmbej (generic, 3.7 TFLOP through the oracle's loop nest): 64 sampled (m, b) rows,
    each a full (e, j) slab of 256 x 48 entries.
The error is relative Frobenius of the contraction part (W - W_initial) against the
oracle's contraction on the compared rows.

Tolerances (beta = bits per slice: 8 for BF16/FP16, 7 for int8; u = 2^(-beta*s) for s slices;
u = 2^-kb for the CRT scheme, kb = ein_split_crt_bits(moduli, K); "+refine" adds one
iterative-refinement correction and is held to the FP64 tolerance):
  FP64 <= 1e-13; FP32 <= 1e-5;
  split: <= max(4 * u * (rho_A + rho_B), 1e-13), where rho_X is the mean over the rows of
  op(X) (a row runs along the contracted index) of max|row| / rms(row). Each row is scaled
  by its largest power of two before slicing, so an element carries an absolute error of
  about u * max|row|; summed with random signs over the contraction this is a relative error
  of about u * (rho_A + rho_B) in C. rho is about 1.7 for U(-1,1) data, about 4-5 for the
  normal CCD data over 2304-65536 contracted elements and larger for the stressed set.
  The headline modes (BF16x5, FP16x5, int8x6, CRTx13) must also meet the 1e-11 target on
  the CCD and U(-1,1) data.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from einsums_amd import workloads
from oracle import sampled

pytestmark = pytest.mark.gpu

O, V = 48, 256
SPLIT = [(st, ns) for st in ("bf16", "fp16", "int8") for ns in (3, 4, 5, 6)] + [("crt", 12), ("crt", 13), ("crt+refine", 13)]
MODES = [("fp64", 0), ("fp32", 0)] + SPLIT


HEADLINE = {("bf16", 5), ("fp16", 5), ("int8", 6), ("crt", 13)}


def rho(X, contracted) -> float:
    """Mean over the rows of X (a row runs over the `contracted` axes) of max|row| / rms(row)."""
    keep = [d for d in range(X.dim()) if d not in contracted]
    R = X.permute(*keep, *contracted).reshape(-1, int(np.prod([X.shape[d] for d in contracted])))
    ratios = R.abs().amax(1) / R.square().mean(1).sqrt()
    return float(ratios[torch.isfinite(ratios)].mean())


def tolerance(mode: str, ns: int, rho_sum: float, k: int) -> float:
    if mode == "fp64" or mode.endswith("+refine"):
        return 1e-13
    if mode == "fp32":
        return 1e-5
    if mode == "crt":
        from einsums_amd import core

        return max(4.0 * 2.0 ** -core.crt_bits(ns, k) * rho_sum, 1e-13)
    beta = 7 if mode == "int8" else 8
    return max(4.0 * 2.0 ** (-beta * ns) * rho_sum, 1e-13)


def mode_id(m):
    return m[0] if not m[1] else f"{m[0]}x{m[1]}"


# ----------------------------------------------------------------- CCD o=48, v=256
class CCDL:
    def __init__(self):
        self.I = workloads.ccd_inputs_device(O, V, seed=11)
        self.t_h = self.I.t_oovv.cpu().numpy()
        self.g_h = self.I.g_oovv.cpu().numpy()
        self.ref = {}
        self.rows = {}
        self.rhos = {}

    def rho_sum(self, name):
        """rho of both operands of the contraction, each along its contracted labels."""
        if name not in self.rhos:
            _, Ci, An, Ai, Bn, Bi, *_ = SPECS[name]
            arrs = {"t_oovv": self.I.t_oovv, "g_oovv": self.I.g_oovv}
            links = [c for c in Ai if c in Bi and c not in Ci]
            self.rhos[name] = rho(arrs[An], [Ai.index(c) for c in links]) + rho(arrs[Bn], [Bi.index(c) for c in links])
        return self.rhos[name]

    def oracle_rows(self, name):
        if name not in self.ref:
            n = sampled.row_count(name, O, V)
            rows = np.arange(n, dtype=np.int64) if name == "Wmnij" else workloads.sample_rows(n, 64)
            self.rows[name] = rows
            self.ref[name] = sampled.ROWS_OF[name](self.t_h, self.g_h, rows).reshape(len(rows), -1)
        return self.rows[name], self.ref[name]


@pytest.fixture(scope="module")
def ccdl(oracle):
    state = CCDL()
    yield state
    del state
    torch.cuda.empty_cache()


SPECS = {c[0]: c for c in workloads.CCD_CONTRACTIONS}


@pytest.mark.parametrize("mode", MODES, ids=mode_id)
@pytest.mark.parametrize("name", ["Wmnij", "Wmbej", "Wabef"])
def test_ccd_o48_v256(core, ccdl, name, mode):
    st, ns = mode
    _, Ci, An, Ai, Bn, Bi, cp, abp, init = SPECS[name]
    I = ccdl.I
    arrs = {"t_oovv": I.t_oovv, "g_oovv": I.g_oovv, "g_oooo": I.g_oooo, "g_vvvv": I.g_vvvv, "g_ovvo": I.g_ovvo}
    A, B, W0 = arrs[An], arrs[Bn], arrs[init]
    if st == "fp32":
        A, B, W = A.float(), B.float(), W0.float()
    else:
        W = W0.clone()
    base = W.clone() if name != "Wabef" else None
    alg = core.einsum(Ci, W, Ai, A, Bi, B, C_prefactor=cp, AB_prefactor=abp, split=ns,
                      slice_type=st if ns else "bf16")
    assert core.algorithm_name(alg) == ("GENERIC" if name == "Wmbej" else "GEMM")
    rows, ref = ccdl.oracle_rows(name)
    nrows = sampled.row_count(name, O, V)
    r = torch.from_numpy(rows).cuda()
    if base is None:  # Wabef: W0 rounded to W's precision, on the sampled rows only
        b = W0.reshape(nrows, -1)[r].to(W.dtype).double()
    else:
        b = base.reshape(nrows, -1)[r].double()
    got = (W.reshape(nrows, -1)[r].double() - b).cpu().numpy()
    err = sampled.rel_err(got, ref)
    kdim = {"Wmnij": V * V, "Wabef": O * O, "Wmbej": O * V}[name]
    tol = tolerance(st, ns, ccdl.rho_sum(name), kdim)
    print(f"{name} o={O} v={V} {mode_id(mode)}: rel err {err:.3e} over {len(rows)} rows (tol {tol:.1e})")
    del W, base
    assert err <= tol
    if (st, ns) in HEADLINE:
        assert err <= 1e-11


# ------------------------------------------------------------------- DGEMM sweep
DG_MODES = [("fp64", 0), ("fp32", 0), ("bf16", 5), ("fp16", 5), ("int8", 6), ("crt", 13), ("crt+refine", 13)]
DG_CASES = [(n, dist, m) for n in (4096, 8192, 16384) for dist in ("uniform", "stressed") for m in DG_MODES]


class DGEMMCache:
    key = None
    data = None


def dgemm_data(oracle, n, dist):
    if DGEMMCache.key != (n, dist):
        DGEMMCache.data = None
        torch.cuda.empty_cache()
        g = torch.Generator(device="cuda").manual_seed(1000 + n + (1 if dist == "stressed" else 0))

        def make():
            x = torch.rand((n, n), dtype=torch.float64, device="cuda", generator=g).mul_(2).sub_(1)
            if dist == "stressed":
                x.mul_(torch.exp2(torch.randint(-20, 21, (n, n), device="cuda", generator=g).double()))
            return x

        A, B = make(), make()
        Ah, Bh = A.cpu().numpy(), B.cpu().numpy()
        if n <= 8192:
            rows = np.arange(n, dtype=np.int64)
            ref = np.zeros(n * n)
            # row-major C = A B is column-major C^T = B^T A^T
            oracle.gemm_colmajor("N", "N", n, n, n, 1.0, Bh.ravel(), n, Ah.ravel(), n, 0.0, ref, n)
        else:
            rows = workloads.sample_rows(n, 64)
            As = np.ascontiguousarray(Ah[rows])
            ref = np.zeros(len(rows) * n)
            oracle.gemm_colmajor("N", "N", n, len(rows), n, 1.0, Bh.ravel(), n, As.ravel(), n, 0.0, ref, n)
        DGEMMCache.data = (A, B, rows, ref.reshape(len(rows), n))
        DGEMMCache.key = (n, dist)
    return DGEMMCache.data


@pytest.mark.parametrize("n,dist,mode", DG_CASES, ids=[f"{n}-{d}-{mode_id(m)}" for n, d, m in DG_CASES])
def test_dgemm_sweep(core, oracle, n, dist, mode):
    st, ns = mode
    A, B, rows, ref = dgemm_data(oracle, n, dist)
    if st == "fp32":
        C = torch.empty((n, n), dtype=torch.float32, device="cuda")
        core.gemm("N", "N", 1.0, A.float(), B.float(), 0.0, C)
    else:
        C = torch.empty((n, n), dtype=torch.float64, device="cuda")
        if ns:
            core.gemm_split("N", "N", 1.0, A, B, 0.0, C, nslices=ns, slice_type=st)
        else:
            core.gemm("N", "N", 1.0, A, B, 0.0, C)
    got = C[torch.from_numpy(rows).cuda()].double().cpu().numpy()
    err = sampled.rel_err(got, ref)
    tol = tolerance(st, ns, rho(A, [1]) + rho(B, [0]), n)
    print(f"DGEMM n={n} {dist} {mode_id(mode)}: rel err {err:.3e} over {len(rows)} rows (tol {tol:.1e})")
    del C
    assert err <= tol
    if (st, ns) in HEADLINE and dist == "uniform":
        assert err <= 1e-11
    if (n, dist, mode) == DG_CASES[-1]:
        DGEMMCache.data = None
        torch.cuda.empty_cache()
