"""Benchmark of the einsum path on one MI355X.

Workload (one "step"): the CCD W intermediates at o=24, v=96 in FP64 with inputs resident
in HBM -- Wmnij += 0.25 t_ijef g_mnef (GEMM), Wabef += 0.25 g_mnef t_mnab (GEMM),
Wmbej += -0.5 t_jnfb g_mnef (generic, non-GEMM layout), E = 0.25 t.g (DOT) and the
permute ijab -> iajb. `value` is the step's algorithmic FP64 FLOP / step time (GFLOP/s).

Prints one JSON line with the roofline of the dominant kernel (Wabef, FP64 MFMA bound,
timed with HIP events on the stream the kernels run on), the split-precision block (the
same W intermediates at o=48, v=256 as plain FP64, int8 / BF16 / FP16 slices and FP32,
timed on the GPU; their errors against the oracle on sampled rows come from the CPU leg)
and the CPU baseline (the oracle's restatement of the reference CPU einsum timed on this
host).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--o", type=int, default=24)
    p.add_argument("--v", type=int, default=96)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--traffic", type=str, default=str(ROOT / "profiles" / "gemm_dense_wabef_o24_v96_traffic.txt"),
                   help="rocprof summary holding the measured HBM bytes per Wabef launch (o=24, v=96 only)")
    p.add_argument("--split", type=int, default=1,
                   help="1: add the split-precision vs FP64 block at o=48, v=256 (checked against the oracle in "
                        "the CPU leg), 0: skip it")
    p.add_argument("--hbm", type=int, default=1,
                   help="1: add the HBM-bound kernels at sizes above the 256 MB cache (nbf=96, o=48 v=256) "
                        "to the line, 0: skip them")
    p.add_argument("--kernel-trace", type=str, default=str(ROOT / "profiles" / "ccd_step_o24_v96_kernel_trace.txt"),
                   help="rocprof kernel-trace summary of this bench holding the Wabef kernel's average duration")
    p.add_argument("--grouped", action="store_true",
                   help="only run the batched uneven-block benchmarks (one JSON line) and exit")
    p.add_argument("--precision-sweep", action="store_true",
                   help="only run the split-precision vs FP64 GEMM sweep (one JSON line per case) and exit")
    p.add_argument("--sweep-cases", default="dgemm,wmnij,wabef,wmbej",
                   help="cases of the precision sweep: any of dgemm, wmnij, wabef, wmbej (comma separated)")
    p.add_argument("--bandwidth", action="store_true",
                   help="only run the HBM-bound kernels (HF J/K, dot, permute; cold and warm cache) and exit")
    p.add_argument("--tune", type=str, default="", metavar="PATH",
                   help="run the CCD workloads once (o=24 v=96 and o=48 v=256, FP64) in GEMM tuning mode, save "
                        "the tuning table to PATH and exit")
    p.add_argument("--tuning-table", type=str, default=str(ROOT / "einsums_amd" / "gemm_tuning_mi355x.txt"),
                   help="GEMM tuning table loaded before the timed steps (skipped when the file is absent)")
    p.add_argument("--out-of-core", type=str, default="", metavar="DIR",
                   help="only run the Wabef update with C in HBM, in pinned host memory and in an HDF5 disk "
                        "tensor under DIR (o=24 v=96 and o=32 v=128; one JSON line each) and exit")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                   help="only run a plain device copy of BYTES bytes (counter calibration) and exit")
    return p.parse_args()


def calibrate_copy(nbytes: int) -> None:
    """A plain copy kernel of known size (identity permute, 16-byte vectors) for scaling
    the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    import einsums_amd.core as core

    n = nbytes // 8
    src = torch.rand(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        core.permute(0.0, "i", dst, 1.0, "i", src)
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    print(json.dumps({"calibrate_copy_bytes_read": n * 8, "calibrate_copy_bytes_written": n * 8, "launches": 3}))


def rocprof_figure(path, flop):
    """The Wabef kernel's average duration in the committed rocprofv3 kernel trace of this
    bench (line `Wabef_kernel_avg_us: <us>`), turned into the same roofline fraction as the
    live HIP-event figure. The trace comes from a separate run; the file names it."""
    p = Path(path)
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("Wabef_kernel_avg_us:"):
            us = float(line.split(":")[1])
            tf = flop / (us * 1e-6) / 1e12
            return {"file": str(p.relative_to(ROOT)) if p.is_relative_to(ROOT) else str(p), "avg_us": us,
                    "achieved": round(tf, 3), "frac": round(tf / FP64_MFMA_PEAK_TFLOPS, 4)}
    return None


def host_cpu_info():
    """(threads to use, CPU model). A one-GPU MI355X host share is 16 CPUs whatever `nproc`
    reports, so the oracle uses at most 16 threads (fewer where fewer are available)."""
    cores = min(len(os.sched_getaffinity(0)), 16)
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def cpu_leg(I, o, v, flops, large, hf=None, budget_s=20.0, max_reps=20, warmups=3):
    """The bench's CPU leg: everything here runs the oracle (the restated reference CPU
    einsum, oracle/), which the GPU legs never touch.

    1. CPU baseline: one full step (Wmnij, Wabef, Wmbej, E, permute) at (o, v) on the
       host's cores. The host BLAS is resolved before timing; W is restored from a pristine
       copy before every repetition, untimed; `warmups` untimed steps, then repetitions
       until `max_reps` or `budget_s` seconds of timed work, whichever comes first.
       Reports GFLOP/s of the median step, with the minimum alongside.
    2. Checker of the split-precision block: the oracle's contraction on the sampled rows of
       the o=48, v=256 W intermediates, against which every GPU mode's rows are compared.
    3. Checker of the HF block: the oracle's G = 2J - K at nbf = 48."""
    cores, model = host_cpu_info()
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import numpy as np

    from oracle import pyoracle, sampled

    pyoracle.set_threads(cores)
    blas = os.path.basename(pyoracle.blas_name())  # dlopen of the host BLAS, outside the timing
    W0 = {"Wmnij": I.g_oooo, "Wabef": I.g_vvvv, "Wmbej": I.g_ovvo}
    W = {k: a.copy() for k, a in W0.items()}
    E = np.zeros(())
    P = np.empty((o, v, o, v))

    def step():
        pyoracle.einsum(1.0, "mnij", W["Wmnij"], 0.25, "ijef", I.t_oovv, "mnef", I.g_oovv)
        pyoracle.einsum(1.0, "abef", W["Wabef"], 0.25, "mnef", I.g_oovv, "mnab", I.t_oovv)
        pyoracle.einsum(1.0, "mbej", W["Wmbej"], -0.5, "jnfb", I.t_oovv, "mnef", I.g_oovv)
        pyoracle.einsum(0.0, "", E, 0.25, "ijab", I.t_oovv, "ijab", I.g_oovv)
        pyoracle.permute(0.0, "iajb", P, 1.0, "ijab", I.t_oovv)

    def restore():
        for k in W:
            np.copyto(W[k], W0[k])

    for _ in range(warmups):
        restore()
        step()
    times = []
    while len(times) < max_reps and sum(times) < budget_s:
        restore()
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    total = sum(flops.values())
    med, mn = float(np.median(times)), float(min(times))
    base = {"value": round(total / med / 1e9, 3), "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": f"full steps (Wmnij, Wabef, Wmbej, E, permute) at o={o} v={v}: {warmups} warm-ups, "
                      f"{len(times)} timed reps, median {med:.3f} s (min {mn:.3f} s, "
                      f"{total / mn / 1e9:.1f} GFLOP/s); CPU {model}; host BLAS {blas}; "
                      f"OMP_PROC_BIND={os.environ['OMP_PROC_BIND']}"}
    if hf is not None:
        g_h, D_h = hf.pop("_inputs")
        ref = np.zeros_like(D_h)
        pyoracle.einsum(0.0, "pq", ref, 2.0, "pqrs", g_h, "rs", D_h)
        pyoracle.einsum(1.0, "pq", ref, -1.0, "prqs", g_h, "rs", D_h)
        for key in ("two_einsum_calls", "fused_jk"):
            got = hf[key].pop("_got")
            hf[key]["rel_err_vs_oracle"] = float(f"{sampled.rel_err(got, ref):.3e}")
    if large is not None:
        t_h, g_h = large.pop("t_host"), large.pop("g_host")
        for name, row in large["contractions"].items():
            rows = row.pop("_rows")
            ref = sampled.ROWS_OF[name](t_h, g_h, rows).reshape(len(rows), -1)
            for key, got in row.pop("_got").items():
                row[key]["rel_err_vs_oracle"] = float(f"{sampled.rel_err(got, ref):.3e}")
        for key in large["total"]:
            large["total"][key]["max_rel_err_vs_oracle"] = max(
                row[key]["rel_err_vs_oracle"] for row in large["contractions"].values())
        # the fastest split mode whose three results all meet 1e-11 against the oracle
        ok = [(v["speedup"], k) for k, v in large["total"].items()
              if k not in ("fp64", "fp32") and v["max_rel_err_vs_oracle"] <= 1e-11]
        if ok:
            sp, k = max(ok)
            large["best_at_1e-11"] = {"mode": k, "speedup": sp,
                                      "max_rel_err_vs_oracle": large["total"][k]["max_rel_err_vs_oracle"]}
    return base


def hf_block(core, nbf=48, nocc=12, reps=20):
    """Hartree-Fock G = 2J - K at nbf = 48 (the HF workload: 8-fold symmetric g of 42.5 MB,
    D from a QR): the reference's two einsum calls (GEMV + generic, g read twice) and the
    fused J/K build (g read once), each timed cold (a 512 MB scratch read before every
    call) and warm (back to back) with HIP events. HBM fraction = algorithmic bytes of the
    fused build (g + D + G) / time / 8 TB/s. The result is compared with the oracle in
    the CPU leg."""
    import numpy as np
    import torch

    from einsums_amd import workloads

    g_h, D_h = workloads.hf_inputs(nbf, nocc, seed=2024)
    g = torch.from_numpy(g_h).cuda()
    D = torch.from_numpy(D_h).cuda()
    G = torch.zeros((nbf, nbf), dtype=torch.float64, device="cuda")
    pj = core.compile_plan("pq", "pqrs", "rs")
    pk = core.compile_plan("pq", "prqs", "rs")
    scratch = torch.zeros(64 << 20, dtype=torch.float64, device="cuda")

    def two_calls():
        pj.execute(0.0, G, 2.0, g, D)
        pk.execute(1.0, G, -1.0, g, D)

    def fused():
        core.fock_jk(g, D, G)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        warm = e0.elapsed_time(e1) / reps
        cold = []
        for _ in range(reps):
            scratch.sum()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            cold.append(a.elapsed_time(b))
        return warm, float(np.median(cold))

    nbytes = 8 * (nbf ** 4 + 2 * nbf ** 2)
    out = {"workload": f"HF G = 2J - K, nbf={nbf} (J: pq <- pqrs,rs GEMV; K: pq <- prqs,rs generic)",
           "bytes_fused": nbytes}
    for name, fn in (("two_einsum_calls", two_calls), ("fused_jk", fused)):
        warm, cold = timed(fn)
        fn()
        out[name] = {"cold_us": round(cold * 1e3, 2), "warm_us": round(warm * 1e3, 2),
                     "cold_hbm_frac_of_fused_bytes": round(nbytes / (cold * 1e-3) / 1e12 / 8.0, 3),
                     "_got": G.cpu().numpy().copy()}
    out["speedup_cold"] = round(out["two_einsum_calls"]["cold_us"] / out["fused_jk"]["cold_us"], 3)
    out["_inputs"] = (g_h, D_h)
    del g, D, G, scratch
    torch.cuda.empty_cache()
    return out


SPLIT_MODES = (("crt", "auto"), ("crt", 13), ("crt+refine", 13), ("int8", 6), ("bf16", 5), ("fp16", 5))


def split_key(st, ns):
    """Name of a split mode in the JSON: "crtx13", "crtx13+refine", "int8x6", "crt-auto" (the
    fewest moduli keeping 40 bits at each contraction's length); the plain modes ("fp64",
    "fp32") keep their names."""
    base, plus, flag = st.partition("+")
    if ns == "auto":
        return f"{base}-auto{plus}{flag}"
    return st if not ns else f"{base}x{ns}{plus}{flag}"


def split_large(core, o=48, v=256, reps=3):
    """Split precision vs plain FP64 on the CCD W intermediates at o=48, v=256: Wmnij (GEMM
    2304 x 2304 x 65536), Wabef (GEMM 65536 x 65536 x 2304, 34 GB output) and Wmbej
    (generic, 12288^3 after packing), each `W += prefactor * contraction` with inputs
    resident in HBM. Every mode is timed with HIP events over `reps` calls after one warm-up
    (W accumulates; the timing does not depend on W's values). A separate call on a fresh
    copy of W gives the rows that the CPU leg compares with the oracle: 64 sampled rows of
    each intermediate (workloads.sample_rows, seed 97)."""
    import torch

    from einsums_amd import workloads

    I = workloads.ccd_inputs_device(o, v, seed=11)
    specs = {c[0]: c for c in workloads.CCD_CONTRACTIONS}
    arrs = {"t_oovv": I.t_oovv, "g_oovv": I.g_oovv, "g_oooo": I.g_oooo, "g_vvvv": I.g_vvvv, "g_ovvo": I.g_ovvo}
    nrows = {"Wmnij": o * o, "Wabef": v * v, "Wmbej": o * v}
    modes = [("fp64", 0)] + list(SPLIT_MODES) + [("fp32", 0)]

    key = split_key

    def timed(fn):
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {"workload": f"CCD W intermediates o={o} v={v}, W += prefactor*contraction (beta=1), FP64 in HBM",
           "rows_checked": "64 sampled rows of each intermediate against the oracle (CPU leg)",
           "contractions": {}, "total": {}}
    for name in ("Wmnij", "Wabef", "Wmbej"):
        _, Ci, An, Ai, Bn, Bi, cp, abp, init = specs[name]
        A, B, W0 = arrs[An], arrs[Bn], arrs[init]
        flop = {"Wmnij": 2.0 * o ** 4 * v ** 2, "Wabef": 2.0 * v ** 4 * o ** 2, "Wmbej": 2.0 * o ** 3 * v ** 3}[name]
        rows = workloads.sample_rows(nrows[name], 64)
        r = torch.from_numpy(rows).cuda()
        row = {"gflop": round(flop / 1e9, 1), "_rows": rows, "_got": {}}
        plan = core.compile_plan(Ci, Ai, Bi)
        for st, ns in modes:
            if st == "fp32":
                A_, B_, W = A.float(), B.float(), W0.float()
            else:
                A_, B_, W = A, B, W0.clone()
            base = W0.reshape(nrows[name], -1)[r].to(W.dtype).double()
            kw = {"split": ns, "slice_type": st} if ns else {}
            plan.execute(cp, W, abp, A_, B_, **kw)
            row["_got"][key(st, ns)] = (W.reshape(nrows[name], -1)[r].double() - base).cpu().numpy()
            ms = timed(lambda: plan.execute(cp, W, abp, A_, B_, **kw))
            row[key(st, ns)] = {"ms": round(ms, 3), "tflops": round(flop / ms / 1e9, 2)}
            del W, A_, B_
            torch.cuda.empty_cache()
        for st, ns in modes[1:]:
            row[key(st, ns)]["speedup"] = round(row["fp64"]["ms"] / row[key(st, ns)]["ms"], 3)
        out["contractions"][name] = row
    fp64_total = sum(out["contractions"][n]["fp64"]["ms"] for n in out["contractions"])
    for st, ns in modes:
        tot = sum(out["contractions"][n][key(st, ns)]["ms"] for n in out["contractions"])
        out["total"][key(st, ns)] = {"ms": round(tot, 3), "speedup": round(fp64_total / tot, 3)}
    out["t_host"] = I.t_oovv.cpu().numpy()
    out["g_host"] = I.g_oovv.cpu().numpy()
    del I, arrs
    torch.cuda.empty_cache()
    return out


def precision_sweep(sizes=(1024, 2048, 4096, 8192, 16384), reps=3, cases=("dgemm", "wmnij", "wabef", "wmbej")):
    """Split-precision vs plain FP64 GEMM: DGEMM n x n x n with U(-1,1) and stressed
    U(-1,1)*2^e (e uniform in [-20, 20]) data, and the Wabef GEMM at o=48, v=256
    (65536 x 65536 x 2304, 34 GB output). Errors: relative Frobenius against the FP64
    result on the GPU, and against numpy on 64 sampled rows (host BLAS, FP64)."""
    import numpy as np
    import torch

    import einsums_amd.core as core

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def case(label, m, n, k, make):
        g = torch.Generator(device="cuda").manual_seed(1234)
        A, B = make(m, k, g), make(k, n, g)
        C64 = torch.empty((m, n), dtype=torch.float64, device="cuda")
        t64 = timed(lambda: core.gemm("N", "N", 1.0, A, B, 0.0, C64))
        rows = torch.linspace(0, m - 1, 64, device="cuda").long()
        ref_rows = A[rows].cpu().numpy() @ B.cpu().numpy()
        row = {"case": label, "m": m, "n": n, "k": k, "fp64_ms": round(t64, 3),
               "fp64_tflops": round(2.0 * m * n * k / t64 / 1e9, 2),
               "fp64_rel_err_rows_vs_host": float(f"{np.linalg.norm(C64[rows].cpu().numpy() - ref_rows) / np.linalg.norm(ref_rows):.3e}")}
        Cs = torch.empty_like(C64)
        A32, B32 = A.float(), B.float()
        C32 = torch.empty((m, n), dtype=torch.float32, device="cuda")
        t32 = timed(lambda: core.gemm("N", "N", 1.0, A32, B32, 0.0, C32))
        row["fp32"] = {"ms": round(t32, 3), "speedup": round(t64 / t32, 3),
                       "rel_err_vs_fp64": float(f"{float(torch.linalg.norm(C32.double() - C64) / torch.linalg.norm(C64)):.3e}")}
        del A32, B32, C32
        for st, ns in SPLIT_MODES:
            ts = timed(lambda: core.gemm_split("N", "N", 1.0, A, B, 0.0, Cs, nslices=ns, slice_type=st))
            err = float(torch.linalg.norm(Cs - C64) / torch.linalg.norm(C64))
            err_rows = float(np.linalg.norm(Cs[rows].cpu().numpy() - ref_rows) / np.linalg.norm(ref_rows))
            row[split_key(st, ns)] = {"ms": round(ts, 3), "speedup": round(t64 / ts, 3), "rel_err_vs_fp64": float(f"{err:.3e}"),
                                 "rel_err_rows_vs_host": float(f"{err_rows:.3e}")}
        del A, B, C64, Cs
        torch.cuda.empty_cache()
        print(json.dumps(row), flush=True)
        return row

    uniform = lambda r, c, g: torch.rand((r, c), dtype=torch.float64, device="cuda", generator=g) * 2 - 1

    def stressed(r, c, g):
        x = uniform(r, c, g)
        e = torch.randint(-20, 21, (r, c), device="cuda", generator=g).double()
        return x * torch.exp2(e)

    out = []
    if "dgemm" in cases:
        for n in sizes:
            out.append(case(f"dgemm uniform n={n}", n, n, n, uniform))
            out.append(case(f"dgemm stressed n={n}", n, n, n, stressed))
    # o=48, v=256: Wmnij (o^2 x o^2, k = v^2) and Wabef (v^2 x v^2, k = o^2) as GEMMs
    if "wmnij" in cases:
        out.append(case("Wmnij o=48 v=256", 2304, 2304, 65536, lambda r, c, g: uniform(r, c, g) * 0.1))
    if "wabef" in cases:
        out.append(case("Wabef o=48 v=256", 65536, 65536, 2304, lambda r, c, g: uniform(r, c, g) * 0.1))
    if "wmbej" in cases:
        out.append(wmbej_case(timed))
    return out


def bandwidth_bench(reps=20, quiet=False, large_only=False):
    """HBM-bound kernels against the 8 TB/s peak: the HF J (GEMV) and K (generic) builds at
    nbf = 48 and 96, G = 2J - K both as the two einsum calls and as the fused J/K build, the CCD energy dot and the ijab->iajb permute at o=24, v=96 and
    o=48, v=256, and a plain copy. Each is timed warm (back-to-back calls, operands may sit
    in the 256 MB infinity cache) and cold (a 512 MB scratch read between calls, timed
    separately: a scratch write would leave dirty lines whose write-back then competes with
    the timed call); bytes are the algorithmic ones: every operand element read once, the
    output written once (read too when beta != 0). quiet: return the rows without printing
    them; large_only: only the sizes above the 256 MB cache (nbf = 96, o=48 v=256)."""
    import numpy as np
    import torch

    import einsums_amd.core as core

    scratch = torch.zeros(64 << 20, dtype=torch.float64, device="cuda")  # 512 MB

    def measure(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        warm = e0.elapsed_time(e1) / reps
        cold = []
        for _ in range(reps):
            scratch.sum()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            torch.cuda.synchronize()
            cold.append(a.elapsed_time(b))
        return warm, float(np.median(cold))

    out = []

    def report(name, nbytes, fn, alg=None):
        warm, cold = measure(fn)
        row = {"case": name, "bytes": int(nbytes), "cold_us": round(cold * 1e3, 2), "cold_TBs": round(nbytes / cold / 1e9, 3),
               "cold_frac": round(nbytes / cold / 1e9 / 8.0, 3), "warm_us": round(warm * 1e3, 2),
               "warm_TBs": round(nbytes / warm / 1e9, 3)}
        if alg is not None:
            row["algorithm"] = alg
        if not quiet:
            print(json.dumps(row), flush=True)
        out.append(row)

    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    rng = np.random.default_rng(3)
    for nbf in ((96,) if large_only else (48, 96)):
        g = dev(rng.standard_normal((nbf,) * 4))
        D = dev(rng.standard_normal((nbf, nbf)))
        J = torch.empty((nbf, nbf), dtype=torch.float64, device="cuda")
        K = torch.empty_like(J)
        pj = core.compile_plan("pq", "pqrs", "rs")
        pk = core.compile_plan("pq", "prqs", "rs")
        nb = 8 * (nbf ** 4 + 2 * nbf ** 2)
        report(f"HF J nbf={nbf} (pq <- pqrs, rs)", nb, lambda: pj.execute(0.0, J, 1.0, g, D))
        out[-1]["algorithm"] = core.algorithm_name(pj.last_algorithm)
        report(f"HF K nbf={nbf} (pq <- prqs, rs)", nb, lambda: pk.execute(0.0, K, 1.0, g, D))
        out[-1]["algorithm"] = core.algorithm_name(pk.last_algorithm)

        def two_calls():
            pj.execute(0.0, J, 2.0, g, D)
            pk.execute(1.0, J, -1.0, g, D)

        # G = 2J - K: the reference's two einsum calls (g read twice), then the fused build
        report(f"HF G = 2J - K nbf={nbf}, two einsum calls", 2 * nb + 8 * nbf ** 2, two_calls, "GEMV + GENERIC")
        report(f"HF G = 2J - K nbf={nbf}, fused J/K build", nb, lambda: core.fock_jk(g, D, K), "FOCK_JK (one pass over g)")
        del g, D, J, K
    for (o, v) in (((48, 256),) if large_only else ((24, 96), (48, 256))):
        t = torch.rand((o, o, v, v), dtype=torch.float64, device="cuda")
        gg = torch.rand((o, o, v, v), dtype=torch.float64, device="cuda")
        E = torch.zeros((), dtype=torch.float64, device="cuda")
        P = torch.empty((o, v, o, v), dtype=torch.float64, device="cuda")
        pe = core.compile_plan("", "ijab", "ijab")
        n = o * o * v * v
        report(f"E dot o={o} v={v}", 16 * n, lambda: pe.execute(0.0, E, 0.25, t, gg), "DOT")
        report(f"permute ijab->iajb o={o} v={v}", 16 * n, lambda: core.permute(0.0, "iajb", P, 1.0, "ijab", t), "PERMUTE")
        # unit-stride axes differ (b in A, f in C): the LDS-tiled transpose
        P2 = P.view(o, v, o, v)
        report(f"permute jnfb->jbnf o={o} v={v} (tiled transpose)", 16 * n,
               lambda: core.permute(0.0, "jbnf", P2, 1.0, "jnfb", t), "PERMUTE")
        report(f"permute jnfb->jbnf o={o} v={v}, beta=1", 24 * n,
               lambda: core.permute(1.0, "jbnf", P2, 1.0, "jnfb", t), "PERMUTE")
        if o == 48:
            report(f"copy {8 * n} B", 16 * n, lambda: core.permute(0.0, "ijab", P.view(o, o, v, v), 1.0, "ijab", t),
                   "PERMUTE (identity)")
        del t, gg, E, P
        torch.cuda.empty_cache()
    return out


def wmbej_case(timed, o=48, v=256):
    """Wmbej = t_jnfb g_mnef at o=48, v=256 (a generic contraction: 12288^3 after packing the
    operands into GEMM layout) through the einsum plan, plain FP64 vs FP32 operands vs the
    split-precision modes. Errors: relative Frobenius against the FP64 result, and against a
    host sum over the links for 64 sampled entries."""
    import numpy as np
    import torch

    import einsums_amd.core as core

    g = torch.Generator(device="cuda").manual_seed(4321)
    t = (torch.rand((o, o, v, v), dtype=torch.float64, device="cuda", generator=g) * 2 - 1) * 0.1
    gg = (torch.rand((o, o, v, v), dtype=torch.float64, device="cuda", generator=g) * 2 - 1) * 0.1
    plan = core.compile_plan("mbej", "jnfb", "mnef")
    W64 = torch.empty((o, v, v, o), dtype=torch.float64, device="cuda")
    t64 = timed(lambda: plan.execute(0.0, W64, 1.0, t, gg))
    flop = 2.0 * o ** 3 * v ** 3
    rng = np.random.default_rng(5)
    idx = [tuple(int(rng.integers(0, x)) for x in (o, v, v, o)) for _ in range(64)]
    th, gh = t.cpu().numpy(), gg.cpu().numpy()
    ref = np.array([np.sum(th[j, :, :, b] * gh[m, :, e, :]) for (m, b, e, j) in idx])

    def sample_err(W):
        got = np.array([float(W[m, b, e, j]) for (m, b, e, j) in idx])
        return float(f"{np.linalg.norm(got - ref) / np.linalg.norm(ref):.3e}")

    row = {"case": f"Wmbej o={o} v={v}", "m": o * v, "n": o * v, "k": o * v, "fp64_ms": round(t64, 3),
           "fp64_tflops": round(flop / t64 / 1e9, 2), "fp64_rel_err_rows_vs_host": sample_err(W64.cpu())}
    t32, g32 = t.float(), gg.float()
    W32 = torch.empty((o, v, v, o), dtype=torch.float32, device="cuda")
    ms32 = timed(lambda: plan.execute(0.0, W32, 1.0, t32, g32))
    row["fp32"] = {"ms": round(ms32, 3), "speedup": round(t64 / ms32, 3),
                   "rel_err_vs_fp64": float(f"{float(torch.linalg.norm(W32.double() - W64) / torch.linalg.norm(W64)):.3e}")}
    del t32, g32, W32
    Ws = torch.empty_like(W64)
    for st, ns in SPLIT_MODES:
        ts = timed(lambda: plan.execute(0.0, Ws, 1.0, t, gg, split=ns, slice_type=st))
        err = float(torch.linalg.norm(Ws - W64) / torch.linalg.norm(W64))
        row[split_key(st, ns)] = {"ms": round(ts, 3), "speedup": round(t64 / ts, 3), "rel_err_vs_fp64": float(f"{err:.3e}"),
                             "rel_err_rows_vs_host": sample_err(Ws.cpu())}
    del t, gg, W64, Ws
    torch.cuda.empty_cache()
    print(json.dumps(row), flush=True)
    return row


def out_of_core_bench(dirpath: str, reps: int = 3) -> None:
    """Wabef += 0.25 t_mnab g_mnef (FP64, beta = 1) with C resident in HBM, in pinned host
    memory (core.gemm_host_c, panels over PCIe) and in an HDF5 disk tensor of shape
    (v, v, v, v) under `dirpath` (disk.gemm_disk_c, panels disk -> host -> device and back).
    The time of each is the median of `reps` calls after one warm-up; `c_gbs` is C's read
    plus write bytes over that time. The streamed results are compared with the HBM one."""
    import numpy as np
    import torch

    import einsums_amd.core as core
    from einsums_amd import disk

    for o, v in ((24, 96), (32, 128)):
        rng = np.random.default_rng(3)
        t = torch.from_numpy(rng.uniform(-1, 1, (o * o, v * v))).cuda()
        g = torch.from_numpy(rng.uniform(-1, 1, (o * o, v * v))).cuda()
        W0 = torch.from_numpy(rng.uniform(-1, 1, (v * v, v * v)))
        c_bytes = 2 * W0.numel() * 8
        flop = 2.0 * (v * v) ** 2 * o * o

        def timed(fn):
            fn()
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                fn()
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            return float(np.median(ts))

        Wd = W0.cuda()
        plan = core.compile_plan("abef", "mnab", "mnef")
        t4, g4 = t.view(o, o, v, v), g.view(o, o, v, v)
        W4 = Wd.view(v, v, v, v)
        t_hbm = timed(lambda: plan.execute(1.0, W4, 0.25, t4, g4))
        Wd.copy_(W0.cuda())
        plan.execute(1.0, W4, 0.25, t4, g4)
        ref = Wd.cpu().numpy()
        Wh = torch.empty_like(W0).pin_memory()

        def host_call():
            core.gemm_host_c("T", "N", 0.25, t, g, 1.0, Wh)

        t_host = timed(host_call)
        Wh.copy_(W0)
        host_call()
        err_host = float(np.linalg.norm(Wh.numpy() - ref) / np.linalg.norm(ref))
        path = Path(dirpath) / f"wabef_o{o}_v{v}.h5"
        if path.exists():
            path.unlink()
        Wk = disk.DiskTensor(path, "Wabef", (v, v, v, v))

        def disk_call():
            disk.gemm_disk_c("T", "N", 0.25, t, g, 1.0, Wk)

        Wk.put(W0.numpy().reshape(v, v, v, v))
        t_disk = timed(disk_call)
        Wk.put(W0.numpy().reshape(v, v, v, v))
        disk_call()
        err_disk = float(np.linalg.norm(Wk.get().reshape(v * v, v * v) - ref) / np.linalg.norm(ref))
        path.unlink()
        row = {"case": f"Wabef o={o} v={v} FP64 beta=1", "c_mb": round(W0.numel() * 8 / 1e6, 1),
               "hbm_ms": round(t_hbm * 1e3, 2), "hbm_tflops": round(flop / t_hbm / 1e12, 2),
               "host_ms": round(t_host * 1e3, 2), "host_c_gbs": round(c_bytes / t_host / 1e9, 2),
               "host_rel_err_vs_hbm": float(f"{err_host:.2e}"),
               "disk_ms": round(t_disk * 1e3, 2), "disk_c_gbs": round(c_bytes / t_disk / 1e9, 2),
               "disk_rel_err_vs_hbm": float(f"{err_disk:.2e}"), "disk_dir": str(dirpath)}
        print(json.dumps({"out_of_core": row}), flush=True)
        del t, g, W0, Wd, Wh


def grouped_bench(reps=5):
    """Batched uneven blocks: (1) per-irrep CCD W intermediates of 8 irreps (o_h in [2,24],
    v_h in [8,96]) as one grouped plan vs one einsum call per block (the reference's block
    loop); (2) 4096 GEMMs with m, n, k log-uniform in [4,512], FP64 and FP32, autotuned
    grouped plan vs per-GEMM calls; (3) a CCD iteration (grouped W intermediates, energy,
    permute) captured in a hipGraph and replayed."""
    import numpy as np
    import torch

    import einsums_amd.core as core
    from einsums_amd import blocks

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    rng = np.random.default_rng(2024)
    o = [int(x) for x in rng.integers(2, 25, 8)]
    v = [int(x) for x in rng.integers(8, 97, 8)]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    terms, flop = [], 0.0
    for h in range(8):
        oo, vv = o[h], v[h]
        t = dev(rng.uniform(-0.01, 0.01, (oo, oo, vv, vv)))
        g = dev(rng.uniform(-1, 1, (oo, oo, vv, vv)))
        terms += [("mnij", dev(rng.uniform(-1, 1, (oo,) * 4)), "ijef", t, "mnef", g, 1.0, 0.25),
                  ("abef", dev(rng.uniform(-1, 1, (vv,) * 4)), "mnef", g, "mnab", t, 1.0, 0.25),
                  ("mbej", dev(rng.uniform(-1, 1, (oo, vv, vv, oo))), "jnfb", t, "mnef", g, 1.0, -0.5)]
        flop += 2.0 * (oo ** 4 * vv ** 2 + vv ** 4 * oo ** 2 + oo ** 3 * vv ** 3)
    plan = blocks.GroupedEinsum(terms, autotune=True)
    t_grouped = timed(plan.execute)

    def loop():
        for ci, C, ai, A, bi, B, cp, abp in terms:
            core.einsum(ci, C, ai, A, bi, B, C_prefactor=cp, AB_prefactor=abp)

    t_loop = timed(loop)
    res = {"ccd_irreps": {"o_h": o, "v_h": v, "gflop": round(flop / 1e9, 3), "grouped_ms": round(t_grouped, 4),
                          "per_block_loop_ms": round(t_loop, 4), "speedup": round(t_loop / t_grouped, 2),
                          "grouped_tflops": round(flop / t_grouped / 1e9, 2), "info": plan.info()}}

    # CCD iteration in a hipGraph: grouped W intermediates + energy + permute of irrep 0
    s = torch.cuda.Stream()
    E = torch.zeros((), dtype=torch.float64, device="cuda")
    t0, g0 = terms[0][3], terms[0][5]
    P = torch.empty((t0.shape[0], t0.shape[2], t0.shape[1], t0.shape[3]), dtype=torch.float64, device="cuda")

    def iteration():
        plan.execute(s)
        core.einsum("", E, "ijab", t0, "ijab", g0, AB_prefactor=0.25, stream=s)
        core.permute(0.0, "iajb", P, 1.0, "ijab", t0, stream=s)

    with torch.cuda.stream(s):
        iteration()
    s.synchronize()
    graph = blocks.Graph(iteration, s)
    with torch.cuda.stream(s):
        t_eager = timed(iteration)
        t_graph = timed(graph.replay)
    res["ccd_iteration_graph"] = {"eager_ms": round(t_eager, 4), "graph_replay_ms": round(t_graph, 4)}

    for dtype in (torch.float64, torch.float32):
        rng = np.random.default_rng(7)
        shapes = np.exp(rng.uniform(np.log(4), np.log(512), (4096, 3))).astype(int)
        gt, gflop = [], 0.0
        for (m, n, k) in shapes:
            A = torch.rand((m, k), dtype=dtype, device="cuda")
            B = torch.rand((k, n), dtype=dtype, device="cuda")
            C = torch.empty((m, n), dtype=dtype, device="cuda")
            gt.append(("ij", C, "ik", A, "kj", B, 0.0, 1.0))
            gflop += 2.0 * m * n * k
        gp = blocks.GroupedEinsum(gt, autotune=True)
        tg = timed(gp.execute)

        def loop2():
            for ci, C, ai, A, bi, B, cp, abp in gt:
                core.einsum(ci, C, ai, A, bi, B, C_prefactor=cp, AB_prefactor=abp)

        tl = timed(loop2)
        res[f"gemm4096_{'f64' if dtype == torch.float64 else 'f32'}"] = {
            "gflop": round(gflop / 1e9, 3), "grouped_ms": round(tg, 3), "per_gemm_loop_ms": round(tl, 3),
            "speedup": round(tl / tg, 2), "grouped_tflops": round(gflop / tg / 1e9, 2), "info": gp.info()}
    return res


def tune(path: str) -> None:
    """GEMM tuning pass: the FP64 GEMM-mapped and packed-generic contractions of the CCD step
    at o=24, v=96 and o=48, v=256, each run once in tuning mode; the table is saved."""
    import torch

    import einsums_amd.core as core
    from einsums_amd import workloads

    for o, v in ((24, 96), (48, 256)):
        I = workloads.ccd_inputs_device(o, v, seed=11, with_vvvv=True)
        arrs = {"t_oovv": I.t_oovv, "g_oovv": I.g_oovv, "g_oooo": I.g_oooo, "g_vvvv": I.g_vvvv, "g_ovvo": I.g_ovvo}
        with core.gemm_tuning():
            for name, Ci, An, Ai, Bn, Bi, cp, abp, init in workloads.CCD_CONTRACTIONS:
                W = arrs[init]
                core.einsum(Ci, W, Ai, arrs[An], Bi, arrs[Bn], C_prefactor=cp, AB_prefactor=abp)
                torch.cuda.synchronize()
                print(json.dumps({"tuned": name, "o": o, "v": v, "entries": int(core.lib.ein_gemm_tuning_count())}),
                      flush=True)
        del I, arrs
        torch.cuda.empty_cache()
    core.save_gemm_tuning(path)
    print(json.dumps({"saved": path, "entries": int(core.lib.ein_gemm_tuning_count())}))


def main():
    args = parse()
    if args.tune:
        tune(args.tune)
        return
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.precision_sweep:
        precision_sweep(cases=tuple(args.sweep_cases.split(",")))
        return
    if args.bandwidth:
        bandwidth_bench()
        return
    if args.out_of_core:
        out_of_core_bench(args.out_of_core)
        return
    if args.grouped:
        print(json.dumps({"grouped": grouped_bench()}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only: the contraction path has no multi-GPU "
                                   "sharding (the output-tile sharding it would use is described, not implemented)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    import einsums_amd.core as core
    from einsums_amd import workloads

    if args.tuning_table and Path(args.tuning_table).exists():
        core.load_gemm_tuning(args.tuning_table)
    o, v = args.o, args.v
    I = workloads.ccd_inputs(o, v, seed=7)
    flops = workloads.ccd_flops(o, v)
    nbytes = workloads.ccd_bytes(o, v)
    dev = lambda a: torch.from_numpy(a).to("cuda")
    t, g = dev(I.t_oovv), dev(I.g_oovv)
    W = {"Wmnij": dev(I.g_oooo), "Wabef": dev(I.g_vvvv), "Wmbej": dev(I.g_ovvo)}
    E = torch.zeros((), dtype=torch.float64, device="cuda")
    P = torch.empty((o, v, o, v), dtype=torch.float64, device="cuda")
    plans = {
        "Wmnij": (core.compile_plan("mnij", "ijef", "mnef"), t, g, 0.25, 1.0),
        "Wabef": (core.compile_plan("abef", "mnef", "mnab"), g, t, 0.25, 1.0),
        "Wmbej": (core.compile_plan("mbej", "jnfb", "mnef"), t, g, -0.5, 1.0),
        "E": (core.compile_plan("", "ijab", "ijab"), t, g, 0.25, 0.0),
    }
    names = ["Wmnij", "Wabef", "Wmbej", "E", "permute"]
    ev = {n: [] for n in names}

    def step(record):
        for n in names:
            if record:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            if n == "permute":
                core.permute(0.0, "iajb", P, 1.0, "ijab", t)
            else:
                plan, A, B, abp, cp = plans[n]
                C = E if n == "E" else W[n]
                plan.execute(cp, C, abp, A, B)
            if record:
                e1.record()
                ev[n].append((e0, e1))

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    # timed region: K steps bracketed by device synchronisation
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = {n: float(np.mean([a.elapsed_time(b) for a, b in ev[n]])) for n in names}  # ms
    ms_per_step = wall * 1e3 / args.steps
    total_flop = sum(flops.values())
    algs = {n: core.algorithm_name(plans[n][0].last_algorithm) for n in names if n in plans}

    dom = "Wabef"
    achieved = flops[dom] / (per[dom] * 1e-3) / 1e12
    traffic = None
    if args.traffic and Path(args.traffic).exists() and (o, v) == (24, 96):
        # FETCH_SIZE + WRITE_SIZE of the Wabef kernel, each scaled by a 1 GiB copy's reading
        for line in Path(args.traffic).read_text().splitlines():
            if line.startswith("Wabef_bytes_per_launch:"):
                traffic = int(line.split(":")[1])
    kernels = {}
    for n in names:
        gf = flops[n] / (per[n] * 1e-3) / 1e9 if flops[n] else 0.0
        gbs = nbytes[n] / (per[n] * 1e-3) / 1e9
        bound = "mfma" if n in ("Wmnij", "Wabef", "Wmbej") else "hbm"
        frac = gf / 1e3 / FP64_MFMA_PEAK_TFLOPS if bound == "mfma" else gbs / HBM_PEAK_GBS
        kernels[n] = {"ms": round(per[n], 4), "gflops": round(gf, 1), "gbs": round(gbs, 1), "bound": bound,
                      "frac": round(frac, 4), "algorithm": algs.get(n, "PERMUTE")}
    out = {
        "metric": "einsum GFLOP/s & % of MFMA/HBM peak; split-prec vs plain FP64 at <=1e-11 rel err",
        "value": round(total_flop / (ms_per_step * 1e-3) / 1e9, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"CCD W intermediates o={o} v={v} FP64: Wmnij, Wabef, Wmbej, E, permute ijab->iajb",
                   "o": o, "v": v, "step_gflop": round(total_flop / 1e9, 3)},
        "roofline": {"bound": "mfma", "kernel": "Wabef (gemm_dense_kernel, FP64 MFMA)", "achieved": round(achieved, 3),
                     "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4),
                     "traffic": traffic},
        "kernels": kernels,
    }
    prof = rocprof_figure(args.kernel_trace, flops[dom]) if (o, v) == (24, 96) else None
    if prof is not None:
        out["roofline"]["rocprof"] = prof
    hf = hf_block(core)
    out["hf"] = hf
    if args.hbm:
        # the HBM-bound kernels above the 256 MB cache, cold and warm (--bandwidth has all sizes)
        rows = bandwidth_bench(reps=10, quiet=True, large_only=True)
        torch.cuda.empty_cache()
        out["hbm"] = {"workload": "HBM-bound kernels above the 256 MB cache: HF J/K/fused G at nbf=96, E dot, "
                                  "permutes and a copy at o=48 v=256 (10 cold calls each, 512 MB scratch read between)",
                      "peak_TBs": HBM_PEAK_GBS / 1e3,
                      "cases": {r["case"]: {"cold_us": r["cold_us"], "cold_frac": r["cold_frac"],
                                            "warm_frac": round(r["warm_TBs"] / 8.0, 3)} for r in rows}}
    large = None
    if args.split:
        large = split_large(core)
        out["split"] = large
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(I, o, v, flops, large, hf)
    else:
        hf.pop("_inputs")
        for key in ("two_einsum_calls", "fused_jk"):
            hf[key].pop("_got")
        if large is not None:
            for k in ("t_host", "g_host"):
                large.pop(k)
            for row in large["contractions"].values():
                row.pop("_rows")
                row.pop("_got")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
