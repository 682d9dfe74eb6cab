"""RHF + MP2 on water/STO-3G, written once over an einsum backend.

Follows the reference test QChemTests.cpp:185-415 step by step (same einsum calls,
same index strings, same convergence rule). S^-1/2 uses numpy on the host; the Fock
diagonalisation uses numpy too unless the backend brings its own ``syev`` (the GPU
backend's ``device_syev=True`` keeps it on the device, as the reference's syev keeps it in
LAPACK).
"""
from __future__ import annotations

import numpy as np


class NumpyOracleBackend:
    """CPU oracle backend (numpy arrays, oracle einsum)."""

    def __init__(self, oracle):
        self.o = oracle

    def zeros(self, shape):
        return np.zeros(shape)

    def from_np(self, a):
        return np.array(a, dtype=np.float64)

    def to_np(self, a):
        return np.array(a)

    def view(self, a, idx):
        return a[idx]

    def einsum(self, cp, ci, C, abp, ai, A, bi, B):
        return self.o.einsum(cp, ci, C, abp, ai, A, bi, B)


class TorchGpuBackend:
    """GPU backend: torch tensors on the device, einsums_amd compile_plan/execute."""

    def __init__(self, core, torch, device_syev=False):
        self.core = core
        self.torch = torch
        self.algorithms = []
        if device_syev:
            self.syev = self._syev

    def _syev(self, F):
        """Eigenvalues (host, for the MP2 denominators) and eigenvectors (device) of F."""
        vecs = F.clone()
        vals = self.torch.empty(F.shape[0], dtype=F.dtype, device=F.device)
        self.core.syev(vecs, vals)
        return vals.cpu().numpy(), vecs

    def zeros(self, shape):
        return self.torch.zeros(shape, dtype=self.torch.float64, device="cuda")

    def from_np(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")

    def to_np(self, a):
        self.torch.cuda.synchronize()
        return a.cpu().numpy()

    def view(self, a, idx):
        return a[idx]

    def einsum(self, cp, ci, C, abp, ai, A, bi, B):
        plan = self.core.compile_plan(ci, ai, bi)
        plan.execute(cp, C, abp, A, B)
        self.algorithms.append((ci, ai, bi, plan.last_algorithm))
        return plan.last_algorithm


def rhf_mp2(be, S, T, V, TEI, nocc=5, enuc=8.002367061810450):
    nbf = S.shape[0]
    H_np = T + V
    w, U = np.linalg.eigh(S)
    X_np = (U * w**-0.5) @ U.T  # pow(S, -0.5)
    H, X, G = be.from_np(H_np), be.from_np(X_np), be.from_np(TEI)
    temp1, Ft, C, D = be.zeros((nbf, nbf)), be.zeros((nbf, nbf)), be.zeros((nbf, nbf)), be.zeros((nbf, nbf))

    def diagonalize(F):
        be.einsum(0.0, "ij", temp1, 1.0, "ik", F, "kj", X)
        be.einsum(0.0, "ij", Ft, 1.0, "kj", temp1, "ki", X)
        if hasattr(be, "syev"):
            return be.syev(Ft)
        evals, evecs = np.linalg.eigh(be.to_np(Ft))
        return evals, be.from_np(evecs)

    def density(Ct, first):
        if first:
            be.einsum(0.0, "ij", C, 1.0, "kj", Ct, "ik", X)
        else:
            be.einsum(0.0, "ij", C, 1.0, "kj", Ct, "ki", X)
        Cocc = be.view(C, (slice(None), slice(0, nocc)))
        be.einsum(0.0, "ij", D, 1.0, "im", Cocc, "jm", Cocc)

    def energy(F):
        E = be.zeros(())
        be.einsum(0.0, "", E, 1.0, "ij", D, "ij", F)
        be.einsum(1.0, "", E, 1.0, "ij", D, "ij", H)
        return float(be.to_np(E)) + enuc

    F = be.from_np(H_np)
    evals, Ct = diagonalize(F)
    density(Ct, True)
    e0, e1, drms, cycles = energy(F), 0.0, 1.0, 0
    while abs(e0 - e1) > 1e-10 and drms > 1e-6 and cycles < 50:
        cycles += 1
        e1 = e0
        D_prev = be.to_np(D).copy()
        F = be.from_np(H_np)
        be.einsum(1.0, "ij", F, 2.0, "kl", D, "ijkl", G)
        be.einsum(1.0, "ij", F, -1.0, "kl", D, "ikjl", G)
        evals, Ct = diagonalize(F)
        density(Ct, False)
        e0 = energy(F)
        diff = be.from_np(be.to_np(D) - D_prev)
        rms = be.zeros(())
        be.einsum(0.0, "", rms, 1.0, "ij", diff, "ij", diff)
        drms = np.sqrt(float(be.to_np(rms))) / 7.0

    # MP2: AO -> MO transformation of the integrals, then the pair energies
    t1, t2 = be.zeros((nbf,) * 4), be.zeros((nbf,) * 4)
    be.einsum(0.0, "pqrs", t1, 1.0, "mqrs", G, "mp", C)
    be.einsum(0.0, "pqrs", t2, 1.0, "pmrs", t1, "mq", C)
    be.einsum(0.0, "pqrs", t1, 1.0, "pqms", t2, "mr", C)
    Gmo = be.zeros((nbf,) * 4)
    be.einsum(0.0, "pqrs", Gmo, 1.0, "pqrm", t1, "ms", C)
    o, v = slice(0, nocc), slice(nocc, nbf)
    iajb = be.view(Gmo, (o, v, o, v))  # strided view, as TEI(Range, Range, Range, Range)
    eo, ev = evals[:nocc], evals[nocc:]
    denom = 1.0 / (eo[:, None, None, None] + eo[None, None, :, None] - ev[None, :, None, None] - ev[None, None, None, :])
    scale = be.from_np(denom)
    amps = be.zeros((nocc, nbf - nocc, nocc, nbf - nocc))
    be.einsum(0.0, "iajb", amps, 1.0, "iajb", iajb, "iajb", scale)
    e_a, e_b = be.zeros(()), be.zeros(())
    be.einsum(0.0, "", e_a, 2.0, "iajb", amps, "iajb", iajb)
    be.einsum(0.0, "", e_b, -1.0, "iajb", amps, "ibja", iajb)
    emp2 = float(be.to_np(e_a)) + float(be.to_np(e_b))
    return {"E_HF": e0, "E_MP2": emp2, "E_total": e0 + emp2, "cycles": cycles}
