"""Synthetic FCIDUMP generators for the CDFCI benchmark and test workloads.

The files written here use the FCIDUMP dialect that the reference parser
accepts (src/input.h:93-285): a namelist header ending with ``&END`` on its
own line, then ``value i a j b`` lines with 1-based spatial indices, C-style
exponents, unique 8-fold two-electron integrals ``(ia|jb)``, one-electron
integrals ``h_ia`` (``j = b = 0``) and a final core-energy line (all zero
indices).  No orbital energies are written, so the reference's HF rule
(src/hamiltonian.h:447-492) takes the lowest spin-orbitals in index order.

Molecule-like generator (all draws from ``numpy.random.default_rng(seed)``,
PCG64, in exactly this order):

1. off-diagonal one-electron integrals: ``norb*(norb-1)/2`` draws of
   ``N(0, 1)`` in row-major order over ``p < q``, scaled by ``h_offdiag_std``
   (std 0.05, times ``h_offdiag_scale``); ``h_qp = h_pq``;
   diagonal ``h_pp = h_diag_a + h_diag_b * p`` (``p`` 0-based);
2. for each factor ``L = 0 .. rank-1``: ``norb*(norb+1)/2`` draws of
   ``N(0, 1)`` in row-major order over ``p <= q``, scaled by
   ``eri_std / (1 + |p - q|)`` (std 0.3), giving a symmetric ``B^L``;
3. ``(pq|rs) = sum_L B^L_pq B^L_rs`` (8-fold symmetric, positive
   semidefinite as a matrix over pairs).

The CDFCI update takes ``sqrt(-lambda)`` of the subspace eigenvalue
(src/solver.h:524,587-591), so the Hamiltonian, core energy included, must
have a negative spectrum.  ``core="auto"`` keeps the core energy 0 when the
closed-shell HF energy is below -5 Ha and otherwise shifts it to
``-ceil(E_HF + 10)``; the chosen value is written to the file and returned.
"""

from __future__ import annotations

import hashlib
import math
import os
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "SyntheticSpec",
    "CONFIGS",
    "molecule_integrals",
    "hubbard_integrals",
    "hf_energy_closed_shell",
    "write_fcidump",
    "write_input_json",
    "generate",
]


@dataclass
class SyntheticSpec:
    """One synthetic workload: integrals plus the solver settings."""

    name: str
    norb: int
    nelec: int
    ms2: int = 0
    rank: int = 14
    seed: int = 0
    h_diag_a: float = -3.0
    h_diag_b: float = 0.4
    h_offdiag_std: float = 0.05
    h_offdiag_scale: float = 1.0
    eri_std: float = 0.3
    core: object = "auto"          # float or "auto"
    kind: str = "molecule"         # "molecule" or "hubbard"
    hubbard_t: float = 1.0
    hubbard_u: float = 4.0
    solver: dict = field(default_factory=dict)


# The five workloads of the benchmark description (norb = spatial orbitals).
CONFIGS = {
    "c1": SyntheticSpec("c1", norb=7, nelec=10, rank=14, seed=0, core=9.0,
                        solver=dict(num_iterations=20000, report_interval=1000,
                                    num_of_coordinate=8, estimate=8,
                                    z_threshold=0.0, max_memory=0.005)),
    "c2": SyntheticSpec("c2", norb=16, nelec=16, rank=32, seed=1,
                        solver=dict(num_iterations=200000, report_interval=1000,
                                    num_of_coordinate=8, estimate=8,
                                    z_threshold=1e-7, max_memory=8.0)),
    "c3_hubbard": SyntheticSpec("c3_hubbard", norb=12, nelec=12, kind="hubbard",
                                solver=dict(num_iterations=200000, report_interval=1000,
                                            num_of_coordinate=8, estimate=8,
                                            z_threshold=0.0, max_memory=1.0)),
    "c3_synth": SyntheticSpec("c3_synth", norb=12, nelec=12, rank=24, seed=2,
                              solver=dict(num_iterations=200000, report_interval=1000,
                                          num_of_coordinate=8, estimate=8,
                                          z_threshold=0.0, max_memory=1.0)),
    "c4": SyntheticSpec("c4", norb=24, nelec=24, rank=48, seed=3, h_offdiag_scale=0.5,
                        solver=dict(num_iterations=100000, report_interval=1000,
                                    num_of_coordinate=64, estimate=64,
                                    z_threshold=1e-6, max_memory=64.0)),
    "c5": SyntheticSpec("c5", norb=42, nelec=48, rank=84, seed=4, h_diag_a=-5.0,
                        h_diag_b=0.25,
                        solver=dict(num_iterations=1000000, report_interval=1000,
                                    num_of_coordinate=256, estimate=256,
                                    z_threshold=1e-5, max_memory=200.0)),
}


def molecule_integrals(spec: SyntheticSpec):
    """Return (h[norb,norb], eri[norb,norb,norb,norb]) in chemists' order."""
    n = spec.norb
    rng = np.random.default_rng(spec.seed)
    h = np.zeros((n, n))
    iu = np.triu_indices(n, 1)
    off = rng.normal(0.0, 1.0, size=len(iu[0])) * (spec.h_offdiag_std * spec.h_offdiag_scale)
    h[iu] = off
    h = h + h.T
    h[np.arange(n), np.arange(n)] = spec.h_diag_a + spec.h_diag_b * np.arange(n)
    iu0 = np.triu_indices(n, 0)
    std = spec.eri_std / (1.0 + np.abs(iu0[0] - iu0[1]))
    B = np.zeros((spec.rank, n, n))
    for L in range(spec.rank):
        v = rng.normal(0.0, 1.0, size=len(iu0[0])) * std
        B[L][iu0] = v
        B[L] = B[L] + np.triu(B[L], 1).T
    eri = np.einsum("lpq,lrs->pqrs", B, B, optimize=True)
    # Make the array bitwise 8-fold symmetric (fl(a+b) == fl(b+a)), so that
    # in-memory checks agree with what the FCIDUMP round trip gives.
    eri = 0.5 * (eri + eri.transpose(1, 0, 2, 3))
    eri = 0.5 * (eri + eri.transpose(0, 1, 3, 2))
    eri = 0.5 * (eri + eri.transpose(2, 3, 0, 1))
    return h, eri


def hubbard_integrals(spec: SyntheticSpec):
    """Periodic 1D Hubbard ring in the site basis: h_{i,i+-1} = -t, (ii|ii) = U."""
    n = spec.norb
    h = np.zeros((n, n))
    for i in range(n):
        j = (i + 1) % n
        h[i, j] = h[j, i] = -spec.hubbard_t
    eri = np.zeros((n, n, n, n))
    for i in range(n):
        eri[i, i, i, i] = spec.hubbard_u
    return h, eri


def hf_energy_closed_shell(h, eri, nalpha, nbeta, core=0.0):
    """Energy of the determinant with the lowest nalpha/nbeta spatial orbitals."""
    ea = list(range(nalpha))
    eb = list(range(nbeta))
    e = core + sum(h[p, p] for p in ea) + sum(h[p, p] for p in eb)
    for i, p in enumerate(ea):
        for q in ea[i + 1:]:
            e += eri[p, p, q, q] - eri[p, q, q, p]
    for i, p in enumerate(eb):
        for q in eb[i + 1:]:
            e += eri[p, p, q, q] - eri[p, q, q, p]
    for p in ea:
        for q in eb:
            e += eri[p, p, q, q]
    return float(e)


def _auto_core(h, eri, nalpha, nbeta):
    e0 = hf_energy_closed_shell(h, eri, nalpha, nbeta, 0.0)
    if e0 < -5.0:
        return 0.0
    return -float(math.ceil(e0 + 10.0))


def write_fcidump(path, h, eri, nelec, ms2, core, tol=0.0):
    """Write an RHF FCIDUMP with unique 8-fold ERIs (i>=a, j>=b, ia>=jb)."""
    n = h.shape[0]
    lines = []
    lines.append(" &FCI NORB=%d,NELEC=%d,MS2=%d," % (n, nelec, ms2))
    lines.append("  ORBSYM=" + ",".join(["1"] * n) + ",")
    lines.append("  ISYM=1,")
    lines.append("  UHF=.FALSE.,")
    lines.append(" &END")
    fmt = "%25.17E %4d %4d %4d %4d"
    for i in range(n):
        for a in range(i + 1):
            ia = i * (i + 1) // 2 + a
            for j in range(n):
                for b in range(j + 1):
                    jb = j * (j + 1) // 2 + b
                    if jb > ia:
                        continue
                    v = eri[i, a, j, b]
                    if abs(v) > tol:
                        lines.append(fmt % (v, i + 1, a + 1, j + 1, b + 1))
    for i in range(n):
        for a in range(i + 1):
            v = h[i, a]
            if abs(v) > tol:
                lines.append(fmt % (v, i + 1, a + 1, 0, 0))
    lines.append(fmt % (core, 0, 0, 0, 0))
    text = "\n".join(lines) + "\n"
    with open(path, "w") as f:
        f.write(text)
    return hashlib.sha256(text.encode()).hexdigest()


def write_input_json(path, fcidump_path, solver: dict, threshold=0.0):
    import json
    cd = {
        "coordinate_pick": "gcd_grad",
        "coordinate_update": "eig",
        "z_threshold_search": False,
    }
    cd.update(solver)
    doc = {"hamiltonian": {"fcidump_path": fcidump_path, "threshold": threshold},
           "solver": {"cdfci": cd}}
    with open(path, "w") as f:
        json.dump(doc, f, indent=4)


def generate(name_or_spec, outdir, **solver_overrides):
    """Write FCIDUMP + input.json for a named workload; return a summary dict."""
    spec = CONFIGS[name_or_spec] if isinstance(name_or_spec, str) else name_or_spec
    os.makedirs(outdir, exist_ok=True)
    if spec.kind == "hubbard":
        h, eri = hubbard_integrals(spec)
    else:
        h, eri = molecule_integrals(spec)
    na = (spec.nelec + spec.ms2) // 2
    nb = (spec.nelec - spec.ms2) // 2
    core = _auto_core(h, eri, na, nb) if spec.core == "auto" else float(spec.core)
    fcid = os.path.join(outdir, "FCIDUMP")
    sha = write_fcidump(fcid, h, eri, spec.nelec, spec.ms2, core)
    solver = dict(spec.solver)
    solver.update(solver_overrides)
    inp = os.path.join(outdir, "input.json")
    write_input_json(inp, fcid, solver)
    return {"name": spec.name, "fcidump": fcid, "input": inp, "sha256": sha, "core": core,
            "hf_energy": hf_energy_closed_shell(h, eri, na, nb, core), "h": h, "eri": eri}


if __name__ == "__main__":  # pragma: no cover
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("name", choices=sorted(CONFIGS))
    ap.add_argument("outdir")
    args = ap.parse_args()
    info = generate(args.name, args.outdir)
    print(info["fcidump"], info["sha256"], "core", info["core"], "E_HF", info["hf_energy"])
