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

Unrestricted variant (``uhf=True``): the draws above give the alpha
integrals; then, from the same generator, ``norb*(norb+1)/2`` draws of
``N(0, 1)`` over ``p <= q`` scaled by ``uhf_delta`` give a symmetric
``D`` with ``h_beta = h_alpha + D``, and for each factor ``L`` another
``norb*(norb+1)/2`` draws scaled by ``uhf_delta * eri_std / (1 + |p - q|)``
give ``B^L_beta = B^L + E^L``.  The blocks are
``(pq|rs)_{st} = sum_L B^L_{s,pq} B^L_{t,rs}`` for spins s, t, which keeps the
spin-orbital tensor positive semidefinite.  Such a Hamiltonian is written in
the reference's unrestricted dialect (src/input.h:125-127,203-237): NORB
counts spin-orbitals, ``UHF=.TRUE.``, and every index is a 1-based
spin-orbital, ``2p + 1`` the alpha and ``2p + 2`` the beta function of
orbital ``p``.

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
    "uhf_integrals",
    "hf_energy_uhf",
    "write_fcidump_uhf",
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
    uhf: bool = False              # write the unrestricted dialect (spin-resolved integrals)
    uhf_delta: float = 0.05        # size of the beta - alpha differences (uhf only)
    rhf_as_uhf: bool = False       # restricted integrals written in the unrestricted dialect
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
    # unrestricted (spin-resolved) integrals: open shell with 7 orbitals,
    # closed shell with 12, and the norb=7 molecule written in the
    # unrestricted dialect (every block equal)
    "u1": SyntheticSpec("u1", norb=7, nelec=9, ms2=1, rank=14, seed=5, uhf=True,
                        solver=dict(num_iterations=20000, report_interval=1000,
                                    num_of_coordinate=8, estimate=8,
                                    z_threshold=0.0, max_memory=0.005)),
    "u3": SyntheticSpec("u3", norb=12, nelec=12, rank=24, seed=6, uhf=True,
                        solver=dict(num_iterations=200000, report_interval=1000,
                                    num_of_coordinate=8, estimate=8,
                                    z_threshold=0.0, max_memory=1.0)),
    "c1_as_uhf": SyntheticSpec("c1_as_uhf", norb=7, nelec=10, rank=14, seed=0, core=9.0, rhf_as_uhf=True,
                               solver=dict(num_iterations=20000, report_interval=1000,
                                           num_of_coordinate=8, estimate=8,
                                           z_threshold=0.0, max_memory=0.005)),
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


def uhf_integrals(spec: SyntheticSpec):
    """Return (h_alpha, h_beta, eri_aa, eri_bb, eri_ab) in chemists' order
    (eri_ab[p,q,r,s] = (p_a q_a|r_b s_b)); the alpha part equals
    molecule_integrals(spec)."""
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
    D = np.zeros((n, n))
    D[iu0] = rng.normal(0.0, 1.0, size=len(iu0[0])) * spec.uhf_delta
    D = D + np.triu(D, 1).T
    hb = h + D
    Bb = np.zeros_like(B)
    for L in range(spec.rank):
        E = np.zeros((n, n))
        E[iu0] = rng.normal(0.0, 1.0, size=len(iu0[0])) * (spec.uhf_delta * std)
        Bb[L] = B[L] + E + np.triu(E, 1).T

    def block(X, Y):
        e = np.einsum("lpq,lrs->pqrs", X, Y, optimize=True)
        e = 0.5 * (e + e.transpose(1, 0, 2, 3))
        return 0.5 * (e + e.transpose(0, 1, 3, 2))

    eaa, ebb, eab = block(B, B), block(Bb, Bb), block(B, Bb)
    eaa = 0.5 * (eaa + eaa.transpose(2, 3, 0, 1))
    ebb = 0.5 * (ebb + ebb.transpose(2, 3, 0, 1))
    return h, hb, eaa, ebb, eab


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


def hf_energy_uhf(ha, hb, eaa, ebb, eab, nalpha, nbeta, core=0.0):
    """Energy of the determinant with the lowest nalpha/nbeta orbitals of
    each spin, from spin-resolved integrals."""
    ea, eb = list(range(nalpha)), list(range(nbeta))
    e = core + sum(ha[p, p] for p in ea) + sum(hb[p, p] for p in eb)
    for i, p in enumerate(ea):
        for q in ea[i + 1:]:
            e += eaa[p, p, q, q] - eaa[p, q, q, p]
    for i, p in enumerate(eb):
        for q in eb[i + 1:]:
            e += ebb[p, p, q, q] - ebb[p, q, q, p]
    for p in ea:
        for q in eb:
            e += eab[p, p, q, q]
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


def write_fcidump_uhf(path, ha, hb, eaa, ebb, eab, nelec, ms2, core, tol=0.0):
    """Write spin-resolved integrals in the unrestricted dialect: NORB =
    2 n spin-orbitals, 1-based spin-orbital indices (2p + 1 alpha, 2p + 2
    beta); same-spin blocks as unique 8-fold (ia|jb), the alpha-beta block
    as (i_a a_a|j_b b_b) with i >= a and j >= b."""
    n = ha.shape[0]
    al = lambda p: 2 * p + 1  # noqa: E731
    be = lambda p: 2 * p + 2  # noqa: E731
    lines = [" &FCI NORB=%d,NELEC=%d,MS2=%d," % (2 * n, nelec, ms2),
             "  ORBSYM=" + ",".join(["1"] * (2 * n)) + ",", "  ISYM=1,", "  UHF=.TRUE.,", " &END"]
    fmt = "%25.17E %4d %4d %4d %4d"
    for eri, f in ((eaa, al), (ebb, be)):
        for i in range(n):
            for a in range(i + 1):
                ia = i * (i + 1) // 2 + a
                for j in range(n):
                    for b in range(j + 1):
                        if j * (j + 1) // 2 + b > ia:
                            continue
                        v = eri[i, a, j, b]
                        if abs(v) > tol:
                            lines.append(fmt % (v, f(i), f(a), f(j), f(b)))
    for i in range(n):
        for a in range(i + 1):
            for j in range(n):
                for b in range(j + 1):
                    v = eab[i, a, j, b]
                    if abs(v) > tol:
                        lines.append(fmt % (v, al(i), al(a), be(j), be(b)))
    for hs, f in ((ha, al), (hb, be)):
        for i in range(n):
            for a in range(i + 1):
                v = hs[i, a]
                if abs(v) > tol:
                    lines.append(fmt % (v, f(i), f(a), 0, 0))
    lines.append(fmt % (core, 0, 0, 0, 0))
    text = "\n".join(lines) + "\n"
    with open(path, "w") as fh:
        fh.write(text)
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
    na = (spec.nelec + spec.ms2) // 2
    nb = (spec.nelec - spec.ms2) // 2
    fcid = os.path.join(outdir, "FCIDUMP")
    if spec.uhf:
        blocks = uhf_integrals(spec)
        e0 = hf_energy_uhf(*blocks, na, nb, 0.0)
        core = (0.0 if e0 < -5.0 else -float(math.ceil(e0 + 10.0))) if spec.core == "auto" else float(spec.core)
        sha = write_fcidump_uhf(fcid, *blocks, spec.nelec, spec.ms2, core)
        h, eri, e_hf = None, None, hf_energy_uhf(*blocks, na, nb, core)
    else:
        if spec.kind == "hubbard":
            h, eri = hubbard_integrals(spec)
        else:
            h, eri = molecule_integrals(spec)
        core = _auto_core(h, eri, na, nb) if spec.core == "auto" else float(spec.core)
        if spec.rhf_as_uhf:
            blocks = (h, h, eri, eri, eri)
            sha = write_fcidump_uhf(fcid, *blocks, spec.nelec, spec.ms2, core)
        else:
            blocks = None
            sha = write_fcidump(fcid, h, eri, spec.nelec, spec.ms2, core)
        e_hf = hf_energy_closed_shell(h, eri, na, nb, core)
    solver = dict(spec.solver)
    solver.update(solver_overrides)
    inp = os.path.join(outdir, "input.json")
    write_input_json(inp, fcid, solver)
    return {"name": spec.name, "fcidump": fcid, "input": inp, "sha256": sha, "core": core,
            "hf_energy": e_hf, "h": h, "eri": eri, "uhf_blocks": blocks}


if __name__ == "__main__":  # pragma: no cover
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("name", choices=sorted(CONFIGS))
    ap.add_argument("outdir")
    args = ap.parse_args()
    info = generate(args.name, args.outdir)
    print(info["fcidump"], info["sha256"], "core", info["core"], "E_HF", info["hf_energy"])
