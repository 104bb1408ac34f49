"""Pins the oracle's energy gradient (the adjoint of the variational energy
with respect to the integrals) without using its code:

* the spin-summed 1-RDM and 2-RDM of the same vector by second quantisation
  (tests/ed.py), folded onto the unique FCIDUMP parameters;
* Euler's identity E = core + sum theta dE/dtheta against the dense
  second-quantised Hamiltonian;
* Hellmann-Feynman: at the exact ground state the gradient equals central
  finite differences (step 1e-5) of the exact ground-state energy with the
  integral perturbed over its whole 8-fold (or 2-fold) orbit.
"""

import numpy as np
import pytest

import ed
from oracle_ctypes import Oracle
from mcdfci_amd import synth


def _system(path):
    norb, nelec, ms2, h, eri, core = ed.read_fcidump(path)
    na, nb = (nelec + ms2) // 2, (nelec - ms2) // 2
    dets, H = ed.hamiltonian(h, eri, core, na, nb)
    return norb, na, nb, h, eri, core, dets, H


@pytest.mark.parametrize("density", [1.0, 0.3])
def test_gradient_equals_folded_rdms(workloads, density):
    path = workloads("c1")["fcidump"]
    norb, na, nb, h, eri, core, dets, H = _system(path)
    rng = np.random.default_rng(7)
    keep = rng.random(len(dets)) < density
    keep[0] = True
    c = np.where(keep, rng.normal(size=len(dets)), 0.0)
    o = Oracle(path)
    e, gh, ge = o.energy_gradient(dets[keep].reshape(-1, 1), c[keep] * 3.0)
    g1, G2 = ed.rdm12(dets, c, norb)
    rh, re = ed.fold_gradient(g1, G2)
    assert np.allclose(gh, rh, atol=1e-12, rtol=0)
    assert np.allclose(ge, re, atol=1e-12, rtol=0)
    # only canonical positions are populated
    for p, q, r, s in zip(*np.nonzero(ge)):
        assert (p, q, r, s) == ed.canonical(p, q, r, s)
    # Euler: E(c) = c.Hc / c.c
    assert abs(e - c @ H @ c / (c @ c)) < 1e-11
    # traces of the RDMs: tr g = N, sum_pr G_pprr = N (N - 1)
    n = na + nb
    assert abs(np.trace(g1) - n) < 1e-12
    assert abs(np.einsum("pprr->", G2) - n * (n - 1)) < 1e-11


def _perturbed_ground(h, eri, core, na, nb, which, idx, step):
    h2, e2 = h.copy(), eri.copy()
    if which == "h":
        p, q = idx
        h2[p, q] += step
        if p != q:
            h2[q, p] += step
    else:
        for t in ed.orbit(*idx):
            e2[t] += step
    _, H = ed.hamiltonian(h2, e2, core, na, nb)
    return np.linalg.eigvalsh(H)[0]


@pytest.mark.parametrize("name", ["c1", "hub6"])
def test_hellmann_feynman_finite_differences(workloads, tmp_path, name):
    if name == "hub6":
        spec = synth.SyntheticSpec("hub6", norb=6, nelec=6, kind="hubbard")
        h, eri = synth.hubbard_integrals(spec)
        path = str(tmp_path / "FCIDUMP")
        synth.write_fcidump(path, h, eri, 6, 0, -20.0)
    else:
        path = workloads(name)["fcidump"]
    norb, na, nb, h, eri, core, dets, H = _system(path)
    w, V = np.linalg.eigh(H)
    c = V[:, 0]
    o = Oracle(path)
    e, gh, ge = o.energy_gradient(dets.reshape(-1, 1), c)
    assert abs(e - w[0]) < 1e-11
    rng = np.random.default_rng(3)
    params = [("h", (0, 0)), ("h", (1, 0)), ("h", (norb - 1, 2)), ("e", (0, 0, 0, 0)), ("e", (1, 0, 1, 0)),
              ("e", (2, 1, 1, 0)), ("e", (3, 2, 1, 0))]
    for _ in range(3):
        p, q, r, s = rng.integers(0, norb, 4)
        params.append(("e", ed.canonical(p, q, r, s)))
    step = 1e-5
    for which, idx in params:
        ep = _perturbed_ground(h, eri, core, na, nb, which, idx, step)
        em = _perturbed_ground(h, eri, core, na, nb, which, idx, -step)
        fd = (ep - em) / (2 * step)
        an = gh[idx] if which == "h" else ge[idx]
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (which, idx, fd, an)
