"""The adjoint on the workloads the gradient is specified on -- the
half-filled 12-site Hubbard ring (t = 1, U = 4, periodic, with its negative
core shift) and the synthetic norb-12 molecule (seed 2, ERI rank 24), each a
sector of 853,776 determinants -- and, converged, on smaller Hubbard rings.

On the norb-12 workloads, after long plain FP64 runs on the device (tens of
thousands of iterations, several hundred thousand determinants with x != 0):
  * the device gradient (mcdfci_energy_gradient) equals the oracle's FP64
    1-RDM/2-RDM contraction of the same wavefunction bit for bit (terms of
    src/hamiltonian.h:252-410), and its energy equals the FP64 Rayleigh
    quotient evaluated by the column generator (mcdfci_energy_fp64);
  * at that fixed wavefunction the energy is linear in the integrals, so
    central differences (step 1e-5) of mcdfci_energy_fp64 under perturbed
    integrals -- a different code path (Slater-Condon columns) -- reproduce
    the gradient along one random symmetric direction, and, on the Hubbard
    ring, dE/dU = sum_i dE/d(ii|ii), to 1e-8 relative.
Converging these 853,776-dimensional runs to the ~1e-12 Ha that central
differences of re-converged energies need takes about 3e5 iterations per
solve on the device (measured: the energy still moves by 2e-5 Ha per 2000
iterations after 6e4 iterations at k = 64), so the Hellmann-Feynman check
with re-converged energies runs on the 6- and 8-site rings, whose runs reach
the exact ground state (dense second quantisation): central differences
(step 1e-5) of the device's re-converged energy along a random symmetric
direction and in U match the gradient to 1e-6 relative.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import ed
from mcdfci_amd import Cdfci, synth
from oracle_ctypes import Oracle
from torch_adjoint import symmetrise

STEP = 1e-5


def _neel(n):
    """Alpha electrons on the even sites, beta on the odd ones: every electron
    can hop, so the first column is long enough for the first pick."""
    d = sum(1 << (2 * p) for p in range(0, n, 2)) | sum(1 << (2 * p + 1) for p in range(1, n, 2))
    return np.array([d], np.uint64)


# (k, iterations, reference determinant) of the long runs
RUNS = {"c3_synth": (64, 24000, None), "c3_hubbard": (16, 60000, _neel(12))}


@pytest.fixture(scope="module")
def long_run(workloads):
    out = {}

    def get(name):
        if name not in out:
            info = workloads(name)
            k, iters, ref = RUNS[name]
            g = Cdfci(info["fcidump"], num_of_coordinate=k, estimate=k, z_threshold=0.0, max_memory=0.05,
                      threshold=0.0)
            g.init(ref)
            g.set_graph(20)
            g.step(iters)
            r = g.report()
            print(name, "after", iters, "iterations: E", r["energy"], "|x|_0", r["x_size"], flush=True)
            out[name] = (info, g)
        return out[name]

    return get


def _integrals(info):
    norb, nelec, ms2, h, eri, core = ed.read_fcidump(info["fcidump"])
    return h, eri, core


@pytest.mark.parametrize("name", ["c3_synth", "c3_hubbard"])
def test_c3_gradient_bit_exact_with_oracle_rdm(name, long_run):
    print("case", name, flush=True)
    info, g = long_run(name)
    e, gh, ge = g.energy_gradient()
    dets, x, _ = g.table()
    keep = x != 0.0
    assert keep.sum() > 100000
    o = Oracle(info["fcidump"])
    eo, oh, oe = o.energy_gradient(dets[keep], x[keep])
    assert e == eo
    assert np.array_equal(gh, oh)
    assert np.array_equal(ge, oe)
    e64 = g.energy_fp64()
    assert abs(e - e64) <= 1e-11 * max(1.0, abs(e))
    assert abs(e - g.report()["energy"]) < 1e-9  # z = Hx at z_threshold = 0


@pytest.mark.parametrize("name", ["c3_synth", "c3_hubbard"])
def test_c3_gradient_matches_fixed_wavefunction_differences(name, long_run):
    print("case", name, flush=True)
    info, g = long_run(name)
    h, eri, core = _integrals(info)
    _, gh, ge = g.energy_gradient()
    rng = np.random.default_rng(23)
    vh, ve = symmetrise(rng.normal(size=h.shape), rng.normal(size=eri.shape) * 0.1)
    an = float((gh * vh).sum() + (ge * ve).sum())
    ep = g.energy_fp64((h + STEP * vh, eri + STEP * ve, core))
    em = g.energy_fp64((h - STEP * vh, eri - STEP * ve, core))
    fd = (ep - em) / (2 * STEP)
    print(name, "directional derivative: adjoint", an, "central difference", fd, flush=True)
    assert abs(fd - an) <= 1e-8 * max(1.0, abs(an)), (fd, an)
    if name == "c3_hubbard":
        n = h.shape[0]
        an_u = float(sum(ge[i, i, i, i] for i in range(n)))  # dE/dU, U = (ii|ii) on every site
        du = np.zeros_like(eri)
        for i in range(n):
            du[i, i, i, i] = 1.0
        fd_u = (g.energy_fp64((h, eri + STEP * du, core)) - g.energy_fp64((h, eri - STEP * du, core))) / (2 * STEP)
        print("dE/dU: adjoint", an_u, "central difference", fd_u, flush=True)
        assert 0.0 < an_u < n / 4.0  # double occupancies of a repulsive half-filled ring
        assert abs(fd_u - an_u) <= 1e-8 * max(1.0, abs(an_u)), (fd_u, an_u)


def _ring(n, u):
    spec = synth.SyntheticSpec("ring", norb=n, nelec=n, kind="hubbard", hubbard_u=u)
    h, eri = synth.hubbard_integrals(spec)
    return h, eri


def _converged(n, h, eri, core, k, iters):
    c = Cdfci(ham=(n, n, 0, core, h, eri), num_of_coordinate=k, estimate=k, z_threshold=0.0, max_memory=0.05,
              threshold=0.0)
    c.init(_neel(n))
    c.set_graph(20)
    c.step(iters)
    return c


@pytest.mark.parametrize("n,k,iters", [(6, 8, 20000), (8, 16, 40000)])
def test_ring_gradient_matches_reconverged_differences(n, k, iters):
    print("case ring", n, flush=True)
    core = -10.0
    h, eri = _ring(n, 4.0)
    dets, H = ed.hamiltonian(h, eri, core, n // 2, n // 2)
    e0 = np.linalg.eigvalsh(H)[0]
    c = _converged(n, h, eri, core, k, iters)
    e = c.report()["energy"]
    assert abs(e - e0) < 1e-11, (e, e0)
    eg, gh, ge = c.energy_gradient()
    rng = np.random.default_rng(5)
    vh, ve = symmetrise(rng.normal(size=h.shape), rng.normal(size=eri.shape) * 0.1)
    an = float((gh * vh).sum() + (ge * ve).sum())
    ep = _converged(n, h + STEP * vh, eri + STEP * ve, core, k, iters).report()["energy"]
    em = _converged(n, h - STEP * vh, eri - STEP * ve, core, k, iters).report()["energy"]
    fd = (ep - em) / (2 * STEP)
    print("ring", n, "directional: adjoint", an, "central difference", fd, flush=True)
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)
    an_u = float(sum(ge[i, i, i, i] for i in range(n)))
    vals = []
    for sgn in (1, -1):
        hu, eu = _ring(n, 4.0 + sgn * STEP)
        vals.append(_converged(n, hu, eu, core, k, iters).report()["energy"])
    fd_u = (vals[0] - vals[1]) / (2 * STEP)
    print("ring", n, "dE/dU: adjoint", an_u, "central difference", fd_u, flush=True)
    assert abs(fd_u - an_u) <= 1e-6 * max(1.0, abs(an_u)), (fd_u, an_u)
