"""The device adjoint (mcdfci_energy_gradient) against the oracle and
against finite differences.

Exact parity: after the same trajectory the device and the oracle hold the
same table, and both accumulate the gradient terms as exact fixed-point
integers, so every entry and the energy agree bit for bit.  Finite
differences: at a converged wavefunction the gradient is the derivative of
the ground-state energy (Hellmann-Feynman); central differences with step
1e-5 of the exact ground-state energy (dense second-quantised H) and of the
device solver's own converged energy agree to 1e-6 relative.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

import ed
from mcdfci_amd import Cdfci
from oracle_ctypes import Oracle
from torch_adjoint import CdfciEnergy, symmetrise


def _path(name, workloads, h2o_fcidump):
    return h2o_fcidump if name == "h2o" else workloads(name)["fcidump"]


@pytest.mark.parametrize("name,k,iters", [("c1", 8, 600), ("c3_synth", 8, 200), ("c3_hubbard", 4, 200),
                                          ("h2o", 8, 300),
                                          # 42 orbitals: two-word determinants
                                          ("c5", 16, 20)])
def test_gradient_bit_exact(name, k, iters, workloads, h2o_fcidump):
    print("case", name, k, iters, flush=True)
    path = _path(name, workloads, h2o_fcidump)
    big = name == "c5"
    g = Cdfci(path, num_of_coordinate=k, estimate=k if big else 8, max_memory=8.0 if big else 0.05, threshold=0.0,
              z_threshold=1e-5 if big else 0.0)
    g.init()
    g.step(iters)
    e, gh, ge = g.energy_gradient()
    dets, x, _ = g.table()
    keep = x != 0.0
    o = Oracle(path)
    eo, oh, oe = o.energy_gradient(dets[keep], x[keep])
    assert e == eo
    assert np.array_equal(gh, oh)
    assert np.array_equal(ge, oe)
    # the variational energy of the table is the solver's energy (z = Hx)
    assert abs(e - g.report()["energy"]) < 1e-10


def test_converged_gradient_matches_exact_finite_differences(workloads):
    path = workloads("c1")["fcidump"]
    norb, nelec, ms2, h, eri, core = ed.read_fcidump(path)
    na, nb = (nelec + ms2) // 2, (nelec - ms2) // 2
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.005, threshold=0.0)
    g.init()
    g.step(20000)
    e, gh, ge = g.energy_gradient()
    step = 1e-5
    for which, idx in [("h", (0, 0)), ("h", (3, 1)), ("e", (0, 0, 0, 0)), ("e", (2, 1, 1, 0)),
                       ("e", (4, 3, 2, 0)), ("e", (6, 5, 6, 5))]:
        vals = []
        for sgn in (1, -1):
            h2, e2 = h.copy(), eri.copy()
            if which == "h":
                h2[idx] += sgn * step
                if idx[0] != idx[1]:
                    h2[idx[::-1]] += sgn * step
            else:
                for t in ed.orbit(*idx):
                    e2[t] += sgn * step
            _, H = ed.hamiltonian(h2, e2, core, na, nb)
            vals.append(np.linalg.eigvalsh(H)[0])
        fd = (vals[0] - vals[1]) / (2 * step)
        an = gh[idx] if which == "h" else ge[idx]
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (which, idx, fd, an)


def test_autograd_function_directional_derivative(workloads):
    """d/de E(theta + e V) at e = 0 by autograd equals the central difference
    of the device solver's converged energy (step 1e-5) along a random
    symmetric direction V."""
    path = workloads("c1")["fcidump"]
    norb, nelec, ms2, h, eri, core = ed.read_fcidump(path)
    rng = np.random.default_rng(11)
    vh, ve = symmetrise(rng.normal(size=h.shape), rng.normal(size=eri.shape) * 0.1)
    ht = torch.tensor(h, requires_grad=True)
    et = torch.tensor(eri, requires_grad=True)
    E = CdfciEnergy.apply(ht, et, nelec, ms2, core, 20000)
    E.backward()
    an = float((ht.grad * torch.from_numpy(vh)).sum() + (et.grad * torch.from_numpy(ve)).sum())
    step = 1e-5
    ep = float(CdfciEnergy.apply(torch.tensor(h + step * vh), torch.tensor(eri + step * ve), nelec, ms2, core,
                                 20000))
    em = float(CdfciEnergy.apply(torch.tensor(h - step * vh), torch.tensor(eri - step * ve), nelec, ms2, core,
                                 20000))
    fd = (ep - em) / (2 * step)
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (fd, an)
