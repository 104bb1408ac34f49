"""torch.autograd.Function over the device CDFCI solver (test harness).

forward(h, eri) symmetrises the integrals (h -> (h + h^T)/2, eri -> the
mean over its 8-fold orbit), runs the FP64 device solver for a fixed number
of iterations and returns the variational energy; backward maps the device
gradient with respect to the unique parameters (mcdfci_energy_gradient)
back onto the input tensors: a unique parameter is the mean of the |orbit|
tensor entries it is built from, so each entry receives dE/dtheta / |orbit|.
"""

from __future__ import annotations

import itertools

import numpy as np
import torch

from mcdfci_amd import Cdfci


def symmetrise(h, eri):
    h = 0.5 * (h + h.T)
    e = eri
    e = (e + e.transpose(1, 0, 2, 3) + e.transpose(0, 1, 3, 2) + e.transpose(1, 0, 3, 2)
         + e.transpose(2, 3, 0, 1) + e.transpose(3, 2, 0, 1) + e.transpose(2, 3, 1, 0)
         + e.transpose(3, 2, 1, 0)) / 8.0
    return h, e


def unfold(gh, ge):
    """Per-entry gradients of the symmetrised inputs from unique-parameter
    gradients at canonical positions."""
    n = gh.shape[0]
    oh = np.zeros_like(gh)
    for p in range(n):
        for q in range(p + 1):
            if p == q:
                oh[p, p] = gh[p, p]
            else:
                oh[p, q] = oh[q, p] = 0.5 * gh[p, q]
    oe = np.zeros_like(ge)
    for p, q, r, s in zip(*np.nonzero(ge)):
        orb = {(p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
               (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)}
        for t in orb:
            oe[t] = ge[p, q, r, s] / len(orb)
    return oh, oe


class CdfciEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h, eri, nelec: int, ms2: int, core: float, iterations: int, k: int = 8):
        hs, es = symmetrise(h.detach().cpu().numpy().astype(np.float64),
                            eri.detach().cpu().numpy().astype(np.float64))
        norb = hs.shape[0]
        solver = Cdfci(ham=(norb, nelec, ms2, core, hs, es), num_of_coordinate=k, estimate=k,
                       z_threshold=0.0, max_memory=0.05, threshold=0.0)
        solver.init()
        solver.step(iterations)
        e, gh, ge = solver.energy_gradient()
        solver.close()
        oh, oe = unfold(gh, ge)
        ctx.save_for_backward(torch.from_numpy(oh), torch.from_numpy(oe))
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        oh, oe = ctx.saved_tensors
        return g * oh, g * oe, None, None, None, None, None
