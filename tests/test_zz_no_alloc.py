"""mcdfci_eig_update_noalloc, the allocation-free form of the subspace
update, against mcdfci_eig_update (bit for bit) through the C ABI: two
calls in a row on the same device array, the second with a smaller k, and
the size limit."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from mcdfci_amd import Cdfci


def _problem(k, seed, first):
    rng = np.random.default_rng(seed)
    Hs = rng.normal(size=(k, k)) * 0.05
    Hs = Hs + Hs.T
    Hs[np.arange(k), np.arange(k)] = -20.0 + rng.normal(size=k)
    x = rng.normal(size=k) * 0.3
    z = -20.0 * x + rng.normal(size=k) * 0.1
    xx = float(x @ x) if first else float(x @ x) + 3.0
    xz = float(np.round(-20.0 * xx, 6))
    s = np.array([xx, 0.0, 0.9, 0.0, xz, 0.0])
    return np.ascontiguousarray(Hs.reshape(-1)), x, z, s


def test_noalloc_eig_update_matches_the_allocating_entry(workloads):
    g = Cdfci(workloads("c1")["fcidump"], num_of_coordinate=8, estimate=8, max_memory=0.01)
    lib = g.lib
    for k, first in ((200, False), (33, False), (9, True)):
        print("case noalloc k", k, "first", first, flush=True)
        Hs, x, z, s = _problem(k, 100 + k, first)
        sa, sb = s.copy(), s.copy()
        dxa, dxb = np.zeros(k), np.full(k, np.nan)
        assert lib.mcdfci_eig_update(g.h, k, Hs, x, z, sa, dxa) == 0
        assert lib.mcdfci_eig_update_noalloc(g.h, k, Hs, x, z, sb, dxb) == 0
        assert np.array_equal(dxa, dxb)
        assert np.array_equal(sa, sb)
    # a second context on the same device shares the array
    h = Cdfci(workloads("c1")["fcidump"], num_of_coordinate=8, estimate=8, max_memory=0.01)
    Hs, x, z, s = _problem(17, 5, False)
    sa, sb = s.copy(), s.copy()
    dxa, dxb = np.zeros(17), np.zeros(17)
    assert lib.mcdfci_eig_update(g.h, 17, Hs, x, z, sa, dxa) == 0
    assert lib.mcdfci_eig_update_noalloc(h.h, 17, Hs, x, z, sb, dxb) == 0
    assert np.array_equal(dxa, dxb) and np.array_equal(sa, sb)
    # beyond the scratch size: a clear error, nothing launched
    k = 257
    Hs, x, z, s = _problem(k, 7, False)
    rc = lib.mcdfci_eig_update_noalloc(g.h, k, Hs, x, z, s, np.zeros(k))
    assert rc == 1
    assert b"[1, 256]" in lib.mcdfci_last_error(g.h)
    g.close()
    h.close()
