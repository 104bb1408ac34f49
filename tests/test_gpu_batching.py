"""Batching additions: HIP-graph replay of captured iterations and the
online autotuner.  Both are result-neutral, so every check is exact: graph
replays against direct launches, and autotuned runs (which pass through
every scatter and pick launch form) against the oracle.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from mcdfci_amd import Cdfci, McdfciError
from oracle_ctypes import Oracle


def _same_state(a, b):
    ra, rb = a.report(), b.report()
    for key in ("energy", "x_size", "z_size", "cols_size", "updates", "scale", "xx", "dx_max", "iterations"):
        assert ra[key] == rb[key], (key, ra[key], rb[key])
    da, xa, za = a.table()
    db, xb, zb = b.table()
    assert np.array_equal(da, db)
    assert np.array_equal(xa, xb)
    assert np.array_equal(za, zb)


@pytest.mark.parametrize("name,k,M,z_thr,mem,G", [
    ("c3_synth", 8, 8, 0.0, 0.05, 5),
    ("c4", 64, 64, 1e-6, 8.0, 4),
])
def test_graph_replay_matches_direct_launches(name, k, M, z_thr, mem, G, workloads):
    print("case", name, k, M, G, flush=True)
    path = workloads(name)["fcidump"]
    kw = dict(num_of_coordinate=k, estimate=M, z_threshold=z_thr, max_memory=mem, threshold=0.0)
    a = Cdfci(path, **kw)
    b = Cdfci(path, **kw)
    a.init()
    b.init()
    b.set_graph(G)
    # several replays with host synchronisation and direct launches between
    # them (the remainder of each step runs uncaptured)
    for n in (1, 2 * G + 3, G, 3 * G - 1):
        a.step(n)
        b.step(n)
        _same_state(a, b)
    # a re-initialisation reallocates the iteration buffers: the graph is
    # captured again
    a.init()
    b.init()
    a.step(2 * G)
    b.step(2 * G)
    _same_state(a, b)
    a.close()
    b.close()


@pytest.mark.parametrize("name,k,M,z_thr,mem", [
    ("c3_synth", 8, 8, 0.0, 0.05),      # one-workgroup and merge-tree picks both eligible
    ("c4", 64, 64, 1e-6, 8.0),
    ("c5", 16, 16, 1e-5, 8.0),          # two-word determinants
])
def test_autotune_keeps_the_trajectory(name, k, M, z_thr, mem, workloads):
    print("case", name, k, M, flush=True)
    path = workloads(name)["fcidump"]
    g = Cdfci(path, num_of_coordinate=k, estimate=M, z_threshold=z_thr, max_memory=mem, threshold=0.0)
    g.init()
    choice = g.autotune(2)
    print("choice", choice, flush=True)
    assert choice["elements_per_thread"] in (1, 2, 4, 8, 16)
    assert choice["iterations"] == 2 * 5 * (2 if k * M <= 4096 else 1)
    g.set_graph(3)
    g.step(7)
    n = choice["iterations"] + 7
    o = Oracle(path)
    o.solver_init(k, M, z_thr, g.report()["max_size"], nthreads=8)
    o.step(n)
    rg, ro = g.report(), o.report()
    assert rg["iterations"] == n
    assert rg["energy"] == ro["energy"]
    for key in ("x_size", "z_size", "cols_size", "updates", "scale", "xx", "dx_max"):
        assert rg[key] == ro[key], (key, rg[key], ro[key])
    gd, gx, gz = g.table()
    od, ox, oz, _ = o.table()
    assert np.array_equal(gd, od)
    assert np.array_equal(gx, ox)
    assert np.array_equal(gz, oz)
    g.close()


@pytest.mark.parametrize("name,k,M,z_thr,mem,iters", [
    # one-workgroup merge tree in LDS (k <= 64, M <= 64)
    ("c3_synth", 8, 8, 0.0, 0.05, 150),
    ("c3_synth", 5, 8, 0.0, 0.05, 120),      # odd number of lists
    ("c3_synth", 16, 16, 0.0, 0.05, 80),
    ("c3_synth", 8, 32, 0.0, 0.05, 80),      # L = k < M
    ("c3_synth", 33, 64, 0.0, 0.05, 30),     # 33 lists of up to 33
    ("c4", 64, 64, 1e-6, 8.0, 4),
    ("c5", 16, 16, 1e-5, 8.0, 8),            # two-word determinants
    ("c5", 64, 64, 1e-5, 8.0, 3),
    # per-level merge launches (M above 64, or k above 64)
    ("c3_synth", 8, 128, 0.0, 0.05, 60),
    ("c3_synth", 96, 96, 0.0, 0.05, 12),
])
def test_tree_pick_keeps_the_trajectory(name, k, M, z_thr, mem, iters, workloads):
    """The merge-tree pick pinned with mcdfci_set_launch for a whole run
    gives the oracle's trajectory (the oracle sorts all candidates at once)."""
    print("case", name, k, M, iters, flush=True)
    path = workloads(name)["fcidump"]
    g = Cdfci(path, num_of_coordinate=k, estimate=M, z_threshold=z_thr, max_memory=mem, threshold=0.0)
    g.init()
    g.set_launch(4, True)
    o = Oracle(path)
    o.solver_init(k, M, z_thr, g.report()["max_size"], nthreads=8)
    for n in (1, iters - 1):
        g.step(n)
        o.step(n)
        rg, ro = g.report(), o.report()
        assert rg["energy"] == ro["energy"]
        for key in ("x_size", "z_size", "cols_size", "updates", "scale", "xx", "dx_max"):
            assert rg[key] == ro[key], (key, rg[key], ro[key])
    gd, gx, gz = g.table()
    od, ox, oz, _ = o.table()
    assert np.array_equal(gd, od)
    assert np.array_equal(gx, ox)
    assert np.array_equal(gz, oz)
    g.close()


def test_set_launch_refusals(workloads):
    path = workloads("c3_synth")["fcidump"]
    g = Cdfci(path, num_of_coordinate=128, estimate=64, max_memory=0.05, threshold=0.0)
    with pytest.raises(McdfciError):
        g.set_launch(4, False)  # 128 * 64 candidates need the merge tree
    with pytest.raises(McdfciError):
        g.set_launch(3, True)
    g.set_launch(16, True)
    g.close()


def test_access_ceiling_then_reinit(workloads):
    """The access-ceiling probe overwrites the table: stepping needs a fresh
    init, after which the run matches one that never probed."""
    path = workloads("c3_synth")["fcidump"]
    kw = dict(num_of_coordinate=8, estimate=8, max_memory=0.05, threshold=0.0)
    a = Cdfci(path, **kw)
    b = Cdfci(path, **kw)
    for mode in (0, 1, 2):
        assert a.access_ceiling(mode, 1 << 16) > 0.0
    with pytest.raises(McdfciError):
        a.step(1)
    a.init()
    b.init()
    a.step(40)
    b.step(40)
    _same_state(a, b)
    a.close()
    b.close()
