"""Parity of the HIP path (through the C ABI) with the CPU oracle.

The device path and the oracle implement the same canonical arithmetic, so
every comparison here is exact (==), not a tolerance: column determinants,
signs and values; H_sub; the subspace update; and whole trajectories --
energy, |x|_0, |z|_0, scale, and the full wavefunction table (keys, x, z)
after hundreds of iterations.  Mixed precision is checked against the FP64
path with the tolerance stated in each test.
"""

import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import mcdfci_amd
from mcdfci_amd import Cdfci, McdfciError, PREC_MIXED, UPDATE_LS
from oracle_ctypes import Oracle


def _path(name, workloads, h2o_fcidump):
    return h2o_fcidump if name == "h2o" else workloads(name)["fcidump"]


def _dict(dets, vals):
    return dict(zip(map(tuple, dets.tolist()), vals.tolist()))


# c5 (42 orbitals, 84 spin-orbitals) runs on two-word determinants.
@pytest.mark.parametrize("name", ["c1", "c3_synth", "c3_hubbard", "h2o", "c2", "c4", "c5"])
def test_columns_bit_exact(name, workloads, h2o_fcidump):
    path = _path(name, workloads, h2o_fcidump)
    print("case", name, flush=True)
    o = Oracle(path)
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.01, threshold=0.0)
    hf = g.hartree_fock()
    assert np.array_equal(hf, o.hf())
    od, ov = o.column(hf)
    rng = np.random.default_rng(2)
    seeds = [hf] + [od[i] for i in rng.choice(len(od), min(6, len(od)), replace=False)]
    for d in seeds:
        gd, gv = g.column(d)
        od2, ov2 = o.column(d)
        A, B = _dict(gd, gv), _dict(od2, ov2)
        assert len(gd) == len(od2)
        assert A.keys() == B.keys()
        assert all(A[key] == B[key] for key in A)
        assert np.array_equal(gd[0], np.asarray(d).reshape(-1))


@pytest.mark.parametrize("name", ["c1", "c3_synth", "h2o", "c5"])
def test_submatrix_bit_exact(name, workloads, h2o_fcidump):
    path = _path(name, workloads, h2o_fcidump)
    o = Oracle(path)
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.01, threshold=0.0)
    od, _ = o.column(o.hf())
    rng = np.random.default_rng(4)
    dets = od[rng.choice(len(od), min(40, len(od)), replace=False)]
    H = g.submatrix(dets)
    for i in range(len(dets)):
        for j in range(i, len(dets)):
            ref = o.diagonal(dets[i]) if i == j else o.off_diagonal(dets[i], dets[j])
            assert H[i, j] == ref


@pytest.mark.parametrize("k", [1, 2, 3, 8, 33, 64, 89, 90, 128, 200, 256])
def test_eig_update_bit_exact(k, workloads):
    g = Cdfci(workloads("c1")["fcidump"], num_of_coordinate=8, estimate=8, max_memory=0.01)
    o = Oracle(workloads("c1")["fcidump"])
    rng = np.random.default_rng(k)
    Hs = rng.normal(size=(k, k)) * 0.05
    Hs = Hs + Hs.T
    Hs[np.arange(k), np.arange(k)] = -20.0 + rng.normal(size=k)
    Hs = np.ascontiguousarray(Hs.reshape(-1))
    for first in (False, True):
        x = rng.normal(size=k) * 0.3
        z = -20.0 * x + rng.normal(size=k) * 0.1
        xx = float(x @ x) if first else float(x @ x) + 3.0
        xz = float(np.round(-20.0 * xx, 6))
        dxg, xxg, scg = g.eig_update(Hs, x, z, (xx, 0.0), (0.9, 0.0), (xz, 0.0))
        s = np.array([xx, 0.0, 0.9, 0.0, xz, 0.0])
        dxo = np.zeros(k)
        assert o.lib.oc_eig_update(k, Hs, x, z, s, dxo, np.zeros(3)) == 0
        assert np.array_equal(dxg, dxo)
        assert xxg == (s[0], s[1]) and scg == (s[2], s[3])


def _run_pair(path, k, M, z_thr, iters, max_memory=0.05, checkpoints=(1, 10), update=0, precision="fp64",
              dump_limit=1 << 22):
    g = Cdfci(path, num_of_coordinate=k, estimate=M, z_threshold=z_thr, max_memory=max_memory, threshold=0.0,
              coordinate_update=update, precision=PREC_MIXED if precision == "mixed" else 0)
    g.init()
    o = Oracle(path, precision=precision)
    o.solver_init(k, M, z_thr, g.report()["max_size"], nthreads=8)
    o.set_update(update)
    done = 0
    for n in [c for c in checkpoints if c < iters] + [iters]:
        step = n - done
        if step <= 0:
            continue
        g.step(step)
        o.step(step)
        done = n
        rg, ro = g.report(), o.report()
        assert rg["energy"] == ro["energy"], (done, rg["energy"], ro["energy"])
        for key in ("x_size", "z_size", "cols_size", "updates", "scale", "xx", "dx_max"):
            assert rg[key] == ro[key], (key, done, rg[key], ro[key])
    # the whole table, z exact: an order-independent digest always, the
    # sorted dump too while it is small enough to sort quickly
    assert g.table_digest() == o.table_digest()
    if rg["z_size"] <= dump_limit:
        gd, gx, gz = g.table()
        od, ox, oz, _ = o.table()
        assert np.array_equal(gd, od)
        assert np.array_equal(gx, ox)
        assert np.array_equal(gz, oz)
    cd, cx, cz = g.candidates()
    kd, kx, kz = o.candidates()
    assert sorted(map(tuple, cd.tolist())) == sorted(map(tuple, kd.tolist()))
    return g, o


@pytest.mark.parametrize("name,k,M,z_thr,iters", [
    ("c1", 8, 8, 0.0, 600),
    ("c1", 1, 4, 0.0, 300),
    ("c1", 3, 20, 0.0, 300),
    ("c3_synth", 8, 8, 0.0, 200),
    ("c3_synth", 16, 16, 1e-4, 150),
    ("c3_hubbard", 4, 8, 0.0, 200),
    ("h2o", 8, 8, 0.0, 300),
    ("c2", 8, 8, 1e-7, 60),
    # k * M above one workgroup's sort: per-column sorts and a merge tree in the pick
    ("c3_synth", 128, 128, 0.0, 60),
    # chunk lists above one workgroup's sort: streaming merge of the column top-M
    ("c2", 16, 1024, 1e-7, 12),
    # the norb-24 molecule at k = M = 256
    ("c4", 256, 256, 1e-6, 4),
    # 29 chunk lists of 256 per column: the column merge's head selection,
    # with and without the tails of the lists whose heads reach the top M
    ("c4", 32, 256, 1e-6, 30),
    # the same with lists kept shorter than the merge window (M = 200 < 256)
    ("c4", 16, 200, 1e-6, 12),
    # 29 lists of 128: one sort of 3712 entries per column (four per thread)
    ("c4", 16, 128, 1e-6, 12),
    # 42 orbitals: two-word determinants
    ("c5", 8, 16, 1e-5, 12),
    ("c5", 64, 64, 1e-5, 4),
    # two words with the top-256 selection (estimate 65..256)
    ("c5", 16, 128, 1e-5, 4),
    ("c5", 32, 256, 1e-5, 3),
])
def test_trajectory_bit_exact(name, k, M, z_thr, iters, workloads, h2o_fcidump):
    print("case", name, k, M, z_thr, iters, flush=True)
    path = _path(name, workloads, h2o_fcidump)
    mem = 8.0 if name in ("c2", "c4", "c5") else 0.05
    _run_pair(path, k, M, z_thr, iters, max_memory=mem)


@pytest.mark.parametrize("name,M,z_thr,iters", [
    ("c1", 8, 0.0, 600),
    ("c3_synth", 8, 0.0, 300),
    ("h2o", 8, 0.0, 300),
    ("c2", 8, 1e-7, 100),
    ("c5", 16, 1e-5, 12),   # two-word determinants
])
def test_line_search_trajectory_bit_exact(name, M, z_thr, iters, workloads, h2o_fcidump):
    """The line-search coordinate update (one coordinate per iteration)."""
    print("case ls", name, M, z_thr, iters, flush=True)
    path = _path(name, workloads, h2o_fcidump)
    mem = 8.0 if name in ("c2", "c5") else 0.05
    g, o = _run_pair(path, 1, M, z_thr, iters, max_memory=mem, update=UPDATE_LS)
    assert g.report()["scale"] == 1.0


def test_line_search_needs_one_coordinate(workloads):
    with pytest.raises(McdfciError):
        Cdfci(workloads("c1")["fcidump"], num_of_coordinate=2, estimate=8, max_memory=0.01,
              coordinate_update=UPDATE_LS)


def test_c1_reaches_exact_ground_state(workloads):
    import ed
    path = workloads("c1")["fcidump"]
    norb, nelec, ms2, h, eri, core = ed.read_fcidump(path)
    _, H = ed.hamiltonian(h, eri, core, nelec // 2, nelec // 2)
    e0 = np.linalg.eigvalsh(H)[0]
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.005, threshold=0.0)
    g.init()
    g.step(20000)
    assert abs(g.report()["energy"] - e0) < 1e-10


def test_table_full_matches_oracle(workloads):
    path = workloads("c3_synth")["fcidump"]
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.0005, threshold=0.0)
    g.init()
    o = Oracle(path)
    o.solver_init(8, 8, 0.0, g.report()["max_size"], nthreads=2)
    with pytest.raises(McdfciError) as ei:
        for _ in range(50):
            g.step(1)
            g.report()
    assert ei.value.code == 3
    rc = 0
    for _ in range(50):
        rc = o.step(1)
        if rc:
            break
    assert rc == 1


def test_too_few_candidates_matches_oracle(workloads):
    """The Hubbard HF column has 5 elements; k = 8 distinct picks are impossible
    (the reference would read past its priority queue, src/solver.h:218-223)."""
    path = workloads("c3_hubbard")["fcidump"]
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.05, threshold=0.0)
    g.init()
    g.step(1)
    with pytest.raises(McdfciError) as ei:
        g.report()
    assert ei.value.code == 6
    o = Oracle(path)
    o.solver_init(8, 8, 0.0, 1e6, nthreads=1)
    assert o.step(1) == 3


def test_positive_spectrum_is_an_error(tmp_path, workloads):
    from mcdfci_amd import synth
    spec = synth.SyntheticSpec("pos", norb=6, nelec=6, rank=8, seed=5, core=500.0)
    info = synth.generate(spec, str(tmp_path))
    g = Cdfci(info["fcidump"], num_of_coordinate=4, estimate=4, max_memory=0.01, threshold=0.0)
    g.init()
    g.step(2)
    with pytest.raises(McdfciError) as ei:
        g.report()
    assert ei.value.code == 5


# ---------------------------------------------------------------- mixed precision
# fp32 integral gathers and off-diagonal column elements, fp64 diagonal,
# H_sub, z, energy and eigen-update: bit-exact with the oracle's mixed form.
@pytest.mark.parametrize("name", ["c1", "c3_synth", "h2o", "c4", "c5"])
def test_mixed_columns_bit_exact(name, workloads, h2o_fcidump):
    path = _path(name, workloads, h2o_fcidump)
    print("case mixed column", name, flush=True)
    o = Oracle(path, precision="mixed")
    o64 = Oracle(path)
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.01, threshold=0.0, precision=PREC_MIXED)
    hf = g.hartree_fock()
    od, _ = o.column(hf)
    rng = np.random.default_rng(3)
    seeds = [hf] + [od[i] for i in rng.choice(len(od), min(4, len(od)), replace=False)]
    differ = 0
    for d in seeds:
        A = _dict(*g.column(d))
        B = _dict(*o.column(d))
        assert A.keys() == B.keys()
        assert all(A[key] == B[key] for key in A)
        F = _dict(*o64.column(d))
        differ += sum(1 for key in A if A[key] != F.get(key))
    assert differ > 0  # the device really evaluates these in fp32


@pytest.mark.parametrize("name,k,M,z_thr,iters", [
    ("c1", 8, 8, 0.0, 400),
    ("c3_synth", 8, 8, 0.0, 150),
    ("h2o", 8, 8, 0.0, 200),
    # the norb-24 molecule at the three batch sizes it is measured at
    ("c4", 8, 8, 1e-6, 40),
    ("c4", 64, 64, 1e-6, 8),
    ("c4", 256, 256, 1e-6, 3),
    # 42 orbitals, two-word determinants, k = M = 256
    ("c5", 256, 256, 1e-5, 2),
])
def test_mixed_trajectory_bit_exact(name, k, M, z_thr, iters, workloads, h2o_fcidump):
    print("case mixed", name, k, M, z_thr, iters, flush=True)
    path = _path(name, workloads, h2o_fcidump)
    mem = 8.0 if name in ("c4", "c5") else 0.05
    cps = (1,) if iters <= 3 else (1, 4)
    _run_pair(path, k, M, z_thr, iters, max_memory=mem, checkpoints=cps, precision="mixed")


def test_mixed_energy_error_vs_fp64_rayleigh_quotient(workloads):
    """The energy a mixed run reports against the FP64 Rayleigh quotient of
    the same x (mcdfci_energy_gradient evaluates c.Hc / c.c with the fp64
    integrals): on the norb-12 synthetic after 300 iterations the two agree
    to the fp32 element error, well inside 1e-5 Ha."""
    path = workloads("c3_synth")["fcidump"]
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.05, threshold=0.0, precision=PREC_MIXED)
    g.init()
    g.step(300)
    e_mixed = g.report()["energy"]
    e64, _, _ = g.energy_gradient()
    e64b = g.energy_fp64()
    print("mixed", e_mixed, "fp64(x)", e64, e64b, "diff", e_mixed - e64, flush=True)
    assert abs(e64b - e64) < 1e-11
    assert abs(e_mixed - e64) < 1e-5
    assert e_mixed != e64


@pytest.mark.parametrize("prec", [0, PREC_MIXED])
def test_energy_fp64_is_the_dense_rayleigh_quotient(prec, workloads):
    """mcdfci_energy_fp64 equals x.Hx / x.x with H the second-quantised fp64
    matrix and x the device table, for both precisions of the run."""
    import ed
    path = workloads("c1")["fcidump"]
    norb, nelec, ms2, h, eri, core = ed.read_fcidump(path)
    dets, H = ed.hamiltonian(h, eri, core, nelec // 2, nelec // 2)
    idx = {int(d): i for i, d in enumerate(dets)}
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.005, threshold=0.0, precision=prec)
    g.init()
    g.step(120)
    td, tx, _ = g.table()
    x = np.zeros(len(dets))
    for d, a in zip(td[:, 0], tx):
        x[idx[int(d)]] = a
    ref = float(x @ H @ x / (x @ x))
    e = g.energy_fp64()
    assert abs(e - ref) < 1e-12, (e, ref)
    if prec == 0:
        assert abs(e - g.report()["energy"]) < 1e-11  # z = Hx exactly at z_threshold = 0


def test_driver_end_to_end(workloads, tmp_path):
    info = workloads("c1")
    import json
    doc = json.load(open(info["input"]))
    doc["solver"]["cdfci"].update(num_iterations=2000, report_interval=500)
    inp = tmp_path / "input.json"
    inp.write_text(json.dumps(doc))
    r = subprocess.run([mcdfci_amd.DRIVER_PATH, str(inp)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in r.stdout.splitlines() if l[:1].isdigit() and len(l.split()) == 9]
    assert [int(x[0]) for x in rows] == [500, 1000, 1500, 2000]
    o = Oracle(info["fcidump"])
    o.solver_init(8, 8, 0.0, 103546.0, nthreads=2)
    o.step(2000)
    assert rows[-1][1] == "%.10f" % o.report()["energy"]
    assert "[LOG] max residual" in r.stdout
    # unsupported pick: the reference's exit(-1)
    doc["solver"]["cdfci"]["coordinate_pick"] = "random"
    inp.write_text(json.dumps(doc))
    r = subprocess.run([mcdfci_amd.DRIVER_PATH, str(inp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 255 and "is not implemented" in r.stdout


def test_driver_line_search(workloads, tmp_path):
    """coordinate_update "ls" with num_of_coordinate 1 through the driver
    (the sCDFCI inputs of the reference's experiments)."""
    info = workloads("c1")
    import json
    doc = json.load(open(info["input"]))
    doc["solver"]["cdfci"].update(num_iterations=3000, report_interval=1000, coordinate_update="ls",
                                  num_of_coordinate=1, estimate=8)
    inp = tmp_path / "input.json"
    inp.write_text(json.dumps(doc))
    r = subprocess.run([mcdfci_amd.DRIVER_PATH, str(inp)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in r.stdout.splitlines() if l[:1].isdigit() and len(l.split()) == 9]
    assert [int(x[0]) for x in rows] == [1000, 2000, 3000]
    o = Oracle(info["fcidump"])
    o.solver_init(1, 8, 0.0, 103546.0, nthreads=1)
    o.set_update(1)
    o.step(3000)
    assert rows[-1][1] == "%.10f" % o.report()["energy"]


def test_driver_z_threshold_search_restarts(workloads, tmp_path):
    """z_threshold_search (src/solver.h:802-809,881-888; run_wrapper.h:101-119):
    the table (2^19 slots, max_size 414187) passes 70 % of max_size at a
    report while still growing, so the solve restarts from Hartree-Fock with
    the threshold doubled, twice (1e-4 -> 2e-4 -> 4e-4); the final solve's
    rows equal the oracle run at 4e-4."""
    info = workloads("c3_synth")
    import json
    doc = json.load(open(info["input"]))
    doc["solver"]["cdfci"].update(num_iterations=6000, report_interval=1000, num_of_coordinate=8, estimate=8,
                                  z_threshold=1e-4, z_threshold_search=True, max_memory=0.02)
    inp = tmp_path / "input.json"
    inp.write_text(json.dumps(doc))
    r = subprocess.run([mcdfci_amd.DRIVER_PATH, str(inp)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    restarts = [l for l in lines if l.startswith("z_threshold_search: z_threshold is too small")]
    assert restarts == [
        "z_threshold_search: z_threshold is too small. Increase z_threhold to 2.00e-04 and restart.",
        "z_threshold_search: z_threshold is too small. Increase z_threhold to 4.00e-04 and restart.",
    ]
    last = max(i for i, l in enumerate(lines) if l.startswith("z_threshold_search"))
    rows = [l.split() for l in lines[last:] if l[:1].isdigit() and len(l.split()) == 9]
    assert [int(x[0]) for x in rows] == [1000, 2000, 3000, 4000, 5000, 6000]
    o = Oracle(info["fcidump"])
    o.solver_init(8, 8, 4e-4, float(int(0.79 * 2 ** 19)), nthreads=8)
    for row in rows:
        assert o.step(1000) == 0
        ro = o.report()
        assert row[1] == "%.10f" % ro["energy"]
        assert int(row[4]) == int(ro["z_size"])


@pytest.mark.parametrize("name,k,M,z_thr,iters,bits", [
    ("c3_synth", 8, 8, 0.0, 120, 54),   # folds through pass 2's list
    ("c4", 64, 64, 1e-6, 4, 53),        # the list overflows: whole-table sweeps
])
def test_z_word_fold_keeps_the_trajectory(name, k, M, z_thr, iters, bits, workloads, monkeypatch):
    """A table z is A * 2^53 + B with single-atomic adds to B; B is folded
    into A once it passes 2^61.  Lowering the fold point to 2^54 / 2^53
    (MCDFCI_ZNORM_BITS) makes folds constant, through the slot list and
    through the whole-table sweep that replaces it when it overflows: the
    trajectory and the table stay bit-exact with the oracle."""
    print("case fold", name, k, bits, flush=True)
    monkeypatch.setenv("MCDFCI_ZNORM_BITS", str(bits))
    mem = 8.0 if name == "c4" else 0.05
    _run_pair(workloads(name)["fcidump"], k, M, z_thr, iters, max_memory=mem, checkpoints=(1, 3))


@pytest.mark.parametrize("tails", [0, 1])
def test_merge_fallback_keeps_the_trajectory(tails, workloads, monkeypatch):
    """The column merge at M = 256 (29 chunk lists of 256 per column) selects
    from the lists' first 64 entries in LDS and merges in the rest of the
    lists that can still reach the column's top M, up to seven of them; with
    more it falls back to the full tournament over the staging buffer.  On
    this workload no column needs more than two, so MCDFCI_MERGE_TAILS lowers
    the seven (0: every column with such a list falls back, 1: those with two)
    and the trajectory and the table stay bit-exact with the oracle."""
    print("case merge fallback", tails, flush=True)
    monkeypatch.setenv("MCDFCI_MERGE_TAILS", str(tails))
    _run_pair(workloads("c4")["fcidump"], 32, 256, 1e-6, 12, max_memory=8.0)
