"""Unrestricted (spin-resolved) integrals on the HIP path, against the CPU
oracle, bit for bit: the reference's UHF FCIDUMP dialect
(src/input.h:125-127,203-237) read by the device library, columns and the
subspace matrix from the per-spin tables, and whole trajectories (reports
and the full wavefunction table) in both precisions, with one- and two-word
determinants.  The oracle's unrestricted path is pinned against the
reference's own parser and Hamiltonian and against second quantisation
(test_oracle_vs_reference.py, test_oracle_uhf.py).
"""

import json
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import mcdfci_amd
from mcdfci_amd import Cdfci, McdfciError, PREC_MIXED, synth
from oracle_ctypes import Oracle

# 34 orbitals per spin (68 spin-orbitals): two-word determinants
W2U = synth.SyntheticSpec("w2u", norb=34, nelec=10, rank=10, seed=8, uhf=True)


@pytest.fixture(scope="module")
def w2u(tmp_path_factory):
    return synth.generate(W2U, str(tmp_path_factory.mktemp("w2u")))


def _path(name, workloads, w2u):
    return w2u["fcidump"] if name == "w2u" else workloads(name)["fcidump"]


def _dict(dets, vals):
    return dict(zip(map(tuple, dets.tolist()), vals.tolist()))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", ["u1", "u3", "c1_as_uhf", "w2u"])
def test_uhf_columns_bit_exact(name, precision, workloads, w2u):
    path = _path(name, workloads, w2u)
    print("case uhf column", name, precision, flush=True)
    o = Oracle(path, precision=precision)
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.01, threshold=0.0,
              precision=PREC_MIXED if precision == "mixed" else 0)
    assert g.uhf
    hf = g.hartree_fock()
    assert np.array_equal(hf, o.hf())
    od, _ = o.column(hf)
    rng = np.random.default_rng(5)
    seeds = [hf] + [od[i] for i in rng.choice(len(od), min(8, len(od)), replace=False)]
    for d in seeds:
        gd, gv = g.column(d)
        A, B = _dict(gd, gv), _dict(*o.column(d))
        assert len(gd) == len(B)
        assert A.keys() == B.keys()
        assert all(A[key] == B[key] for key in A)


@pytest.mark.parametrize("name", ["u1", "u3", "w2u"])
def test_uhf_submatrix_bit_exact(name, workloads, w2u):
    path = _path(name, workloads, w2u)
    o = Oracle(path)
    g = Cdfci(path, num_of_coordinate=8, estimate=8, max_memory=0.01, threshold=0.0)
    od, _ = o.column(o.hf())
    rng = np.random.default_rng(6)
    dets = od[rng.choice(len(od), min(40, len(od)), replace=False)]
    H = g.submatrix(dets)
    for i in range(len(dets)):
        for j in range(i, len(dets)):
            ref = o.diagonal(dets[i]) if i == j else o.off_diagonal(dets[i], dets[j])
            assert H[i, j] == ref


def _run_pair(path, k, M, z_thr, iters, max_memory, precision="fp64"):
    g = Cdfci(path, num_of_coordinate=k, estimate=M, z_threshold=z_thr, max_memory=max_memory, threshold=0.0,
              precision=PREC_MIXED if precision == "mixed" else 0)
    g.init()
    o = Oracle(path, precision=precision)
    o.solver_init(k, M, z_thr, g.report()["max_size"], nthreads=8)
    done = 0
    for n in [c for c in (1, 10) if c < iters] + [iters]:
        g.step(n - done)
        o.step(n - done)
        done = n
        rg, ro = g.report(), o.report()
        assert rg["energy"] == ro["energy"], (done, rg["energy"], ro["energy"])
        for key in ("x_size", "z_size", "cols_size", "updates", "scale", "xx", "dx_max"):
            assert rg[key] == ro[key], (key, done, rg[key], ro[key])
    assert g.table_digest() == o.table_digest()
    gd, gx, gz = g.table()
    od, ox, oz, _ = o.table()
    assert np.array_equal(gd, od) and np.array_equal(gx, ox) and np.array_equal(gz, oz)
    return g, o


@pytest.mark.parametrize("name,k,M,z_thr,iters,precision", [
    ("u1", 8, 8, 0.0, 600, "fp64"),
    ("u1", 3, 20, 0.0, 300, "mixed"),
    ("u3", 8, 8, 0.0, 200, "fp64"),
    ("u3", 64, 64, 1e-4, 40, "fp64"),
    ("u3", 16, 16, 1e-4, 100, "mixed"),
    ("w2u", 16, 32, 1e-5, 12, "fp64"),
    ("w2u", 64, 64, 1e-5, 4, "mixed"),
])
def test_uhf_trajectory_bit_exact(name, k, M, z_thr, iters, precision, workloads, w2u):
    print("case uhf", name, k, M, z_thr, iters, precision, flush=True)
    _run_pair(_path(name, workloads, w2u), k, M, z_thr, iters, 0.05 if name == "u1" else 1.0, precision)


def test_restricted_file_in_unrestricted_dialect_is_the_same_run(workloads):
    """Equal alpha and beta integrals in the unrestricted dialect: the same
    trajectory and table as the restricted file, on the device."""
    runs = []
    for name in ("c1", "c1_as_uhf"):
        g = Cdfci(workloads(name)["fcidump"], num_of_coordinate=8, estimate=8, max_memory=0.05, threshold=0.0)
        g.init()
        g.step(400)
        runs.append((g.report(), g.table_digest(), g.uhf))
    (ra, da, ua), (rb, db, ub) = runs
    assert not ua and ub
    assert ra["energy"] == rb["energy"] and ra["x_size"] == rb["x_size"] and ra["z_size"] == rb["z_size"]
    assert np.array_equal(da, db)


def test_uhf_reaches_the_exact_ground_state(workloads):
    import ed
    spec = synth.CONFIGS["u1"]
    blocks = synth.uhf_integrals(spec)
    na, nb = (spec.nelec + spec.ms2) // 2, (spec.nelec - spec.ms2) // 2
    dets, H = ed.hamiltonian_uhf(*blocks, 0.0, na, nb)
    e0 = np.linalg.eigvalsh(H)[0]
    g = Cdfci(workloads("u1")["fcidump"], num_of_coordinate=8, estimate=8, max_memory=0.005, threshold=0.0)
    g.init()
    g.step(6000)
    assert abs(g.report()["energy"] - e0) < 1e-10
    # the FP64 Rayleigh quotient of the table's x, against the dense matrix
    idx = {int(d): i for i, d in enumerate(dets)}
    td, tx, _ = g.table()
    x = np.zeros(len(dets))
    for d, a in zip(td[:, 0], tx):
        x[idx[int(d)]] = a
    assert abs(g.energy_fp64() - float(x @ H @ x / (x @ x))) < 1e-12
    with pytest.raises(McdfciError) as ei:
        g.energy_gradient()
    assert ei.value.code == 7


def test_uhf_in_memory_integrals_equal_the_file(workloads):
    """mcdfci_create_uhf with the generator's arrays gives the file's run."""
    spec = synth.CONFIGS["u1"]
    info = workloads("u1")
    ha, hb, eaa, ebb, eab = synth.uhf_integrals(spec)
    a = Cdfci(info["fcidump"], num_of_coordinate=8, estimate=8, max_memory=0.005, threshold=0.0)
    b = Cdfci(ham_uhf=(spec.norb, spec.nelec, spec.ms2, info["core"], ha, hb, eaa, ebb, eab),
              num_of_coordinate=8, estimate=8, max_memory=0.005, threshold=0.0)
    for g in (a, b):
        g.init()
        g.step(200)
    # the file's 18 significant digits give back every double, and the
    # generator's blocks hold each symmetric orbit bitwise equal
    assert a.report()["energy"] == b.report()["energy"]
    assert np.array_equal(a.table_digest(), b.table_digest())


def test_driver_on_an_unrestricted_file(workloads, tmp_path):
    info = workloads("u1")
    doc = json.load(open(info["input"]))
    doc["solver"]["cdfci"].update(num_iterations=1000, report_interval=500)
    inp = tmp_path / "input.json"
    inp.write_text(json.dumps(doc))
    r = subprocess.run([mcdfci_amd.DRIVER_PATH, str(inp)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Hartree Fock: unrestricted" in r.stdout
    assert "Number of spin-orbitals: 14" in r.stdout
    # report rows (the HF orbital list of 9 electrons also has 9 fields)
    rows = [l.split() for l in r.stdout.splitlines()
            if l[:1].isdigit() and len(l.split()) == 9 and "." in l.split()[1]]
    assert [int(x[0]) for x in rows] == [500, 1000]
    o = Oracle(info["fcidump"])
    o.solver_init(8, 8, 0.0, 103546.0, nthreads=2)
    o.step(1000)
    assert rows[-1][1] == "%.10f" % o.report()["energy"]
