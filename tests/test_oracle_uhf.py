"""Pins the oracle on unrestricted (spin-resolved) integrals, the FCIDUMP
dialect of src/input.h:125-127,203-237 (UHF=.TRUE.: NORB counts
spin-orbitals, every index is a spin-orbital, 2p + 1 alpha and 2p + 2 beta,
1-based), with mathematics independent of its code:

  * second quantisation over the spin-resolved Hamiltonian
    H = sum_{pq,s} h^s_pq a+_ps a_qs
      + 1/2 sum_{pqrs,st} (pq|rs)_st a+_ps a+_rt a_st a_qs + E_core
    gives every column of the u1 open shell (5 alpha, 4 beta in 7
    orbitals), value and sign;
  * the solver reaches that matrix's lowest eigenvalue;
  * a restricted Hamiltonian written in the unrestricted dialect (every
    block equal) gives the restricted file's columns and trajectory bit
    for bit.

The reference's own parser and Hamiltonian on the same files:
test_oracle_vs_reference.py (u1, u3, c1_as_uhf).  The device's refusal of
spin-changing integrals happens while reading the file, before any device
call, and is checked here through the C ABI.
"""

import ctypes as C
import os

import numpy as np
import pytest

import ed
from mcdfci_amd import synth
from oracle_ctypes import Oracle


def _u1_matrix():
    spec = synth.CONFIGS["u1"]
    ha, hb, eaa, ebb, eab = synth.uhf_integrals(spec)
    na, nb = (spec.nelec + spec.ms2) // 2, (spec.nelec - spec.ms2) // 2
    return ed.hamiltonian_uhf(ha, hb, eaa, ebb, eab, 0.0, na, nb)


def test_uhf_columns_match_second_quantisation(workloads):
    info = workloads("u1")
    assert info["core"] == 0.0
    dets, H = _u1_matrix()
    idx = {int(d): i for i, d in enumerate(dets)}
    o = Oracle(info["fcidump"])
    assert int(o.hf()[0]) in idx  # 5 alpha, 4 beta
    for j, d in enumerate(dets):
        cd, cv = o.column(np.array([d]))
        col = np.zeros(len(dets))
        for dd, v in zip(cd[:, 0], cv):
            col[idx[int(dd)]] = v
        assert np.allclose(col, H[:, j], atol=1e-12, rtol=0)
        assert np.all((np.abs(H[:, j]) > 1e-12) <= (col != 0))


def test_uhf_alpha_and_beta_blocks_differ(workloads):
    """The u1 columns depend on the beta integrals: the same file with the
    beta blocks replaced by the alpha ones gives other values."""
    spec = synth.CONFIGS["u1"]
    ha, hb, eaa, ebb, eab = synth.uhf_integrals(spec)
    na, nb = (spec.nelec + spec.ms2) // 2, (spec.nelec - spec.ms2) // 2
    _, H = _u1_matrix()
    _, H0 = ed.hamiltonian_uhf(ha, ha, eaa, eaa, eaa, 0.0, na, nb)
    assert np.abs(H - H0).max() > 1e-3


def test_uhf_solver_reaches_exact_ground_state(workloads):
    info = workloads("u1")
    _, H = _u1_matrix()
    e0 = np.linalg.eigvalsh(H)[0]
    o = Oracle(info["fcidump"])
    o.solver_init(8, 8, 0.0, 103546.0, nthreads=2)
    o.step(6000)
    assert abs(o.report()["energy"] - e0) < 1e-10


def test_restricted_file_in_unrestricted_dialect_is_the_same_run(workloads):
    a = Oracle(workloads("c1")["fcidump"])
    b = Oracle(workloads("c1_as_uhf")["fcidump"])
    assert np.array_equal(a.hf(), b.hf())
    d, _ = a.column(a.hf())
    for det in [a.hf()] + [d[i] for i in range(0, len(d), 7)]:
        da, va = a.column(det)
        db, vb = b.column(det)
        A = {int(x[0]): v for x, v in zip(da, va)}
        B = {int(x[0]): v for x, v in zip(db, vb)}
        assert A == B
    for o in (a, b):
        o.solver_init(8, 8, 0.0, 103546.0, nthreads=2)
        o.step(300)
    ra, rb = a.report(), b.report()
    assert ra["energy"] == rb["energy"] and ra["x_size"] == rb["x_size"] and ra["z_size"] == rb["z_size"]
    assert a.table_digest() == b.table_digest()


def _write_uhf(path, lines, norb_so=4, nelec=2):
    with open(path, "w") as f:
        f.write(" &FCI NORB=%d,NELEC=%d,MS2=0,\n  UHF=.TRUE.,\n &END\n" % (norb_so, nelec))
        for v, i, a, j, b in lines:
            f.write("%25.17E %4d %4d %4d %4d\n" % (v, i, a, j, b))
        f.write("%25.17E %4d %4d %4d %4d\n" % (0.0, 0, 0, 0, 0))


@pytest.mark.parametrize("lines,norb_so,message", [
    ([(0.5, 1, 1, 1, 1), (0.1, 1, 2, 3, 3)], 4, "spin-changing two-electron"),   # (1a 1b|2a 2a)
    ([(0.5, 1, 1, 1, 1), (-1.0, 1, 2, 0, 0)], 4, "spin-changing one-electron"),  # h(1a, 1b)
    ([(0.5, 1, 1, 1, 1)], 5, "odd number of spin-orbitals"),
])
def test_device_reader_refuses_what_it_cannot_hold(tmp_path, lines, norb_so, message):
    import mcdfci_amd
    lib = mcdfci_amd.load_library()
    path = str(tmp_path / "FCIDUMP")
    _write_uhf(path, lines, norb_so)
    o = mcdfci_amd.Opts(1, 1, 0.0, 0.01, 0.0, 0, 0, 0)
    h = C.c_void_p()
    rc = lib.mcdfci_create_from_fcidump(os.fsencode(path), C.byref(o), C.byref(h))
    assert rc == 1  # MCDFCI_E_INPUT, before any device call
    assert message in lib.mcdfci_last_error(None).decode()
    assert not h.value


def test_zero_spin_changing_integrals_are_accepted_by_the_reader(tmp_path):
    """A spin-changing integral of value 0 adds nothing and is dropped: the
    reader gets past the integrals (a CPU-only box then fails at the
    device, not at the file)."""
    import mcdfci_amd
    lib = mcdfci_amd.load_library()
    path = str(tmp_path / "FCIDUMP")
    _write_uhf(path, [(0.5, 1, 1, 1, 1), (0.0, 1, 2, 3, 3), (0.0, 1, 2, 0, 0), (-1.0, 1, 1, 0, 0)])
    o = mcdfci_amd.Opts(1, 1, 0.0, 0.01, 0.0, 0, 0, 0)
    h = C.c_void_p()
    rc = lib.mcdfci_create_from_fcidump(os.fsencode(path), C.byref(o), C.byref(h))
    if rc == 0:
        lib.mcdfci_destroy(h)
    else:
        msg = lib.mcdfci_last_error(None).decode()
        assert "spin-changing" not in msg and "Can not read" not in msg


@pytest.mark.skipif(not __import__("oracle_ctypes").ref_available(), reason="oracle/_ref not built")
def test_two_word_unrestricted_matches_the_reference(tmp_path):
    """34 orbitals per spin (68 spin-orbitals, two words): HF, diagonal and
    columns of the oracle against the reference's HamiltonianMolecule<2>
    reading the same unrestricted file."""
    from oracle_ctypes import RefHam
    spec = synth.SyntheticSpec("w2u", norb=34, nelec=10, rank=10, seed=8, uhf=True)
    info = synth.generate(spec, str(tmp_path))
    o, r = Oracle(info["fcidump"]), RefHam(info["fcidump"], 0.0)
    assert o.nw == 2
    assert np.array_equal(o.hf(), r.hf())
    assert o.diagonal(o.hf()) == r.diagonal(r.hf())
    d, _ = o.column(o.hf())
    hi = [x for x in d if int(x[1]) != 0]
    rng = np.random.default_rng(9)
    seeds = [o.hf()] + [hi[i] for i in rng.choice(len(hi), 10, replace=False)]
    for det in seeds:
        od, ov = o.column(det)
        rd, rv = r.column(det)
        A = {tuple(int(w) for w in x[:2]): v for x, v in zip(od, ov)}
        B = {tuple(int(w) for w in x[:2]): v for x, v in zip(rd, rv)}
        assert A == B
