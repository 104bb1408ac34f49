"""Pins the oracle's mixed-precision column (the device's precision =
mixed): off-diagonal column elements evaluated in fp32 from fp32 roundings
of the integrals, the diagonal and everything downstream in fp64.

The reference has no fp32 form, so the checks come from mathematics:
  * on a dyadic Hamiltonian (every integral a multiple of 2^-10 below 2^5)
    every fp32 rounding and sum is exact, so the mixed column must equal
    the fp64 column bit for bit -- a dropped or extra term, a wrong sign, a
    transposed integral index changes a value;
  * on the synthetic molecule-like Hamiltonians every mixed value lies
    within the fp32 rounding bound of the fp64 value (which the reference's
    own tables pin), the determinant sets agree, and some values differ
    (the mode really evaluates in fp32);
  * the mixed solver converges to the lowest eigenvalue of the fp32-element
    matrix built by second quantisation (within the bound that matrix's
    distance from the fp64 one allows) and keeps z = H_mixed x.
"""

import numpy as np
import pytest

import ed
from mcdfci_amd import synth
from oracle_ctypes import Oracle

U32 = 2.0 ** -24


def _dict(dets, vals):
    return dict(zip(map(tuple, dets.tolist()), vals.tolist()))


def _dyadic(tmp_path, norb=7, nelec=10):
    spec = synth.SyntheticSpec("dy", norb=norb, nelec=nelec, rank=14, seed=11, core=9.0)
    h, eri = synth.molecule_integrals(spec)
    q = 2.0 ** 10
    h = np.round(h * q) / q
    eri = np.round(eri * q) / q
    path = str(tmp_path / "FCIDUMP.dyadic")
    synth.write_fcidump(path, h, eri, nelec, 0, 9.0)
    return path, h, eri


def test_mixed_equals_fp64_on_dyadic_integrals(tmp_path):
    path, h, eri = _dyadic(tmp_path)
    assert np.all(np.abs(h) < 32) and np.all(np.abs(eri) < 32)
    o64 = Oracle(path)
    o32 = Oracle(path, precision="mixed")
    d0 = o64.hf()
    dets, _ = o64.column(d0)
    nonzero = 0
    for d in dets[:60]:
        a = _dict(*o64.column(d))
        b = _dict(*o32.column(d))
        assert a.keys() == b.keys()
        assert all(a[key] == b[key] for key in a)
        nonzero += sum(1 for v in a.values() if v != 0.0)
    assert nonzero > 1000


@pytest.mark.parametrize("name", ["c1", "c3_synth"])
def test_mixed_within_fp32_bound_of_fp64(name, workloads):
    info = workloads(name)
    h, eri = info["h"], info["eri"]
    o64 = Oracle(info["fcidump"])
    o32 = Oracle(info["fcidump"], precision="mixed")
    ne = o64.nelec
    # uniform bounds: a single element sums h and ne table terms (each
    # rounded once, then ne fp32 adds); a double is one or two roundings and
    # one subtraction
    same = np.abs(np.einsum("prkk->prk", eri) - np.einsum("pkkr->prk", eri))
    opp = np.abs(np.einsum("prkk->prk", eri))
    big = np.abs(h) + ne * np.maximum(same, opp).max(axis=2)
    b_single = (ne + 2) * U32 * big.max()
    b_double = 4 * U32 * np.abs(eri).max()
    bound = max(b_single, b_double)
    d0 = o64.hf()
    dets, _ = o64.column(d0)
    rng = np.random.default_rng(1)
    differ = 0
    for d in [d0] + [dets[i] for i in rng.choice(len(dets), 8, replace=False)]:
        a = _dict(*o64.column(d))
        b = _dict(*o32.column(d))
        assert b.keys() <= a.keys()
        # an element drops out only if its fp32 value is exactly zero
        assert all(abs(a[key]) <= bound for key in a.keys() - b.keys())
        diag = tuple(np.asarray(d).reshape(-1).tolist())
        assert a[diag] == b[diag]  # the diagonal stays fp64
        err = max(abs(a[key] - b[key]) for key in b)
        assert err <= bound, (err, bound)
        differ += sum(1 for key in b if a[key] != b[key])
    assert differ > 0


def _mixed_dense(path, o32):
    """The matrix of fp32 column elements over the whole sector, assembled
    from the oracle's mixed columns, and the fp64 matrix by second
    quantisation."""
    norb, nelec, ms2, h, eri, core = ed.read_fcidump(path)
    dets, H = ed.hamiltonian(h, eri, core, (nelec + ms2) // 2, (nelec - ms2) // 2)
    idx = {int(d): i for i, d in enumerate(dets)}
    Hm = np.zeros_like(H)
    for j, d in enumerate(dets):
        cd, cv = o32.column(np.array([d]))
        for dd, v in zip(cd[:, 0], cv):
            Hm[idx[int(dd)], j] = v
    return dets, H, Hm


def test_mixed_solver_converges_to_fp32_matrix(workloads):
    path = workloads("c1")["fcidump"]
    o32 = Oracle(path, precision="mixed")
    dets, H, Hm = _mixed_dense(path, o32)
    assert np.array_equal(Hm, Hm.T)  # fp32 elements are symmetric too
    assert np.abs(Hm - H).max() < 1e-5
    e_m = np.linalg.eigvalsh(Hm)[0]
    e_64 = np.linalg.eigvalsh(H)[0]
    # Weyl: eigenvalues move by at most the spectral norm of the difference
    assert abs(e_m - e_64) <= np.linalg.norm(Hm - H, 2) + 1e-12
    o32.solver_init(8, 8, 0.0, 103546.0, nthreads=2)
    o32.step(20000)
    assert abs(o32.report()["energy"] - e_m) < 1e-10


def test_mixed_invariant_z_equals_Hmixed_x(workloads):
    path = workloads("c1")["fcidump"]
    o32 = Oracle(path, precision="mixed")
    dets, H, Hm = _mixed_dense(path, o32)
    idx = {int(d): i for i, d in enumerate(dets)}
    o32.solver_init(8, 8, 0.0, 103546.0, nthreads=2)
    o32.step(150)
    td, tx, tz, _ = o32.table()
    x = np.zeros(len(dets))
    z = np.zeros(len(dets))
    for d, a, b in zip(td[:, 0], tx, tz):
        x[idx[int(d)]] = a
        z[idx[int(d)]] = b
    # z_threshold = 0: the table keeps z = H_mixed x exactly as the fp64
    # form keeps z = H x
    assert np.allclose(z, Hm @ x, atol=1e-11)
    assert not np.allclose(z, H @ x, atol=1e-11, rtol=0)
