"""The CPU oracle against the reference's own code (oracle/_ref/libref.so,
compiled from /root/reference/src by oracle/Makefile).

Pins: FCIDUMP parsing, the RHF expansion and the unrestricted dialect
(src/input.h:125-127,203-285; the u1 open shell and the c1 molecule
rewritten in spin-orbitals), the excitation
tables, get_column / get_diagonal / get_off_diagonal / get_hartree_fock
(src/hamiltonian.h), CDFCI::get_hamiltonian_submatrix and the gcd_grad pick
(src/solver.h), and WaveFunctionCuckoo::update_x / update_z_and_get_sub at
k = 1, where the reference is deterministic (src/wavefunction.h).
"""

import numpy as np
import pytest

from oracle_ctypes import Oracle, RefHam, ref_available, ref_lib

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (no reference sources)")

CASES = ["c1", "c3_synth", "c3_hubbard", "u1", "u3", "c1_as_uhf"]


def _as_dict(dets, vals):
    return {tuple(int(w) for w in d): v for d, v in zip(dets, vals)}


@pytest.fixture(params=CASES + ["h2o"])
def pair(request, workloads, h2o_fcidump):
    path = h2o_fcidump if request.param == "h2o" else workloads(request.param)["fcidump"]
    return Oracle(path, 0.0), RefHam(path, 0.0), path


def test_hartree_fock_and_diagonal(pair):
    o, r, _ = pair
    assert np.array_equal(o.hf(), r.hf())
    assert o.diagonal(o.hf()) == r.diagonal(r.hf())


def test_columns_bit_exact(pair):
    o, r, _ = pair
    rng = np.random.default_rng(7)
    seeds = [o.hf()]
    rd, _ = r.column(o.hf())
    seeds += [rd[i] for i in rng.choice(len(rd), min(25, len(rd)), replace=False)]
    for d in seeds:
        od, ov = o.column(d)
        rdd, rv = r.column(d)
        assert len(od) == len(rdd)
        A, B = _as_dict(od, ov), _as_dict(rdd, rv)
        assert A.keys() == B.keys()
        assert all(A[key] == B[key] for key in A)  # bit-exact values and signs
        assert od[0].tolist() == list(d.reshape(-1)[: o.nw])  # diagonal first


def test_off_diagonal_and_submatrix(pair):
    o, r, _ = pair
    rng = np.random.default_rng(11)
    rd, _ = r.column(o.hf())
    picks = rd[rng.choice(len(rd), min(12, len(rd)), replace=False)]
    H = r.submatrix(picks)
    k = len(picks)
    for i in range(k):
        for j in range(k):
            v = o.diagonal(picks[i]) if i == j else o.off_diagonal(picks[i], picks[j])
            assert v == H[i, j]
    # also pairs at excitation level 3 and above give 0
    far = rng.choice(len(rd), 4)
    for a in far:
        c2, _ = r.column(rd[a])
        for b in rng.choice(len(c2), 5):
            assert o.off_diagonal(o.hf(), c2[b]) == r.off_diagonal(r.hf(), c2[b])


def test_pick_matches_reference(workloads):
    """gcd_grad pick on candidate lists with duplicates and exact ties."""
    path = workloads("c1")["fcidump"]
    o = Oracle(path)
    lib = ref_lib()
    rng = np.random.default_rng(3)
    d, _ = o.column(o.hf())
    for trial in range(40):
        n = int(rng.integers(10, 60))
        idx = rng.choice(len(d), n)
        dets = np.ascontiguousarray(d[idx, 0])
        x = np.round(rng.normal(size=len(d)), 3)[idx]
        z = np.round(rng.normal(size=len(d)), 3)[idx]  # duplicates share (x, z)
        if trial % 2:
            z[: n // 2] = np.abs(z[: n // 2])  # create exact ties with x = 0
            x[: n // 2] = 0.0
        k = int(min(len(np.unique(dets)), rng.integers(1, 9)))
        xx, sc = float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.5, 2.0))
        out_r = np.zeros(k, np.uint64)
        lib.ref_pick(1, n, dets, x, z, xx, sc, k, out_r, np.zeros(k), np.zeros(k))
        out_o = np.zeros(k, np.uint64)
        got = o.lib.oc_pick(o.h, n, dets, x, z, np.array([xx, 0.0, sc, 0.0]), k, out_o)
        assert got == k
        assert out_o.tolist() == out_r.tolist()


@pytest.mark.parametrize("name,z_thr", [("c1", 0.0), ("c3_synth", 0.0), ("c3_synth", 1e-3), ("u1", 0.0),
                                         ("u3", 1e-3)])
def test_wavefunction_k1_matches_reference(workloads, name, z_thr):
    """update_x + update_z_and_get_sub with one coordinate per step."""
    path = workloads(name)["fcidump"]
    o = Oracle(path)
    r = RefHam(path)
    lib = ref_lib()
    M = 4000  # every present element of a column is a candidate: no boundary ties
    wf = lib.ref_wf_create(1, 1 << 24, M)
    o.lib.oc_wf_reset(o.h, 1, M, z_thr, float(1 << 24), 2)
    rng = np.random.default_rng(5)
    det = o.hf()
    sc = np.zeros(8)
    for step in range(12):
        dx = np.array([1.0 if step == 0 else float(rng.normal() * 0.3)])
        xx, scale = float(rng.uniform(1.0, 2.0)), float(rng.uniform(0.6, 1.4))
        if step == 0:
            xx, scale = 1.0, 1.0
        lib.ref_wf_scalars(wf, 0, sc)
        lib.ref_wf_scalars(wf, 1, np.array([xx, scale, sc[2]]))
        o.lib.oc_wf_set_scalars(o.h, np.array([xx, 0.0, scale, 0.0]), sc[2])
        lib.ref_wf_update_x(wf, 1, np.array(det, np.uint64), dx)
        cd, cv = r.column(det)
        off = np.array([0, len(cd)], np.int64)
        cap = len(cd) + 8
        rdets, rx, rz = np.zeros(cap, np.uint64), np.zeros(cap), np.zeros(cap)
        nr = lib.ref_wf_update_z(wf, 1, off, np.ascontiguousarray(cd[:, 0]), cv, np.array(det, np.uint64), dx,
                                 z_thr, rdets, rx, rz, cap)
        o.lib.oc_wf_update(o.h, 1, np.array(det, np.uint64), dx)
        od, ox, oz = o.candidates()
        assert nr == len(od)
        A = {int(a): (b, c) for a, b, c in zip(rdets[:nr], rx[:nr], rz[:nr])}
        B = {int(a): (b, c) for a, b, c in zip(od[:, 0], ox, oz)}
        assert A.keys() == B.keys()
        for key in A:
            assert A[key][0] == B[key][0]  # x bit-exact
            assert abs(A[key][1] - B[key][1]) <= 1e-13 * max(1.0, abs(A[key][1]))
        st = np.zeros(5)
        o.lib.oc_wf_state(o.h, st)
        lib.ref_wf_scalars(wf, 0, sc)
        assert st[0] == sc[3] and st[1] == sc[4] and st[2] == sc[5]  # x_size, z_size, updates
        assert abs(st[3] - sc[2]) <= 1e-12 * max(1.0, abs(sc[2]))  # x.z
        # next coordinate: the largest-gradient candidate that differs from det
        key = np.abs(ox * (xx * scale * scale) + oz)
        order = np.argsort(-key, kind="stable")
        nxt = [od[i, 0] for i in order if od[i, 0] != det[0]]
        det = np.array([nxt[0]], np.uint64)
    lib.ref_wf_destroy(wf)
