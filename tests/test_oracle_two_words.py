"""Pins the oracle's two-word determinant path (33 to 63 spatial orbitals,
spin-orbitals 64 and up in word 1), which the norb-42 workload runs on.

Two independent checks:
  * the reference's own code for N = 2 words (oracle/_ref/libref.so built
    from src/ by oracle/Makefile): FCIDUMP parse, get_hartree_fock,
    get_diagonal, get_column, get_off_diagonal, the subspace matrix, the
    gcd_grad pick and WaveFunctionCuckoo::update_x / update_z_and_get_sub
    at k = 1, on a 34-orbital synthetic whose excitations cross the word
    boundary (src/determinant.h:24-152, src/hamiltonian.h:98-492,
    src/solver.h:196-224,637-670, src/wavefunction.h:398-605);
  * second quantisation with Jordan-Wigner signs over all 66 spin-orbitals
    of a 33-orbital Hamiltonian (Python integers, no word split at all):
    sampled columns with electrons on both sides of spin-orbital 64, and
    the exact ground state of the (1 alpha, 1 beta) sector (dimension
    1089), which the oracle's solver must reach.
"""

import numpy as np
import pytest

from mcdfci_amd import synth
from oracle_ctypes import Oracle, RefHam, ref_available, ref_lib

W2 = synth.SyntheticSpec("w2", norb=34, nelec=10, rank=10, seed=7)


@pytest.fixture(scope="module")
def w2(tmp_path_factory):
    return synth.generate(W2, str(tmp_path_factory.mktemp("w2")))


def _words(d):
    return tuple(int(w) for w in np.asarray(d, dtype=np.uint64).reshape(-1)[:2])


def _as_dict(dets, vals):
    return {_words(d): v for d, v in zip(dets, vals)}


def _high_word_seeds(o, n, rng):
    """Determinants two excitations from HF with electrons in word 1."""
    d, _ = o.column(o.hf())
    hi = [x for x in d if int(x[1]) != 0]
    picks = [hi[i] for i in rng.choice(len(hi), min(n, len(hi)), replace=False)]
    out = list(picks)
    for p in picks[:4]:
        c, _ = o.column(p)
        hi2 = [x for x in c if bin(int(x[1])).count("1") >= 2]
        out += [hi2[i] for i in rng.choice(len(hi2), min(3, len(hi2)), replace=False)]
    return out


need_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (no reference sources)")


@pytest.fixture(scope="module")
def pair(w2):
    if not ref_available():
        pytest.skip("oracle/_ref not built")
    return Oracle(w2["fcidump"]), RefHam(w2["fcidump"], 0.0)


@need_ref
def test_two_words_hf_diagonal(pair):
    o, r = pair
    assert o.nw == 2
    assert np.array_equal(o.hf(), r.hf())
    assert o.diagonal(o.hf()) == r.diagonal(r.hf())


@need_ref
def test_two_words_columns_bit_exact(pair):
    o, r = pair
    rng = np.random.default_rng(21)
    seeds = [o.hf()] + _high_word_seeds(o, 8, rng)
    assert sum(1 for s in seeds if int(s[1]) != 0) >= 8
    for d in seeds:
        od, ov = o.column(d)
        rd, rv = r.column(d)
        assert len(od) == len(rd)
        A, B = _as_dict(od, ov), _as_dict(rd, rv)
        assert A.keys() == B.keys()
        assert all(A[key] == B[key] for key in A)
        assert _words(od[0]) == _words(d)
        assert r.diagonal(d) == o.diagonal(d)


@need_ref
def test_two_words_off_diagonal_and_submatrix(pair):
    o, r = pair
    rng = np.random.default_rng(23)
    picks = np.array(_high_word_seeds(o, 10, rng)[:14], dtype=np.uint64)
    H = r.submatrix(picks)
    for i in range(len(picks)):
        for j in range(len(picks)):
            v = o.diagonal(picks[i]) if i == j else o.off_diagonal(picks[i], picks[j])
            assert v == H[i, j]
    c, _ = o.column(picks[0])
    for b in rng.choice(len(c), 40):
        assert o.off_diagonal(picks[0], c[b]) == r.off_diagonal(picks[0], c[b])
        assert o.off_diagonal(o.hf(), c[b]) == r.off_diagonal(r.hf(), c[b])


@need_ref
def test_two_words_pick_matches_reference(pair):
    """gcd_grad pick with duplicates and exact ties among two-word
    determinants.  The reference breaks a tie by the sign of
    (long)(a_w - b_w) from word 0 (src/wavefunction.h:30-36), an order only
    while the word-0 values lie less than 2^63 apart; the candidates here
    keep bit 63 of word 0 clear, as every tie the oracle is asked to match."""
    o, r = pair
    lib = ref_lib()
    rng = np.random.default_rng(29)
    pool = [d for d in _high_word_seeds(o, 30, rng) if int(d[0]) < (1 << 63)]
    c, _ = o.column(o.hf())
    pool += [d for d in c if int(d[0]) < (1 << 63)]
    pool = np.array(pool, dtype=np.uint64)
    for trial in range(40):
        n = int(rng.integers(10, 60))
        idx = rng.choice(len(pool), n)
        dets = np.ascontiguousarray(pool[idx].reshape(-1))
        x = np.round(rng.normal(size=len(pool)), 3)[idx]
        z = np.round(rng.normal(size=len(pool)), 3)[idx]
        if trial % 2:
            z[: n // 2] = np.abs(z[: n // 2])
            x[: n // 2] = 0.0
        k = int(min(len({_words(d) for d in pool[idx]}), rng.integers(1, 9)))
        xx, sc = float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.5, 2.0))
        out_r = np.zeros(2 * k, np.uint64)
        lib.ref_pick(2, n, dets, x, z, xx, sc, k, out_r, np.zeros(k), np.zeros(k))
        out_o = np.zeros(2 * k, np.uint64)
        assert o.lib.oc_pick(o.h, n, dets, x, z, np.array([xx, 0.0, sc, 0.0]), k, out_o) == k
        assert out_o.tolist() == out_r.tolist()


@need_ref
@pytest.mark.parametrize("z_thr", [0.0, 1e-3])
def test_two_words_wavefunction_k1_matches_reference(pair, z_thr):
    """update_x + update_z_and_get_sub with one two-word coordinate per step."""
    o, r = pair
    lib = ref_lib()
    M = 40000  # above the column length: every present element is a candidate, no boundary ties
    wf = lib.ref_wf_create(2, 1 << 22, M)
    o.lib.oc_wf_reset(o.h, 1, M, z_thr, float(1 << 22), 2)
    rng = np.random.default_rng(31)
    det = np.array(o.hf(), np.uint64)
    sc = np.zeros(8)
    for step in range(8):
        dx = np.array([1.0 if step == 0 else float(rng.normal() * 0.3)])
        xx, scale = (1.0, 1.0) if step == 0 else (float(rng.uniform(1.0, 2.0)), float(rng.uniform(0.6, 1.4)))
        lib.ref_wf_scalars(wf, 0, sc)
        lib.ref_wf_scalars(wf, 1, np.array([xx, scale, sc[2]]))
        o.lib.oc_wf_set_scalars(o.h, np.array([xx, 0.0, scale, 0.0]), sc[2])
        lib.ref_wf_update_x(wf, 1, det, dx)
        cd, cv = r.column(det)
        off = np.array([0, len(cd)], np.int64)
        cap = len(cd) + 8
        rdets, rx, rz = np.zeros(2 * cap, np.uint64), np.zeros(cap), np.zeros(cap)
        nr = lib.ref_wf_update_z(wf, 1, off, np.ascontiguousarray(cd.reshape(-1)), cv, det, dx, z_thr,
                                 rdets, rx, rz, cap)
        o.lib.oc_wf_update(o.h, 1, det, dx)
        od, ox, oz = o.candidates()
        assert nr == len(od)
        A = {_words(rdets[2 * i:2 * i + 2]): (rx[i], rz[i]) for i in range(nr)}
        B = {_words(d): (a, b) for d, a, b in zip(od, ox, oz)}
        assert A.keys() == B.keys()
        for key in A:
            assert A[key][0] == B[key][0]
            assert abs(A[key][1] - B[key][1]) <= 1e-13 * max(1.0, abs(A[key][1]))
        st = np.zeros(5)
        o.lib.oc_wf_state(o.h, st)
        lib.ref_wf_scalars(wf, 0, sc)
        assert st[0] == sc[3] and st[1] == sc[4] and st[2] == sc[5]
        assert abs(st[3] - sc[2]) <= 1e-12 * max(1.0, abs(sc[2]))
        # next coordinate: a high-word candidate with a large gradient
        key = np.abs(ox * (xx * scale * scale) + oz)
        order = np.argsort(-key, kind="stable")
        nxt = [od[i] for i in order if _words(od[i]) != _words(det) and int(od[i][1]) != 0]
        det = np.array(nxt[0], np.uint64)
    lib.ref_wf_destroy(wf)


# ---------------------------------------------------------------- second quantisation
def _parity_below(D, s):
    return bin(D & ((1 << s) - 1)).count("1") & 1


def sq_column(J, h, eri, core, ns):
    """{I: <I|H|J>} by applying sum h_pq a+_p a_q + 1/2 sum (pq|rs) a+_p a+_r a_s a_q
    to the determinant J (a Python integer over ns spin-orbitals; the sign of
    each operator is the parity of the occupied spin-orbitals below it)."""
    out = {J: core}
    occ = [s for s in range(ns) if J >> s & 1]
    for q in occ:
        s1 = _parity_below(J, q)
        J1 = J ^ (1 << q)
        for p in range(q & 1, ns, 2):
            if J1 >> p & 1:
                continue
            I = J1 | (1 << p)
            v = h[p >> 1, q >> 1]
            out[I] = out.get(I, 0.0) + (-v if s1 ^ _parity_below(J1, p) else v)
    for q in occ:
        s1 = _parity_below(J, q)
        J1 = J ^ (1 << q)
        for s in occ:
            if s == q:
                continue
            s2 = _parity_below(J1, s)
            J2 = J1 ^ (1 << s)
            for r in range(s & 1, ns, 2):
                if J2 >> r & 1:
                    continue
                s3 = _parity_below(J2, r)
                J3 = J2 | (1 << r)
                for p in range(q & 1, ns, 2):
                    if J3 >> p & 1:
                        continue
                    I = J3 | (1 << p)
                    v = 0.5 * eri[p >> 1, q >> 1, r >> 1, s >> 1]
                    out[I] = out.get(I, 0.0) + (-v if s1 ^ s2 ^ s3 ^ _parity_below(J3, p) else v)
    return out


def _int(d):
    w = _words(d)
    return w[0] | (w[1] << 64)


def _split(J):
    return np.array([J & ((1 << 64) - 1), J >> 64], np.uint64)


def test_two_word_columns_match_second_quantisation(tmp_path):
    spec = synth.SyntheticSpec("sq33", norb=33, nelec=5, ms2=1, rank=8, seed=13)
    info = synth.generate(spec, str(tmp_path))
    o = Oracle(info["fcidump"])
    assert o.nw == 2
    h, eri, core, ns = info["h"], info["eri"], info["core"], 66
    rng = np.random.default_rng(41)
    seeds = [o.hf()]
    # random determinants of the (3 alpha, 2 beta) sector with electrons on
    # both sides of spin-orbital 64 (spatial orbital 32 straddles the words)
    while len(seeds) < 14:
        a = rng.choice(33, 3, replace=False)
        b = rng.choice(33, 2, replace=False)
        J = sum(1 << (2 * int(p)) for p in a) | sum(1 << (2 * int(p) + 1) for p in b)
        if J >> 64:
            seeds.append(_split(J))
    seeds.append(_split((1 << 64) | (1 << 65) | (1 << 62) | (1 << 0) | (1 << 2)))
    for d in seeds:
        J = _int(d)
        ref = sq_column(J, h, eri, core, ns)
        od, ov = o.column(d)
        got = {_int(x): v for x, v in zip(od, ov)}
        assert len(got) == len(od)
        for I, v in got.items():
            assert abs(v - ref.get(I, 0.0)) <= 1e-12, (hex(J), hex(I), v, ref.get(I))
        missing = [I for I, v in ref.items() if abs(v) > 1e-10 and I not in got]
        assert not missing


@pytest.mark.parametrize("na,nb,k", [(2, 0, 8), (1, 1, 16)])
def test_two_word_solver_reaches_exact_ground_state(tmp_path, na, nb, k):
    """Two electrons in 33 orbitals, every determinant with an electron in or
    next to word 1: the oracle's solver must reach the lowest eigenvalue of
    the second-quantised H of the sector (dimension 528 or 1089).  In the
    (1 alpha, 1 beta) sector the lowest state is a singlet 3.2 mHa below the
    triplet; at k = 4 or 8 the iteration settles on the triplet (an
    eigenvector the exact spin symmetry of H keeps it on once the picks
    split its pairs), at k = 16 on the singlet, as here."""
    spec = synth.SyntheticSpec("ed33", norb=33, nelec=na + nb, ms2=na - nb, rank=6, seed=17, h_diag_a=-1.0,
                               h_diag_b=0.05)
    info = synth.generate(spec, str(tmp_path))
    h, eri, core = info["h"], info["eri"], info["core"]
    from itertools import combinations
    dets = sorted(sum(1 << (2 * p) for p in a) | sum(1 << (2 * q + 1) for q in b)
                  for a in combinations(range(33), na) for b in combinations(range(33), nb))
    idx = {d: i for i, d in enumerate(dets)}
    H = np.zeros((len(dets), len(dets)))
    for j, J in enumerate(dets):
        for I, v in sq_column(J, h, eri, core, 66).items():
            H[idx[I], j] += v
    assert np.allclose(H, H.T, atol=1e-12)
    e0 = np.linalg.eigvalsh(H)[0]
    assert e0 < 0
    o = Oracle(info["fcidump"])
    # the lowest orbitals of each spin (with MS2 != 0 and no orbital energies
    # the reference's Hartree-Fock rule, src/hamiltonian.h:469-492, fills
    # only nelec spin-orbitals in index order and would leave one out)
    ref = sum(1 << (2 * p) for p in range(na)) | sum(1 << (2 * q + 1) for q in range(nb))
    o.solver_init(k, k, 0.0, 1e6, nthreads=2, ref=np.array([ref, 0], np.uint64))
    o.step(6000)
    e = o.report()["energy"]
    assert abs(e - e0) < 1e-9, (e, e0)
