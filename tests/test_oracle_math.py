"""Pins the oracle with mathematics that does not depend on its code:
second-quantised exact diagonalisation, closed forms, LAPACK (numpy) eigen
decompositions, the algorithm's own invariants, and exact rational
arithmetic for the numeric building blocks."""

import itertools
import os
from fractions import Fraction

import numpy as np
import pytest

import ed
from oracle_ctypes import Oracle, oracle_lib
from mcdfci_amd import synth


def _ed_matrix(path):
    norb, nelec, ms2, h, eri, core = ed.read_fcidump(path)
    na, nb = (nelec + ms2) // 2, (nelec - ms2) // 2
    return ed.hamiltonian(h, eri, core, na, nb)


def test_columns_match_second_quantisation(workloads):
    """Every column element (value and sign) equals <D'|H|D> built from
    creation/annihilation operators in interleaved Jordan-Wigner order."""
    path = workloads("c1")["fcidump"]
    dets, H = _ed_matrix(path)
    idx = {int(d): i for i, d in enumerate(dets)}
    o = Oracle(path)
    for j, d in enumerate(dets):
        cd, cv = o.column(np.array([d]))
        col = np.zeros(len(dets))
        for dd, v in zip(cd[:, 0], cv):
            col[idx[int(dd)]] = v
        assert np.allclose(col, H[:, j], atol=1e-13, rtol=0)
        # column support = nonzero pattern of H (no element missing)
        assert np.all((np.abs(H[:, j]) > 1e-12) <= (col != 0))


def test_c1_converges_to_exact_ground_state(workloads):
    path = workloads("c1")["fcidump"]
    dets, H = _ed_matrix(path)
    e0 = np.linalg.eigvalsh(H)[0]
    o = Oracle(path)
    o.solver_init(8, 8, 0.0, 103546.0, nthreads=2)
    o.step(20000)
    r = o.report()
    assert abs(r["energy"] - e0) < 1e-10


def test_invariants_z_equals_Hx(workloads):
    """With z_threshold = 0 the table keeps z = H x, xx = sum x^2 and
    x.z = sum x z (the reference's debug invariant, src/wavefunction.h:640-667)."""
    path = workloads("c1")["fcidump"]
    dets, H = _ed_matrix(path)
    idx = {int(d): i for i, d in enumerate(dets)}
    o = Oracle(path)
    o.solver_init(8, 8, 0.0, 103546.0, nthreads=2)
    o.step(150)
    td, tx, tz, _ = o.table()
    x = np.zeros(len(dets))
    z = np.zeros(len(dets))
    for d, a, b in zip(td[:, 0], tx, tz):
        x[idx[int(d)]] = a
        z[idx[int(d)]] = b
    assert np.allclose(z, H @ x, atol=1e-11)
    r = o.report()
    assert abs(r["xx"] - x @ x) < 1e-10 * (x @ x)
    assert abs(r["xz"] - x @ z) < 1e-10 * abs(x @ z)
    assert abs(r["energy"] - (x @ H @ x) / (x @ x)) < 1e-11


def test_two_site_hubbard_closed_form(tmp_path):
    """H2 Hubbard dimer: E0 = (U - sqrt(U^2 + 16 t^2)) / 2 (+ core)."""
    t, U, core = 1.0, 4.0, -10.0
    spec = synth.SyntheticSpec("hub2", norb=2, nelec=2, kind="hubbard", hubbard_t=t, hubbard_u=U, core=core)
    h, eri = synth.hubbard_integrals(spec)
    # periodic ring of two sites double-counts the bond; use a single bond
    h = np.array([[0.0, -t], [-t, 0.0]])
    path = tmp_path / "FCIDUMP"
    synth.write_fcidump(str(path), h, eri, 2, 0, core)
    o = Oracle(str(path))
    o.solver_init(2, 4, 0.0, 1000.0, nthreads=1)
    o.step(200)
    exact = core + (U - np.sqrt(U * U + 16 * t * t)) / 2
    assert abs(o.report()["energy"] - exact) < 1e-12


def test_small_hubbard_ring_matches_ed(tmp_path):
    spec = synth.SyntheticSpec("hub6", norb=6, nelec=6, kind="hubbard")
    h, eri = synth.hubbard_integrals(spec)
    path = tmp_path / "FCIDUMP"
    synth.write_fcidump(str(path), h, eri, 6, 0, -20.0)
    dets, H = _ed_matrix(str(path))
    e0 = np.linalg.eigvalsh(H)[0]
    o = Oracle(str(path))
    o.solver_init(4, 8, 0.0, 1e5, nthreads=1)
    o.step(8000)
    assert abs(o.report()["energy"] - e0) < 1e-9


@pytest.mark.parametrize("n", [1, 2, 3, 5, 9, 17, 33, 65, 129, 257])
def test_lowest_eigpair_vs_lapack(n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    A = A + A.T
    A[np.arange(n), np.arange(n)] -= 5.0 + np.arange(n) * 0.1
    lib = oracle_lib()
    import ctypes as C
    lam = C.c_double()
    qh, ql = np.zeros(n), np.zeros(n)
    lib.oc_lowest_eig(n, np.ascontiguousarray(A.reshape(-1)), C.byref(lam), qh, ql)
    w, V = np.linalg.eigh(A)
    assert abs(lam.value - w[0]) <= 1e-12 * max(1.0, abs(w[0]))
    v = V[:, 0] * np.sign(V[np.argmax(np.abs(V[:, 0]) > 1e-300), 0])
    lead = np.nonzero(qh)[0][0]
    v = v * np.sign(v[lead])
    assert np.allclose(qh, v, atol=1e-10)
    assert abs(np.dot(qh, qh) - 1.0) < 1e-14


def _lowest(A, n):
    import ctypes as C
    lam = C.c_double()
    qh, ql = np.zeros(n), np.zeros(n)
    oracle_lib().oc_lowest_eig(n, np.ascontiguousarray(A.reshape(-1)), C.byref(lam), qh, ql)
    return lam.value, qh, ql


def test_lanczos_finds_eigenvalue_hidden_from_start_vector():
    """The lowest eigenvector lives in a block that the start vector e_0 does
    not touch (an exact symmetry); the fixed perturbation of the start vector
    still finds it, as dsyevr would."""
    rng = np.random.default_rng(3)
    n1, n2 = 20, 30
    B1 = rng.normal(size=(n1, n1)) * 0.3
    B2 = rng.normal(size=(n2, n2)) * 0.3
    A = np.zeros((n1 + n2, n1 + n2))
    A[:n1, :n1] = B1 + B1.T - 2.0 * np.eye(n1)
    A[n1:, n1:] = B2 + B2.T - 6.0 * np.eye(n2)
    w, V = np.linalg.eigh(A)
    assert np.all(np.abs(V[:n1, 0]) == 0.0)
    lam, qh, ql = _lowest(A, n1 + n2)
    assert abs(lam - w[0]) < 1e-12 * abs(w[0])
    assert np.linalg.norm(A @ qh - w[0] * qh) < 1e-12 * np.abs(A).sum(1).max()


def test_lanczos_degenerate_and_breakdown():
    # doubly degenerate lowest eigenvalue: any vector of the eigenspace
    rng = np.random.default_rng(5)
    n = 24
    Qm, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.concatenate([[-3.0, -3.0], np.linspace(-1.0, 2.0, n - 2)])
    A = (Qm * ev) @ Qm.T
    A = (A + A.T) / 2
    lam, qh, _ = _lowest(A, n)
    assert abs(lam + 3.0) < 1e-12
    assert np.linalg.norm(A @ qh + 3.0 * qh) < 1e-12
    assert abs(qh @ qh - 1.0) < 1e-14
    # breakdown after one step: A = c I keeps the Krylov space one-dimensional
    A = -2.5 * np.eye(7)
    lam, qh, _ = _lowest(A, 7)
    assert abs(lam + 2.5) < 1e-15 and abs(qh @ qh - 1.0) < 1e-15


def test_eig_update_on_solver_subproblems_vs_numpy(workloads):
    """The subspace update inside a real run (norb-12 molecule, k = 64, the
    Lanczos step count carried from solve to solve) agrees with the numpy
    restatement of src/solver.h:333-609 on each iteration's inputs."""
    import ctypes as C
    path = workloads("c3_synth")["fcidump"]
    o = Oracle(path)
    k = 64
    o.solver_init(k, k, 0.0, 1e7, nthreads=4)
    lib = o.lib
    lib.oc_last_step.restype = C.c_int
    lib.oc_last_step.argtypes = [C.c_void_p] * 7
    for it in range(8):
        assert o.step(1) == 0
        dets = np.zeros(2 * k, np.uint64)
        x, z, dx = np.zeros(k), np.zeros(k), np.zeros(k)
        Hs, s6 = np.zeros(k * k), np.zeros(6)
        lib.oc_last_step(o.h, *[a.ctypes.data_as(C.c_void_p) for a in (dets, x, z, Hs, s6, dx)])
        xx = s6[0] + s6[1]
        hi, lo = s6[4:6].view(np.int64)
        xz = float((int(hi) << 64) + (int(lo) & ((1 << 64) - 1))) * 2.0 ** -64
        dx_ref, xx_ref, _ = _reference_update_numpy(Hs.reshape(k, k), x, z, xx, xz)
        if it == 0:
            # first-iteration branch: x_new = v sqrt(-lambda) with v's global
            # sign arbitrary (canonical here, LAPACK's choice in the reference)
            sgn = np.sign((x + dx) @ (x + dx_ref))
            dx_ref = sgn * (x + dx_ref) - x
        assert np.allclose(dx, dx_ref, atol=1e-10, rtol=0), it
        assert abs(o.report()["xx"] - xx_ref) <= 1e-12 * xx_ref


def _reference_update_numpy(Hs, x, z, xx, xz):
    """The subspace update of src/solver.h:333-609 written with numpy.eigh."""
    k = len(x)
    Hsym = np.triu(Hs) + np.triu(Hs, 1).T
    Hx = Hsym @ x
    yy = xx - x @ x
    if abs(yy) > 1e-14:
        y1 = np.sqrt(yy)
        A = np.zeros((k + 1, k + 1))
        A[0, 0] = (xz - 2 * (x @ z) + x @ Hx) / yy
        A[0, 1:] = A[1:, 0] = (z - Hx) / y1
        A[1:, 1:] = Hsym
        w, V = np.linalg.eigh(A)
        v = V[:, 0]
        lam = -w[0]
        gam = v[0] / y1
        dx = v[1:] / gam - x
        sc = gam * np.sqrt(lam)
        return dx, lam / sc / sc, abs(sc)
    w, V = np.linalg.eigh(Hsym)
    v = V[:, 0]
    return v * np.sqrt(-w[0]) - x, -w[0], 1.0


# k + 1 > 90 sums sixteen column classes (the multi-workgroup form), below four
@pytest.mark.parametrize("k", [1, 2, 8, 33, 64, 128, 256])
def test_eig_update_vs_numpy_and_identities(k):
    rng = np.random.default_rng(100 + k)
    Hs = rng.normal(size=(k, k)) * 0.1
    Hs = Hs + Hs.T
    Hs[np.arange(k), np.arange(k)] = -10.0 + rng.normal(size=k)
    x = rng.normal(size=k) * 0.2
    z = -10.0 * x + rng.normal(size=k) * 0.05
    xx = float(x @ x) + 1.5
    xz = -10.0 * xx
    s = np.array([xx, 0.0, 1.0, 0.0, xz, 0.0])
    dx = np.zeros(k)
    lib = oracle_lib()
    rc = lib.oc_eig_update(k, np.ascontiguousarray(Hs.reshape(-1)), x, z, s, dx, np.zeros(3))
    assert rc == 0
    dx_ref, xx_ref, sc_ref = _reference_update_numpy(Hs, x, z, xx, xz)
    assert np.allclose(dx, dx_ref, atol=1e-11)
    assert abs(s[0] - xx_ref) < 1e-11 * xx_ref
    assert abs(abs(s[2]) - sc_ref) < 1e-11
    # identity: |x_new|^2 = yy + sum (x + dx)^2 = xx_new
    yy = xx - x @ x
    assert abs(yy + np.sum((x + dx) ** 2) - s[0]) < 1e-10 * s[0]


def test_eig_update_first_iteration_branch():
    k = 8
    rng = np.random.default_rng(9)
    Hs = rng.normal(size=(k, k)) * 0.1
    Hs = Hs + Hs.T
    Hs[np.arange(k), np.arange(k)] = -5.0
    x = np.zeros(k)
    x[0] = 1.0
    z = Hs[:, 0].copy()
    s = np.array([1.0, 0.0, 1.0, 0.0, Hs[0, 0], 0.0])
    dx = np.zeros(k)
    assert oracle_lib().oc_eig_update(k, np.ascontiguousarray(Hs.reshape(-1)), x, z, s, dx, np.zeros(3)) == 0
    w, V = np.linalg.eigh(Hs)
    v = V[:, 0] * np.sign(V[0, 0])
    assert np.allclose(dx, v * np.sqrt(-w[0]) - x, atol=1e-12)
    assert abs(s[0] + w[0]) < 1e-12 and s[2] == 1.0


def test_eig_update_rejects_positive_spectrum():
    k = 3
    Hs = np.eye(k) * 2.0
    x = np.array([1.0, 0, 0])
    s = np.array([1.0, 0.0, 1.0, 0.0, 2.0, 0.0])
    assert oracle_lib().oc_eig_update(k, np.ascontiguousarray(Hs.reshape(-1)), x, Hs[:, 0].copy(), s,
                                      np.zeros(k), np.zeros(3)) == 1


def _fx(v):
    return Fraction(v)


def test_double_double_against_exact_rationals():
    lib = oracle_lib()
    rng = np.random.default_rng(0)
    out = np.zeros(8)
    for _ in range(400):
        a = float(rng.normal() * 10.0 ** rng.integers(-3, 4))
        al = a * 1e-17 * rng.normal()
        b = float(rng.normal() * 10.0 ** rng.integers(-3, 4))
        bl = b * 1e-17 * rng.normal()
        lib.oc_dd_ops(a, al, b, bl, out)
        A, B = _fx(a) + _fx(al), _fx(b) + _fx(bl)
        for got, exact in ((out[0:2], A + B), (out[2:4], A * B), (out[4:6], A / B)):
            err = abs(_fx(got[0]) + _fx(got[1]) - exact)
            assert err <= abs(exact) * Fraction(1, 2 ** 100) + Fraction(1, 2 ** 1000)
        if a > 0:
            got = _fx(out[6]) + _fx(out[7])
            assert abs(got * got - A) <= A * Fraction(1, 2 ** 98)


def test_fixed_point_round_trip_exact():
    lib = oracle_lib()
    rng = np.random.default_rng(1)
    o = np.zeros(2, np.int64)
    for _ in range(2000):
        d = float(rng.normal() * 10.0 ** rng.integers(-22, 17))
        lib.oc_to_fixed(d, o)
        v = (int(o[0]) << 64) | (int(o[1]) & ((1 << 64) - 1))
        if v >= 1 << 127:
            v -= 1 << 128
        exact = Fraction(d) * 2 ** 64
        # round half even to the nearest integer
        assert abs(Fraction(v) - exact) <= Fraction(1, 2)
        back = lib.oc_fixed_to_double(int(o[0]), int(o[1]))
        assert abs(_fx(back) - Fraction(v, 2 ** 64)) <= abs(Fraction(v, 2 ** 64)) * Fraction(1, 2 ** 51)


# ---------------------------------------------------------------- line search
def _ls_closed_form(h, x, z, xx):
    """CoordinateUpdateLS::operator() (src/solver.h:254-314) in numpy float64:
    Cardano / trigonometric root, then the Newton polish."""
    dA = -h
    p1 = xx - x * x - dA
    q = z + dA * x
    p3, q2 = p1 / 3, q / 2
    d = p3 * p3 * p3 + q2 * q2
    if d >= 0:
        r = np.sqrt(d)
        rt = np.cbrt(-q2 + r) + np.cbrt(-q2 - r)
    else:
        r = np.sqrt(-d)
        if q2 >= 0:
            rt = 2 * np.sqrt(-p3) * np.cos((np.arctan2(-r, -q2) - 2 * np.pi) / 3.0)
        else:
            rt = 2 * np.sqrt(-p3) * np.cos(np.arctan2(r, -q2) / 3.0)
    dx = rt - x

    def newton(dx):
        return dx - (dx * (dx * (dx + 3 * x) + (3 * x * x + p1)) + p1 * x + q + x * x * x) / \
            (dx * (3 * dx + 6 * x) + 3 * x * x + p1)

    dxn = newton(dx)
    it = 0
    while abs((dxn - dx) / dx) > 1e-12 and it < 10:
        dx, dxn = dxn, newton(dxn)
        it += 1
    return dx


def _ls_cases():
    rng = np.random.default_rng(11)
    cases = []
    for _ in range(400):
        xx = float(rng.uniform(0.2, 40.0))
        x = float(rng.normal() * np.sqrt(xx) * 0.5)
        h = float(rng.uniform(-30.0, 5.0))
        z = float(rng.normal() * 10.0)
        cases.append((h, x, z, xx))
    # three real roots of either sign of q, one real root, q = 0, x = 0
    cases += [(-20.0, 0.3, 1.5, 1.0), (-20.0, 0.3, -1.5, 1.0), (5.0, 0.1, 0.2, 1.0),
              (-10.0, 0.0, 0.0, 2.0), (-9.0, 1.0, -9.0, 1.0), (-3.0, 0.0, 0.7, 1.0)]
    return cases


def test_line_search_root_matches_closed_form():
    """The restated root (Newton from outside all roots, IEEE operations only)
    is the root the reference's closed form selects, to 1e-12 relative."""
    o = oracle_lib()
    for h, x, z, xx in _ls_cases():
        out = np.zeros(3)
        o.oc_ls_update(h, x, z, xx, 0.0, out)
        ref = _ls_closed_form(h, x, z, xx)
        assert abs(out[0] - ref) <= 1e-12 * max(1.0, abs(ref)), (h, x, z, xx, out[0], ref)
        t = x + out[0]
        p1, q = xx - x * x + h, z - h * x
        # a root, and the lowest of f along the coordinate (f' = t^3 + p1 t + q)
        assert abs(t ** 3 + p1 * t + q) <= 1e-10 * max(1.0, abs(t) ** 3 + abs(p1 * t) + abs(q))
        roots = np.roots([1.0, 0.0, p1, q])
        real = roots[np.abs(roots.imag) < 1e-9].real
        f = lambda s: s ** 4 / 4 + p1 * s * s / 2 + q * s
        assert f(t) <= min(f(r) for r in real) + 1e-9 * max(1.0, abs(f(t)))
        # xx bookkeeping: |x|^2 - x^2 + (x + dx)^2
        assert abs(out[1] - (xx - x * x + t * t)) <= 1e-13 * max(1.0, xx)


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "_ref",
                                                    "libref.so")), reason="oracle/_ref not built")
def test_line_search_matches_reference_code():
    from oracle_ctypes import ref_lib
    r = ref_lib()
    o = oracle_lib()
    for h, x, z, xx in _ls_cases():
        out = np.zeros(3)
        o.oc_ls_update(h, x, z, xx, 0.0, out)
        ref = r.ref_ls_update(h, x, z, xx)
        assert abs(out[0] - ref) <= 1e-12 * max(1.0, abs(ref)), (h, x, z, xx, out[0], ref)


def test_line_search_solver_converges_and_keeps_invariants(workloads):
    """k = 1 line-search iterations with |x|^2 and scale = 1 maintained:
    z = H x, xx = sum x^2, and the energy decreases to the ED ground state."""
    path = workloads("c1")["fcidump"]
    dets, H = _ed_matrix(path)
    e0 = np.linalg.eigvalsh(H)[0]
    idx = {int(d): i for i, d in enumerate(dets)}
    o = Oracle(path)
    o.solver_init(1, 8, 0.0, 103546.0, nthreads=1)
    o.set_update(1)
    last = 0.0
    for n in range(6):
        o.step(500)
        r = o.report()
        assert r["error"] == 0
        assert r["scale"] == 1.0
        assert r["energy"] <= last + 1e-12 or n == 0
        last = r["energy"]
    td, tx, tz, _ = o.table()
    x = np.zeros(len(dets))
    z = np.zeros(len(dets))
    for d, a, b in zip(td[:, 0], tx, tz):
        x[idx[int(d)]] = a
        z[idx[int(d)]] = b
    assert np.allclose(z, H @ x, atol=1e-11)
    assert abs(r["xx"] - x @ x) < 1e-10 * (x @ x)
    o.step(20000)
    assert abs(o.report()["energy"] - e0) < 1e-8


def test_line_search_takes_one_coordinate(workloads):
    o = Oracle(workloads("c1")["fcidump"])
    o.solver_init(4, 8, 0.0, 103546.0, nthreads=1)
    with pytest.raises(RuntimeError):
        o.set_update(1)


def _mix64(x):
    m = (1 << 64) - 1
    x ^= x >> 33
    x = (x * 0xff51afd7ed558ccd) & m
    x ^= x >> 33
    x = (x * 0xc4ceb9fe1a85ec53) & m
    x ^= x >> 33
    return x


def test_table_digest_definition(workloads):
    """The oracle's table digest is the documented function of the dumped
    entries (key words, x bits with 0 for a zero x, exact 128-bit z)."""
    import struct
    path = workloads("c1")["fcidump"]
    o = Oracle(path)
    o.solver_init(8, 8, 0.0, 103546.0, nthreads=2)
    o.step(40)
    td, tx, tz, zf = o.table()
    m = (1 << 64) - 1
    ks = es = 0
    for d, x, (zh, zl) in zip(td[:, 0].tolist(), tx.tolist(), zf.tolist()):
        e = _mix64((d + 0x9e3779b97f4a7c15) & m)
        e = _mix64(e ^ 0)
        ks = (ks + e) & m
        xb = 0 if x == 0.0 else struct.unpack("<Q", struct.pack("<d", x))[0]
        e = _mix64(e ^ xb)
        e = _mix64(e ^ (zh & m))
        es = (es + _mix64(e ^ (zl & m))) & m
    assert o.table_digest() == (len(td), ks, es)
