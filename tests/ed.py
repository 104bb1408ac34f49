"""Exact diagonalisation by second quantisation (test infrastructure).

Independent of the Slater-Condon tables: the Hamiltonian
    H = sum_{pq,s} h_pq a+_{ps} a_{qs}
      + 1/2 sum_{pqrs,st} (pq|rs) a+_{ps} a+_{rt} a_{st} a_{qs} + E_core
is applied operator by operator to every determinant of the (n_alpha,
n_beta) sector, with Jordan-Wigner signs taken in the interleaved
spin-orbital order (spin-orbital 2p = alpha p, 2p+1 = beta p), which is the
ordering of the reference's determinant bit strings (src/input.h:242-243).
"""

from __future__ import annotations

import itertools

import numpy as np


def read_fcidump(path):
    """Minimal RHF FCIDUMP reader returning (norb, nelec, ms2, h, eri, core)."""
    with open(path) as f:
        lines = f.read().splitlines()
    header = []
    i = 0
    while "&END" not in lines[i] and "/" not in lines[i]:
        header.append(lines[i])
        i += 1
    hdr = " ".join(header).replace(" ", "")
    import re
    norb = int(re.search(r"NORB=(\d+)", hdr).group(1))
    nelec = int(re.search(r"NELEC=(\d+)", hdr).group(1))
    ms2 = int(re.search(r"MS2=(\d+)", hdr).group(1))
    h = np.zeros((norb, norb))
    eri = np.zeros((norb, norb, norb, norb))
    core = 0.0
    for line in lines[i + 1:]:
        t = line.split()
        if len(t) < 5:
            continue
        v = float(t[0])
        p, q, r, s = (int(x) for x in t[1:5])
        if p and q and r and s:
            p, q, r, s = p - 1, q - 1, r - 1, s - 1
            for a, b, c, d in ((p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                               (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)):
                eri[a, b, c, d] = v
        elif q:
            h[p - 1, q - 1] = v
            h[q - 1, p - 1] = v
        elif p:
            pass
        else:
            core = v
    return norb, nelec, ms2, h, eri, core


def basis(norb, na, nb):
    dets = []
    for ao in itertools.combinations(range(norb), na):
        for bo in itertools.combinations(range(norb), nb):
            d = 0
            for p in ao:
                d |= 1 << (2 * p)
            for p in bo:
                d |= 1 << (2 * p + 1)
            dets.append(d)
    return np.array(sorted(dets), dtype=np.uint64)


def _below(s):
    return np.uint64((1 << s) - 1)


def _annihilate(D, valid, sign, s):
    bit = np.uint64(1 << s)
    valid = valid & ((D & bit) != 0)
    par = np.bitwise_count(D & _below(s)) & 1
    sign = np.where(par == 1, -sign, sign)
    return D ^ bit, valid, sign


def _create(D, valid, sign, s):
    bit = np.uint64(1 << s)
    valid = valid & ((D & bit) == 0)
    par = np.bitwise_count(D & _below(s)) & 1
    sign = np.where(par == 1, -sign, sign)
    return D | bit, valid, sign


def hamiltonian(h, eri, core, na, nb, tol=0.0):
    """Dense H over the (na, nb) sector; returns (dets, H)."""
    return _hamiltonian((h, h), lambda sp, st: eri, core, na, nb, tol)


def hamiltonian_uhf(ha, hb, eaa, ebb, eab, core, na, nb, tol=0.0):
    """Dense H from spin-resolved integrals: h of spin s, and (pq|rs) of
    spins (s, t) = eaa, ebb, eab[p,q,r,s] (alpha pair first) and
    eab[r,s,p,q] for (beta, alpha)."""
    eba = np.ascontiguousarray(eab.transpose(2, 3, 0, 1))
    blocks = {(0, 0): eaa, (1, 1): ebb, (0, 1): eab, (1, 0): eba}
    return _hamiltonian((ha, hb), lambda sp, st: blocks[(sp, st)], core, na, nb, tol)


def _hamiltonian(hs, block, core, na, nb, tol):
    n = hs[0].shape[0]
    dets = basis(n, na, nb)
    dim = len(dets)
    H = np.zeros((dim, dim))
    H[np.arange(dim), np.arange(dim)] += core
    cols = np.arange(dim)

    def scatter(Dn, valid, sign, coef):
        idx = np.searchsorted(dets, Dn[valid])
        np.add.at(H, (idx, cols[valid]), coef * sign[valid])

    for sp in (0, 1):
        h = hs[sp]
        for p in range(n):
            for q in range(n):
                if abs(h[p, q]) <= tol:
                    continue
                D = dets.copy()
                valid = np.ones(dim, bool)
                sign = np.ones(dim)
                D, valid, sign = _annihilate(D, valid, sign, 2 * q + sp)
                D, valid, sign = _create(D, valid, sign, 2 * p + sp)
                scatter(D, valid, sign, h[p, q])
    for sp in (0, 1):
        for st in (0, 1):
            eri = block(sp, st)
            for p, q, r, s in itertools.product(range(n), repeat=4):
                v = eri[p, q, r, s]
                if abs(v) <= tol:
                    continue
                D = dets.copy()
                valid = np.ones(dim, bool)
                sign = np.ones(dim)
                D, valid, sign = _annihilate(D, valid, sign, 2 * q + sp)
                D, valid, sign = _annihilate(D, valid, sign, 2 * s + st)
                D, valid, sign = _create(D, valid, sign, 2 * r + st)
                D, valid, sign = _create(D, valid, sign, 2 * p + sp)
                scatter(D, valid, sign, 0.5 * v)
    return dets, H


def ground_state(path):
    norb, nelec, ms2, h, eri, core = read_fcidump(path)
    na, nb = (nelec + ms2) // 2, (nelec - ms2) // 2
    dets, H = hamiltonian(h, eri, core, na, nb)
    w, v = np.linalg.eigh(H)
    return dets, H, w[0], v[:, 0]


def _apply(dets, vec, ops):
    """Apply a string of (kind, spin-orbital) operators, rightmost first, to
    the vector vec over dets; returns (target dets, coefficients) of the
    surviving terms."""
    D = dets.copy()
    valid = np.ones(len(dets), bool)
    sign = np.ones(len(dets))
    for kind, s in reversed(ops):
        if kind == "a":
            D, valid, sign = _annihilate(D, valid, sign, s)
        else:
            D, valid, sign = _create(D, valid, sign, s)
    return D[valid], (sign * vec)[valid]


def _expect(dets, c, ops):
    Dn, coef = _apply(dets, c, ops)
    idx = np.searchsorted(dets, Dn)
    ok = (idx < len(dets))
    ok[ok] = dets[idx[ok]] == Dn[ok]
    return float(np.dot(c[idx[ok]], coef[ok]))


def rdm12(dets, c, norb):
    """Spin-summed 1-RDM g[p,q] = sum_s <a+_ps a_qs> and 2-RDM (chemists'
    order) G[p,q,r,s] = sum_{st} <a+_ps a+_rt a_st a_qs> of the normalised
    vector c over the sorted determinant list dets, by second quantisation."""
    c = np.asarray(c, dtype=np.float64) / np.linalg.norm(c)
    g = np.zeros((norb, norb))
    G = np.zeros((norb, norb, norb, norb))
    for p in range(norb):
        for q in range(norb):
            for sp in (0, 1):
                g[p, q] += _expect(dets, c, [("c", 2 * p + sp), ("a", 2 * q + sp)])
    for p, q, r, s in itertools.product(range(norb), repeat=4):
        v = 0.0
        for sp in (0, 1):
            for st in (0, 1):
                v += _expect(dets, c, [("c", 2 * p + sp), ("c", 2 * r + st), ("a", 2 * s + st), ("a", 2 * q + sp)])
        G[p, q, r, s] = v
    return g, G


def orbit(p, q, r, s):
    """The distinct index tuples of the 8-fold symmetry orbit of (pq|rs)."""
    return sorted({(p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                   (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)})


def canonical(p, q, r, s):
    a, b = max(p, q), min(p, q)
    c, d = max(r, s), min(r, s)
    if a * (a + 1) // 2 + b < c * (c + 1) // 2 + d:
        a, b, c, d = c, d, a, b
    return a, b, c, d


def fold_gradient(g, G):
    """dE/dtheta for the unique parameters of E = sum h g + 1/2 sum (pq|rs) G
    + core, at the canonical positions (p >= q; canonical (pq|rs))."""
    n = g.shape[0]
    gh = np.zeros((n, n))
    for p in range(n):
        for q in range(p + 1):
            gh[p, q] = g[p, q] + g[q, p] if p != q else g[p, p]
    ge = np.zeros((n, n, n, n))
    for p, q, r, s in itertools.product(range(n), repeat=4):
        if (p, q, r, s) == canonical(p, q, r, s):
            ge[p, q, r, s] = 0.5 * sum(G[t] for t in orbit(p, q, r, s))
    return gh, ge
