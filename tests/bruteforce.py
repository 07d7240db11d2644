"""Independent dense Hamiltonian for tiny systems, by explicit second quantization.

Determinants are bit masks over 2L spin-orbitals in the reference's interleaved order (alpha
orbital p -> 2p, beta p -> 2p+1); |D> = a+_{p1} a+_{p2} ... |0> with p1 < p2 < ..., so
a_p and a+_p carry (-1)^{#occupied spin-orbitals below p}.  H = I0 + sum h_pq a+_p a_q +
1/2 sum (pq|rs) a+_p a+_r a_s a_q over same-spin pairs.  Nothing here shares code or formulas
with the oracle's Slater-Condon rules: it applies the operators one by one.
"""
from __future__ import annotations

import numpy as np


def _sign_below(det: int, p: int) -> int:
    return -1 if bin(det & ((1 << p) - 1)).count("1") & 1 else 1


def annihilate(det, p):
    if not (det >> p) & 1:
        return None, 0
    return det & ~(1 << p), _sign_below(det, p)


def create(det, p):
    if (det >> p) & 1:
        return None, 0
    return det | (1 << p), _sign_below(det, p)


def interleave(a: int, b: int, L: int) -> int:
    d = 0
    for p in range(L):
        if (a >> p) & 1:
            d |= 1 << (2 * p)
        if (b >> p) & 1:
            d |= 1 << (2 * p + 1)
    return d


def dense_hamiltonian(h1, eri, ecore, alpha, beta, L, drop_exchange=False, flip_mixed=False):
    """H in the determinant basis index(ia, ib) = ia * len(beta) + ib.  The two flags plant
    mistakes (used to show the comparisons are sensitive to them)."""
    dets = [interleave(int(a), int(b), L) for a in alpha for b in beta]
    index = {d: k for k, d in enumerate(dets)}
    n = len(dets)
    H = np.zeros((n, n))
    so = range(2 * L)
    for j, dj in enumerate(dets):
        H[j, j] += ecore
        for p in so:
            for q in so:
                if p % 2 != q % 2:
                    continue
                v = h1[p // 2, q // 2]
                if v == 0.0:
                    continue
                d1, s1 = annihilate(dj, q)
                if d1 is None:
                    continue
                d2, s2 = create(d1, p)
                if d2 is None or d2 not in index:
                    continue
                H[index[d2], j] += v * s1 * s2
        for p in so:
            for q in so:
                if p % 2 != q % 2:
                    continue
                for r in so:
                    for s in so:
                        if r % 2 != s % 2:
                            continue
                        v = eri[p // 2, q // 2, r // 2, s // 2]
                        if drop_exchange and p % 2 == r % 2 and p != r:
                            continue
                        d1, s1 = annihilate(dj, q)
                        if d1 is None:
                            continue
                        d2, s2 = annihilate(d1, s)
                        if d2 is None:
                            continue
                        d3, s3 = create(d2, r)
                        if d3 is None:
                            continue
                        d4, s4 = create(d3, p)
                        if d4 is None or d4 not in index:
                            continue
                        sg = s1 * s2 * s3 * s4
                        if flip_mixed and p % 2 != r % 2 and d4 != dj:
                            sg = -sg
                        H[index[d4], j] += 0.5 * v * sg
    return H
