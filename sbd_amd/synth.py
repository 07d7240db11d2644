"""Seeded synthetic molecular Hamiltonians and string sets for the selected-basis engine.

Every workload this package is measured or tested on is generated here, from fixed seeds,
with numpy's PCG64 generator, so the same instance is rebuilt bit-for-bit on any host.

Integral recipe (one per instance):
  * h_pp = sorted Uniform[-2, 1];  h_pq = N(0, 0.05) * exp(-|p-q|/4) for p != q, symmetric.
  * (pq|rs) = sum_{k<R} L^k_pq L^k_rs with every L^k symmetric, entries N(0, 0.1)
    (positive semi-definite, 8-fold symmetric); core energy 0.

String sets:
  * full space: every n-electron string of L orbitals (numeric order);
  * Fermi sampling: n orbitals drawn without replacement with weights
    1 / (1 + exp((eps_p - mu) / T)), eps = sorted h_pp, mu = eps[n-1]; strings are
    deduplicated in draw order until the requested count is reached, then sorted.
    Weighted sampling without replacement is done with the Gumbel-top-k construction,
    which draws the same distribution as sequential weighted draws.

Strings are uint64 bit masks, bit p = orbital p (the reference's `from_string` maps the
rightmost character to bit 0, include/sbd/framework/bit_manipulation.h:158-172), and a sorted
array is in the reference's numeric order (bit_manipulation.h:26-43, :324-330).

The integrals are exposed in the reference's own storage (chemistry/basic/makeintegrals.h:10-55):
  * ``fcidump_entries()`` -> the FCIDump integral list (value, i, j, k, l), 1-based;
  * ``packed_eri()``       -> twoInt::store, pair index ij = max(I,J)(max+1)/2 + min,
                              store[a(a+1)/2 + b] with a = max(ij, kl), b = min
                              (chemistry/basic/integrals.h:37-42).
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

__all__ = [
    "Integrals", "Instance", "make_integrals", "full_space", "fermi_strings", "as_words",
    "make_window", "embed",
    "pack_eri", "unpack_eri", "fcidump_entries", "make_instance", "WORKLOADS",
]


@dataclass
class Integrals:
    norb: int
    h1: np.ndarray          # (L, L) symmetric, spatial orbitals
    eri: np.ndarray         # (L, L, L, L) chemists' (pq|rs), 8-fold symmetric
    ecore: float = 0.0

    def packed_eri(self) -> np.ndarray:
        return pack_eri(self.eri)

    def fcidump_entries(self):
        return fcidump_entries(self)


@dataclass
class Instance:
    name: str
    ints: Integrals
    alpha: np.ndarray       # uint64 strings, sorted, unique; (n, 2) words above 64 orbitals
    beta: np.ndarray
    n_alpha_el: int
    n_beta_el: int
    meta: dict = field(default_factory=dict)

    @property
    def norb(self) -> int:
        return self.ints.norb

    @property
    def ndet(self) -> int:
        return int(self.alpha.shape[0]) * int(self.beta.shape[0])


def make_integrals(norb: int, rank: int, seed: int) -> Integrals:
    rng = np.random.default_rng(seed)
    L = norb
    diag = np.sort(rng.uniform(-2.0, 1.0, size=L))
    off = rng.normal(0.0, 0.05, size=(L, L))
    p = np.arange(L)
    decay = np.exp(-np.abs(p[:, None] - p[None, :]) / 4.0)
    up = np.triu(off * decay, 1)
    h1 = up + up.T + np.diag(diag)
    a = rng.normal(0.0, 0.1, size=(rank, L, L))
    lk = np.triu(a) + np.transpose(np.triu(a, 1), (0, 2, 1))
    eri = np.einsum("kpq,krs->pqrs", lk, lk, optimize=True)
    # exact 8-fold symmetry: every element takes the value of its canonical representative
    # (the einsum is symmetric only up to the rounding of its summation order)
    eri = unpack_eri(pack_eri(eri), L)
    return Integrals(norb=L, h1=h1, eri=eri, ecore=0.0)


def full_space(norb: int, nel: int) -> np.ndarray:
    out = [sum(1 << o for o in occ) for occ in itertools.combinations(range(norb), nel)]
    return np.array(sorted(out), dtype=np.uint64)


def fermi_strings(norb: int, nel: int, count: int, eps: np.ndarray, temperature: float,
                  seed: int, batch: int | None = None) -> np.ndarray:
    """Fermi-weighted strings, ``count`` unique ones (see module docstring).  Up to 64 orbitals
    a sorted uint64 array; above, a (count, 2) uint64 array of (low word, high word) rows
    sorted by numeric value."""
    n_possible = 1
    for k in range(nel):
        n_possible = n_possible * (norb - k) // (k + 1)
    if count > n_possible:
        raise ValueError(f"{count} strings requested but only {n_possible} exist")
    rng = np.random.default_rng(seed)
    eps = np.sort(np.asarray(eps, dtype=np.float64))
    mu = eps[nel - 1]
    x = np.clip((eps - mu) / temperature, -700.0, 700.0)
    logw = -np.log1p(np.exp(x))
    weights = np.exp(logw)
    if norb <= 64:
        pw = np.uint64(1) << np.arange(norb, dtype=np.uint64)
    batch = batch or max(4 * count, 4096)
    seen: dict[int, None] = {}
    while len(seen) < count:
        u = rng.random(size=(batch, norb))
        g = -np.log(-np.log(np.clip(u, 1e-300, 1.0 - 1e-16)))
        keys = np.log(weights)[None, :] + g
        occ = np.argpartition(-keys, nel - 1, axis=1)[:, :nel]
        if norb <= 64:
            masks = np.bitwise_or.reduce(pw[occ], axis=1).tolist()
        else:
            masks = [sum(1 << int(p) for p in row) for row in occ]
        for m in masks:
            if m not in seen:
                seen[m] = None
                if len(seen) == count:
                    break
    keys = sorted(seen.keys())
    if norb <= 64:
        return np.array(keys, dtype=np.uint64)
    return np.array([[m & 0xFFFFFFFFFFFFFFFF, m >> 64] for m in keys], dtype=np.uint64)


def _pair(i, j):
    a, b = max(i, j), min(i, j)
    return a * (a + 1) // 2 + b


def pack_eri(eri: np.ndarray) -> np.ndarray:
    """Full (pq|rs) -> the reference's twoInt::store (integrals.h:37-42)."""
    L = eri.shape[0]
    npair = L * (L + 1) // 2
    idx_i, idx_j = np.tril_indices(L)            # i >= j, pair index = i(i+1)/2 + j
    pair_vals = eri[idx_i[:, None], idx_j[:, None], idx_i[None, :], idx_j[None, :]]
    a, b = np.tril_indices(npair)
    out = np.empty(npair * (npair + 1) // 2, dtype=np.float64)
    out[a * (a + 1) // 2 + b] = pair_vals[a, b]
    return out


def unpack_eri(store: np.ndarray, norb: int) -> np.ndarray:
    L = norb
    p = np.arange(L)
    pq = np.maximum(p[:, None], p[None, :])
    pq = pq * (pq + 1) // 2 + np.minimum(p[:, None], p[None, :])
    a = np.maximum(pq[:, :, None, None], pq[None, None, :, :])
    b = np.minimum(pq[:, :, None, None], pq[None, None, :, :])
    return store[a * (a + 1) // 2 + b]


def fcidump_entries(ints: Integrals):
    """The FCIDump integral list: (value, i, j, k, l) 1-based, one line per unique element,
    in the order a FCIDUMP file lists them (two-electron, one-electron, core)."""
    L = ints.norb
    vals, idx = [], []
    for i in range(L):
        for j in range(i + 1):
            for k in range(L):
                for l in range(k + 1):
                    if _pair(i, j) >= _pair(k, l):
                        vals.append(ints.eri[i, j, k, l])
                        idx.append((i + 1, j + 1, k + 1, l + 1))
    for i in range(L):
        for j in range(i + 1):
            vals.append(ints.h1[i, j])
            idx.append((i + 1, j + 1, 0, 0))
    vals.append(ints.ecore)
    idx.append((0, 0, 0, 0))
    return np.asarray(vals, dtype=np.float64), np.asarray(idx, dtype=np.int32)


# Named workloads.  Seeds: integrals 0, alpha strings 1, beta strings 2 (plus the
# reduced "family" instances used by parity tests and the CPU baseline sample).
WORKLOADS = {
    # full C(12,4) x C(12,4) space, rank-24 integrals: 245,025 determinants
    "c1": dict(norb=12, nel=4, rank=24, strings=None, temperature=None),
    # L=36, 8a8b, rank-72, 20,000 x 20,000 Fermi-sampled strings: 4e8 determinants
    "c2": dict(norb=36, nel=8, rank=72, strings=20000, temperature=0.3),
    # L=48, 12a12b, rank-96, 10,000 x 10,000 Fermi-sampled strings: 1e8 determinants
    "c3": dict(norb=48, nel=12, rank=96, strings=10000, temperature=0.3),
}


def make_instance(name: str, strings: int | None = None, seed_ints: int = 0,
                  seed_alpha: int = 1, seed_beta: int = 2) -> Instance:
    """Build a named workload; ``strings`` overrides the per-spin string count (a reduced
    instance of the same family: same integrals, same sampler, fewer strings)."""
    w = WORKLOADS[name]
    L, nel, rank = w["norb"], w["nel"], w["rank"]
    ints = make_integrals(L, rank, seed_ints)
    n = strings if strings is not None else w["strings"]
    if w["temperature"] is None and strings is None:
        alpha = full_space(L, nel)
        beta = alpha.copy()
    else:
        eps = np.diag(ints.h1)
        T = w["temperature"] if w["temperature"] is not None else 0.3
        alpha = fermi_strings(L, nel, n, eps, T, seed_alpha)
        beta = fermi_strings(L, nel, n, eps, T, seed_beta)
    return Instance(name=name if strings is None else f"{name}_{n}", ints=ints,
                    alpha=alpha, beta=beta, n_alpha_el=nel, n_beta_el=nel,
                    meta=dict(seed_ints=seed_ints, seed_alpha=seed_alpha,
                              seed_beta=seed_beta, rank=rank))


def as_words(masks, norb: int) -> np.ndarray:
    """Python-int string masks -> the engine's layout: a uint64 array up to 64 orbitals, else
    (n, 2) rows of (low word, high word)."""
    masks = [int(m) for m in masks]
    if norb <= 64:
        return np.array(masks, dtype=np.uint64)
    return np.array([[m & 0xFFFFFFFFFFFFFFFF, m >> 64] for m in masks], dtype=np.uint64).reshape(-1, 2)


def masks_of(strings: np.ndarray) -> list:
    """The inverse of as_words."""
    if strings.ndim == 1:
        return [int(m) for m in strings.tolist()]
    return [int(lo) | (int(hi) << 64) for lo, hi in strings.tolist()]


def make_window(norb: int, lo: int, hi: int, n_alpha_el: int, n_beta_el: int, rank: int,
                seed_ints: int = 0, name: str = "window") -> Instance:
    """Every string of each spin with its electrons in orbitals [lo, hi) (the full space of the
    window, sorted), with the recipe's integrals of all norb orbitals.  A window across orbital
    64 puts the excitations of the strings across the two words of a 128-orbital string."""
    ints = make_integrals(norb, rank, seed_ints)
    strings = lambda nel: as_words([int(m) << lo for m in full_space(hi - lo, nel).tolist()], norb)
    return Instance(name=name, ints=ints, alpha=strings(n_alpha_el), beta=strings(n_beta_el),
                    n_alpha_el=n_alpha_el, n_beta_el=n_beta_el,
                    meta=dict(seed_ints=seed_ints, rank=rank, window=[lo, hi]))


def embed(inst: Instance, norb: int, shift: int, seed_ints: int = 99) -> Instance:
    """The same problem with orbital p renamed p + shift among norb orbitals: the integrals of
    the orbitals [shift, shift + L0) are inst's, all others come from the recipe (seed_ints) and
    never enter, since no string occupies them.  Its H_ii, sigma, energies and RDMs (at the
    renamed indices) are those of inst."""
    L0 = inst.norb
    if shift + L0 > norb:
        raise ValueError("the renamed orbitals must fit in norb")
    big = make_integrals(norb, 8, seed_ints)
    blk = slice(shift, shift + L0)
    h1 = big.h1.copy()
    h1[blk, blk] = inst.ints.h1
    eri = big.eri.copy()
    eri[blk, blk, blk, blk] = inst.ints.eri
    ints = Integrals(norb=norb, h1=h1, eri=eri, ecore=inst.ints.ecore)
    move = lambda st: as_words([m << shift for m in masks_of(st)], norb)
    return Instance(name=f"{inst.name}_at{shift}", ints=ints, alpha=move(inst.alpha),
                    beta=move(inst.beta), n_alpha_el=inst.n_alpha_el, n_beta_el=inst.n_beta_el,
                    meta=dict(inst.meta, embed_shift=shift, embed_norb=norb))


def make_custom(norb: int, n_alpha_el: int, n_beta_el: int, rank: int,
                n_alpha: int | None = None, n_beta: int | None = None,
                temperature: float = 0.3, seed_ints: int = 0, seed_alpha: int = 1,
                seed_beta: int = 2, name: str = "custom") -> Instance:
    """Any size of the same recipe: ``n_alpha``/``n_beta`` None means the full string space."""
    ints = make_integrals(norb, rank, seed_ints)
    eps = np.diag(ints.h1)
    alpha = (full_space(norb, n_alpha_el) if n_alpha is None else
             fermi_strings(norb, n_alpha_el, n_alpha, eps, temperature, seed_alpha))
    beta = (full_space(norb, n_beta_el) if n_beta is None else
            fermi_strings(norb, n_beta_el, n_beta, eps, temperature, seed_beta))
    return Instance(name=name, ints=ints, alpha=alpha, beta=beta, n_alpha_el=n_alpha_el,
                    n_beta_el=n_beta_el, meta=dict(seed_ints=seed_ints, seed_alpha=seed_alpha,
                                                   seed_beta=seed_beta, rank=rank))


def make_batch_instances(count: int = 64, seed: int = 0, dets=(1e3, 1e6)) -> list:
    """Many small problems of the same recipe: norb uniform in {10..20}, norb // 3 electrons per
    spin, rank-(2 norb) integrals with a seed of their own, and Fermi-sampled strings (equal
    counts per spin) giving a log-uniform determinant count in [dets[0], dets[1]]; where that
    needs more than half of a spin's string space, the full space is used instead (weighted
    sampling of nearly every string would wait on the rarest ones)."""
    from math import comb
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        L = int(rng.integers(10, 21))
        ne = L // 3
        target = float(np.exp(rng.uniform(np.log(dets[0]), np.log(dets[1]))))
        full = comb(L, ne)
        ns = max(2, round(np.sqrt(target)))
        n = None if 2 * ns > full else ns
        s = 1000 + 3 * (seed * count + k)
        out.append(make_custom(L, ne, ne, 2 * L, n, n, seed_ints=s, seed_alpha=s + 1,
                               seed_beta=s + 2, name=f"batch{seed}_{k}"))
    return out
