"""The engine's two-word string helpers (sbd_amd/csrc/common.h: Str<W>, count_range, first,
drop_first, sign_single, parity_ref, sign_double, det_sign, the packed excitation fields), run on
the host through tests/cpp/str_words (no GPU).

Pinned three ways: (1) brute force on Python integers of what each helper counts (set bits,
lowest set bits, set bits in a range, and the basis sign as the parity of the number of pairs
(p in x, q in y) with q < p, counted pair by pair); (2) on strings of one word, the two-word
build equals the one-word build; (3) renaming every orbital p -> p + s (s up to 64, so strings
and index ranges cross from the first word into the second) leaves every count and sign
unchanged and moves the lowest set bits by s.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "str_words")


@pytest.fixture(scope="module")
def run():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "str_words"], check=True)

    def go(cases):
        text = "".join(f"{x & (2**64 - 1)} {x >> 64} {y & (2**64 - 1)} {y >> 64} {' '.join(map(str, idx))}\n"
                       for x, y, idx in cases)
        out = subprocess.run([BIN], input=text, capture_output=True, text=True, check=True).stdout
        return [list(map(int, ln.split())) for ln in out.splitlines()]
    return go


def rand_string(rng, lo, hi, nel):
    return sum(1 << int(p) for p in rng.choice(np.arange(lo, hi), nel, replace=False))


def cases_in(rng, lo, hi, n):
    """Strings and indices inside orbitals [lo, hi)."""
    out = []
    for _ in range(n):
        # x holds at least two orbitals: first(drop_first(x)) is defined for non-empty strings only
        x, y = rand_string(rng, lo, hi, rng.integers(2, 9)), rand_string(rng, lo, hi, rng.integers(1, 9))
        i, a = sorted(rng.choice(np.arange(lo, hi), 2, replace=False).tolist())
        I, J, A, B = rng.choice(np.arange(lo, hi), 4, replace=False).tolist()
        I, J = sorted((I, J))
        A, B = sorted((A, B))
        out.append((x, y, (i, a, I, J, A, B)))
    return out


def pairs_sign(x, y):
    n = sum(1 for p in range(128) if (x >> p) & 1 for q in range(p) if (y >> q) & 1)
    return -1 if n % 2 else 1


def test_brute_force_definitions(run):
    rng = np.random.default_rng(0)
    cases = cases_in(rng, 0, 128, 400) + cases_in(rng, 56, 72, 200)  # the second set straddles bit 64
    for (x, y, (i, a, I, J, A, B)), r in zip(cases, run(cases)):
        popc, first, second, cnt, ss, pr, sd, ds = r[:8]
        assert popc == bin(x).count("1")
        assert first == (x & -x).bit_length() - 1
        rest = x & (x - 1)
        assert second == (rest & -rest).bit_length() - 1
        assert cnt == bin((x >> i) & ((1 << (a - i)) - 1)).count("1")
        assert ds == pairs_sign(x, y)
        assert r[8:12] == [i, a, (i + a) & 1, a * 128 + i]


def test_one_word_build_agrees(run):
    rng = np.random.default_rng(1)
    cases = cases_in(rng, 0, 64, 400)
    for r in run(cases):
        assert len(r) == 20 and r[:8] == r[12:20]


@pytest.mark.parametrize("shift", [7, 30, 58, 64])
def test_renaming_orbitals_across_the_word_boundary(run, shift):
    rng = np.random.default_rng(2 + shift)
    base = cases_in(rng, 0, 64, 300)
    moved = [(x << shift, y << shift, tuple(v + shift for v in idx)) for x, y, idx in base]
    for r0, r1 in zip(run(base), run(moved)):
        assert r1[0] == r0[0]                          # set bits
        assert r1[1] == r0[1] + shift                  # lowest set bit
        assert r1[3:8] == r0[3:8]                      # range count, the three signs, det_sign
