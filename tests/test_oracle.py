"""The CPU oracle, pinned.

* against the compiled reference's golden fixtures (tests/golden/, written by
  make_golden.py): excitation lists and integer task order exactly, H_ii and sigma bit for
  bit, Davidson energies to 1e-9 Ha, RDMs;
* against what the mathematics fixes, independently of the reference: a dense Hamiltonian
  built by explicit second quantization (tests/bruteforce.py), the closed-form excitation
  counts of a full string space, a dense eigensolver, and the RDM sum rules;
* and the brute-force comparison is shown to fail for planted mistakes (a dropped
  same-spin interaction, a flipped mixed-spin sign).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from sbd_amd import synth

import bruteforce as BF

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
# w70_3a3b: 70 orbitals, every 3-electron string of orbitals 59..68 per spin (two 64-bit words
# per string here, four 20-bit words in the reference)
CASES = ["l6_2a2b", "l8_3a2b", "c2fam_120x110", "c2fam_300x280", "c3fam_200x220", "w70_3a3b"]


def load(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return meta, z


_systems = {}


def system(name):
    if name not in _systems:
        meta, z = load(name)
        if meta["recipe"] == "sbd_amd.synth.make_window":
            lo, hi = meta["window"]
            inst = synth.make_window(meta["norb"], lo, hi, meta["n_alpha_el"], meta["n_beta_el"],
                                     meta["rank"], seed_ints=meta["seed_ints"])
        else:
            inst = synth.make_custom(meta["norb"], meta["n_alpha_el"], meta["n_beta_el"], meta["rank"],
                                     meta["n_alpha"], meta["n_beta"], seed_ints=meta["seed_ints"],
                                     seed_alpha=meta["seed_alpha"], seed_beta=meta["seed_beta"])
        _systems[name] = (meta, z, inst, O.System.from_instance(inst))
    return _systems[name]


def check_vector(z, prefix, v, exact=True, rtol=0.0):
    if f"{prefix}_full" in z:
        ref = z[f"{prefix}_full"]
        assert ref.shape == v.shape
        if exact:
            assert np.array_equal(ref, v)
        else:
            assert np.max(np.abs(ref - v)) <= rtol * max(np.max(np.abs(ref)), 1.0)
    else:
        idx, val = z[f"{prefix}_idx"], z[f"{prefix}_val"]
        if exact:
            assert np.array_equal(v[idx], val)
        else:
            assert np.max(np.abs(v[idx] - val)) <= rtol * max(np.max(np.abs(val)), 1.0)
    tol = 0.0 if exact else rtol
    assert abs(v.sum() - float(z[f"{prefix}_sum"])) <= max(tol, 1e-9) * max(1.0, np.abs(v).sum())
    assert abs(v @ v - float(z[f"{prefix}_sumsq"])) <= max(tol, 1e-12) * max(1.0, v @ v)


@pytest.mark.parametrize("name", CASES)
def test_excitation_lists_match_reference(name):
    meta, z, inst, so = system(name)
    assert list(z["tasks"]) == [2, 1, 0]   # helper.h:449-500 at one rank
    for tag, conn in (("sa", so.conn_a[0:2]), ("da", so.conn_a[2:4]),
                      ("sb", so.conn_b[0:2]), ("db", so.conn_b[2:4])):
        off, idx = conn
        assert np.array_equal(np.diff(off), z[f"conn_{tag}_cnt"])
        assert np.array_equal(idx, z[f"conn_{tag}_idx"])


@pytest.mark.parametrize("name", CASES)
def test_hdiag_bit_exact(name):
    meta, z, inst, so = system(name)
    check_vector(z, "hii", so.hdiag(), exact=True)


@pytest.mark.parametrize("name", CASES)
def test_sigma_bit_exact(name):
    meta, z, inst, so = system(name)
    hii = so.hdiag()
    for s in meta["sigma_seeds"]:
        x = np.random.default_rng(s).standard_normal(so.ndet)
        check_vector(z, f"sigma{s}", so.mult(x, hii), exact=True)


# c2fam_300x280's Davidson takes ~25 s here; its energy is pinned on the GPU side
@pytest.mark.parametrize("name", ["l6_2a2b", "l8_3a2b", "c2fam_120x110", "c3fam_200x220"])
def test_davidson_energy_and_rdm(name):
    meta, z, inst, so = system(name)
    r = so.davidson(max_it=100, num_block=10, eps=1e-8)
    assert abs(r["energy"] - float(z["energy"][0])) < 1e-9
    # the early trace is rounding-insensitive; the later one depends on reduction order
    assert np.allclose(r["trace"][:3], z["trace"][:3], rtol=0, atol=1e-9)
    one, two = so.correlation(r["W"])
    assert np.max(np.abs(one.ravel() - z["one"])) < 1e-6
    ref_two = np.zeros(int(z["two_size"][0]))
    ref_two[z["two_idx"]] = z["two_val"]
    mask = np.zeros(ref_two.size, bool)
    mask[z["two_idx"]] = True
    assert np.max(np.abs(two.ravel()[mask] - ref_two[mask])) < 1e-6
    assert abs(two.sum() - float(z["two_sum"][0])) < 1e-6 * max(1.0, abs(float(z["two_sum"][0])))
    assert np.max(np.abs(so.occupation(r["W"]) - z["density"])) < 1e-6


@pytest.mark.parametrize("name", ["c2fam_300x280", "c3fam_200x220", "w70_3a3b"])
def test_rdm_of_seeded_vector_matches_reference(name):
    """Correlation and OccupationDensity of the same seeded vector as the compiled reference:
    36 orbitals (two 20-bit words per string), 48 (three) and 70 (four)."""
    meta, z, inst, so = system(name)
    W = np.random.default_rng(meta["rdmvec_seed"]).standard_normal(so.ndet)
    W /= np.linalg.norm(W)
    one, two = so.correlation(W)
    one, two = one.ravel(), two.ravel()
    assert np.max(np.abs(one - z["rdmvec_one"])) < 1e-15
    assert np.max(np.abs(two[z["rdmvec_two_idx"]] - z["rdmvec_two_val"])) < 1e-15
    assert int(np.count_nonzero(np.abs(two) > 1e-14)) == int(z["rdmvec_two_nnz"][0])
    scale = float(np.sum(np.abs(two)))
    assert abs(two @ np.cos(0.1 * np.arange(two.size)) - float(z["rdmvec_two_wsum"][0])) < 1e-14 * scale
    # the reference's occupation sums run over OpenMP threads: same terms, another order
    assert np.max(np.abs(so.occupation(W) - z["rdmvec_density"])) < 1e-12


def test_davidson_trace_two_word_strings():
    """The 70-orbital fixture's first Davidson steps (its full run, 499 steps, takes minutes
    here; its energy is pinned on the GPU side, test_gpu_golden.py)."""
    meta, z, inst, so = system("w70_3a3b")
    r = so.davidson(max_it=1, num_block=10, eps=1e-8)
    assert r["steps"] == 10
    assert np.allclose(r["trace"][:3], z["trace"][:3], rtol=0, atol=1e-9)


def test_c1_fixture_consistency():
    """The C1 fixture (reference Davidson, 245,025 determinants) obeys the closed-form checks
    the oracle is held to: trace of the 1-RDM = electron counts; energy identity."""
    z = np.load(os.path.join(GOLDEN, "c1.npz"), allow_pickle=False)
    inst = synth.make_instance("c1")
    L = inst.norb
    one = z["one"].reshape(2, L, L, order="C")
    assert abs(np.trace(one[0].reshape(L, L)) - 4) < 1e-10
    assert abs(np.trace(one[1].reshape(L, L)) - 4) < 1e-10
    two = np.zeros(int(z["two_size"][0]))
    two[z["two_idx"]] = z["two_val"]
    e = rdm_energy(inst, z["one"].reshape(2, L * L), two.reshape(4, L ** 4))
    assert abs(e - float(z["energy"][0])) < 1e-8


# ----------------------------------------------------------------- independent pins

def rdm_energy(inst, one, two):
    """E = I0 + sum_ij h_ij (g_a + g_b)[i + L j] + 1/2 sum (ik|jl) G[i + L j + L^2 k + L^3 l]
    (the reference sample's self-check, samples/.../main.cc:166-216)."""
    L = inst.norb
    h, eri = inst.ints.h1, inst.ints.eri
    g = one[0].reshape(L, L, order="F") + one[1].reshape(L, L, order="F")   # g[i, j]
    e1 = np.sum(h * g)
    G = two.sum(axis=0).reshape(L, L, L, L, order="F")   # G[i, j, k, l]
    e2 = 0.5 * np.einsum("ikjl,ijkl->", eri, G)
    return inst.ints.ecore + e1 + e2


TINY = [(4, 2, 1, 3), (4, 2, 2, 2), (5, 2, 2, 3), (5, 3, 1, 4)]


def dense_oracle(so):
    n = so.ndet
    H = np.zeros((n, n))
    hii = so.hdiag()
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        H[:, j] = so.mult(e, hii)
    return H


@pytest.mark.parametrize("case", TINY)
def test_hamiltonian_matches_second_quantization(case):
    L, na, nb, R = case
    inst = synth.make_custom(L, na, nb, R, None, None, seed_ints=21)
    so = O.System.from_instance(inst)
    H = dense_oracle(so)
    B = BF.dense_hamiltonian(inst.ints.h1, inst.ints.eri, inst.ints.ecore, inst.alpha, inst.beta, L)
    assert np.max(np.abs(H - B)) < 1e-12


@pytest.mark.parametrize("flag", ["drop_exchange", "flip_mixed"])
def test_bruteforce_comparison_catches_planted_mistakes(flag):
    L, na, nb, R = 5, 2, 2, 3
    inst = synth.make_custom(L, na, nb, R, None, None, seed_ints=21)
    H = dense_oracle(O.System.from_instance(inst))
    B = BF.dense_hamiltonian(inst.ints.h1, inst.ints.eri, inst.ints.ecore, inst.alpha, inst.beta, L,
                             **{flag: True})
    assert np.max(np.abs(H - B)) > 1e-3


@pytest.mark.parametrize("L,n", [(6, 2), (8, 3), (12, 4)])
def test_full_space_excitation_counts_closed_form(L, n):
    s = synth.full_space(L, n)
    off_s, idx_s, off_d, idx_d = O.excitations(s, L)
    assert np.all(np.diff(off_s) == n * (L - n))
    assert np.all(np.diff(off_d) == math.comb(n, 2) * math.comb(L - n, 2))
    # every listed ket really is at popcount distance 2 / 4, ascending
    for i in range(0, s.size, max(1, s.size // 25)):
        ks = idx_s[off_s[i]:off_s[i + 1]]
        kd = idx_d[off_d[i]:off_d[i + 1]]
        assert np.all(np.diff(ks) > 0) and np.all(np.diff(kd) > 0)
        assert all(bin(int(s[i]) ^ int(s[k])).count("1") == 2 for k in ks)
        assert all(bin(int(s[i]) ^ int(s[k])).count("1") == 4 for k in kd)


def test_hdiag_closed_form():
    inst = synth.make_custom(8, 3, 2, 6, None, None, seed_ints=4)
    so = O.System.from_instance(inst)
    h, eri = inst.ints.h1, inst.ints.eri
    got = so.hdiag()
    k = 0
    for a in inst.alpha:
        for b in inst.beta:
            oa = [p for p in range(8) if (int(a) >> p) & 1]
            ob = [p for p in range(8) if (int(b) >> p) & 1]
            e = sum(h[p, p] for p in oa + ob)
            for occ in (oa, ob):
                for i in occ:
                    for j in occ:
                        if i < j:
                            e += eri[i, i, j, j] - eri[i, j, j, i]
            e += sum(eri[i, i, j, j] for i in oa for j in ob)
            assert abs(got[k] - e) < 1e-12
            k += 1


@pytest.mark.parametrize("case", [(6, 2, 2, 4), (6, 3, 2, 5)])
def test_davidson_equals_dense_ground_state(case):
    L, na, nb, R = case
    inst = synth.make_custom(L, na, nb, R, None, None, seed_ints=8)
    so = O.System.from_instance(inst)
    e0 = np.linalg.eigvalsh(dense_oracle(so))[0]
    r = so.davidson(max_it=200, num_block=10, eps=1e-9)
    assert abs(r["energy"] - e0) < 1e-9


def test_rdm_sum_rules():
    inst = synth.make_custom(6, 3, 2, 4, None, None, seed_ints=9)
    so = O.System.from_instance(inst)
    W = so.davidson(max_it=200, num_block=10, eps=1e-10)["W"]
    one, two = so.correlation(W)
    L, N = 6, {0: 3, 1: 2}
    g = [one[s].reshape(L, L, order="F") for s in range(2)]
    G = [two[k].reshape(L, L, L, L, order="F") for k in range(4)]   # G[i,j,k,l]
    for s in range(2):
        assert abs(np.trace(g[s]) - N[s]) < 1e-10
        assert np.max(np.abs(g[s] - g[s].T)) < 1e-10
    for s in range(2):
        for t in range(2):
            # sum_j G^{st}_{i j k j} = (N_t - delta_st) g^s_{ik}
            lhs = np.einsum("ijkj->ik", G[s + 2 * t])
            assert np.max(np.abs(lhs - (N[t] - (s == t)) * g[s])) < 1e-10
    # same-spin antisymmetry: G^{ss}_{ijkl} = -G^{ss}_{jikl}
    for s in range(2):
        assert np.max(np.abs(G[3 * s] + G[3 * s].transpose(1, 0, 2, 3))) < 1e-10
    e = rdm_energy(inst, one, two)
    assert abs(e - float(W @ so.mult(W))) < 1e-9


def test_davidson_deadline_stops_after_one_step():
    """max_time forces the reference's deadline branch (davidson.h:880-902): one step, and W is
    still a normalized Ritz vector."""
    inst = synth.make_custom(6, 2, 2, 4, None, None, seed_ints=3)
    so = O.System.from_instance(inst)
    r = so.davidson(max_it=10, num_block=10, eps=1e-12, max_time=1e-9)
    assert r["steps"] == 1
    assert abs(np.linalg.norm(r["W"]) - 1.0) < 1e-12
