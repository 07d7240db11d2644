"""The HIP engine against the compiled reference's golden fixtures (tests/golden/).

Lists bit-exact; H_ii and sigma to 1e-12 relative; Davidson energies to 1e-9 Ha; RDMs of the
converged vector to 1e-6 (both solves stop at a 1e-8 residual); RDMs of a seeded vector to
1e-12 (36, 48 and 70 orbitals).  w70_3a3b: 70 orbitals, every 3-electron string of orbitals
59..68 per spin, so the strings and their excitations cross from the first 64-bit word into
the second (the reference keeps strings of any length in bit_length-bit words).
"""
import json
import os

import numpy as np
import pytest

from sbd_amd import synth
from sbd_amd._lib import Engine

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = ["l6_2a2b", "l8_3a2b", "c2fam_120x110", "c2fam_300x280", "c3fam_200x220", "w70_3a3b"]


def load(name):
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)
    return json.loads(str(z["meta"])), z


def build(meta):
    if meta["recipe"] == "sbd_amd.synth.make_window":
        lo, hi = meta["window"]
        inst = synth.make_window(meta["norb"], lo, hi, meta["n_alpha_el"], meta["n_beta_el"], meta["rank"],
                                 seed_ints=meta["seed_ints"])
    else:
        inst = synth.make_custom(meta["norb"], meta["n_alpha_el"], meta["n_beta_el"], meta["rank"],
                                 meta["n_alpha"], meta["n_beta"], seed_ints=meta["seed_ints"],
                                 seed_alpha=meta["seed_alpha"], seed_beta=meta["seed_beta"])
    return inst, Engine.from_instance(inst)


def compare(z, prefix, v, rtol):
    scale = max(float(np.max(np.abs(v))), 1.0)
    if f"{prefix}_full" in z:
        assert np.max(np.abs(v - z[f"{prefix}_full"])) <= rtol * scale
    else:
        assert np.max(np.abs(v[z[f"{prefix}_idx"]] - z[f"{prefix}_val"])) <= rtol * scale
    assert abs(v @ v - float(z[f"{prefix}_sumsq"])) <= 10 * rtol * max(1.0, v @ v)


@pytest.mark.parametrize("name", CASES)
def test_golden_lists_hdiag_sigma(name):
    print(f"[golden {name}]", flush=True)
    meta, z = load(name)
    inst, eng = build(meta)
    for spin, (ts, td) in enumerate((("sa", "da"), ("sb", "db"))):
        off_s, idx_s, off_d, idx_d = eng.excitations(spin)
        assert np.array_equal(np.diff(off_s), z[f"conn_{ts}_cnt"])
        assert np.array_equal(idx_s, z[f"conn_{ts}_idx"])
        assert np.array_equal(np.diff(off_d), z[f"conn_{td}_cnt"])
        assert np.array_equal(idx_d, z[f"conn_{td}_idx"])
    compare(z, "hii", eng.hdiag(), 1e-12)
    for s in meta["sigma_seeds"]:
        x = np.random.default_rng(s).standard_normal(eng.ndet)
        compare(z, f"sigma{s}", eng.sigma(x), 1e-12)


@pytest.mark.parametrize("name", CASES)
def test_golden_davidson_and_rdm(name):
    print(f"[golden davidson {name}]", flush=True)
    meta, z = load(name)
    inst, eng = build(meta)
    r = eng.davidson(max_iteration=100, num_block=10, eps=1e-8)
    assert r["converged"]
    assert abs(r["energy"] - float(z["energy"][0])) < 1e-9
    one, two = eng.rdm(r["W"])
    assert np.max(np.abs(one.ravel() - z["one"])) < 1e-6
    assert np.max(np.abs(two.ravel()[z["two_idx"]] - z["two_val"])) < 1e-6
    assert np.max(np.abs(eng.occupation(r["W"]) - z["density"])) < 1e-6


@pytest.mark.parametrize("name", ["c2fam_300x280", "c3fam_200x220", "w70_3a3b"])
def test_golden_rdm_seeded_vector(name):
    """Correlation and OccupationDensity of one normalised seeded vector: every stored 2-RDM
    entry (the largest 150,000 or all), the whole of the 1-RDM and density, and whole-array
    checksums of the 2-RDM (a cosine-weighted sum moves with any single entry)."""
    print(f"[golden rdm {name}]", flush=True)
    meta, z = load(name)
    inst, eng = build(meta)
    W = np.random.default_rng(meta["rdmvec_seed"]).standard_normal(eng.ndet)
    W /= np.linalg.norm(W)
    one, two = eng.rdm(W)
    one, two = one.ravel(), two.ravel()
    assert np.max(np.abs(one - z["rdmvec_one"])) < 1e-12
    assert np.max(np.abs(two[z["rdmvec_two_idx"]] - z["rdmvec_two_val"])) < 1e-12
    assert two.size == int(z["rdmvec_two_size"][0])
    scale = float(np.sum(np.abs(two)))
    assert abs(two.sum() - float(z["rdmvec_two_sum"][0])) < 1e-12 * scale
    assert abs(two @ two - float(z["rdmvec_two_sumsq"][0])) < 1e-12 * max(1.0, float(z["rdmvec_two_sumsq"][0]))
    w = two @ np.cos(0.1 * np.arange(two.size))
    assert abs(w - float(z["rdmvec_two_wsum"][0])) < 1e-12 * scale
    assert np.max(np.abs(eng.occupation(W) - z["rdmvec_density"])) < 1e-12


def test_golden_c1_davidson_rdm_energy_identity():
    """C1: full C(12,4)^2 space, 245,025 determinants, reference Davidson to 1e-8."""
    print("[golden c1]", flush=True)
    z = np.load(os.path.join(GOLDEN, "c1.npz"), allow_pickle=False)
    inst = synth.make_instance("c1")
    eng = Engine.from_instance(inst)
    r = eng.davidson(max_iteration=200, num_block=10, eps=1e-8)
    assert r["converged"]
    assert abs(r["energy"] - float(z["energy"][0])) < 1e-9
    assert abs(r["steps"] - z["trace"].size) <= 2
    one, two = eng.rdm(r["W"])
    assert np.max(np.abs(one.ravel() - z["one"])) < 1e-6
    assert np.max(np.abs(two.ravel()[z["two_idx"]] - z["two_val"])) < 1e-6
    s = np.sign(r["W"][z["W_idx"]] @ z["W_val"])
    assert np.max(np.abs(s * r["W"][z["W_idx"]] - z["W_val"])) < 1e-6
