"""Strings of more than 64 orbitals (two 64-bit words) against the one-word path.

A problem whose orbitals are renamed p -> p + shift inside a larger orbital set (the integrals
of the renamed block copied, all others random and never entered, since no string occupies
them) has the same H_ii, sigma, energies and RDMs, the RDMs at the renamed indices and zero
elsewhere.  Every summation keeps its order under the renaming (occupied orbitals ascending,
singles and doubles in ascending ket order), so the two-word build must agree with the
one-word build bit for bit.  The shift puts the occupied orbitals across bit 64, so signs,
excitation words and spectator sums cross the word boundary.  The one-word path itself is
pinned against the compiled reference's fixtures (test_gpu_golden.py), and so is one
70-orbital window problem (w70_3a3b).
"""
import numpy as np
import pytest

from sbd_amd import synth
from sbd_amd._lib import Engine

pytestmark = pytest.mark.gpu

ACC_FP64, ACC_TWOSUM = 0, 1  # sbdx_set_fp32_accumulation modes


def block_sigma(eng, prec, X):
    """sbdx_sigma_block_device of the interleaved block X (ndet x k) on the GPU."""
    import torch
    dt = torch.float64 if prec == 0 else torch.float32
    dx = torch.tensor(X, dtype=dt, device="cuda").contiguous()
    dy = torch.empty_like(dx)
    eng.sigma_block_device(prec, X.shape[1], dx.data_ptr(), dy.data_ptr())
    return dy.cpu().numpy()


@pytest.fixture(scope="module")
def pair():
    # 40 orbitals, 8 + 8 electrons, Fermi-sampled strings: their occupations sit in orbitals
    # 0..~20, renamed to 58..~78 among 98
    print("[wide pair 40 -> 98 at 58]", flush=True)
    base = synth.make_custom(40, 8, 8, 80, 300, 280, seed_ints=0)
    big = synth.embed(base, 98, 58)
    assert big.alpha.shape == (300, 2)
    # strings hold orbitals on both sides of bit 64
    assert (big.alpha[:, 0] != 0).all() and (big.alpha[:, 1] != 0).all()
    return Engine.from_instance(base), Engine.from_instance(big)


def test_lists_and_hdiag_identical(pair):
    e0, e1 = pair
    for spin in (0, 1):
        for a, b in zip(e0.excitations(spin), e1.excitations(spin)):
            assert np.array_equal(a, b)
    assert np.array_equal(e0.hdiag(), e1.hdiag())


@pytest.mark.parametrize("acc", ["fp64", "fp32", "fp32_twosum"])
def test_sigma_identical(pair, acc):
    pytest.importorskip("torch")
    e0, e1 = pair
    print(f"[wide sigma {acc}]", flush=True)
    for e in (e0, e1):
        e.set_fp32_accumulation(ACC_TWOSUM if acc == "fp32_twosum" else ACC_FP64)
    try:
        for seed in (0, 1):
            x = np.random.default_rng(seed).standard_normal(e0.ndet)
            if acc == "fp64":
                assert np.array_equal(e0.sigma(x), e1.sigma(x))
            else:
                X = x[:, None]
                assert np.array_equal(block_sigma(e0, 1, X), block_sigma(e1, 1, X))
    finally:
        for e in (e0, e1):
            e.set_fp32_accumulation(ACC_FP64)


@pytest.mark.parametrize("k", [2, 4, 8])
@pytest.mark.parametrize("prec", [0, 1])
def test_block_sigma_identical(pair, prec, k):
    pytest.importorskip("torch")
    e0, e1 = pair
    print(f"[wide block sigma prec={prec} K={k}]", flush=True)
    X = np.random.default_rng(5 + k).standard_normal((e0.ndet, k))
    assert np.array_equal(block_sigma(e0, prec, X), block_sigma(e1, prec, X))


@pytest.mark.parametrize("k,variant", [(1, 25616), (2, 25608), (4, 25604)])
@pytest.mark.parametrize("prec", [0, 1])
def test_integral_rows_from_global_memory(pair, prec, k, variant):
    """Above 64 orbitals the default variants stage each integral row in LDS; the 256-thread
    variants read it from global memory instead.  Both give the one-word build's sigma."""
    pytest.importorskip("torch")
    e0, e1 = pair
    print(f"[wide global-V prec={prec} K={k} variant={variant}]", flush=True)
    X = np.random.default_rng(20 + k).standard_normal((e0.ndet, k))
    e1.set_mix_variant(prec, k, variant)
    try:
        assert np.array_equal(block_sigma(e0, prec, X), block_sigma(e1, prec, X))
    finally:
        e1.set_mix_variant(prec, k, 0)


def test_davidson_identical(pair):
    e0, e1 = pair
    print("[wide davidson]", flush=True)
    r0 = e0.davidson(max_iteration=100, num_block=10, eps=1e-8)
    r1 = e1.davidson(max_iteration=100, num_block=10, eps=1e-8)
    assert r0["converged"] and r1["converged"]
    assert r0["steps"] == r1["steps"]
    assert r0["energy"] == r1["energy"]


def test_rdm_and_occupation_at_renamed_indices():
    """12 orbitals, 3 + 3 electrons, full space, renamed to 58..69 among 70."""
    print("[wide rdm 12 -> 70 at 58]", flush=True)
    base = synth.make_custom(12, 3, 3, 8, None, None, seed_ints=4)
    big = synth.embed(base, 70, 58)
    e0, e1 = Engine.from_instance(base), Engine.from_instance(big)
    W = np.random.default_rng(3).standard_normal(e0.ndet)
    W /= np.linalg.norm(W)
    o0, t0 = e0.rdm(W)
    o1, t1 = e1.rdm(W)
    L0, L, s = 12, 70, 58
    blk = slice(s, s + L0)
    o0, o1 = o0.reshape(2, L0, L0), o1.reshape(2, L, L)
    t0, t1 = t0.reshape(4, L0, L0, L0, L0), t1.reshape(4, L, L, L, L)
    assert np.max(np.abs(o1[:, blk, blk] - o0)) < 1e-13
    assert np.max(np.abs(t1[:, blk, blk, blk, blk] - t0)) < 1e-13
    o1[:, blk, blk] = 0.0  # nothing outside the renamed block
    t1[:, blk, blk, blk, blk] = 0.0
    assert not o1.any() and not t1.any()
    d0, d1 = e0.occupation(W), e1.occupation(W)
    assert np.max(np.abs(d1[2 * s:2 * (s + L0)] - d0)) < 1e-13
    d1[2 * s:2 * (s + L0)] = 0.0
    assert not d1.any()


def test_two_word_input_at_64_orbitals_or_fewer():
    """sbdx_create_words with two words per string but at most 64 orbitals (high words zero)
    builds the one-word context: lists, H_ii and sigma equal sbdx_create's."""
    print("[wide two-word input, 10 orbitals]", flush=True)
    inst = synth.make_custom(10, 3, 3, 8, 60, 50, seed_ints=3)
    e0 = Engine.from_instance(inst)
    wide = lambda s: np.stack([s, np.zeros_like(s)], axis=1)
    e1 = Engine(inst.norb, wide(inst.alpha), wide(inst.beta), *ref_integrals(inst))
    for spin in (0, 1):
        for a, b in zip(e0.excitations(spin), e1.excitations(spin)):
            assert np.array_equal(a, b)
    assert np.array_equal(e0.hdiag(), e1.hdiag())
    x = np.random.default_rng(4).standard_normal(e0.ndet)
    assert np.array_equal(e0.sigma(x), e1.sigma(x))


def ref_integrals(inst):
    """(I0, I1, I2) of an instance in the reference's storage, as Engine.from_instance passes them."""
    L = inst.norb
    I1 = np.zeros((2 * L, 2 * L))
    I1[0::2, 0::2] = inst.ints.h1
    I1[1::2, 1::2] = inst.ints.h1
    return inst.ints.ecore, I1.ravel(), inst.ints.packed_eri()
