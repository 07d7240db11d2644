"""The seeded workload generator: deterministic, and the recipe it states."""
import math
import os

import numpy as np

from sbd_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_integrals_recipe():
    ints = synth.make_integrals(12, 24, seed=0)
    h, eri = ints.h1, ints.eri
    assert np.allclose(h, h.T)
    d = np.diag(h)
    assert np.all(np.diff(d) >= 0) and d.min() >= -2 and d.max() <= 1
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.allclose(eri, eri.transpose(perm), atol=1e-14)
    M = eri.reshape(144, 144)
    assert np.linalg.eigvalsh(M).min() > -1e-10           # sum of squares of L^k


def test_generator_deterministic():
    a = synth.make_instance("c2", strings=500)
    b = synth.make_instance("c2", strings=500)
    assert np.array_equal(a.alpha, b.alpha) and np.array_equal(a.beta, b.beta)
    assert np.array_equal(a.ints.eri, b.ints.eri)
    assert not np.array_equal(a.alpha, a.beta)


def test_strings():
    s = synth.full_space(12, 4)
    assert s.size == math.comb(12, 4) == 495 and np.all(np.diff(s.astype(np.int64)) > 0)
    inst = synth.make_instance("c3", strings=700)
    for t in (inst.alpha, inst.beta):
        assert t.size == 700 and np.unique(t).size == 700
        assert np.all(np.diff(t.astype(np.int64)) > 0)
        assert all(bin(int(x)).count("1") == 12 for x in t)
        assert int(t.max()) < (1 << 48)


def test_packed_eri_round_trip():
    ints = synth.make_integrals(7, 5, seed=3)
    store = ints.packed_eri()
    B = 7 * 8 // 2
    assert store.size == B * (B + 1) // 2
    assert np.array_equal(synth.unpack_eri(store, 7), ints.eri)


def test_fcidump_entries_cover_store():
    ints = synth.make_integrals(5, 4, seed=1)
    vals, idx = ints.fcidump_entries()
    two = idx[:, 2] > 0
    B = 15
    assert two.sum() == B * (B + 1) // 2
    assert ((idx[:, 2] == 0) & (idx[:, 0] > 0)).sum() == 15


def test_batch_instances_are_seeded_and_in_range():
    a = synth.make_batch_instances(12, seed=4, dets=(1e3, 1e5))
    b = synth.make_batch_instances(12, seed=4, dets=(1e3, 1e5))
    for x, y in zip(a, b):
        assert x.norb == y.norb and np.array_equal(x.alpha, y.alpha) and np.array_equal(x.beta, y.beta)
        assert 10 <= x.norb <= 20 and x.n_alpha_el == x.norb // 3 == x.n_beta_el
        assert x.meta["rank"] == 2 * x.norb
        pc = [bin(int(s)).count("1") for s in x.alpha]
        assert set(pc) == {x.n_alpha_el}
        assert x.ndet >= 4
    assert len({x.meta["seed_ints"] for x in a}) == 12


def test_bench_solver_variants_are_valid_options():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    from sbd_amd._lib import DavidsonOpts
    fields = {f for f, _ in DavidsonOpts._fields_}
    for name, opts in bench.SOLVERS.items():
        # "mlp" (network file) and "fp32_acc" (sbdx_set_fp32_accumulation) are set on the
        # engine by the benchmark, not passed as Davidson options
        assert set(opts) - {"mlp", "fp32_acc"} <= fields, name
        assert opts.get("fp32_acc", 0) in (0, 1), name
        if "mlp" in opts:
            assert os.path.exists(os.path.join(ROOT, opts["mlp"])), name


def test_words_round_trip_and_window_strings():
    """as_words / masks_of invert each other; a window's strings are exactly the full space of
    the window's orbitals, two words each above 64 orbitals."""
    rng = np.random.default_rng(0)
    masks = [int(rng.integers(1, 2**62)) | (int(rng.integers(0, 2**62)) << 64) for _ in range(20)]
    assert synth.masks_of(synth.as_words(masks, 128)) == masks
    w = synth.make_window(70, 59, 69, 3, 3, 8, seed_ints=6)
    got = synth.masks_of(w.alpha)
    assert w.alpha.shape == (120, 2) and got == sorted(got)
    assert all(bin(m).count("1") == 3 and m >> 69 == 0 and m & ((1 << 59) - 1) == 0 for m in got)
    assert len(set(got)) == 120  # C(10, 3)


def test_embed_renames_orbitals():
    """embed: strings shifted by the renaming, the integrals of the renamed block copied."""
    base = synth.make_custom(10, 3, 3, 8, 60, 50, seed_ints=3)
    big = synth.embed(base, 70, 58)
    assert synth.masks_of(big.alpha) == [m << 58 for m in synth.masks_of(base.alpha)]
    assert synth.masks_of(big.beta) == [m << 58 for m in synth.masks_of(base.beta)]
    blk = slice(58, 68)
    assert np.array_equal(big.ints.h1[blk, blk], base.ints.h1)
    assert np.array_equal(big.ints.eri[blk, blk, blk, blk], base.ints.eri)
    assert big.ndet == base.ndet


def test_one_word_sampling_is_unchanged():
    """The sampled strings behind the committed fixtures keep their one-word form and values:
    the first three alpha strings of the 36-orbital 120 x 110 case (tests/golden/
    c2fam_120x110.npz was written from this sample)."""
    inst = synth.make_custom(36, 8, 8, 72, 120, 110, seed_ints=0)
    assert inst.alpha.ndim == 1 and inst.alpha.dtype == np.uint64
    assert int(inst.alpha[0]) == 255 and int(inst.alpha[1]) == 383 and int(inst.alpha[2]) == 447
