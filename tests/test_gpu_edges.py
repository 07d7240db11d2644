"""Edge cases of the string spaces, through the C ABI, against the CPU oracle: an empty spin
(no electrons: one string, no excitations), a filled spin (every orbital occupied), one
electron per spin, a single string per spin (one determinant), strings with no excitation
partner in the sampled set, and 64 orbitals (the largest string width).  Excitation lists
bit-exact; H_ii and sigma 1e-12 relative in FP64, 2e-6 in FP32 (K = 1 and 2)."""
import numpy as np
import pytest

import oracle as O
from sbd_amd import synth
from sbd_amd._lib import Engine

pytestmark = pytest.mark.gpu


def _strings(L, ne, picks):
    return np.array(sorted(set(sum(1 << p for p in occ) for occ in picks)), dtype=np.uint64)


def _inst(L, alpha, beta, na_el, nb_el, rank=None, seed=5):
    ints = synth.make_integrals(L, rank or 2 * L, seed)
    return synth.Instance(name="edge", ints=ints, alpha=alpha, beta=beta, n_alpha_el=na_el,
                          n_beta_el=nb_el, meta={})


def _cases():
    yield "empty_alpha", _inst(6, np.array([0], np.uint64), synth.full_space(6, 2), 0, 2)
    yield "filled_beta", _inst(5, synth.full_space(5, 2), np.array([31], np.uint64), 2, 5)
    yield "one_electron_each", _inst(7, synth.full_space(7, 1), synth.full_space(7, 1), 1, 1)
    yield "single_determinant", _inst(8, np.array([0b1011], np.uint64), np.array([0b0111], np.uint64), 3, 3)
    # strings three or more orbital moves apart have no single or double partner
    yield "isolated_strings", _inst(12, _strings(12, 3, [(0, 1, 2), (6, 7, 8), (9, 10, 11), (0, 1, 3)]),
                                    _strings(12, 3, [(0, 1, 2), (3, 4, 5), (0, 4, 5)]), 3, 3)
    rng = np.random.default_rng(0)
    picks_a = [tuple(rng.choice(64, 4, replace=False)) for _ in range(40)] + [(60, 61, 62, 63)]
    picks_b = [tuple(rng.choice(64, 3, replace=False)) for _ in range(30)] + [(0, 1, 63)]
    yield "sixty_four_orbitals", _inst(64, _strings(64, 4, picks_a), _strings(64, 3, picks_b), 4, 3, rank=8)


CASES = list(_cases())


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def pair(request):
    name, inst = request.param
    print(f"[edge case {name}]", flush=True)
    return name, O.System.from_instance(inst), Engine.from_instance(inst)


def test_lists_hdiag_sigma(pair):
    name, so, eng = pair
    for spin, conn in ((0, so.conn_a), (1, so.conn_b)):
        off_s, idx_s, off_d, idx_d = eng.excitations(spin)
        assert np.array_equal(off_s, conn[0]) and np.array_equal(idx_s, conn[1])
        assert np.array_equal(off_d, conn[2]) and np.array_equal(idx_d, conn[3])
    ref_h = so.hdiag()
    assert np.max(np.abs(eng.hdiag() - ref_h)) <= 1e-12 * max(1.0, np.max(np.abs(ref_h)))
    x = np.random.default_rng(1).standard_normal(so.ndet)
    ref = so.mult(x)
    scale = max(1e-300, np.max(np.abs(ref)))
    assert np.max(np.abs(eng.sigma(x) - ref)) <= 1e-12 * scale


@pytest.mark.parametrize("k", [1, 2])
def test_fp32_blocks(pair, k):
    torch = pytest.importorskip("torch")
    name, so, eng = pair
    X = np.random.default_rng(2 + k).standard_normal((so.ndet, k))
    dx = torch.tensor(X, dtype=torch.float32, device="cuda").contiguous()
    dy = torch.empty_like(dx)
    eng.sigma_block_device(1, k, dx.data_ptr(), dy.data_ptr())
    Y = dy.cpu().numpy().astype(np.float64).reshape(so.ndet, k)
    for v in range(k):
        ref = so.mult(X[:, v])
        assert np.max(np.abs(Y[:, v] - ref)) <= 2e-6 * max(1e-300, np.max(np.abs(ref))), v


def test_davidson_lowest(pair):
    name, so, eng = pair
    ref = so.davidson(max_it=200, num_block=10, eps=1e-9)
    if so.ndet == 1:
        got = eng.davidson(max_iteration=5, num_block=2, eps=1e-9)
    else:
        got = eng.davidson(max_iteration=200, num_block=min(10, max(2, so.ndet)), eps=1e-9)
    assert abs(got["energy"] - ref["energy"]) < 1e-9
