"""Stored Hamiltonian (sbdx_store_hamiltonian; the reference's method 1, qcham.h:10-312 and
mult.h:13-123): H as CSR in HBM, every sigma build a sparse product with it.

The element count must equal N + the exact off-diagonal count of the on-the-fly build; sigma
must match the oracle to 1e-12 relative (FP64; FP32 2e-6), for single vectors and fused
blocks; the Davidson energy the oracle's to 1e-9 Ha; new integrals drop the stored H.
"""
import numpy as np
import pytest

import oracle as O
from sbd_amd import synth
from sbd_amd._lib import Engine

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = [
    ("l8_3a2b_full", (8, 3, 2, 6, None, None)),
    ("l10_3a3b_samp", (10, 3, 3, 8, 60, 50)),
    ("c2fam_120x110", (36, 8, 8, 72, 120, 110)),
]


@pytest.fixture(scope="module", params=CASES, ids=[c[0] for c in CASES])
def stored(request):
    name, (L, na, nb, R, Na, Nb) = request.param
    print(f"[case {name}]", flush=True)
    inst = synth.make_custom(L, na, nb, R, Na, Nb, seed_ints=3)
    eng = Engine.from_instance(inst)
    nnz = eng.store_hamiltonian(True)
    return name, inst, O.System.from_instance(inst), eng, nnz


def test_element_count(stored):
    name, inst, so, eng, nnz = stored
    sa, da = eng.excitation_counts(0)
    sb, db = eng.excitation_counts(1)
    off = inst.beta.size * (sa + da) + inst.alpha.size * (sb + db) + sa * sb
    assert nnz == so.ndet + off


def test_sigma(stored):
    name, inst, so, eng, nnz = stored
    x = np.random.default_rng(1).standard_normal(so.ndet)
    ref = so.mult(x)
    got = eng.sigma(x)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    dx = torch.tensor(x, dtype=torch.float32, device="cuda")
    dy = torch.empty_like(dx)
    eng.sigma_device(1, 1, dx.data_ptr(), dy.data_ptr())
    got32 = dy.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(got32 - ref)) <= 2e-6 * np.max(np.abs(ref))


@pytest.mark.parametrize("k", [2, 4])
def test_block_sigma(stored, k):
    name, inst, so, eng, nnz = stored
    X = np.random.default_rng(k).standard_normal((so.ndet, k))
    dx = torch.tensor(X, device="cuda").contiguous()
    dy = torch.empty_like(dx)
    eng.sigma_block_device(0, k, dx.data_ptr(), dy.data_ptr())
    Y = dy.cpu().numpy()
    for v in range(k):
        ref = so.mult(X[:, v])
        assert np.max(np.abs(Y[:, v] - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_davidson(stored):
    name, inst, so, eng, nnz = stored
    ref = so.davidson(max_it=300, num_block=10, eps=1e-8)
    got = eng.davidson(max_iteration=300, num_block=10, eps=1e-8)
    assert got["converged"] and abs(got["energy"] - ref["energy"]) < 1e-9
    got1 = eng.davidson(method=1, precision=1, block=2, num_block=8, max_iteration=300, eps=1e-8)
    assert got1["converged"] and abs(got1["energy"] - ref["energy"]) < 1e-9


def test_new_integrals_drop_the_stored_h(stored):
    """H is linear in (I0, I1, I2): after doubling them, sigma must double (a stored H kept
    from the old integrals would not)."""
    name, inst, so, eng, nnz = stored
    L = inst.norb
    I1 = np.zeros((2 * L, 2 * L))
    I1[0::2, 0::2] = inst.ints.h1
    I1[1::2, 1::2] = inst.ints.h1
    eng.set_integrals(2.0 * inst.ints.ecore, 2.0 * I1.ravel(), 2.0 * inst.ints.packed_eri())
    x = np.random.default_rng(2).standard_normal(so.ndet)
    ref = 2.0 * so.mult(x)
    assert np.max(np.abs(eng.sigma(x) - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert eng.store_hamiltonian(False) == 0
    eng.set_integrals(inst.ints.ecore, I1.ravel(), inst.ints.packed_eri())


def test_stored_above_64_orbitals():
    """Two-word strings: the stored H of the 70-orbital window problem (every 3-electron string
    of orbitals 59..68 per spin) gives the on-the-fly build's sigma (itself pinned against the
    compiled reference, tests/test_gpu_golden.py w70_3a3b), and the Davidson energy of the
    reference's fixture."""
    print("[stored w70]", flush=True)
    import json
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "w70_3a3b.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    lo, hi = meta["window"]
    inst = synth.make_window(meta["norb"], lo, hi, 3, 3, meta["rank"], seed_ints=meta["seed_ints"])
    fly, sto = Engine.from_instance(inst), Engine.from_instance(inst)
    sto.store_hamiltonian(True)
    for seed in (0, 1):
        x = np.random.default_rng(seed).standard_normal(inst.ndet)
        ref = fly.sigma(x)
        assert np.max(np.abs(sto.sigma(x) - ref)) <= 1e-12 * np.max(np.abs(ref))
    r = sto.davidson(max_iteration=100, num_block=10, eps=1e-8)
    assert r["converged"] and abs(r["energy"] - float(z["energy"][0])) < 1e-9
