"""Block thick-restart Davidson (method 1): FP64 and mixed precision, one or several roots.

The exact spectrum of the small cases comes from a dense eigensolve of H assembled column by
column from the oracle's FP64 sigma (numpy eigh); the larger case compares the solver's
energies with the oracle's Davidson energy.  Energies must match to 1e-9 Ha in FP64 and, after
the FP64 refinement, in mixed precision too; returned vectors must have FP64 residuals (oracle
sigma) below the solver's tolerance.
"""
import numpy as np
import pytest

import oracle as O
from sbd_amd import synth
from sbd_amd._lib import Engine, SbdxError

pytestmark = pytest.mark.gpu

SMALL = [
    ("l6_2a2b_full", (6, 2, 2, 4, None, None)),
    ("l8_3a2b_full", (8, 3, 2, 6, None, None)),
    ("l10_3a3b_samp", (10, 3, 3, 8, 60, 50)),
]


@pytest.fixture(scope="module", params=SMALL, ids=[c[0] for c in SMALL])
def small(request):
    name, (L, na, nb, R, Na, Nb) = request.param
    print(f"[case {name}]", flush=True)
    inst = synth.make_custom(L, na, nb, R, Na, Nb, seed_ints=5)
    so = O.System.from_instance(inst)
    hii = so.hdiag()
    n = so.ndet
    H = np.empty((n, n))
    e = np.zeros(n)
    for j in range(n):
        e[j] = 1.0
        H[:, j] = so.mult(e, hii)
        e[j] = 0.0
    evals = np.linalg.eigvalsh(0.5 * (H + H.T))
    return name, so, Engine.from_instance(inst), H, evals


def residuals(H, W, E):
    return [float(np.linalg.norm(H @ W[v] - E[v] * W[v])) for v in range(len(E))]


@pytest.mark.parametrize("block", [1, 2, 4])
def test_lowest_root_fp64(small, block):
    name, so, eng, H, evals = small
    r = eng.davidson(method=1, block=block, nroots=1, num_block=8, max_iteration=200, eps=1e-9)
    assert r["converged"], r
    assert abs(r["energy"] - evals[0]) < 1e-9, (r["energy"], evals[0])
    W = r["W"].reshape(1, -1)
    assert abs(np.linalg.norm(W[0]) - 1.0) < 1e-9
    assert residuals(H, W, [evals[0]])[0] < 1e-8


@pytest.mark.parametrize("nroots", [2, 4, 8])
@pytest.mark.parametrize("prec", [0, 1])
def test_several_roots(small, nroots, prec):
    name, so, eng, H, evals = small
    if so.ndet < 4 * nroots:
        pytest.skip("too few determinants")
    r = eng.davidson(method=1, nroots=nroots, num_block=8, max_iteration=300, eps=1e-8, precision=prec)
    assert r["converged"], r
    W = r["W"].reshape(nroots, -1)
    E = np.array([W[v] @ H @ W[v] for v in range(nroots)])
    assert np.max(np.abs(E - evals[:nroots])) < 1e-9, (E, evals[:nroots])
    assert abs(r["energy"] - evals[0]) < 1e-9
    assert max(residuals(H, W, E)) < 1e-7
    G = W @ W.T
    assert np.max(np.abs(G - np.eye(nroots))) < 1e-8


def test_mixed_precision_refines_to_fp64(small):
    """FP32 corrections with the FP64 anchor reach the FP64 energy to 1e-9 Ha and an FP64
    residual below 1e-8, beyond what FP32 arithmetic alone can represent (|E| eps_32 ~ 1e-6)."""
    name, so, eng, H, evals = small
    r = eng.davidson(method=1, precision=1, nroots=1, num_block=8, max_iteration=300, eps=1e-9)
    assert r["converged"], r
    assert abs(r["energy"] - evals[0]) < 1e-9
    W = r["W"].reshape(1, -1)
    assert residuals(H, W, [evals[0]])[0] < 1e-8


def test_mixed_precision_twosum_refines_to_fp64(small):
    """The mixed solver with compensated FP32 (two-sum) sigma builds reaches the FP64 energy."""
    name, so, eng, H, evals = small
    eng.set_fp32_accumulation(1)
    try:
        r = eng.davidson(method=1, precision=1, block=2, num_block=8, max_iteration=300, eps=1e-9)
    finally:
        eng.set_fp32_accumulation(0)
    assert r["converged"], r
    assert abs(r["energy"] - evals[0]) < 1e-9


def test_time_to_reference_energy(small):
    name, so, eng, H, evals = small
    r = eng.davidson(method=1, precision=1, block=2, num_block=8, max_iteration=300, eps=1e-12,
                     energy_ref=evals[0], energy_tol=1e-6, stop_at_tol=1)
    assert r["steps_to_tol"] > 0 and r["seconds_to_tol"] >= 0
    assert abs(r["energy"] - evals[0]) < 1e-6
    assert not r["converged"] or r["residual"] < 1e-12


def test_time_to_reference_energy_method0(small):
    """The reference's algorithm stops at the reference energy too, with a normalised vector."""
    name, so, eng, H, evals = small
    r = eng.davidson(method=0, num_block=10, max_iteration=300, eps=1e-12,
                     energy_ref=evals[0], energy_tol=1e-6, stop_at_tol=1)
    assert r["steps_to_tol"] > 0 and r["steps"] == r["steps_to_tol"]
    assert abs(r["energy"] - evals[0]) < 1e-6
    assert abs(np.linalg.norm(r["W"]) - 1.0) < 1e-9


def test_method0_rejects_block_options(small):
    name, so, eng, H, evals = small
    with pytest.raises(SbdxError):
        eng.davidson(method=0, nroots=2)
    with pytest.raises(SbdxError):
        eng.davidson(method=0, precision=1)
    with pytest.raises(SbdxError):
        eng.davidson(method=1, nroots=9)


@pytest.mark.parametrize("prec", [0, 1])
def test_block_matches_oracle_davidson_c2_family(prec):
    """84,000 determinants of the 36-orbital family: the block solver (2 vectors) and the
    oracle's Davidson (the reference algorithm) converge to the same energy."""
    print("[case c2fam_300x280 block]", flush=True)
    inst = synth.make_custom(36, 8, 8, 72, 300, 280, seed_ints=3)
    so = O.System.from_instance(inst)
    ref = so.davidson(max_it=400, num_block=10, eps=1e-9)
    eng = Engine.from_instance(inst)
    r = eng.davidson(method=1, precision=prec, block=2, num_block=10, max_iteration=400, eps=1e-9)
    assert r["converged"], r
    assert abs(r["energy"] - ref["energy"]) < 1e-9, (r["energy"], ref["energy"])
    W = r["W"]
    s = so.mult(W)
    assert np.linalg.norm(s - r["energy"] * W) < 1e-8


@pytest.mark.parametrize("prec,acc", [(0, 0), (1, 0), (1, 1)], ids=["fp64", "mixed", "mixed_twosum"])
def test_block_matches_reference_c3_family(prec, acc):
    """44,000 determinants of the 48-orbital family (the mixed-precision workload's family): the
    block solver, FP64 or mixed (FP32 corrections with either FP32 accumulation, FP64 anchor),
    reaches the compiled reference's Davidson energy of the committed fixture to 1e-9 Ha.  This
    basis splits into blocks of H (its 200 alpha strings form components of 81, 5, 3, ...
    strings under single and double excitations), so the solver's extra start vectors must lie
    in determinant 0's block, where the reference's Davidson stays."""
    import json
    import os
    print(f"[case c3fam_200x220 block prec={prec} acc={acc}]", flush=True)
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "c3fam_200x220.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    inst = synth.make_custom(48, 12, 12, 96, meta["n_alpha"], meta["n_beta"], seed_ints=meta["seed_ints"],
                             seed_alpha=meta["seed_alpha"], seed_beta=meta["seed_beta"])
    eng = Engine.from_instance(inst)
    eng.set_fp32_accumulation(acc)
    r = eng.davidson(method=1, precision=prec, block=2, num_block=10, max_iteration=400, eps=1e-9)
    assert r["converged"], r
    assert abs(r["energy"] - float(z["energy"][0])) < 1e-9, (r["energy"], float(z["energy"][0]))
