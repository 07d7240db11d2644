"""Batched lockstep Davidson (sbdx_batch_*): many uneven problems, every lockstep step one
replayed HIP graph.

Every problem's energy must equal its own single-problem solve (block solver, same options) to
1e-9 Ha, and the smallest ones the oracle's Davidson; the returned vectors must be normalised
eigenvectors (oracle sigma residual).  The graphs are replayed many times per solve with eager
host work and synchronisations in between, and a second solve of the same batch must give the
same energies (replays against fresh eager state).
"""
import numpy as np
import pytest

import oracle as O
from sbd_amd import synth
from sbd_amd._lib import Batch, Engine, SbdxError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problems():
    insts = synth.make_batch_instances(count=6, seed=3, dets=(200, 3e4))
    print("[batch] dets:", [i.ndet for i in insts], flush=True)
    return insts, [Engine.from_instance(i) for i in insts]


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_matches_single_solves(problems, prec):
    insts, engs = problems
    opts = dict(num_block=8, max_iteration=300, eps=1e-8, precision=prec)
    b = Batch(engs)
    res = b.davidson(want_vectors=True, **opts)
    again = b.davidson(**opts)
    for inst, eng, r, r2 in zip(insts, engs, res, again):
        single = eng.davidson(method=1, **opts)
        assert r["converged"] and single["converged"], (inst.name, r, single["residual"])
        assert abs(r["energy"] - single["energy"]) < 1e-9, inst.name
        assert abs(r2["energy"] - r["energy"]) < 1e-9, inst.name
        W = r["W"]
        assert abs(np.linalg.norm(W) - 1.0) < 1e-10
        if inst.ndet <= 5000:
            so = O.System.from_instance(inst)
            ref = so.davidson(max_it=300, num_block=10, eps=1e-9)
            assert abs(r["energy"] - ref["energy"]) < 1e-9, inst.name
            s = so.mult(W)
            assert np.linalg.norm(s - r["energy"] * W) < 1e-7
    b.close()


def test_batch_rejects_unsupported_options(problems):
    insts, engs = problems
    b = Batch(engs[:2])
    with pytest.raises(SbdxError):
        b.davidson(nroots=2)
    with pytest.raises(SbdxError):
        b.davidson(num_block=40)
    b.close()
