"""Bitwise reproducibility of sigma and of the solvers.

The reference owns each sigma row in one thread and visits its connections in a fixed order
(chemistry/tpb/mult.h:490-597), so its sigma does not change from run to run.  The engine's
passes are owner-computes too, and the mixed pass sums every column's entries in ascending ket
order, each entry over the row's alpha singles in ascending order: a build gives the same bits
on every run and for every kernel variant (tile shape, threads, stages), and so does a Davidson
solve (same trace, same step count).
"""
import numpy as np
import pytest

from sbd_amd import synth
from sbd_amd._lib import Engine

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# variants of the mixed kernel (1000000 (stages - 1) + 100 threads + entries per thread) per
# block width; the same lists as the autotuner's candidates
VARIANTS = {1: [25632, 25616, 51216, 102408, 1102408, 3102408, 1051216, 1025616],
            2: [102404, 25608, 102408, 25616, 51216, 1102404, 3102404, 1051208]}


@pytest.fixture(scope="module")
def eng():
    print("[determinism c2fam 300x280]", flush=True)
    return Engine.from_instance(synth.make_custom(36, 8, 8, 72, 300, 280, seed_ints=3))


def build(eng, prec, k, x):
    y = torch.empty_like(x)
    eng.sigma_block_device(prec, k, x.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("prec,k", [(0, 1), (0, 2), (1, 1), (1, 2)])
def test_sigma_bitwise_over_runs_and_variants(eng, prec, k):
    dt = torch.float64 if prec == 0 else torch.float32
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(eng.ndet * k, dtype=dt, device="cuda", generator=g)
    first = build(eng, prec, k, x)
    assert np.array_equal(first, build(eng, prec, k, x))
    for var in VARIANTS[k]:
        eng.set_mix_variant(prec, k, var)
        assert np.array_equal(first, build(eng, prec, k, x)), var
    eng.set_mix_variant(prec, k, 0)  # back to the default


def test_two_sum_bitwise(eng):
    x = torch.randn(eng.ndet, dtype=torch.float32, device="cuda",
                    generator=torch.Generator(device="cuda").manual_seed(6))
    eng.set_fp32_accumulation(1)
    try:
        a = build(eng, 1, 1, x)
        assert np.array_equal(a, build(eng, 1, 1, x))
    finally:
        eng.set_fp32_accumulation(0)


@pytest.mark.parametrize("opts", [dict(method=0, num_block=10),
                                  dict(method=1, precision=1, block=2, num_block=8)],
                         ids=["fp64_reference_algorithm", "mixed_block2"])
def test_davidson_trace_bitwise(eng, opts):
    a = eng.davidson(max_iteration=300, eps=1e-8, **opts)
    b = eng.davidson(max_iteration=300, eps=1e-8, **opts)
    assert a["converged"] and b["converged"]
    assert a["steps"] == b["steps"]
    assert np.array_equal(a["trace"], b["trace"])
    assert a["energy"] == b["energy"]
    assert np.array_equal(a["W"], b["W"])
