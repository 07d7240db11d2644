"""sbdx_sigma_device_noalloc (the sigma build for callers that cannot allocate device memory)
against sbdx_sigma_device, through the C ABI: bit-identical output, on two calls in a row -- the
second on a smaller problem, so the shared scratch holds the first call's larger contents --
for FP64 and FP32, and the host-side refusals (too large for the scratch, FP32 tables not
built)."""
import numpy as np
import pytest

from sbd_amd import synth
from sbd_amd._lib import Engine, SbdxError, SBDX_FP32, SBDX_FP64

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = [  # (L, alpha electrons, beta electrons, rank, alpha strings, beta strings): larger first
    (36, 8, 8, 72, 300, 280),
    (10, 3, 3, 8, 60, 50),
]


def reference_and_noalloc(eng, prec, seed):
    dt = torch.float64 if prec == SBDX_FP64 else torch.float32
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn(eng.ndet, dtype=dt, device="cuda", generator=g)
    ref = torch.empty_like(x)
    eng.sigma_device(prec, 1, x.data_ptr(), ref.data_ptr())
    got = torch.full_like(x, float("nan"))
    eng.sigma_device_noalloc(prec, x.data_ptr(), got.data_ptr())
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("prec", [SBDX_FP64, SBDX_FP32], ids=["fp64", "fp32"])
def test_noalloc_equals_sigma_device_on_two_calls(prec):
    engines = []
    for i, case in enumerate(CASES):
        print(f"[no-alloc prec={prec} case={case}]", flush=True)
        eng = Engine.from_instance(synth.make_custom(*case, seed_ints=3))
        engines.append(eng)  # keep both contexts alive: the second call follows the first
        ref, got = reference_and_noalloc(eng, prec, seed=10 + i)
        assert np.array_equal(ref, got), np.max(np.abs(ref - got))


def test_noalloc_refusals():
    eng = Engine.from_instance(synth.make_custom(10, 3, 3, 8, 60, 50, seed_ints=4))
    x = torch.ones(eng.ndet, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    with pytest.raises(SbdxError) as e:  # FP32 tables are built by the first FP32 sigma_device
        eng.sigma_device_noalloc(SBDX_FP32, x.data_ptr(), y.data_ptr())
    assert "FP32" in str(e.value)
    # 1500 x 1500 strings = 2.25e6 determinants: 4 x 8 bytes each exceeds the 64 MiB scratch
    big = Engine.from_instance(synth.make_custom(20, 4, 4, 8, 1500, 1500, seed_ints=5))
    xb = torch.zeros(big.ndet, dtype=torch.float64, device="cuda")
    yb = torch.empty_like(xb)
    with pytest.raises(SbdxError) as e:
        big.sigma_device_noalloc(SBDX_FP64, xb.data_ptr(), yb.data_ptr())
    assert "too large" in str(e.value)
