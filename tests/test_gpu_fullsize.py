"""Size-independent checks of the HIP sigma at the benchmark's full size (4e8 determinants),
where the CPU oracle cannot build a whole sigma:

* one column H e_k equals, entry by entry, the oracle's Slater-Condon element for every
  determinant connected to k (and is zero elsewhere);
* H is symmetric: x.(H y) = y.(H x);
* FP32 sigma agrees with FP64 sigma to FP32 accuracy.
"""
import numpy as np
import pytest

import oracle as O
from sbd_amd import synth
from sbd_amd._lib import Engine

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def c2():
    print("[fullsize c2 setup]", flush=True)
    inst = synth.make_instance("c2")
    return inst, Engine.from_instance(inst)


def popdist(s, x):
    v = np.bitwise_xor(s, np.uint64(x))
    return np.array([bin(int(t)).count("1") for t in v])


def test_column_matches_slater_condon(c2):
    inst, eng = c2
    na, nb, L = inst.alpha.size, inst.beta.size, inst.norb
    vals, idx = inst.ints.fcidump_entries()
    I0, I1, I2, Dm, Xm = O.setup_integrals(L, vals, idx)
    ints = O.Ints(L, I0, I1.ctypes.data, I2.ctypes.data, Dm.ctypes.data, Xm.ctypes.data)
    lib = O.lib()
    rng = np.random.default_rng(11)
    for ka, kb in [(0, 0), (int(rng.integers(na)), int(rng.integers(nb)))]:
        k = ka * nb + kb
        e = torch.zeros(eng.ndet, dtype=torch.float64, device="cuda")
        e[k] = 1.0
        col = torch.empty_like(e)
        eng.sigma_device(0, 1, e.data_ptr(), col.data_ptr())
        col = col.cpu().numpy()
        da, db = popdist(inst.alpha, inst.alpha[ka]), popdist(inst.beta, inst.beta[kb])
        ra = np.flatnonzero(da <= 4)
        rb = np.flatnonzero(db <= 4)
        ket = np.zeros(O.DW, np.uint64)
        lib.sbdo_det_from_ab(inst.alpha[ka:ka + 1], inst.beta[kb:kb + 1], L, ket)
        expect = {}
        for ia in ra:
            for ib in rb:
                if da[ia] + db[ib] > 4:
                    continue
                bra = np.zeros(O.DW, np.uint64)
                lib.sbdo_det_from_ab(inst.alpha[ia:ia + 1], inst.beta[ib:ib + 1], L, bra)
                expect[ia * nb + ib] = lib.sbdo_hij(bra, ket, ints)
        nz = np.flatnonzero(col)
        assert set(nz.tolist()) <= set(expect.keys())
        keys = np.array(sorted(expect.keys()))
        ref = np.array([expect[i] for i in keys])
        assert np.max(np.abs(col[keys] - ref)) <= 1e-12 * np.max(np.abs(ref))
        print(f"column {k}: {keys.size} connected determinants checked", flush=True)


def test_symmetry_and_fp32(c2):
    inst, eng = c2
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(eng.ndet, dtype=torch.float64, device="cuda", generator=g)
    y = torch.randn(eng.ndet, dtype=torch.float64, device="cuda", generator=g)
    hx, hy = torch.empty_like(x), torch.empty_like(y)
    eng.sigma_device(0, 1, x.data_ptr(), hx.data_ptr())
    eng.sigma_device(0, 1, y.data_ptr(), hy.data_ptr())
    a, b = float(torch.dot(y, hx)), float(torch.dot(x, hy))
    assert abs(a - b) <= 1e-10 * abs(a)
    del y, hy
    x32 = x.float()
    h32 = torch.empty_like(x32)
    eng.sigma_device(1, 1, x32.data_ptr(), h32.data_ptr())
    err = float((h32.double() - hx).abs().max() / hx.abs().max())
    assert err < 1e-5, err
