"""The drop-in C++ header API (include/sbd_amd/sbd.h): a C++ program drives sbd::tpb::diag,
sbd::makeQChamDiagTerms, sbd::mult and sbd::Correlation through it, as the reference's sample
driver does, and its outputs are checked against the CPU oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from sbd_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "shim_diag")


def test_cpp_drop_in_api(tmp_path):
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/shim_diag not built (needs the image's MPI)")
    inst = synth.make_custom(10, 3, 3, 8, 60, 50, seed_ints=3)
    vals, idx = inst.ints.fcidump_entries()
    path = str(tmp_path / "inst.bin")
    O.write_instance(path, inst.norb, 3, 3, inst.alpha, inst.beta, vals, idx)
    out = str(tmp_path / "shim")
    r = subprocess.run([BIN, path, out, "1e-8"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    so = O.System.from_instance(inst)
    ref = so.davidson(max_it=200, num_block=10, eps=1e-8)
    assert abs(res["energy"] - ref["energy"]) < 1e-9
    assert " Energy = " in r.stdout and "Davidson iteration" in r.stdout
    hii = np.fromfile(out + ".hii")
    assert np.max(np.abs(hii - so.hdiag())) < 1e-12 * np.abs(hii).max()
    x = np.cos(0.37 * np.arange(so.ndet))
    y = np.fromfile(out + ".multy")
    want = 1.0 + so.mult(x)
    assert np.max(np.abs(y - want)) < 1e-12 * np.abs(want).max()
    one, two = so.correlation(ref["W"])
    assert np.max(np.abs(np.fromfile(out + ".one") - one.ravel())) < 1e-6
    assert np.max(np.abs(np.fromfile(out + ".two") - two.ravel())) < 1e-6
    assert np.max(np.abs(np.fromfile(out + ".density") - so.occupation(ref["W"]))) < 1e-6
    assert res["carryover"] >= 1
    wf = out + ".wf000000"
    assert os.path.getsize(wf) == 3 * 8 + (60 + 50) * 8 + so.ndet * 8


def test_cpp_drop_in_with_the_block_mixed_solver(tmp_path):
    """The same driver with --amd_solver block --amd_precision fp32: the drop-in's Davidson
    runs the block mixed-precision solver and reaches the same energy."""
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/shim_diag not built (needs the image's MPI)")
    inst = synth.make_custom(10, 3, 3, 8, 60, 50, seed_ints=3)
    vals, idx = inst.ints.fcidump_entries()
    path = str(tmp_path / "inst.bin")
    O.write_instance(path, inst.norb, 3, 3, inst.alpha, inst.beta, vals, idx)
    out = str(tmp_path / "shim")
    r = subprocess.run([BIN, path, out, "1e-8", "--amd_solver", "block", "--amd_precision", "fp32",
                        "--amd_block", "2"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    ref = O.System.from_instance(inst).davidson(max_it=200, num_block=10, eps=1e-8)
    assert abs(res["energy"] - ref["energy"]) < 1e-9
    assert "(mixed)" in r.stdout


def test_cpp_restart_from_saved_wavefunction(tmp_path):
    """SaveWavefunction then LoadWavefunction (restart.h:20-67, :414-657): a second diag started
    from the first one's file converges at once to the same energy; a basis with extra strings
    gets the saved amplitudes on the common determinants and zeros elsewhere."""
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/shim_diag not built (needs the image's MPI)")
    inst = synth.make_custom(10, 3, 3, 8, 60, 50, seed_ints=3)
    vals, idx = inst.ints.fcidump_entries()
    path = str(tmp_path / "inst.bin")
    O.write_instance(path, inst.norb, 3, 3, inst.alpha, inst.beta, vals, idx)
    first = str(tmp_path / "first")
    r1 = subprocess.run([BIN, path, first, "1e-8"], capture_output=True, text=True, timeout=240)
    assert r1.returncode == 0, r1.stderr[-2000:]
    e1 = json.loads(r1.stdout.strip().splitlines()[-1])["energy"]
    second = str(tmp_path / "second")
    r2 = subprocess.run([BIN, path, second, "1e-8", "--load", first + ".wf"], capture_output=True,
                        text=True, timeout=240)
    assert r2.returncode == 0, r2.stderr[-2000:]
    e2 = json.loads(r2.stdout.strip().splitlines()[-1])["energy"]
    assert abs(e1 - e2) < 1e-9
    steps1 = r1.stdout.count("Davidson iteration")
    steps2 = r2.stdout.count("Davidson iteration")
    assert steps2 < steps1, (steps1, steps2)
    saved = np.fromfile(first + ".wf000000", dtype=np.uint64)
    na, nb, ln = (int(v) for v in saved[:3])
    assert (na, nb, ln) == (60, 50, 1)


def test_cpp_drop_in_stored_hamiltonian(tmp_path):
    """--method 1 (the reference's stored-Hamiltonian method): the drop-in stores H as CSR on
    the GPU and its Davidson reaches the same energy."""
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/shim_diag not built (needs the image's MPI)")
    inst = synth.make_custom(10, 3, 3, 8, 60, 50, seed_ints=3)
    vals, idx = inst.ints.fcidump_entries()
    path = str(tmp_path / "inst.bin")
    O.write_instance(path, inst.norb, 3, 3, inst.alpha, inst.beta, vals, idx)
    out = str(tmp_path / "shim")
    r = subprocess.run([BIN, path, out, "1e-8", "--method", "1"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    ref = O.System.from_instance(inst).davidson(max_it=200, num_block=10, eps=1e-8)
    assert abs(res["energy"] - ref["energy"]) < 1e-9
    assert "Stored Hamiltonian elements" in r.stdout


def test_cpp_drop_in_api_above_64_orbitals(tmp_path):
    """The driver on the 70-orbital window problem of the compiled reference's fixture
    (tests/golden/w70_3a3b.npz: every 3-electron string of orbitals 59..68 per spin, so the
    header packs each bit_length-20 string into two 64-bit words): energy, H_ii and the RDMs of
    the converged vector against the reference's own outputs."""
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/shim_diag not built (needs the image's MPI)")
    from sbd_amd._lib import Engine
    z = np.load(os.path.join(ROOT, "tests", "golden", "w70_3a3b.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    lo, hi = meta["window"]
    inst = synth.make_window(meta["norb"], lo, hi, 3, 3, meta["rank"], seed_ints=meta["seed_ints"])
    vals, idx = inst.ints.fcidump_entries()
    path = str(tmp_path / "inst.bin")
    O.write_instance(path, inst.norb, 3, 3, inst.alpha, inst.beta, vals, idx)
    out = str(tmp_path / "shim")
    print("[shim w70]", flush=True)
    r = subprocess.run([BIN, path, out, "1e-8"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["ndet"] == inst.ndet
    assert abs(res["energy"] - float(z["energy"][0])) < 1e-9
    hii = np.fromfile(out + ".hii")
    assert np.max(np.abs(hii[z["hii_idx"]] - z["hii_val"])) < 1e-12 * np.abs(hii).max()
    x = np.cos(0.37 * np.arange(inst.ndet))
    want = 1.0 + Engine.from_instance(inst).sigma(x)
    assert np.max(np.abs(np.fromfile(out + ".multy") - want)) < 1e-14 * np.abs(want).max()
    assert np.max(np.abs(np.fromfile(out + ".one") - z["one"])) < 1e-6
    two = np.fromfile(out + ".two")
    assert two.size == 4 * 70 ** 4
    assert np.max(np.abs(two[z["two_idx"]] - z["two_val"])) < 1e-6
    assert np.max(np.abs(np.fromfile(out + ".density") - z["density"])) < 1e-6
